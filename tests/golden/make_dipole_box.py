"""Writes tests/golden/dipole_box_mesh.npz: the ISM-V2 mesh file of the
reference's shallow-water dipole run (examples/shallowwater/dipole/run/
box.mesh: 489 corners, 936 edges, 448 elements, straight sides on the unit
square, scaled by 4e5 in both directions by the run's shallowwater.params),
the mesh of the reference's GlobalTimeDerivative profile at N = 10
(doc/shallowwater/ShallowWater-Projects.tex:316-329).  Input data only; the
npz holds the file's bytes (uint8) so it travels without the reference.

    python tests/golden/make_dipole_box.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/examples/shallowwater/dipole/run/box.mesh"


def main():
    with open(SRC, "rb") as f:
        data = f.read()
    dst = os.path.join(HERE, "dipole_box_mesh.npz")
    np.savez_compressed(dst, mesh=np.frombuffer(data, dtype=np.uint8))
    print(dst, os.path.getsize(dst))


if __name__ == "__main__":
    main()
