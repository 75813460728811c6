"""Writes tests/golden/ref_specmesh_annulus.npz: an ISM-V2 mesh file of a
curved quarter annulus (tests/specmesh_gen.py; rotated corner orderings,
one clockwise element) and what the reference's own ReadSpecMeshFile_QuadMesh
(src/geom/QuadMeshClass.f90:2100-2465, compiled unmodified by oracle/ref.mk
into oracle/_ref/ref_specmesh) built from it at N = 3 on Gauss nodes and
N = 5 on Gauss-Lobatto nodes, and at N = 4 on Gauss nodes followed by its
ScaleTheMesh(2, 3) (QuadMeshClass.f90:1691-1714).

The fixture holds the file's text (uint8) and, per case '<N>_<quad>_'
('<N>_<quad>_s_' when scaled), the
edge table, corners, node types, corner positions and the element geometry
in the reference's element-last shapes.

    make -f oracle/ref.mk && python tests/golden/make_ref_specmesh.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_run as R  # noqa: E402
import specmesh_gen as G  # noqa: E402

NR, NT, GDEG, SEED, MIRROR = 3, 4, 4, 1, (5,)
# (N, quadrature, ScaleTheMesh factors or None)
CASES = ((3, 1, None), (5, -1, None), (4, 1, (2.0, 3.0)))


def tag(N, quad, scale):
    return f"{N}_{quad}_" + ("" if scale is None else "s_")


def main():
    text = G.quarter_annulus(NR, NT, GDEG, seed=SEED, mirror=MIRROR)
    out = {"mesh": np.frombuffer(text.encode(), dtype=np.uint8)}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "annulus.mesh")
        with open(path, "w") as f:
            f.write(text)
        for N, quad, scale in CASES:
            ref = R.run_specmesh(path, N, quad, scale=scale)
            for k, v in ref.items():
                out[tag(N, quad, scale) + k] = np.asarray(v)
    dst = os.path.join(HERE, "ref_specmesh_annulus.npz")
    np.savez_compressed(dst, **out)
    print(dst, os.path.getsize(dst))


if __name__ == "__main__":
    main()
