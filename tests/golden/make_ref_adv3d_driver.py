"""Writes tests/golden/refdriver_adv3d.npz: a run of the reference's own 3-D
advection programs (src/highend/advection3d/Advection3D_InitialConditions.f90
and Advection3D_Driver.f90, compiled unmodified against the reference's
modules by oracle/ref.mk into oracle/_ref/ref_adv3d_ic and ref_adv3d_driver)
with adv3d_driver_advection.params (2 tracers, N = 5, roll-off cutoff 3, a
3 x 2 x 2 mesh scaled to 1 x 1.2 x 0.8, 20 steps of 2e-3, dumps every 10):
the initial pickup and Tecplot file the initial-condition program writes
(the pickup is the input of the drop-in's run), the driver's final pickup,
its Tecplot dumps at iterations 0, 10 and 20 and the mesh file, read as
numbers by tests/cgsem_outputs.numeric_rows.

    make -f oracle/ref.mk && python tests/golden/make_ref_adv3d_driver.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))

from cgsem_outputs import numeric_rows  # noqa: E402


def _bytes(path):
    with open(path, "rb") as f:
        return np.frombuffer(f.read(), np.uint8)


def main():
    ref = os.path.join(ROOT, "oracle", "_ref")
    out = {}
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(HERE, "adv3d_driver_advection.params"), os.path.join(d, "advection.params"))
        subprocess.run([os.path.join(ref, "ref_adv3d_ic")], cwd=d, check=True, capture_output=True, timeout=300)
        out["pickup0"] = _bytes(os.path.join(d, "Advection.0000000000.pickup"))
        out["tecinit"] = numeric_rows(os.path.join(d, "Advection.init.tec"), 8)
        subprocess.run([os.path.join(ref, "ref_adv3d_driver")], cwd=d, check=True, capture_output=True, timeout=300)
        out["pickup20"] = _bytes(os.path.join(d, "Advection.0000000020.pickup"))
        for it in (0, 10, 20):
            out[f"tec{it}"] = numeric_rows(os.path.join(d, f"Advection.{it:010d}.tec"), 8)
        out["mesh"] = numeric_rows(os.path.join(d, "mesh.tec"), 13)
    path = os.path.join(HERE, "refdriver_adv3d.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
