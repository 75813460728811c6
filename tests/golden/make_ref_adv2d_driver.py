"""Writes tests/golden/refdriver_adv2d.npz: a run of the reference's own 2-D
advection-diffusion programs (src/highend/advection-diffusion2d/
Advection_InitialConditions.f90 and Advection_Driver.f90, compiled
unmodified against the reference's modules by oracle/ref.mk into
oracle/_ref/ref_adv2d_ic and ref_adv2d_driver) with
adv2d_driver_advection.params (2 tracers, kappa 1e-3, N = 5 on a 4 x 3 mesh
scaled to 2 x 1.5, 20 steps of 1e-4, dumps every 10): the initial pickup the
initial-condition program writes (the input of the drop-in's run), the
driver's final pickup, its Tecplot dumps at iterations 0, 10 and 20 and the
mesh file, read as numbers by tests/cgsem_outputs.numeric_rows.

    make -f oracle/ref.mk && python tests/golden/make_ref_adv2d_driver.py
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


def main():
    ref = os.path.join(ROOT, "oracle", "_ref")
    out = {}
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(HERE, "adv2d_driver_advection.params"), os.path.join(d, "advection.params"))
        subprocess.run([os.path.join(ref, "ref_adv2d_ic")], cwd=d, check=True, capture_output=True, timeout=300)
        out["pickup0"] = np.frombuffer(open(os.path.join(d, "Advection.0000000000.pickup"), "rb").read(), np.uint8)
        subprocess.run([os.path.join(ref, "ref_adv2d_driver")], cwd=d, check=True, capture_output=True, timeout=300)
        out["pickup20"] = np.frombuffer(open(os.path.join(d, "Advection.0000000020.pickup"), "rb").read(), np.uint8)
        for it in (0, 10, 20):
            out[f"tec{it}"] = numeric_rows(os.path.join(d, f"Advection.{it:010d}.tec"), 10)
        out["mesh"] = numeric_rows(os.path.join(d, "mesh.tec"), 7)
    path = os.path.join(HERE, "refdriver_adv2d.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
