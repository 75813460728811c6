"""Writes tests/golden/ref_cgsem_paraboloid.npz: the output files of the
reference's own CGSEM example driver (examples/cgsem-elliptic2d/paraboloid,
CGsemElliptic2D_Driver.f90, compiled unmodified against the reference's
modules by oracle/ref.mk into oracle/_ref/ref_cgsem_paraboloid) run with the
example's runtime.params (cgsem_paraboloid_runtime.params here): the
operator and preconditioner matrices (Omat.txt, PCmat.txt), their masks
(omask.tec, mask.tec), the solution on the plotting points (CGsemSol.tec)
and the residual history (residual.curve), read by tests/cgsem_outputs.py.

    make -f oracle/ref.mk && python tests/golden/make_ref_cgsem_paraboloid.py
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

from cgsem_outputs import read_run  # noqa: E402


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_cgsem_paraboloid")
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(HERE, "cgsem_paraboloid_runtime.params"), os.path.join(d, "runtime.params"))
        subprocess.run([exe], cwd=d, check=True, capture_output=True, timeout=300)
        out = read_run(d)
    path = os.path.join(HERE, "ref_cgsem_paraboloid.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
