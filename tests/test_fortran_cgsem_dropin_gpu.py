"""The reference's own CGSEM example driver (examples/cgsem-elliptic2d/
paraboloid/localmods/CGsemElliptic2D_Driver.f90), compiled unchanged against
the drop-in modules (self_amd/fortran/cgsem_dropin.f90 with sw_dropin.f90)
by oracle/ref.mk into oracle/_ref/dropin_cgsem_paraboloid, run on the GPU
with the example's runtime.params (N = 10 on one element, FEM preconditioner
with 20 inner iterations to 1e-8, tolerance 1e-8), against every output file
of the same driver compiled against the reference's own modules (fixture
tests/golden/ref_cgsem_paraboloid.npz, made by
tests/golden/make_ref_cgsem_paraboloid.py):
 * Omat.txt (the operator on the 81 free nodes) and PCmat.txt (the N = 1
   preconditioner on the sub-mesh's free nodes, in the reference's sub-mesh
   order): to 1e-12 of the largest entry;
 * omask.tec, mask.tec: masks equal, positions to 1e-13;
 * residual.curve: the same number of iterations, the history to 1e-6;
 * CGsemSol.tec: positions and source to 1e-13, the solution to 1e-7 of its
   largest value (both runs stop at a relative residual of 1e-8)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cgsem_outputs import read_run

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_cgsem_paraboloid")
GOLDEN = os.path.join(ROOT, "tests", "golden", "ref_cgsem_paraboloid.npz")


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_reference_cgsem_driver_on_the_drop_in(gpu_lib, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/dropin_cgsem_paraboloid not built (make -f oracle/ref.mk)")
    shutil.copy(os.path.join(ROOT, "tests", "golden", "cgsem_paraboloid_runtime.params"),
                tmp_path / "runtime.params")
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    got = read_run(str(tmp_path))
    ref = dict(np.load(GOLDEN))
    for k in ("Omat", "PCmat"):
        assert got[k].shape == ref[k].shape, k
        e = _rel(got[k], ref[k])
        print(f"{k}: {got[k].shape}, max difference / max entry {e:.3e}", flush=True)
        assert e < 1e-12, (k, e)
    for k in ("omask", "mask"):
        assert got[k].shape == ref[k].shape, k
        assert np.array_equal(got[k][:, 2], ref[k][:, 2]), k
        assert np.max(np.abs(got[k][:, :2] - ref[k][:, :2])) < 1e-13, k
    its, its_ref = int(np.count_nonzero(got["resi"][1:])), int(np.count_nonzero(ref["resi"][1:]))
    print(f"iterations {its} (reference {its_ref})", flush=True)
    assert got["resi"].shape == ref["resi"].shape and its == its_ref
    e = _rel(got["resi"][: its + 1], ref["resi"][: its + 1])
    assert e < 1e-6, e
    s, sr = got["sol"], ref["sol"]
    assert s.shape == sr.shape
    assert np.max(np.abs(s[:, :2] - sr[:, :2])) < 1e-13
    assert _rel(s[:, 3], sr[:, 3]) < 1e-13
    e = _rel(s[:, 2], sr[:, 2])
    print(f"solution on the plotting points: max difference / max {e:.3e}", flush=True)
    assert e < 1e-7, e
