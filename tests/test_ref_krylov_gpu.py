"""The HIP Krylov solvers against the reference's own loops.

tests/golden/ref_krylov_*.npz hold what src/iterativesolve/
IterativeSolvers_Class.f90 (compiled unmodified) computed on dense systems:
ioerr, the residual history and the solution (the oracle reproduces them
bitwise).  The device solvers run the same loops on a device dense operator
(self_linop_dense_create) through the C ABI; they sum in another order, so
the histories agree to 1e-8 relative while above round-off, the iteration
counts and ioerr exactly (for FGMRES within one), the solutions to 1e-9.
"""
import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_krylov_*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_device_solver_matches_reference(gpu_lib, path):
    from self_amd.krylov import LinOp
    g = dict(np.load(path))
    op = LinOp.dense(g["A"], g["b"])
    method = str(g["method"])
    x, resi, iters, ioerr = op.Solve(method, tolerance=float(g["tolerance"]), maxIters=int(g["maxIters"]),
                                     mInner=int(g["mInner"]), sol=g["x0"])
    ref = g["resi"]
    n_ref = np.count_nonzero(ref) - 1
    assert ioerr == int(g["ioerr"])
    assert abs(iters - n_ref) <= (1 if method == "pcgmres" else 0)
    k = min(iters, n_ref, 16)
    assert np.all(np.abs(resi[:k] - ref[:k]) <= 1e-8 * np.abs(ref[:k]) + 1e-13)
    assert np.abs(x - g["x"]).max() <= 1e-9 * max(np.abs(g["x"]).max(), 1.0)
    op.Trash()
