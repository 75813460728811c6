"""The oracle's Krylov loops (orc_krylov.c: CG, PCG, GMRES(m), FGMRES(m))
against the reference's own (src/iterativesolve/IterativeSolvers_Class.f90
compiled unmodified, oracle/ref.mk) on dense systems.

Both sides apply the same operator (column sums in ascending order,
b - A s, s / diag A: oracle/ref_krylov.f90 and orc_dense_*), so the
comparison isolates the loops: ioerr, the residual history and the solution
are bitwise equal, including the reference's stop tests (sqrt(|r|)/r0 < tol
and the absolute sqrt(|r|) > tol failure check of the CG forms, quirk Q5,
which reports -1 here although the loop stopped on the relative test) and a
GMRES run that exhausts its restarts.
"""
import glob
import os

import numpy as np
import pytest

from oracle import pyoracle as po
import ref_run as R

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_krylov_*.npz")))
SOLVERS = {"cg": po.solve_cg, "pcg": po.solve_pcg, "gmres": po.solve_gmres, "pcgmres": po.solve_pcgmres}


def load(path):
    z = np.load(path)
    return {k: z[k] for k in z.files}


def run_oracle(g, **over):
    kw = dict(x0=g["x0"], maxIters=int(g["maxIters"]), m=int(g["mInner"]), tol=float(g["tolerance"]), native=True)
    kw.update(over)
    return SOLVERS[str(g["method"])](g["A"], g["b"], **kw)


def test_fixtures_cover_every_method():
    assert {str(load(p)["method"]) for p in GOLDEN} == set(SOLVERS)


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_oracle_loop_is_bitwise_the_reference(path):
    g = load(path)
    x, resi, ioerr, it = run_oracle(g)
    assert ioerr == int(g["ioerr"])
    assert it == np.count_nonzero(g["resi"]) - 1
    np.testing.assert_array_equal(resi[:len(g["resi"])], g["resi"])
    np.testing.assert_array_equal(x, g["x"])


def test_a_mistake_in_the_oracle_is_caught():
    """A different restart length or a looser tolerance changes the
    history; the numpy operator (another summation order) moves it off the
    bitwise result."""
    g = load([p for p in GOLDEN if "gmres_ns40" in p][0])
    _, resi, _, _ = run_oracle(g, m=int(g["mInner"]) - 1)
    assert not np.array_equal(resi[:len(g["resi"])], g["resi"])
    g = load([p for p in GOLDEN if os.path.basename(p) == "ref_krylov_cg_spd40.npz"][0])
    _, _, _, it = run_oracle(g, tol=10 * float(g["tolerance"]))
    assert it != np.count_nonzero(g["resi"]) - 1
    x, _, _, _ = run_oracle(g, native=False)
    assert not np.array_equal(x, g["x"])


@pytest.mark.skipif(not R.available("krylov"), reason="oracle/_ref/ref_krylov not built (needs the reference sources)")
@pytest.mark.parametrize("method,m", [("cg", 1), ("pcg", 1), ("gmres", 5), ("pcgmres", 7)])
def test_live_reference_run_matches_oracle(method, m):
    rng = np.random.default_rng(77)
    n = 25
    if method in ("cg", "pcg"):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        A = (Q * np.geomspace(1.0, 300.0, n)) @ Q.T
        A = 0.5 * (A + A.T) + np.diag(rng.uniform(0.0, 3.0, n))
    else:
        A = np.eye(n) * 3.0 + rng.standard_normal((n, n)) * 0.35
    b, x0 = rng.standard_normal(n), 0.2 * rng.standard_normal(n)
    g = R.run_krylov(A, b, x0, method, 80, m, 1e-9)
    x, resi, ioerr, it = SOLVERS[method](A, b, x0=x0, maxIters=80, m=m, tol=1e-9, native=True)
    assert ioerr == g["ioerr"]
    np.testing.assert_array_equal(resi[:len(g["resi"])], g["resi"])
    np.testing.assert_array_equal(x, g["x"])
