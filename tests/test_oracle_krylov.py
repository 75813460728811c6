"""Pins of the oracle's Krylov solvers (oracle/orc_krylov.c).

* The reference's own fixture examples/elliptic1d/parabola/run/
  {residual,elliptic}.curve (CG on the 1-D CG-SEM operator, N=5, 10
  elements, tol 1e-8) is committed under tests/golden/ and must be
  reproduced: all 82 residual entries to the 8 printed digits and the
  solution to the printed 17.
* CG/PCG/GMRES/FGMRES against numpy.linalg.solve on small dense systems,
  GMRES residuals monotone within a cycle, CG finite termination.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _read_curve(path, section=None):
    rows = []
    cur = None
    with open(path) as f:
        for line in f:
            t = line.strip()
            if t.startswith("#"):
                cur = t[1:]
                continue
            if t and (section is None or cur == section):
                rows.append([float(v) for v in t.split()])
    return np.array(rows)


def test_elliptic1d_reproduces_reference_fixture():
    x, u, resi, ioerr, iters = po.elliptic1d(N=5, nElems=10, maxIters=1000, tol=1e-8)
    res = _read_curve(os.path.join(GOLDEN, "elliptic1d_residual.curve"))
    sol = _read_curve(os.path.join(GOLDEN, "elliptic1d_elliptic.curve"), "solution")
    src = _read_curve(os.path.join(GOLDEN, "elliptic1d_elliptic.curve"), "source")
    nz = int(np.count_nonzero(res[:, 1]))
    assert nz == 82 and iters == 81          # the CG loop EXITed after iteration 81
    assert np.all(res[82:, 1] == 0.0) and np.all(resi[82:] == 0.0)
    rel = np.abs(resi[:82] - res[:82, 1]) / np.abs(res[:82, 1])
    assert rel.max() < 1e-7                  # 8 printed significant digits
    assert np.abs(x - sol[:, 0]).max() < 1e-16
    assert np.abs(u - sol[:, 1]).max() < 1e-15
    assert np.all(src[:, 1] == 1.0)
    # quirk Q5: the final failure test is the absolute SQRT(rNorm) > TOL, so a
    # run that EXITs on the relative test still reports ioerr = -1 here? No:
    # sqrt(3.8e-19) = 6e-10 < 1e-8, so ioerr keeps its initial -2 (ioerr = 0
    # after EXIT is unreachable in the reference).
    assert ioerr == -2
    # quirk Q10: the masked (not summed) source gives 29/30 of the exact
    # mid-point value -1/8
    mid = u[np.argmin(np.abs(x - 0.5))]
    assert abs(mid - (-0.12083333333333)) < 1e-12


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.linspace(1.0, 20.0, n)
    return (Q * ev) @ Q.T, rng.standard_normal(n)


def _nonsym(n, seed):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) * 0.4
    return A, rng.standard_normal(n)


@pytest.mark.parametrize("solver", [po.solve_cg, po.solve_pcg])
def test_cg_pcg_match_dense_solve(solver):
    A, b = _spd(30, 1)
    x, resi, ioerr, iters = solver(A, b, maxIters=200, tol=1e-12)
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(b) < 1e-10
    assert iters <= 60
    assert resi[0] == pytest.approx(np.linalg.norm(b))


def test_cg_terminates_in_n_steps_on_small_system():
    A, b = _spd(8, 3)
    x, resi, ioerr, iters = po.solve_cg(A, b, maxIters=50, tol=1e-10)
    # the residual collapses by iteration n = 8 (exact-arithmetic termination);
    # the loop itself runs on because of the sqrt-of-norm stop test (quirk Q5)
    assert resi[8] < 1e-9 * resi[0]
    assert np.allclose(A @ x, b, atol=1e-10)


@pytest.mark.parametrize("solver", [po.solve_gmres, po.solve_pcgmres])
def test_gmres_match_dense_solve(solver):
    A, b = _nonsym(40, 2)
    x, resi, ioerr, iters = solver(A, b, maxIters=20, m=10, tol=1e-12)
    assert ioerr == 0
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(b) < 1e-10


def test_gmres_residual_monotone_within_cycle():
    A, b = _nonsym(60, 5)
    m = 8
    x, resi, ioerr, iters = po.solve_gmres(A, b, maxIters=3, m=m, tol=1e-30)
    # resi(l) for l = 1..m is the Givens residual |bhat(i+1)| of cycle 1
    cyc = resi[1:m + 1]
    assert np.all(np.diff(cyc) <= 1e-12 * cyc[0])
    assert cyc[-1] < resi[0]
    assert ioerr == -1  # unreachable tolerance -> not converged


def test_planted_mistake_is_caught_by_dense_check():
    """A wrong sign in the update (planted here in a copy of CG written in
    numpy, not in the oracle) would not solve the system: the dense check
    above has the power to see such a mistake."""
    A, b = _spd(20, 7)
    x = np.zeros(20)
    r = b - A @ x
    v = r.copy()
    for _ in range(40):
        z = A @ v
        a = (r @ r) / (v @ z)
        x = x - a * v          # planted: wrong sign
        rn = r - a * z
        v = rn + (rn @ rn) / (r @ r) * v
        r = rn
    assert np.linalg.norm(x - np.linalg.solve(A, b)) > 1e-3


def test_oracle_cgsem_first_solve_quirk():
    """CGsemElliptic_2D Solve (:1155-1156) takes its first residual from the
    stored solution; on a fresh model its Dirichlet nodes are still zero, so
    the first solve is the homogeneous problem with the Dirichlet values
    pasted afterwards (:1205), and only a second solve includes them.  With
    zero Dirichlet data both solves agree."""
    import math
    N, nX, nY = 4, 4, 3
    def model(d):
        o = po.CGSEM2D(N, nX, nY, 1.0, 1.5, 0.0, tolerance=1e-12, maxIters=2000)
        x, y = o.mesh.x, o.mesh.y
        o.fluxCoeff[:] = 1.0 + 0.5 * np.sin(2 * math.pi * x) * np.sin(2 * math.pi * y)
        o.source[:] = np.sin(math.pi * x) * np.sin(2 * math.pi * y / 1.5) + 0.3 * x * y
        o.s.dirichletValue = d
        return o
    o0 = model(0.0)
    o0.solve(1)
    o = model(0.2)
    o.solve(1)
    first = o.solution.copy()
    o.solve(1)
    second = o.solution.copy()
    inner = np.abs(o0.solution) > 0  # interior nodes of the homogeneous solve
    # first solve: interior equals the homogeneous solution, boundary holds 0.2
    assert np.allclose(first[inner], o0.solution[inner], rtol=0, atol=1e-12)
    assert np.isclose(first.max(), 0.2)
    # second solve: the boundary data enter the interior (u ~ 0.2 + the homogeneous part)
    assert np.abs(second[inner] - first[inner]).min() > 0.1
    # a repeat solve of the homogeneous problem returns at IF (r0 < TOL)
    # (:1162-1164) before copying u back, leaving the stored solution as
    # Residual masked it (:1063); unmasked it is the first solution
    o2 = model(0.0)
    o2.solve(1)
    assert o2.solve(1) == (-2, 0)
    assert not np.array_equal(o2.solution, o0.solution)
    assert np.array_equal(o2.unmask(o2.solution), o0.solution)
