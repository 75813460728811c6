"""GPU parity of the Krylov solvers and the CG-SEM elliptic operator.

* Elliptic1D + device CG reproduces the reference fixture
  examples/elliptic1d/parabola/run/{residual,elliptic}.curve (committed under
  tests/golden/): residual history to the 8 printed digits while it is above
  round-off, solution to 1e-13.
* Dense systems: device CG/PCG/GMRES/FGMRES against the oracle (same
  iteration counts, residual histories to 1e-9 relative) and numpy.
* CG-SEM 2-D: MatrixAction, Residual and the reference PCG against the
  oracle (restatement of CGsemElliptic_2D_Class.f90); Jacobi PCG, GMRES and
  the mixed FP32/FP64 refinement solve reach the same solution.
"""
import math
import os

import numpy as np
import pytest

from conftest import rel_l2
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _curve(name, section=None):
    rows, cur = [], None
    for line in open(os.path.join(GOLDEN, name)):
        t = line.strip()
        if t.startswith("#"):
            cur = t[1:]
            continue
        if t and (section is None or cur == section):
            rows.append([float(v) for v in t.split()])
    return np.array(rows)


def test_elliptic1d_device_cg_reproduces_fixture(gpu_lib):
    from self_amd.krylov import LinOp
    op = LinOp.elliptic1d(N=5, nElems=10)
    sol, resi, iters, ioerr = op.Solve("cg", tolerance=1e-8, maxIters=1000)
    ref = _curve("elliptic1d_residual.curve")
    # iterations 0..24 carry the O(1) residuals; iteration 25 is the
    # cancellation step (0.09 -> 7e-8) whose value is set by round-off
    above = ref[:, 1] > 1e-3
    assert above.sum() == 25
    rel = np.abs(resi[:len(ref)][above] - ref[above, 1]) / ref[above, 1]
    assert rel.max() < 1e-7
    assert resi[25] < 1e-6 and ref[25, 1] < 1e-6
    # the loop stops when sqrt(|r|)/r0 < 1e-8, i.e. once |r| reaches round-off
    assert 60 <= iters <= 120 and resi[iters] < 3e-18 * 1e2
    assert ioerr == -2
    x, u = op.UnMaskSolution(sol)
    sref = _curve("elliptic1d_elliptic.curve", "solution")
    assert np.abs(x - sref[:, 0]).max() < 1e-15
    assert np.abs(u - sref[:, 1]).max() < 1e-13
    op.Trash()


def _nonsym(n, seed):
    rng = np.random.default_rng(seed)
    return np.eye(n) * 4.0 + rng.standard_normal((n, n)) * 0.4, rng.standard_normal(n)


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (Q * np.linspace(1.0, 20.0, n)) @ Q.T, rng.standard_normal(n)


@pytest.mark.parametrize("method", ["cg", "pcg"])
def test_dense_cg_pcg_match_oracle(gpu_lib, method):
    from self_amd.krylov import LinOp
    A, b = _spd(40, 11)
    op = LinOp.dense(A, b)
    # the stop test is sqrt(|r|)/r0 < tol (quirk Q5): 1e-7 stops at |r| ~ 1e-14 r0^2,
    # above round-off, so the iteration count is not decided by rounding noise
    x, resi, iters, ioerr = op.Solve(method, tolerance=1e-7, maxIters=300)
    fn = po.solve_cg if method == "cg" else po.solve_pcg
    xo, resio, ioo, ito = fn(A, b, maxIters=300, tol=1e-7)
    assert iters == ito and ioerr == ioo
    k = min(iters, 30)
    assert np.all(np.abs(resi[:k] - resio[:k]) <= 1e-9 * np.abs(resio[:k]) + 1e-14)
    assert np.linalg.norm(x - np.linalg.solve(A, b)) < 1e-8 * np.linalg.norm(b)
    op.Trash()


@pytest.mark.parametrize("method", ["gmres", "pcgmres"])
@pytest.mark.parametrize("ortho", [0, 1])
def test_dense_gmres_match_oracle(gpu_lib, method, ortho):
    from self_amd.krylov import LinOp
    A, b = _nonsym(60, 12)
    op = LinOp.dense(A, b)
    x, resi, iters, ioerr = op.Solve(method, tolerance=1e-11, maxIters=10, mInner=8, ortho=ortho)
    fn = po.solve_gmres if method == "gmres" else po.solve_pcgmres
    xo, resio, ioo, ito = fn(A, b, maxIters=10, m=8, tol=1e-11)
    assert ioerr == ioo == 0
    assert abs(iters - ito) <= (0 if ortho == 0 else 1)
    k = min(iters, ito, 16)
    assert np.all(np.abs(resi[:k] - resio[:k]) <= 1e-8 * np.abs(resio[:k]) + 1e-13)
    assert np.linalg.norm(x - np.linalg.solve(A, b)) < 1e-9 * np.linalg.norm(b)
    op.Trash()


def test_gmres_not_converged_reports_minus_one(gpu_lib):
    from self_amd.krylov import LinOp
    A, b = _nonsym(50, 13)
    op = LinOp.dense(A, b)
    x, resi, iters, ioerr = op.Solve("gmres", tolerance=1e-30, maxIters=2, mInner=5)
    assert ioerr == -1 and iters == 10
    op.Trash()


# ------------------------------------------------------------------ CG-SEM 2-D
def _cgsem_pair(N=5, nX=4, nY=3, Lx=1.0, Ly=1.5, warp=0.0, kvar=True, dirichlet=0.0):
    from self_amd.krylov import CGsemElliptic2D, CGsemParams
    o = po.CGSEM2D(N, nX, nY, Lx, Ly, warp, tolerance=1e-12, maxIters=4000)
    p = CGsemParams(polyDeg=N, nXElem=nX, nYElem=nY, xScale=Lx, yScale=Ly, tolerance=1e-12,
                    maximumIterates=4000, warp=warp)
    g = CGsemElliptic2D(p)
    x = o.mesh.x
    y = o.mesh.y
    k = 1.0 + 0.5 * np.sin(2 * math.pi * x) * np.sin(2 * math.pi * y) if kvar else np.ones_like(x)
    s = np.sin(math.pi * x / Lx) * np.sin(2 * math.pi * y / Ly) + 0.3 * x * y
    o.fluxCoeff[:] = k
    o.source[:] = s
    o.s.dirichletValue = dirichlet
    shape = (g.nElems, N + 1, N + 1)
    g.SetFluxCoefficient(k.reshape(shape))
    g.SetSource(s.reshape(shape))
    g.SetDirichletBoundaryCondition(np.full(shape, dirichlet))
    return o, g, shape


@pytest.mark.parametrize("warp", [0.0, 0.04])
def test_cgsem_matrix_action_and_residual_match_oracle(gpu_lib, warp):
    o, g, shape = _cgsem_pair(warp=warp, dirichlet=0.3)
    rng = np.random.default_rng(4)
    u = o.mask(rng.standard_normal(o.solution.size))
    Au_o, _ = o.matrix_action(u)
    Au_g = g.MatrixAction(u.reshape(shape))
    assert rel_l2(Au_g.ravel(), Au_o) < 1e-13
    # residual of a solution with Dirichlet values on the boundary
    sol = o.mask(rng.standard_normal(o.solution.size))
    o.solution[:] = sol
    lib = po.lib()
    lib.orc_cgsem2d_set_dirichlet(po.C.byref(o.s))
    g.SetSolution(o.solution.reshape(shape))
    r_o = o.residual()
    r_g = g.Residual()
    assert rel_l2(r_g.ravel(), r_o) < 1e-13
    g.Trash()


@pytest.mark.parametrize("precond", [0, 1])
def test_cgsem_reference_pcg_matches_oracle(gpu_lib, precond):
    o, g, shape = _cgsem_pair(N=5, nX=4, nY=3)
    ioo, ito = o.solve(precond)
    io, it, ms = g.Solve("cgsem_pcg", precond="jacobi" if precond else "none", tolerance=1e-12, maxIters=4000)
    assert io == ioo and abs(it - ito) <= 1
    k = min(it, ito, 15)
    assert np.all(np.abs(g.resi[:k] - o.resi[:k]) <= 1e-9 * o.resi[:k])
    assert rel_l2(g.GetSolution().ravel(), o.solution) < 1e-10
    g.Trash()


def test_cgsem_solvers_agree_and_converge_spectrally(gpu_lib):
    """PCG (Jacobi), GMRES(30) and the FP32/FP64 refinement solve agree, and
    the discrete solution converges spectrally to a manufactured solution."""
    from self_amd.krylov import CGsemElliptic2D, CGsemParams
    errs = []
    for N in (3, 5, 7):
        p = CGsemParams(polyDeg=N, nXElem=4, nYElem=4, tolerance=1e-12, maximumIterates=5000)
        g = CGsemElliptic2D(p)
        x, y = g.mesh.x, g.mesh.y
        ue = np.sin(math.pi * x) * np.sin(math.pi * y)
        g.SetSource(-2 * math.pi ** 2 * ue)
        io, it, _ = g.Solve("cgsem_pcg", precond="jacobi")
        u1 = g.GetSolution()
        errs.append(np.abs(u1 - ue).max())
        g.SetSolution(np.zeros_like(ue))
        io2, it2, _ = g.Solve("pcgmres", maxIters=400, mInner=30, ortho=1, tolerance=1e-12)
        assert io2 == 0
        assert np.abs(g.GetSolution() - u1).max() < 1e-9
        g.SetSolution(np.zeros_like(ue))
        io3, outer, inner, rel, _ = g.SolveMixed(tolerance=1e-10)
        assert io3 == 0 and rel <= 1e-10
        assert np.abs(g.GetSolution() - u1).max() < 1e-8
        g.SetSolution(np.zeros_like(ue))
        io4, updates, its, rel4, _ = g.SolveReliable(tolerance=1e-10)
        assert io4 == 0 and rel4 <= 1e-10 and updates >= 1
        assert np.abs(g.GetSolution() - u1).max() < 1e-8
        g.Trash()
    assert errs[0] > 1e3 * errs[2] and errs[2] < 1e-9, errs


def test_cgsem_solve_is_deterministic(gpu_lib):
    """Two solves of the same problem, with another model's solve in between,
    give bitwise the same history and solution (fixed reduction trees)."""
    o, g, shape = _cgsem_pair(N=5, nX=4, nY=3)
    o2, g2, shape2 = _cgsem_pair(N=3, nX=5, nY=5)
    g.Solve("cgsem_pcg", precond="jacobi")
    r1, s1 = g.resi.copy(), g.GetSolution()
    g2.Solve("cgsem_pcg", precond="jacobi")
    g.SetSolution(np.zeros(shape))
    g.Solve("cgsem_pcg", precond="jacobi")
    assert np.array_equal(r1, g.resi) and np.array_equal(s1, g.GetSolution())
    g.Trash()
    g2.Trash()


def test_cgsem_reliable_updates_reach_the_fp64_solution(gpu_lib):
    """FP32 PCG with FP64 reliable updates: the true FP64 residual (recomputed
    through the model's Residual, not the solver's own bookkeeping) meets the
    tolerance, the solution is the FP64 Jacobi-PCG solution (the device
    solver pinned against the oracle above), the FP32 sequence needs about as
    many iterations as FP64 PCG (it is never restarted), and a repeat solve
    is bitwise the same.  Inhomogeneous Dirichlet values."""
    from self_amd import _lib as L
    _, g, shape = _cgsem_pair(N=5, nX=8, nY=6, dirichlet=0.2)
    io_o, it_o, _ = g.Solve("cgsem_pcg", precond="jacobi", tolerance=1e-12, maxIters=4000)
    assert io_o == -2
    ref = g.GetSolution()
    g.SetSolution(np.zeros(shape))
    r0 = np.linalg.norm(g.Residual())
    io, updates, its, rel, _ = g.SolveReliable(tolerance=1e-12, delta=0.2)
    assert io == 0 and updates >= 5
    true = np.linalg.norm(g.Residual()) / r0
    assert true <= 1e-12 and abs(true - rel) <= 1e-3 * rel
    sol = g.GetSolution()
    assert rel_l2(sol.ravel(), ref.ravel()) < 1e-9
    assert its <= 1.3 * it_o, (its, it_o)
    g.SetSolution(np.zeros(shape))
    io2, updates2, its2, rel2, _ = g.SolveReliable(tolerance=1e-12, delta=0.2)
    assert (io2, updates2, its2, rel2) == (io, updates, its, rel)
    assert np.array_equal(g.GetSolution(), sol)
    # out of FP32 iterations: not converged, the update made at the end still counts
    g.SetSolution(np.zeros(shape))
    io3, updates3, its3, rel3, _ = g.SolveReliable(tolerance=1e-12, delta=0.2, maxIters=40)
    assert io3 == -1 and its3 == 40 and rel3 > 1e-12
    assert abs(np.linalg.norm(g.Residual()) / r0 - rel3) <= 1e-3 * rel3
    # out of updates
    g.SetSolution(np.zeros(shape))
    io4, updates4, _, rel4, _ = g.SolveReliable(tolerance=1e-12, delta=0.2, maxUpdates=2)
    assert io4 == -1 and updates4 == 2 and rel4 > 1e-12
    with pytest.raises(L.SelfError):
        g.SolveReliable(delta=1.0)
    g.Trash()


def test_cgsem_fp64_reliable_updates_stop_on_the_true_residual(gpu_lib):
    """The FP64 form of the reliable-update solve (an FP64 PCG sequence whose
    residual is replaced by the true FP64 residual whenever it has dropped
    by delta): the true residual, recomputed through the model's Residual,
    meets the tolerance; the solution is the FP64 Jacobi-PCG solution to
    1e-10; the sequence needs about as many iterations as the reference's
    loop (it is never restarted); a repeat solve is bitwise the same."""
    _, g, shape = _cgsem_pair(N=5, nX=8, nY=6, dirichlet=0.2)
    io_o, it_o, _ = g.Solve("cgsem_pcg", precond="jacobi", tolerance=1e-12, maxIters=4000)
    ref = g.GetSolution()
    g.SetSolution(np.zeros(shape))
    r0 = np.linalg.norm(g.Residual())
    io, updates, its, rel, _ = g.SolveReliable(tolerance=1e-12, delta=1e-2, inner="fp64")
    true = np.linalg.norm(g.Residual()) / r0
    print(f"reference loop: {it_o} iterations; FP64 reliable: {its} iterations, {updates} updates, "
          f"true residual {true:.3e}", flush=True)
    assert io == 0 and updates >= 2
    assert true <= 1e-12 and abs(true - rel) <= 1e-3 * rel
    sol = g.GetSolution()
    assert rel_l2(sol.ravel(), ref.ravel()) < 1e-10
    assert its <= 1.1 * it_o + 10, (its, it_o)
    g.SetSolution(np.zeros(shape))
    io2, updates2, its2, rel2, _ = g.SolveReliable(tolerance=1e-12, delta=1e-2, inner="fp64")
    assert (io2, updates2, its2, rel2) == (io, updates, its, rel)
    assert np.array_equal(g.GetSolution(), sol)
    g.Trash()
