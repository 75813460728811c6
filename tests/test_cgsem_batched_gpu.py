"""The batched CGSEM PCG (self_cgsem2d_solve_batched): many problems of
different mesh sizes, degrees, coefficients and right-hand sides solved
together, two launches per CG iteration for all of them.  Each problem's
solution, ioerr and iteration count must be bitwise those of its own
single solve (self_cgsem2d_solve with the CGSEM PCG), which the CGSEM tests
pin to the oracle and the compiled reference (tolerance: exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(nX, N, seed, warp=0.0, zero_source=False):
    from self_amd.krylov import CGsemElliptic2D, CGsemParams
    from self_amd.workloads import cgsem_poisson_fields
    m = CGsemElliptic2D(CGsemParams(polyDeg=N, nXElem=nX, nYElem=nX, warp=warp))
    k, f = cgsem_poisson_fields(m.mesh.x, m.mesh.y, seed)
    m.SetFluxCoefficient(k)
    m.SetSource(np.zeros_like(f) if zero_source else -f)
    # nonzero Dirichlet values on a third of the problems
    d = np.sin(3 * m.mesh.x) * np.cos(2 * m.mesh.y) if seed % 3 == 0 else np.zeros_like(k)
    m.SetDirichletBoundaryCondition(d)
    m.SetSolution(np.zeros_like(k))
    return m


CASES = [(4, 3, 11, 0.0), (8, 5, 12, 0.05), (16, 7, 13, 0.0), (8, 2, 14, 0.0), (12, 3, 15, 0.1),
         (16, 5, 16, 0.0), (4, 7, 17, 0.0), (24, 3, 18, 0.02), (8, 9, 19, 0.0), (32, 5, 20, 0.0)]


def _check_against_single(cases, precond, tol, maxIters, zero=()):
    from self_amd.krylov import SolveBatched
    models = [_problem(nX, N, s, w, i in zero) for i, (nX, N, s, w) in enumerate(cases)]
    io, its, ms = SolveBatched(models, precond=precond, tolerance=tol, maxIters=maxIters)
    assert ms > 0
    outs = [m.GetSolution() for m in models]
    for m in models:
        m.Trash()
    for i, (nX, N, s, w) in enumerate(cases):
        m = _problem(nX, N, s, w, i in zero)
        io1, it1, _ = m.Solve("cgsem_pcg", precond=precond, tolerance=tol, maxIters=maxIters)
        u1 = m.GetSolution()
        m.Trash()
        print(f"problem {i}: {nX}^2 N={N} ioerr {io[i]} iterations {its[i]} (single {io1}, {it1})", flush=True)
        assert (io[i], its[i]) == (io1, it1), i
        assert np.array_equal(outs[i], u1), i
    return io, its


def test_batched_equals_single_solves_jacobi(gpu_lib):
    io, its = _check_against_single(CASES, "jacobi", 1e-10, 5000)
    assert all(e == -2 for e in io)  # converged (the reference's ioerr stays -2)
    assert len(set(its)) > 3          # the problems stop at different iterations


def test_batched_identity_iteration_limit_and_zero_source(gpu_lib):
    """No preconditioner; maxIters 45 stops the larger problems (ioerr -1)
    while the small ones converge; a zero source with zero Dirichlet values
    returns at once (r0 < tol: 0 iterations, the reference's early RETURN)."""
    cases = [(4, 3, 21, 0.0), (16, 5, 22, 0.0), (8, 7, 24, 0.0), (4, 2, 25, 0.0), (16, 3, 27, 0.03)]
    io, its = _check_against_single(cases, "none", 1e-8, 45, zero=(3,))
    assert -1 in io and its[3] == 0


def test_batched_workload_sample(gpu_lib):
    """The 64-problem batch (16^2..128^2, N in {3, 5, 7}, seeded right-hand
    sides) with Jacobi to 1e-10: every problem converges, its true residual
    |b - A u| / |b - A 0| computed afresh is within 2e-9 (the loop stops on
    its recursively updated residual, the reference's test, which drifts
    from the true one by rounding: up to 1.0e-9 here), and five seeded
    problems rerun alone give bitwise the same solution and count."""
    from self_amd.krylov import SolveBatched
    from self_amd.workloads import cgsem_batch, cgsem_helmholtz
    models, info = cgsem_batch(64)
    r0 = [np.linalg.norm(m.Residual()) for m in models]
    io, its, ms = SolveBatched(models, precond="jacobi", tolerance=1e-10, maxIters=20000)
    print(f"64 problems: {ms:.1f} ms, iterations {min(its)}..{max(its)}", flush=True)
    assert all(e == -2 for e in io)
    rel = [np.linalg.norm(m.Residual()) / r for m, r in zip(models, r0)]
    assert max(rel) < 2e-9, max(rel)
    rng = np.random.default_rng(7)
    pick = rng.choice(64, size=5, replace=False)
    outs = {int(i): models[int(i)].GetSolution() for i in pick}
    for m in models:
        m.Trash()
    for i in pick:
        nX, N, s = info[int(i)]
        m, _ = cgsem_helmholtz(nX, N, seed=s)
        io1, it1, _ = m.Solve("cgsem_pcg", precond="jacobi", tolerance=1e-10, maxIters=20000)
        assert (io1, it1) == (io[int(i)], its[int(i)]), (int(i), nX, N)
        assert np.array_equal(m.GetSolution(), outs[int(i)]), (int(i), nX, N)
        m.Trash()


def test_batched_rejects_bad_calls(gpu_lib):
    import ctypes as C
    from self_amd import _lib as L
    from self_amd.krylov import SolveBatched
    a, b = _problem(4, 3, 31), _problem(4, 3, 32)
    with pytest.raises(L.SelfError):
        SolveBatched([a, b, a])
    with pytest.raises(KeyError):
        SolveBatched([a, b], precond="fem")
    io, it = (C.c_int * 1)(), (C.c_int * 1)()
    with pytest.raises(L.SelfError):
        L.check(L.lib().self_cgsem2d_solve_batched((C.c_void_p * 1)(a._h), 1, 2, 1e-10, 100, io, it, None))
    with pytest.raises(L.SelfError):
        L.check(L.lib().self_cgsem2d_solve_batched((C.c_void_p * 1)(a._h), 0, 1, 1e-10, 100, io, it, None))
    a.Trash()
    b.Trash()
