"""The oracle's restatement of the hydrostatic primitive equations
(src/highend/hydrostaticprimitives/HydrostaticPrimitivesClass.f90, oracle/
orc_hpe.c) pinned by closed forms.

The reference class does not compile (it uses undeclared names and calls an
undefined barotropic step), so there is no reference output to compare with;
every routine is pinned here by what the mathematics fixes:

* DiagnoseP integrates the hydrostatic balance p_z = b downward from p = 0 at
  the top: for b polynomial of degree <= N-1 in z the column quadrature and
  the face interpolation are exact, so p equals -int_z^0 b dz' to round-off.
* CalculateDivU and DiagnoseW: for polynomial velocities div U is the exact
  divergence and w = w_bottom - int_{-H}^z div U dz' with w_bottom = 0 on a
  flat sea floor.
* GlobalTimeDerivative as a whole: for u = a(x) c(z), v = d(y) c2(z),
  b = beta(x) e(z) + gamma(y) e2(z) with zero depth integral of c and c2,
  constant background buoyancy B and zero background w, the tendencies are
  u_t = beta'(x) E(z), v_t = gamma'(y) E2(z) (E = int_z^0 e),
  b_t = B (a'(x) c(z) + d'(y) c2(z)) at every node, on walled and on
  periodic meshes.  A flipped pressure-gradient, integration or flux sign, a
  missing Jacobian or a transposed metric term changes these.
* A state at rest in hydrostatic balance has zero tendency.
* Discrete conservation: the quadrature-weighted Jacobian-weighted sum of
  each tendency equals minus the boundary-face sum of its numerical flux.
* The RK3 step converges at third order in dt.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def _model(N=4, nX=3, nY=2, nL=3, L=(2.0, 3.0, 0.5), periodic=(False, False), warp=0.0, diagnoseP=True, dt=1e-3):
    return po.HydrostaticPrimitive(N, nX, nY, nL, periodic=periodic, scale=L, warp=warp, dt=dt, diagnoseP=diagnoseP)


def _xyz(m, H):
    n1 = m.N + 1
    sh = (m.nElems, n1, n1, n1)
    return m.mesh.x.reshape(sh), m.mesh.y.reshape(sh), m.mesh.z.reshape(sh) - H


def test_diagnose_p_is_the_hydrostatic_integral():
    H = 0.5
    m = _model(N=5, L=(2.0, 3.0, H))
    x, y, z = _xyz(m, H)
    sol, bg = m.view(m.sol, 5), m.view(m.bg, 5)
    sol[:, 2] = np.cos(x) * (1.0 + 2.0 * z - 3.0 * z ** 3) + y ** 2
    m.call("diagnose_p")
    # p = -int_z^0 b dz'
    P = lambda zz: 1.0 * zz + zz ** 2 - 0.75 * zz ** 4  # noqa: E731  antiderivative of the z-polynomial
    exact = -(np.cos(x) * (P(0.0) - P(z)) + y ** 2 * (0.0 - z))
    assert np.abs(sol[:, 4] - exact).max() < 1e-13 * np.abs(exact).max()
    assert not np.any(bg)


@pytest.mark.parametrize("periodic", [(False, False), (True, True)])
def test_divu_and_diagnose_w_are_exact_for_polynomials(periodic):
    H, N = 0.5, 5
    m = _model(N=N, L=(2.0, 3.0, H), periodic=periodic)
    x, y, z = _xyz(m, H)
    sol = m.view(m.sol, 5)
    sol[:, 0] = (x ** 3 - x) * (1.0 + z)
    sol[:, 1] = (y ** 2) * (z ** 2 - 0.1)
    m.call("traces")
    m.call("divu")
    div = (3 * x ** 2 - 1) * (1.0 + z) + 2 * y * (z ** 2 - 0.1)
    assert np.abs(m.divU.reshape(x.shape) - div).max() < 1e-12 * np.abs(div).max()
    m.call("diagnose_w")
    assert np.abs(m.bottomW).max() == 0.0
    Z0 = -H
    W = -((3 * x ** 2 - 1) * ((z + z ** 2 / 2) - (Z0 + Z0 ** 2 / 2)) +
          2 * y * ((z ** 3 / 3 - 0.1 * z) - (Z0 ** 3 / 3 - 0.1 * Z0)))
    assert np.abs(sol[:, 3] - W).max() < 1e-12 * np.abs(W).max()


def _balanced_tendency_case(m, H, B=0.7):
    x, y, z = _xyz(m, H)
    sol, bg = m.view(m.sol, 5), m.view(m.bg, 5)
    # c, c2 with zero depth integral over [-H, 0]; e, e2 of degree <= N-1
    c = lambda zz: zz + H / 2  # noqa: E731
    c2 = lambda zz: zz ** 2 - H ** 2 / 3  # noqa: E731
    e = lambda zz: 1.0 + zz  # noqa: E731
    E = lambda zz: -(zz + zz ** 2 / 2)  # noqa: E731  int_z^0 e
    e2 = lambda zz: zz ** 2  # noqa: E731
    E2 = lambda zz: -zz ** 3 / 3  # noqa: E731
    a, da = (lambda xx: xx ** 2 - 0.3 * xx), (lambda xx: 2 * xx - 0.3)
    d, dd = (lambda yy: yy ** 3), (lambda yy: 3 * yy ** 2)
    be, dbe = (lambda xx: 0.2 * xx ** 3), (lambda xx: 0.6 * xx ** 2)
    ga, dga = (lambda yy: yy - 0.5 * yy ** 2), (lambda yy: 1.0 - yy)
    sol[:, 0] = a(x) * c(z)
    sol[:, 1] = d(y) * c2(z)
    sol[:, 2] = be(x) * e(z) + ga(y) * e2(z)
    bg[:, 2] = B
    expect = np.stack([dbe(x) * E(z), dga(y) * E2(z), B * (da(x) * c(z) + dd(y) * c2(z))], axis=1)
    return expect


@pytest.mark.parametrize("N", [3, 4, 6])
def test_global_time_derivative_closed_form(N):
    H = 0.5
    m = _model(N=N, nX=3, nY=2, nL=3, L=(2.0, 3.0, H))
    expect = _balanced_tendency_case(m, H)
    m.call("global_time_derivative")
    t = m.view(m.tend, 3)
    for f in range(3):
        err = np.abs(t[:, f] - expect[:, f]).max()
        assert err < 1e-11 * max(1.0, np.abs(expect[:, f]).max()), (f, err)


def test_global_time_derivative_closed_form_periodic():
    """Periodic in x and y: trigonometric horizontal structure is not a
    polynomial, so only the exactness in z is used: a single Fourier mode per
    direction with the vertical structure of the walled case, compared at a
    tolerance set by the horizontal resolution (spectral in N)."""
    H, N = 0.5, 9
    Lx, Ly = 2.0, 3.0
    m = _model(N=N, nX=2, nY=2, nL=2, L=(Lx, Ly, H), periodic=(True, True))
    x, y, z = _xyz(m, H)
    kx, ky = 2 * np.pi / Lx, 2 * np.pi / Ly
    sol, bg = m.view(m.sol, 5), m.view(m.bg, 5)
    B = 0.7
    sol[:, 0] = np.sin(kx * x) * (z + H / 2)
    sol[:, 1] = np.cos(ky * y) * (z ** 2 - H ** 2 / 3)
    sol[:, 2] = np.cos(kx * x) * (1.0 + z) + np.sin(ky * y) * z ** 2
    bg[:, 2] = B
    m.call("global_time_derivative")
    t = m.view(m.tend, 3)
    ut = -kx * np.sin(kx * x) * (-(z + z ** 2 / 2))
    vt = ky * np.cos(ky * y) * (-z ** 3 / 3)
    bt = B * (kx * np.cos(kx * x) * (z + H / 2) - ky * np.sin(ky * y) * (z ** 2 - H ** 2 / 3))
    for f, ex in enumerate((ut, vt, bt)):
        err = np.abs(t[:, f] - ex).max()
        assert err < 1e-5 * np.abs(ex).max(), (f, err)


def test_rest_state_in_hydrostatic_balance_has_zero_tendency():
    H = 0.5
    for periodic in [(False, False), (True, True)]:
        m = _model(N=5, L=(2.0, 3.0, H), periodic=periodic)
        x, y, z = _xyz(m, H)
        sol, bg = m.view(m.sol, 5), m.view(m.bg, 5)
        sol[:, 2] = np.exp(3 * z) + np.sin(7 * z)  # any b(z)
        bg[:, 2] = 1.0 + z
        m.call("global_time_derivative")
        assert np.abs(m.tend).max() < 1e-12, np.abs(m.tend).max()
        assert np.abs(sol[:, 3]).max() == 0.0
        assert np.abs(sol[:, 4]).max() > 0.01  # the pressure was diagnosed


def test_discrete_conservation_identity():
    H = 0.5
    m = _model(N=4, nX=3, nY=3, nL=2, L=(2.0, 3.0, H), periodic=(True, True), warp=0.0)
    rng = np.random.default_rng(3)
    sol, bg = m.view(m.sol, 5), m.view(m.bg, 5)
    sol[:, :3] = rng.standard_normal(sol[:, :3].shape)
    bg[:, 2] = 1.0 + rng.random(bg[:, 2].shape)
    bg[:, 3] = 0.1 * rng.standard_normal(bg[:, 3].shape)
    m.call("global_time_derivative")
    n1 = m.N + 1
    w = m.nodal.qWeight
    W3 = np.einsum("k,j,i->kji", w, w, w)
    J = m.mesh.J.reshape(m.nElems, n1, n1, n1)
    t = m.view(m.tend, 3)
    fl = m.bflux.reshape(m.nElems, 6, 3, n1 * n1)
    W2 = np.einsum("b,a->ba", w, w).ravel()
    faces = m.mesh.faces
    for f in range(3):
        vol = np.sum(W3 * J * t[:, f])
        bnd = 0.0
        for fc in faces:
            if fc[2] < 0:
                bnd += np.sum(W2 * fl[fc[0] - 1, fc[1] - 1, f])
        scale = np.sum(W3 * J * np.abs(t[:, f]))
        assert abs(vol + bnd) < 1e-12 * scale, (f, vol, bnd)
    # no flux through the flat top and bottom for u and v
    for fc in faces:
        if fc[2] < 0:
            assert np.all(fl[fc[0] - 1, fc[1] - 1, :2] == 0.0)


def test_rk3_is_third_order():
    """Self-convergence of ForwardStepRK3 at fixed end time.  A background
    vertical velocity keeps the upwind speed max(|w_in + W|, |w_ex + W|) of
    the buoyancy flux away from its kink at zero, so the right-hand side is
    smooth in the state and the third order shows."""
    H = 0.5

    def run(dt, nsteps):
        m = _model(N=4, nX=2, nY=2, nL=2, L=(2.0, 3.0, H), periodic=(True, True), dt=dt)
        x, y, z = _xyz(m, H)
        sol, bg = m.view(m.sol, 5), m.view(m.bg, 5)
        sol[:, 0] = 0.1 * np.sin(np.pi * x) * (z + H / 2)
        sol[:, 1] = 0.1 * np.cos(2 * np.pi * y / 3) * z
        sol[:, 2] = 0.2 * np.cos(np.pi * x) * (1.0 + z)
        bg[:, 2] = 1.0 + z
        bg[:, 3] = 1.0
        for _ in range(nsteps):
            m.call("forward_step_rk3")
        return m.view(m.sol, 5)[:, :3].copy()

    T = 0.08
    ref = run(T / 256, 256)
    errs = [np.abs(run(T / n, n) - ref).max() for n in (8, 16, 32)]
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 7.0 < r1 < 11.0 and 7.0 < r2 < 11.0, errs


def test_ab2_first_step_uses_zero_previous_tendency():
    """ForwardStepAB2 with a zero previous tendency advances by 3/2 dt of the
    tendency, and stores the tendency for the next step."""
    H = 0.5
    m = _model(N=3, nX=2, nY=2, nL=2, L=(2.0, 3.0, H), periodic=(True, True), dt=0.01)
    rng = np.random.default_rng(1)
    sol = m.view(m.sol, 5)
    sol[:, :3] = rng.standard_normal(sol[:, :3].shape)
    m.view(m.bg, 5)[:, 2] = 1.0
    s0 = sol[:, :3].copy()
    m.call("global_time_derivative")
    t0 = m.view(m.tend, 3).copy()
    m.call("forward_step_ab2")
    assert np.abs(sol[:, :3] - (s0 + 0.015 * t0)).max() < 1e-14 * np.abs(s0).max() * 10
    assert np.array_equal(m.view(m.G, 3), t0)
