"""Closed-form pins of the shallow-water oracle's source terms
(oracle/orc_sw2d.c, Source of ConservativeShallowWaterClass.f90:1662-1694):
Coriolis, linear drag and relaxation, and the geostrophic balance between
Coriolis and the pressure gradient.

On a uniform state over a flat bottom on a periodic affine mesh every flux
divergence vanishes (free stream), so the semi-discrete system reduces at
each node to the linear ODE

    d/dt (U, V) = [[-d/h - tau,  f], [-f, -d/h - tau]] (U, V) + tau (U_r, V_r)
    d/dt H      = -tau (H - H_r)

whose exact solution is a matrix exponential.  RK3 converges to it at third
order.  A flipped Coriolis sign, drag divided by H instead of the bathymetry
depth h, a dropped relaxation target or a wrong relaxation sign each move
the solution by O(1) and fail.  The geostrophic jet (a straight, balanced
nonlinear steady state) pins Coriolis against the pressure term on a
non-uniform state, with spectral convergence of the residual tendency.
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from sw2d_cases import build_oracle_case

G = 9.81


def _uniform_case(N, f, drag, relax_vals, U0, V0, H0, h, dt, g=G, nX=3, nY=3):
    n1 = N + 1
    E = nX * nY
    relax = None
    if relax_vals is not None:
        relax = np.zeros((E, 4, n1, n1))
        for k in range(4):
            relax[:, k] = relax_vals[k]
    orc, info = build_oracle_case(N, nX, nY, True, True, Lx=2.0, Ly=3.0, g=g, dt=dt, drag=drag, relax=relax)
    sol = np.zeros((E, 3, n1, n1))
    sol[:, 0], sol[:, 1], sol[:, 2] = U0, V0, H0
    orc.sol[:] = sol.ravel()
    bathy = np.zeros((E, 3, n1, n1))
    bathy[:, 0] = h
    orc.bathy[:] = bathy.ravel()
    orc.bathymetry_gradient()
    orc.plv[:] = f
    return orc, E, n1


def _exact(t, f, drag, h, tau, Ur, Vr, U0, V0, H0, Hr):
    A = np.array([[-drag / h - tau, f], [-f, -drag / h - tau]])
    b = tau * np.array([Ur, Vr])
    # affine ODE y' = A y + b through the augmented matrix exponential
    M = np.zeros((3, 3))
    M[:2, :2] = A
    M[:2, 2] = b
    y = expm(M * t) @ np.array([U0, V0, 1.0])
    H = Hr + (H0 - Hr) * math.exp(-tau * t)
    return y[0], y[1], H


def _run(orc, nsteps, dt):
    for k in range(nsteps):
        orc.forward_step_rk3(k * dt)


@pytest.mark.parametrize("f,drag,tau", [(0.7, 0.0, 0.0),      # inertial oscillation
                                        (0.0, 0.9, 0.0),      # linear drag, divided by h (not H)
                                        (0.0, 0.0, 0.6),      # relaxation of U, V, H to their targets
                                        (0.7, 0.9, 0.6)])     # all three together
def test_uniform_state_follows_source_ode(f, drag, tau):
    U0, V0, H0, h = 1.3, -0.4, 12.0, 10.0
    Ur, Vr, Hr = 0.25, 0.8, 11.5
    T = 2.0
    errs = []
    for nsteps in (20, 40, 80):
        dt = T / nsteps
        # g = 0: a uniform state has no pressure gradient, and at these
        # deliberately large steps round-off-seeded gravity waves would grow
        orc, E, n1 = _uniform_case(3, f, drag, (Ur, Vr, Hr, tau) if tau else None, U0, V0, H0, h, dt, g=0.0)
        _run(orc, nsteps, dt)
        s = orc.sol.reshape(E, 3, n1, n1)
        if tau:
            ex = _exact(T, f, drag, h, tau, Ur, Vr, U0, V0, H0, Hr)
        else:
            ex = _exact(T, f, drag, h, 0.0, 0.0, 0.0, U0, V0, H0, H0)
        errs.append(max(np.abs(s[:, k] - ex[k]).max() for k in range(3)))
        # every node carries the same value (the flux divergence of a uniform state vanishes)
        assert np.ptp(s[:, 0]) < 1e-12 and np.ptp(s[:, 1]) < 1e-12
    assert errs[-1] < 1e-6, errs
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 6.5 < r1 < 9.5 and 6.5 < r2 < 9.5, errs


def test_uniform_state_tendency_signs():
    """One tendency of a uniform state: the rotation sense (U_t = +fV,
    V_t = -fU), the drag's divisor (bathymetry depth h = 10, not H = 12) and
    the relaxation direction, against independently chosen numbers."""
    # g = 0 keeps the round-off of the (vanishing) pressure divergence out
    orc, E, n1 = _uniform_case(4, 0.5, 2.0, (1.0, -2.0, 13.0, 0.25), 3.0, 4.0, 12.0, 10.0, 1e-3, g=0.0)
    orc.global_time_derivative(0.0)
    t = orc.tend.reshape(E, 3, n1, n1)
    # U_t = 0.5*4 - 2*3/10 - 0.25*(3-1) = 2 - 0.6 - 0.5 = 0.9
    # V_t = -0.5*3 - 2*4/10 - 0.25*(4+2) = -1.5 - 0.8 - 1.5 = -3.8
    # H_t = -0.25*(12-13) = 0.25
    assert np.abs(t[:, 0] - 0.9).max() < 1e-13
    assert np.abs(t[:, 1] + 3.8).max() < 1e-13
    assert np.abs(t[:, 2] - 0.25).max() < 1e-13


def _jet_state(x, y, Lx, Ly, f, g, H0, A):
    """eta(xi), xi = 2 pi (x/Lx + y/Ly); u = (g/f) z x grad(eta): a straight
    geostrophic jet along the level lines of eta.  u . grad = 0 on every
    field, so all nonlinear terms vanish and the state is an exact steady
    solution of the conservative shallow-water equations over a flat bottom."""
    kx, ky = 2 * math.pi / Lx, 2 * math.pi / Ly
    xi = kx * x + ky * y
    eta = A * np.sin(xi)
    deta = A * np.cos(xi)
    u = (g / f) * (-ky * deta)
    v = (g / f) * (kx * deta)
    H = H0 + eta
    return u * H, v * H, H


@pytest.mark.parametrize("warp,final", [(0.0, 1e-5), (0.04, 1e-3)])
def test_geostrophic_jet_residual_converges_spectrally(warp, final):
    f, H0, A = 1.0, 10.0, 0.1
    Lx, Ly = 2.0, 3.0
    res = []
    for N in (3, 5, 7, 9):
        n1 = N + 1
        orc, info = build_oracle_case(N, 4, 4, True, True, Lx=Lx, Ly=Ly, warp=warp, g=G)
        E = 16
        U, V, H = _jet_state(info["x"], info["y"], Lx, Ly, f, G, H0, A)
        orc.sol[:] = np.stack([U, V, H], axis=1).ravel()
        bathy = np.zeros((E, 3, n1, n1))
        bathy[:, 0] = H0
        orc.bathy[:] = bathy.ravel()
        orc.bathymetry_gradient()
        orc.plv[:] = f
        orc.global_time_derivative(0.0)
        t = orc.tend.reshape(E, 3, n1, n1)
        # scale: the Coriolis force on the jet, f |U|
        res.append(np.abs(t[:, :2]).max() / (f * np.abs(U).max()))
    # spectral: each +2 in degree gains well over an order of magnitude
    assert res[0] > 10 * res[1] > 100 * res[2], res
    assert res[-1] < final, res


def test_geostrophic_jet_with_flipped_rotation_is_not_steady():
    """The same jet with f of the other sign (the balance needs f u = -g
    grad eta x z): the residual is of the order of the Coriolis force."""
    f, H0, A = 1.0, 10.0, 0.1
    N, n1, E = 7, 8, 16
    orc, info = build_oracle_case(N, 4, 4, True, True, Lx=2.0, Ly=3.0, g=G)
    U, V, H = _jet_state(info["x"], info["y"], 2.0, 3.0, f, G, H0, A)
    orc.sol[:] = np.stack([U, V, H], axis=1).ravel()
    bathy = np.zeros((E, 3, n1, n1))
    bathy[:, 0] = H0
    orc.bathy[:] = bathy.ravel()
    orc.bathymetry_gradient()
    orc.plv[:] = -f
    orc.global_time_derivative(0.0)
    t = orc.tend.reshape(E, 3, n1, n1)
    assert np.abs(t[:, :2]).max() > 1.5 * f * np.abs(U).max()
