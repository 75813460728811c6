"""Pins the 3-D oracle (oracle/orc_adv3d.c): the roll-off filter, the
adaptive filtering logic and the 3-D advection operator.

  * filter keeps constants and scales each Legendre tensor mode by
    exp(-ln2 (a^2 + b^2 + c^2) / nc^2)  -- mode indexing, cutoff, V / V^-1
  * filter = F1 (x) F1 (x) F1 with F1 built here from the 1-D formula
  * adaptive filtering: energies and limiter branches recomputed in numpy
  * free stream on an affine mesh, mass conservation on a warped periodic
    mesh, translation of a sine, RK3 order (Richardson)
  * inflow and NO_NORMAL_FLOW boundary fluxes in closed form
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po


def _legendre(n, x):
    return np.polynomial.legendre.legval(x, [0] * n + [1])


def test_filter_keeps_constants_and_damps_modes_by_the_rolloff_factor():
    N, nc = 5, 3
    nd = po.Nodal3D(N)
    F = po.rolloff_filter3d(N, nc, nd.nodes, nd.qWeight)
    x = nd.nodes
    one = np.ones((N + 1) ** 3)
    assert np.abs(F @ one - 1.0).max() < 1e-13
    alpha = math.log(2.0) / nc ** 2
    for (a, b, c) in [(1, 0, 0), (0, 2, 0), (0, 0, 3), (1, 2, 0), (2, 1, 3), (5, 0, 1), (5, 5, 5)]:
        mode = np.einsum("k,j,i->kji", _legendre(c, x), _legendre(b, x), _legendre(a, x)).ravel()
        got = F @ mode
        assert np.abs(got - math.exp(-alpha * (a * a + b * b + c * c)) * mode).max() < 1e-12, (a, b, c)


def test_filter_is_the_kronecker_cube_of_the_1d_filter():
    N, nc = 4, 2
    nd = po.Nodal3D(N)
    F = po.rolloff_filter3d(N, nc, nd.nodes, nd.qWeight)
    x, w = nd.nodes, nd.qWeight
    alpha = math.log(2.0) / nc ** 2
    L = np.array([_legendre(a, x) for a in range(N + 1)])          # L[a, i]
    norms = (L ** 2) @ w
    V1 = L * w[None, :] / norms[:, None]                             # modal projection
    F1 = L.T @ np.diag(np.exp(-alpha * np.arange(N + 1) ** 2)) @ V1
    assert np.abs(F - np.kron(F1, np.kron(F1, F1))).max() < 1e-13


def _cube(N, n, periodic=True, warp=0.0, sign=(1, -1, 1)):
    nd = po.Nodal3D(N)
    mesh = po.HexMesh(n, n, n, N, nd.nodes, periodic=(periodic,) * 3, warp=warp, warpSign=sign)
    E, n1 = mesh.nElems, N + 1
    xyz = [getattr(mesh, a).reshape(E, n1, n1, n1) for a in ("x", "y", "z")]
    return nd, mesh, xyz


def test_free_stream_on_an_affine_mesh():
    nd, mesh, (x, y, z) = _cube(4, 3)
    m = po.Advection3D(nd, mesh, nTracers=2, useFilter=False)
    m.c[:] = 2.5
    m.u[:], m.v[:], m.w[:] = 1.0, -0.7, 0.3
    m.global_time_derivative(0.0, 1)
    assert np.abs(m.tend).max() < 1e-11


def test_mass_is_conserved_on_a_warped_periodic_mesh():
    nd, mesh, (x, y, z) = _cube(4, 3, warp=0.04)
    m = po.Advection3D(nd, mesh, nTracers=1, dt=2e-3, useFilter=False)
    E, n1 = mesh.nElems, 5
    m.c[:] = (1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.2 * np.sin(2 * np.pi * z)).ravel()
    m.u[:], m.v[:], m.w[:] = 1.0, 0.7, 0.3
    w3 = np.einsum("k,j,i->kji", nd.qWeight, nd.qWeight, nd.qWeight)
    J = mesh.J.reshape(E, n1, n1, n1)

    def mass():
        return float(np.sum(m.c.reshape(E, n1, n1, n1) * J * w3))

    m0 = mass()
    for s in range(4):
        m.forward_step_rk3(s * 2e-3)
    assert abs(mass() - m0) < 1e-13 * abs(m0)


def test_translation_of_a_sine():
    dt, T = 2e-3, 0.06
    nd, mesh, (x, y, z) = _cube(7, 3)
    m = po.Advection3D(nd, mesh, dt=dt, useFilter=False)
    u = (1.0, 0.7, 0.3)
    m.c[:] = np.sin(2 * np.pi * (x + y - z)).ravel()
    m.u[:], m.v[:], m.w[:] = u
    for s in range(int(round(T / dt))):
        m.forward_step_rk3(s * dt)
    exact = np.sin(2 * np.pi * ((x - u[0] * T) + (y - u[1] * T) - (z - u[2] * T)))
    assert np.abs(m.c - exact.ravel()).max() < 5e-5


def test_rk3_is_third_order_in_time():
    T = 0.048
    sols = []
    for dt in (0.008, 0.004, 0.002):
        nd, mesh, (x, y, z) = _cube(4, 2)
        m = po.Advection3D(nd, mesh, dt=dt, useFilter=False)
        m.c[:] = np.sin(2 * np.pi * x).ravel()
        m.u[:], m.v[:], m.w[:] = 1.0, 0.5, 0.25
        for s in range(int(round(T / dt))):
            m.forward_step_rk3(s * dt)
        sols.append(m.c.copy())
    d1 = np.abs(sols[0] - sols[1]).max()
    d2 = np.abs(sols[1] - sols[2]).max()
    assert 2.8 < math.log2(d1 / d2) < 3.2, (d1, d2)


def _flag_faces(faces, side, flag):
    f = faces.copy()
    f[(f[:, 2] <= 0) & (f[:, 1] == side), 2] = flag
    return f


def test_inflow_boundary_closed_form():
    N, t = 4, 0.6
    nd, mesh, _ = _cube(N, 3, periodic=False)
    faces = _flag_faces(mesh.faces, 4, po.INFLOW_ONE)
    m = po.Advection3D(nd, mesh, nTracers=2, useFilter=False, faces=faces)
    m.u[:] = 1.0
    m.global_time_derivative(t, 1)
    E, n1 = mesh.nElems, N + 1
    tend = m.tend.reshape(E, 2, n1, n1, n1)
    J = mesh.J.reshape(E, n1, n1, n1)
    nLen = mesh.nLen.reshape(E, 6, n1, n1)       # [e][side][b][a]
    expect = np.zeros((E, n1, n1, n1))
    for e1, s1, e2, *_ in faces:
        if e2 == po.INFLOW_ONE:
            e = e1 - 1
            # west face node (a, b) = (p, q): tend[k=q][j=p][i]
            expect[e] = (math.tanh(t) * nLen[e, 3][:, :, None] * nd.lagLeft[None, None, :]
                         / (nd.qWeight[None, None, :] * J[e]))
    assert np.abs(tend[:, 0] - expect).max() < 1e-12 * np.abs(expect).max()
    assert np.abs(tend[:, 1]).max() < 1e-14


def test_no_normal_flow_external_state_is_the_moving_gaussian():
    """c = 0, u = (-1, 0, 0): only eastern wall faces see inflow, where the
    external state is exp(-|x_b - t|^2 / 0.04) at the face node."""
    N, t = 3, 0.3
    nd, mesh, _ = _cube(N, 2, periodic=False)
    m = po.Advection3D(nd, mesh, useFilter=False)
    m.u[:] = -1.0
    m.global_time_derivative(t, 1)
    E, n1 = mesh.nElems, N + 1
    J = mesh.J.reshape(E, n1, n1, n1)
    nLen = mesh.nLen.reshape(E, 6, n1, n1)
    bpos = mesh.bpos.reshape(E, 6, n1, n1, 3)
    expect = np.zeros((E, n1, n1, n1))
    for e1, s1, e2, *_ in mesh.faces:
        if e2 <= 0 and s1 == 2:
            e = e1 - 1
            xb = bpos[e, 1]
            ex = np.exp(-((xb[..., 0] - t) ** 2 + (xb[..., 1] - t) ** 2 + (xb[..., 2] - t) ** 2) / 0.04)
            expect[e] = ex[:, :, None] * nLen[e, 1][:, :, None] * nd.lagRight[None, None, :] / (
                nd.qWeight[None, None, :] * J[e])
    assert np.abs(m.tend.reshape(E, n1, n1, n1) - expect).max() < 1e-12 * np.abs(expect).max()


def test_adaptive_filter_energies_and_limiter_branches():
    N = 4
    nd, mesh, (x, y, z) = _cube(N, 2, warp=0.03)
    m = po.Advection3D(nd, mesh, nTracers=1, nCutoff=2)
    E, n1 = mesh.nElems, N + 1
    rng = np.random.default_rng(7)
    c0 = (np.sin(2 * np.pi * x) + 0.3 * rng.standard_normal(x.shape)).ravel()
    F = po.rolloff_filter3d(N, 2, nd.nodes, nd.qWeight)
    w3 = np.einsum("k,j,i->kji", nd.qWeight, nd.qWeight, nd.qWeight).ravel()
    J = mesh.J.reshape(E, -1)
    cE = c0.reshape(E, -1)
    cf = cE @ F.T
    E1 = np.sum(cf ** 2 * J * w3, axis=1)
    E2 = np.sum((cE - cf) ** 2 * J * w3, axis=1)
    # stage 1 from zero energies: dE2 = E2 > 0 and |dE1|/dE2 = E1/E2 > 1 -> lim = xi, no filtering
    m.c[:] = c0
    m.adaptive_filter(1)
    assert np.allclose(m.E1, E1, rtol=1e-12) and np.allclose(m.E2, E2, rtol=1e-12)
    assert np.all(E1 > E2)
    assert np.allclose(m.lim[:E], E2 / E1, rtol=1e-12)
    assert m.nFiltered == 0 and np.array_equal(m.c, c0)
    # stage 2 with priors that make dE2 > 0 and |dE1| <= dE2: the limiter of
    # stage 1 carries over; scaling it below xi/1.05 forces the filter
    m.E1[:] = E1
    m.E2[:] = 0.5 * E2
    m.lim[:E] = 0.5 * E2 / E1
    m.adaptive_filter(2)
    assert np.allclose(m.lim[E:2 * E], 0.5 * E2 / E1, rtol=1e-12)
    assert m.nFiltered == E
    assert np.abs(m.c - cf.ravel()).max() < 1e-13
    # stage 3 with dE2 < 0: lim = xi of the filtered state, no further filtering
    c1 = m.c.copy()
    m.E2[:] = 1e30
    m.adaptive_filter(3)
    assert m.nFiltered == 0 and np.array_equal(m.c, c1)
    assert np.all(m.lim[2 * E:] == m.E2 / m.E1)
