"""Pins the advection-diffusion oracle (oracle/orc_adv2d.c) with closed forms.

Each check is chosen so that a plausible restatement mistake fails it:
  * exact gradient of a linear field     -- dMatP orientation, metric terms,
                                             central-flux sign, lift weights
  * constant state has zero tendency     -- flux/volume balance, normals
  * mass conservation on a warped mesh   -- edge flux antisymmetry, J weighting
  * translation of a sine at speed u     -- upwind sign, RK3 coefficients
  * diffusion decay exp(-8 pi^2 kappa t) -- kappa sign and BR1 gradient flux
  * relaxation source (r - c) k          -- Source :1829-1847
  * inflow boundary closed form          -- GetExternalState, boundary lift
  * third-order convergence in dt        -- low-storage RK3 (a, b, g)
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from adv2d_cases import build_oracle_case, flag_inflow


def _model(N, nX, periodic=True, Lx=1.0, warp=0.0, kappa=0.0, dt=1e-3, nq=1, edges_fn=None, relax=False):
    nodal = po.NodalStorage2D(N)
    mesh = po.QuadMesh(nX, nX, N, nodal.nodes, periodic, periodic, Lx, Lx, warp)
    edges = mesh.edges if edges_fn is None else edges_fn(mesh.edges)
    m = po.Advection2D(nodal, mesh, nTracers=nq, kappa=kappa, dt=dt, with_relax=relax, edges=edges)
    n1, E = N + 1, mesh.nElems
    return m, nodal, mesh, mesh.x.reshape(E, n1, n1), mesh.y.reshape(E, n1, n1)


def _interior_elements(mesh):
    """Elements none of whose sides lie on a boundary edge."""
    bad = set()
    for e1, s1, e2, s2, st, inc in mesh.edges:
        if e2 <= 0:
            bad.add(e1 - 1)
    return [e for e in range(mesh.nElems) if e not in bad]


def test_gradient_of_linear_field_is_exact_on_interior_elements():
    m, nodal, mesh, x, y = _model(N=5, nX=5, periodic=False, Lx=2.0)
    m.c[:] = (2.0 * x - 3.0 * y + 0.7).ravel()
    m.global_time_derivative(0.0)
    n1 = 6
    gx = m.dcdx.reshape(-1, n1, n1)
    gy = m.dcdy.reshape(-1, n1, n1)
    inner = _interior_elements(mesh)
    assert len(inner) == 9
    assert np.abs(gx[inner] - 2.0).max() < 1e-12
    assert np.abs(gy[inner] + 3.0).max() < 1e-12


def test_transposed_dmatp_breaks_linear_gradient():
    """The reference's stored 2-D dMatP is the transpose of the weak operator;
    with it, the y-gradient of a linear field is no longer exact."""
    nodal = po.NodalStorage2D(5)
    mesh = po.QuadMesh(5, 5, 5, nodal.nodes, False, False, 2.0, 2.0)
    m = po.Advection2D(nodal, mesh, dMatP=nodal.dMatP)
    y = mesh.y.reshape(-1, 6, 6)
    m.c[:] = (-3.0 * y).ravel()
    m.global_time_derivative(0.0)
    inner = _interior_elements(mesh)
    assert np.abs(m.dcdy.reshape(-1, 6, 6)[inner] + 3.0).max() > 1e-2


def test_constant_state_has_zero_tendency():
    m, nodal, mesh, x, y = _model(N=6, nX=4, Lx=1.0, kappa=0.1, nq=2)
    m.c[:] = 1.7
    m.u[:] = 1.3
    m.v[:] = 0.4
    m.global_time_derivative(0.0)
    assert np.abs(m.tend).max() < 1e-10
    assert np.abs(m.dcdx).max() < 1e-10 and np.abs(m.dcdy).max() < 1e-10


def test_free_stream_error_on_a_warped_mesh_converges_spectrally():
    """On a curved mesh the free stream is kept only to the accuracy of the
    mapping, which improves exponentially with N."""
    errs = []
    for N in (4, 8):
        m, nodal, mesh, x, y = _model(N=N, nX=4, warp=0.05)
        m.c[:] = 1.7
        m.u[:] = 1.3
        m.v[:] = 0.4
        m.global_time_derivative(0.0)
        errs.append(np.abs(m.tend).max())
    assert errs[1] < 1e-2 * errs[0], errs


def test_mass_is_conserved_on_a_warped_periodic_mesh():
    orc, info = build_oracle_case(6, 4, 4, warp=0.05, nq=2, kappa=0.02, dt=2e-4)
    nodal, mesh = info["nodal"], info["mesh"]
    n1, E = 7, mesh.nElems
    w2 = np.outer(nodal.qWeight, nodal.qWeight)
    J = mesh.J.reshape(E, n1, n1)

    def mass():
        c = orc.c.reshape(E, 2, n1, n1)
        return [float(np.sum(c[:, f] * J * w2)) for f in range(2)]

    m0 = mass()
    for s in range(5):
        orc.forward_step_rk3(s * 2e-4)
    m1 = mass()
    for a, b in zip(m0, m1):
        assert abs(a - b) < 1e-13 * abs(a)


def test_translation_of_a_sine_matches_the_exact_solution():
    dt, T = 2e-3, 0.1
    m, nodal, mesh, x, y = _model(N=8, nX=6, dt=dt)
    m.c[:] = np.sin(2 * math.pi * (x + 2 * y)).ravel()
    m.u[:] = 1.0
    m.v[:] = -0.5
    nsteps = int(round(T / dt))
    for s in range(nsteps):
        m.forward_step_rk3(s * dt)
    exact = np.sin(2 * math.pi * ((x - T) + 2 * (y + 0.5 * T)))
    assert np.abs(m.c - exact.ravel()).max() < 2e-6


def test_diffusion_decays_at_the_analytic_rate():
    kappa, dt, T = 0.01, 1e-3, 0.2
    m, nodal, mesh, x, y = _model(N=8, nX=4, kappa=kappa, dt=dt)
    c0 = np.sin(2 * math.pi * x) * np.sin(2 * math.pi * y)
    m.c[:] = c0.ravel()
    nsteps = int(round(T / dt))
    for s in range(nsteps):
        m.forward_step_rk3(s * dt)
    exact = c0 * math.exp(-8 * math.pi ** 2 * kappa * T)
    assert np.abs(m.c - exact.ravel()).max() < 1e-6 * np.abs(exact).max() + 1e-9


def test_relaxation_source():
    m, nodal, mesh, x, y = _model(N=4, nX=3, relax=True)
    c0 = 0.3
    m.c[:] = c0
    r = np.cos(2 * math.pi * x) + 0.5
    k = 1.0 + 0.5 * np.sin(2 * math.pi * y)
    m.relax[:] = r.ravel()
    m.rFac[:] = k.ravel()
    m.global_time_derivative(0.0)
    assert np.abs(m.tend - ((r - c0) * k).ravel()).max() < 1e-14


@pytest.mark.parametrize("nq", [1, 2])
def test_inflow_boundary_closed_form(nq):
    """c = 0, u = (1, 0), kappa = 0: only the upwind inflow flux at western
    boundary sides acts, tend = tanh(t) |n| l_W(i) / (w_i J)."""
    N, t = 5, 0.7
    m, nodal, mesh, x, y = _model(N=N, nX=4, periodic=False, nq=nq, edges_fn=flag_inflow)
    m.u[:] = 1.0
    m.v[:] = 0.0
    m.global_time_derivative(t)
    n1, E = N + 1, mesh.nElems
    tend = m.tend.reshape(E, nq, n1, n1)
    J = mesh.J.reshape(E, n1, n1)
    nLen = mesh.nLen.reshape(E, 4, n1)
    expect = np.zeros((E, n1, n1))
    for e1, s1, e2, s2, st, inc in mesh.edges:
        if e2 <= 0 and s1 == 4:
            e = e1 - 1
            expect[e] = (math.tanh(t) * nLen[e, 3][:, None] * nodal.lagWest[None, :]
                         / (nodal.qWeight[None, :] * J[e]))
    assert np.abs(tend[:, 0] - expect).max() < 1e-13 * np.abs(expect).max()
    if nq == 2:
        # tracer 2 enters only through southern sides, where v = 0: no flux
        assert np.abs(tend[:, 1]).max() < 1e-14


def test_rk3_is_third_order_in_time():
    """Richardson: the spatial operator is the same for every dt, so the
    differences of successive halvings shrink by 2^3."""
    T = 0.24
    sols = []
    for dt in (0.004, 0.002, 0.001):
        m, nodal, mesh, x, y = _model(N=6, nX=3, dt=dt, kappa=1e-3)
        m.c[:] = np.sin(2 * math.pi * x).ravel()
        m.u[:] = 1.0
        m.v[:] = 0.5
        for s in range(int(round(T / dt))):
            m.forward_step_rk3(s * dt)
        sols.append(m.c.copy())
    d1 = np.abs(sols[0] - sols[1]).max()
    d2 = np.abs(sols[1] - sols[2]).max()
    order = math.log2(d1 / d2)
    assert 2.8 < order < 3.2, (d1, d2, order)
