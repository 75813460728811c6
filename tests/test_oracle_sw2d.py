"""Pins of the oracle's shallow-water DGSEM (oracle/orc_sw2d.c) by what the
mathematics fixes: flux consistency and antisymmetry, wall reflection, the
lake at rest with a linear bottom (derivative exactness + bathymetry source),
free-stream preservation, exact discrete mass conservation, third-order RK3
convergence, and the mesh geometry (areas, outward unit normals, edge table).
A dropped term, a wrong sign or index, or a transposed operand fails one of
these."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from sw2d_cases import build_oracle_case

G = 9.81


def _phys_flux(q, n):
    U, V, H = q
    f = np.array([U * U / H + 0.5 * G * H * H, U * V / H, U])
    g = np.array([U * V / H, V * V / H + 0.5 * G * H * H, V])
    return f * n[0] + g * n[1]


def test_riemann_consistency_and_antisymmetry():
    rng = np.random.default_rng(0)
    for _ in range(50):
        a = np.array([rng.normal(), rng.normal(), rng.uniform(0.5, 3.0)])
        b = np.array([rng.normal(), rng.normal(), rng.uniform(0.5, 3.0)])
        th = rng.uniform(0, 2 * math.pi)
        n = np.array([math.cos(th), math.sin(th)])
        assert np.allclose(po.riemann(a, a, n, G), _phys_flux(a, n), rtol=1e-13, atol=1e-13)
        assert np.allclose(po.riemann(a, b, n, G), -po.riemann(b, a, -n, G), rtol=1e-13, atol=1e-13)
        # upwinding: the dissipation is -0.5*lambda*(b - a), lambda = max |u_n +- c|
        lam = max(abs(a[:2] @ n / a[2]) + math.sqrt(G * a[2]), abs(b[:2] @ n / b[2]) + math.sqrt(G * b[2]))
        central = 0.5 * (_phys_flux(a, n) + _phys_flux(b, n))
        assert np.allclose(po.riemann(a, b, n, G), central - 0.5 * lam * (b - a), rtol=1e-12, atol=1e-12)


def test_wall_state_reflects_normal_momentum():
    L = po.lib()
    import ctypes as C
    rng = np.random.default_rng(1)
    for _ in range(20):
        th = rng.uniform(0, 2 * math.pi)
        n = np.array([math.cos(th), math.sin(th)])
        q = np.array([rng.normal(), rng.normal(), 2.0])
        ex = np.zeros(3)
        L.orc_sw2d_external_state(po.NO_NORMAL_FLOW, po._p(n), po._p(q), po._p(ex))
        t = np.array([-n[1], n[0]])
        assert abs(ex[:2] @ n + q[:2] @ n) < 1e-14
        assert abs(ex[:2] @ t - q[:2] @ t) < 1e-14
        assert ex[2] == q[2]
        L.orc_sw2d_external_state(po.RADIATION, po._p(n), po._p(q), po._p(ex))
        assert np.all(ex == 0.0)


def test_lake_at_rest_with_linear_bottom_walled():
    """H = h = a + b x + c y, U = V = 0 on a walled affine mesh: the weak
    pressure divergence and the g H grad(h) source cancel exactly (all
    operands are polynomials of degree <= N)."""
    N = 5
    orc, info = build_oracle_case(N, 3, 4, False, False, Lx=2.0, Ly=3.0)
    n1 = N + 1
    E = 12
    x, y = info["x"], info["y"]
    h = 10.0 + 0.3 * x - 0.2 * y
    sol = np.stack([0 * h, 0 * h, h], axis=1)
    orc.sol[:] = sol.ravel()
    orc.bathy[:] = np.stack([h, 0 * h, 0 * h], axis=1).ravel()
    orc.plv[:] = 0.7
    orc.bathymetry_gradient()
    b = orc.bathy.reshape(E, 3, n1, n1)
    assert np.abs(b[:, 1] - 0.3).max() < 1e-12 and np.abs(b[:, 2] + 0.2).max() < 1e-12
    orc.global_time_derivative(0.0)
    t = orc.tend.reshape(E, 3, n1, n1)
    scale = G * 10.0 * 0.3
    assert np.abs(t).max() < 1e-11 * scale


def test_constant_state_is_preserved_periodic_affine():
    N = 7
    orc, info = build_oracle_case(N, 4, 3, True, True)
    E, n1 = 12, 8
    sol = np.zeros((E, 3, n1, n1))
    sol[:, 0], sol[:, 1], sol[:, 2] = 1.3, -0.4, 5.0
    orc.sol[:] = sol.ravel()
    orc.bathy[:] = np.stack([np.full((E, n1, n1), 5.0), np.zeros((E, n1, n1)), np.zeros((E, n1, n1))], 1).ravel()
    orc.plv[:] = 0.0
    orc.bathymetry_gradient()
    orc.global_time_derivative(0.0)
    # round-off scale: |Dhat| * (g H^2 / 2) * |metric| / J summed over a line
    w = orc.nodal.qWeight
    scale = np.abs(orc.nodal.dMatS).sum(axis=1).max() * 0.5 * G * 25.0 * 1.5 / orc.mesh.J.min()
    assert np.abs(orc.tend).max() < 1e-14 * scale


@pytest.mark.parametrize("periodic,warp", [((True, True), 0.0), ((True, True), 0.05), ((False, False), 0.04),
                                           ((True, False), 0.0)])
def test_discrete_mass_conservation(periodic, warp):
    """sum_e sum_ij J w_i w_j dH/dt = 0: edge fluxes cancel pairwise and the
    wall flux of H vanishes for the reflected state."""
    N = 5
    orc, info = build_oracle_case(N, 4, 3, *periodic, warp=warp)
    orc.bathymetry_gradient()
    orc.global_time_derivative(0.0)
    n1 = N + 1
    w = orc.nodal.qWeight
    J = orc.mesh.J.reshape(-1, n1, n1)
    t = orc.tend.reshape(-1, 3, n1, n1)
    total = np.sum(J * w[None, :, None] * w[None, None, :] * t[:, 2])
    scale = np.sum(J * w[None, :, None] * w[None, None, :] * np.abs(t[:, 2]))
    assert abs(total) < 1e-13 * scale


def test_rk3_is_third_order():
    """dH/dt = -tau*H at every node (U = V = 0, constant H, relaxation to
    H_r = 0): ForwardStepRK3 (ConservativeShallowWaterClass.f90:1173-1218)
    must converge with order 3 to H0 exp(-tau t)."""
    N = 3
    errs = []
    for nsteps in (10, 20, 40):
        n1 = N + 1
        E = 9
        relax = np.zeros((E, 4, n1, n1))
        relax[:, 3] = 2.0  # 1/timescale
        dt = 1.0 / nsteps
        # g = 0 removes gravity waves, whose round-off-seeded growth would
        # otherwise dominate at these deliberately large steps
        orc, info = build_oracle_case(N, 3, 3, True, True, dt=dt, relax=relax, g=0.0)
        sol = np.zeros((E, 3, n1, n1))
        sol[:, 2] = 4.0
        orc.sol[:] = sol.ravel()
        orc.bathy[:] = np.stack([np.full((E, n1, n1), 4.0), 0 * sol[:, 0], 0 * sol[:, 0]], 1).ravel()
        orc.plv[:] = 0.0
        for k in range(nsteps):
            orc.forward_step_rk3(k * dt)
        H = orc.sol.reshape(E, 3, n1, n1)[:, 2]
        errs.append(np.abs(H - 4.0 * math.exp(-2.0)).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 7.0 < r1 < 9.5 and 7.0 < r2 < 9.5, errs


def test_mesh_geometry_areas_normals_edges():
    N = 5
    nodal = po.NodalStorage2D(N)
    for px, py, warp in ((True, True, 0.0), (False, False, 0.05), (True, False, 0.03)):
        m = po.QuadMesh(4, 3, N, nodal.nodes, px, py, 2.0, 3.0, warp)
        n1 = N + 1
        w = nodal.qWeight
        J = m.J.reshape(-1, n1, n1)
        assert np.all(J > 0)
        area = np.sum(J * w[None, :, None] * w[None, None, :])
        assert abs(area - 6.0) < 1e-12 * 6.0
        nh = m.nHat.reshape(-1, 4, n1, 2)
        assert np.abs(np.linalg.norm(nh, axis=-1) - 1.0).max() < 1e-14
        x = m.x.reshape(-1, n1, n1)
        y = m.y.reshape(-1, n1, n1)
        xc, yc = x.mean(axis=(1, 2)), y.mean(axis=(1, 2))
        # outward: the normal at a south/north/east/west side node points away from the element centre
        for side, (ii, jj) in enumerate([(slice(None), 0), (-1, slice(None)), (slice(None), -1), (0, slice(None))]):
            px_ = x[:, jj, ii] if side in (0, 2) else x[:, jj, ii]
            py_ = y[:, jj, ii] if side in (0, 2) else y[:, jj, ii]
            d = (px_ - xc[:, None]) * nh[:, side, :, 0] + (py_ - yc[:, None]) * nh[:, side, :, 1]
            assert np.all(d > 0), side
        # every element side belongs to exactly one edge; interior neighbours are the grid neighbours
        seen = np.zeros((m.nElems, 4), dtype=int)
        for e1, s1, e2, s2, start, inc in m.edges:
            seen[e1 - 1, s1 - 1] += 1
            if e2 > 0:
                seen[e2 - 1, abs(s2) - 1] += 1
                assert s2 > 0 and (start, inc) == (1, 1)
                ix1, iy1 = (e1 - 1) % 4, (e1 - 1) // 4
                ix2, iy2 = (e2 - 1) % 4, (e2 - 1) // 4
                if s1 in (2, 4):
                    assert iy1 == iy2 and (ix2 - ix1) % 4 in (1, 3)
                else:
                    assert ix1 == ix2 and (iy2 - iy1) % 3 in (1, 2)
            else:
                assert e2 == po.NO_NORMAL_FLOW
        assert np.all(seen == 1)
