"""GPU parity of the hydrostatic primitive equations (self_amd/csrc/hpe.hip)
against the oracle (oracle/orc_hpe.c, pinned by the closed forms of
tests/test_oracle_hpe.py), through the C ABI.

* One time derivative on walled, periodic and warped stacked meshes, N = 1..7,
  with and without the diagnosed pressure: the diagnosed w and p, the
  divergence of the lateral velocity and the tendency to relative L2 1e-12
  (the device folds the column quadrature into matrices and uses fused
  multiply-adds; the oracle sums in the reference's order).
* Five RK3 steps and three AB2 steps: the prognostic state to 1e-11.
* The closed-form tendency of the balanced polynomial state and the zero
  tendency of a rest state in hydrostatic balance directly on the device.
* At 16 x 16 columns x 8 layers, N = 7: lateral momentum is conserved to
  round-off on a periodic flat mesh (no lateral flux through the flat top
  and bottom), a size-independent property.
"""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [
    dict(N=3, n=(3, 2, 3), periodic=(False, False), warp=0.0, dp=True),
    dict(N=7, n=(2, 2, 2), periodic=(True, True), warp=0.0, dp=True),
    dict(N=5, n=(2, 3, 2), periodic=(True, False), warp=0.03, dp=True),
    dict(N=4, n=(2, 2, 3), periodic=(False, True), warp=0.02, dp=False),
    dict(N=1, n=(4, 3, 2), periodic=(False, False), warp=0.0, dp=True),
    dict(N=6, n=(2, 2, 2), periodic=(True, True), warp=0.04, dp=True),
]
L3 = (2.0, 3.0, 0.5)


def _ids(c):
    return f"N{c['N']}-{'x'.join(map(str, c['n']))}-p{''.join(str(int(p)) for p in c['periodic'])}-w{c['warp']}" + (
        "" if c["dp"] else "-noP")


def _state(x, y, z, seed=5):
    rng = np.random.default_rng(seed)
    E = x.shape[0]
    sol = np.zeros((E, 5) + x.shape[1:])
    sol[:, 0] = 0.3 * np.sin(np.pi * x) * np.cos(np.pi * z) + 0.02 * rng.standard_normal(x.shape)
    sol[:, 1] = 0.2 * np.cos(2 * np.pi * y / 3) * (z - 0.2) + 0.02 * rng.standard_normal(x.shape)
    sol[:, 2] = 0.5 * np.cos(np.pi * x) * np.exp(z) + 0.1 * y + 0.02 * rng.standard_normal(x.shape)
    sol[:, 3] = 0.1 * rng.standard_normal(x.shape)  # overwritten by DiagnoseW
    sol[:, 4] = 0.1 * rng.standard_normal(x.shape)  # overwritten by DiagnoseP when enabled
    bg = np.zeros_like(sol)
    bg[:, 2] = 1.0 + 0.5 * z
    bg[:, 3] = 0.05 * np.sin(2 * np.pi * x)
    return sol, bg


def _pair(case, dt=1e-3):
    from self_amd.hydrostatic import HydrostaticParams, HydrostaticPrimitive
    N, (nX, nY, nL) = case["N"], case["n"]
    sign = (1, -1, 1)
    orc = po.HydrostaticPrimitive(N, nX, nY, nL, periodic=case["periodic"], scale=L3, warp=case["warp"],
                                  warpSign=sign, dt=dt, diagnoseP=case["dp"])
    gpu = HydrostaticPrimitive(HydrostaticParams(polyDeg=N, nXElem=nX, nYElem=nY, nZElem=nL, dt=dt, xScale=L3[0],
                                                 yScale=L3[1], zScale=L3[2], periodic=case["periodic"],
                                                 warp=case["warp"], warpSign=sign, diagnoseP=case["dp"]))
    E, n1 = orc.nElems, N + 1
    assert gpu.nElems == E
    x, y, z = (getattr(orc.mesh, a).reshape(E, n1, n1, n1) for a in ("x", "y", "z"))
    assert np.abs(gpu.mesh.x - x).max() < 1e-14 and np.abs(gpu.mesh.z - z).max() < 1e-14
    sol, bg = _state(x, y, z - L3[2])
    orc.view(orc.sol, 5)[:] = sol
    orc.view(orc.bg, 5)[:] = bg
    gpu.SetSolution(sol)
    gpu.SetBackgroundState(bg)
    return orc, gpu


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_global_time_derivative_matches_oracle(gpu_lib, case):
    orc, gpu = _pair(case)
    print(f"hpe gtd {_ids(case)}", flush=True)
    orc.call("global_time_derivative")
    gpu.GlobalTimeDerivative()
    so, sg = orc.view(orc.sol, 5), gpu.GetSolution()
    for f in (3, 4):
        assert rel_l2(sg[:, f], so[:, f]) < 1e-12, (f, rel_l2(sg[:, f], so[:, f]))
    assert rel_l2(gpu.GetDivU().ravel(), orc.divU) < 1e-12
    to, tg = orc.view(orc.tend, 3), gpu.GetTendency()
    for f in range(3):
        assert rel_l2(tg[:, f], to[:, f]) < 1e-12, (f, rel_l2(tg[:, f], to[:, f]))
    # face values of b, w, p as the oracle's traces
    n1 = case["N"] + 1
    fo = orc.bsol.reshape(orc.nElems, 6, 5, n1, n1)[:, :, 2:]
    assert rel_l2(gpu.GetFaceValues().ravel(), fo.ravel()) < 1e-12
    gpu.Trash()


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[3]], ids=_ids)
def test_rk3_and_ab2_match_oracle(gpu_lib, case):
    orc, gpu = _pair(case, dt=2e-3)
    for _ in range(5):
        orc.call("forward_step_rk3")
    gpu.ForwardStepRK3(5)
    so, sg = orc.view(orc.sol, 5), gpu.GetSolution()
    for f in range(3):
        assert rel_l2(sg[:, f], so[:, f]) < 1e-11, (f, rel_l2(sg[:, f], so[:, f]))
    orc.G[:] = 0.0
    gpu.SetPreviousTendency(np.zeros((orc.nElems, 3) + so.shape[2:]))
    for _ in range(3):
        orc.call("forward_step_ab2")
    gpu.ForwardStepAB2(3)
    so, sg = orc.view(orc.sol, 5), gpu.GetSolution()
    for f in range(3):
        assert rel_l2(sg[:, f], so[:, f]) < 1e-11, (f, rel_l2(sg[:, f], so[:, f]))
    gpu.Trash()


def test_closed_form_tendency_on_the_device(gpu_lib):
    from self_amd.hydrostatic import HydrostaticParams, HydrostaticPrimitive
    H, N = 0.5, 5
    g = HydrostaticPrimitive(HydrostaticParams(polyDeg=N, nXElem=3, nYElem=2, nZElem=3, xScale=2.0, yScale=3.0,
                                               zScale=H))
    x, y, z = g.mesh.x, g.mesh.y, g.mesh.z - H
    sol = np.zeros((g.nElems, 5) + x.shape[1:])
    bg = np.zeros_like(sol)
    B = 0.7
    sol[:, 0] = (x ** 2 - 0.3 * x) * (z + H / 2)
    sol[:, 1] = y ** 3 * (z ** 2 - H ** 2 / 3)
    sol[:, 2] = 0.2 * x ** 3 * (1.0 + z) + (y - 0.5 * y ** 2) * z ** 2
    bg[:, 2] = B
    g.SetSolution(sol)
    g.SetBackgroundState(bg)
    g.GlobalTimeDerivative()
    t = g.GetTendency()
    expect = [0.6 * x ** 2 * (-(z + z ** 2 / 2)), (1.0 - y) * (-z ** 3 / 3),
              B * ((2 * x - 0.3) * (z + H / 2) + 3 * y ** 2 * (z ** 2 - H ** 2 / 3))]
    for f in range(3):
        assert np.abs(t[:, f] - expect[f]).max() < 1e-11 * max(1.0, np.abs(expect[f]).max()), f
    # rest state in hydrostatic balance
    sol[:] = 0.0
    sol[:, 2] = np.exp(3 * z) + np.sin(7 * z)
    bg[:, 2] = 1.0 + z
    g.SetSolution(sol)
    g.SetBackgroundState(bg)
    g.GlobalTimeDerivative()
    assert np.abs(g.GetTendency()).max() < 1e-12
    g.Trash()


def test_full_size_lateral_momentum_conservation(gpu_lib):
    from self_amd.hydrostatic import HydrostaticParams, HydrostaticPrimitive
    N, H = 7, 0.5
    g = HydrostaticPrimitive(HydrostaticParams(polyDeg=N, nXElem=16, nYElem=16, nZElem=8, xScale=2.0, yScale=3.0,
                                               zScale=H, periodic=(True, True), dt=5e-4))
    x, y, z = g.mesh.x, g.mesh.y, g.mesh.z - H
    sol, bg = _state(x, y, z, seed=9)
    g.SetSolution(sol)
    g.SetBackgroundState(bg)
    g.GlobalTimeDerivative()
    t = g.GetTendency()
    w = g.dgStorage.qWeight
    W3 = np.einsum("k,j,i->kji", w, w, w)
    for f in range(2):
        tot = np.sum(W3 * g.mesh.J * t[:, f])
        scale = np.sum(W3 * g.mesh.J * np.abs(t[:, f]))
        assert abs(tot) < 1e-12 * scale, (f, tot, scale)
    g.ForwardStepRK3(10)
    s = g.GetSolution()
    assert np.all(np.isfinite(s))
    # momentum stays where it was to round-off over the steps
    for f in range(2):
        m0 = np.sum(W3 * g.mesh.J * sol[:, f])
        m1 = np.sum(W3 * g.mesh.J * s[:, f])
        assert abs(m1 - m0) < 1e-11 * np.sum(W3 * g.mesh.J * np.abs(sol[:, f])), (f, m0, m1)
    g.Trash()
