"""GPU parity of the 3-D advection path (with the adaptive roll-off filter)
against the oracle (oracle/orc_adv3d.c).

The oracle applies the reference's dense (N+1)^3 x (N+1)^3 filter matrix;
the HIP path applies its three 1-D factors, so filtered states agree to
round-off, and the limiter decisions (threshold tests on energy ratios) are
compared as counts.  Tolerance: relative L2 1e-12 per tracer for one
tendency or three RK3 steps, 1e-11 after 20 steps.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [
    dict(N=3, n=(3, 2, 2), nq=2, periodic=(True, True, True), warp=0.04),
    dict(N=7, n=(2, 2, 2), nq=1, periodic=(True, True, True), warp=0.05),
    dict(N=4, n=(2, 3, 2), nq=2, periodic=(False, False, False), warp=0.03, inflow=True, relax=True),
    dict(N=5, n=(2, 2, 3), nq=1, periodic=(True, False, True), warp=0.0, relax=True),
    dict(N=1, n=(4, 3, 2), nq=1, periodic=(True, True, False), warp=0.0),
    dict(N=2, n=(3, 3, 3), nq=2, periodic=(False, True, True), warp=0.02, filter=False),
]


def _ids(c):
    return (f"N{c['N']}-{'x'.join(map(str, c['n']))}-q{c['nq']}-p{''.join(str(int(p)) for p in c['periodic'])}"
            f"-w{c['warp']}{'-in' if c.get('inflow') else ''}{'-rx' if c.get('relax') else ''}"
            f"{'-nofilter' if c.get('filter') is False else ''}")


def _flag(faces):
    f = np.array(faces, dtype=np.int32).copy()
    b = f[:, 2] <= 0
    f[b & (f[:, 1] == 4), 2] = po.INFLOW_ONE
    f[b & (f[:, 1] == 1), 2] = po.INFLOW_TWO
    return f  # other boundary faces keep NO_NORMAL_FLOW (moving Gaussian)


def _pair(case, dt=2e-3, nc=2):
    from self_amd.advection3d import Advection3D, Advection3DParams
    N, (nX, nY, nZ), nq = case["N"], case["n"], case["nq"]
    useF = case.get("filter", True)
    sign = (1, -1, 1)
    nd = po.Nodal3D(N)
    mesh = po.HexMesh(nX, nY, nZ, N, nd.nodes, periodic=case["periodic"], warp=case["warp"], warpSign=sign)
    faces = _flag(mesh.faces) if case.get("inflow") else mesh.faces
    orc = po.Advection3D(nd, mesh, nTracers=nq, dt=dt, nCutoff=nc, useFilter=useF,
                         with_relax=case.get("relax", False), faces=faces)
    E, n1 = mesh.nElems, N + 1
    x, y, z = (getattr(mesh, a).reshape(E, n1, n1, n1) for a in ("x", "y", "z"))
    rng = np.random.default_rng(11)
    c = np.stack([np.sin(2 * np.pi * (x + f * 0.3)) * np.cos(2 * np.pi * y) + 0.5 * np.cos(2 * np.pi * z)
                  + 0.05 * rng.standard_normal(x.shape) for f in range(nq)], axis=1)
    u = 1.0 + 0.2 * np.sin(2 * np.pi * y)
    v = 0.7 + 0.1 * np.cos(2 * np.pi * z)
    w = 0.3 + 0.1 * np.sin(2 * np.pi * x)
    orc.c[:] = c.ravel()
    orc.u[:], orc.v[:], orc.w[:] = u.ravel(), v.ravel(), w.ravel()
    p = Advection3DParams(polyDeg=N, nXElem=nX, nYElem=nY, nZElem=nZ, nTracers=nq, nCutoff=nc, dt=dt,
                          periodic=case["periodic"], warp=case["warp"], warpSign=sign, useFilter=useF)
    m = Advection3D(p, faces=faces)
    m.SetSolution(c)
    m.SetVelocity(u, v, w)
    if case.get("relax"):
        r = np.cos(2 * np.pi * x)[:, None] * np.ones((1, nq, 1, 1, 1))
        k = (0.5 + 0.25 * np.sin(2 * np.pi * z))[:, None] * np.ones((1, nq, 1, 1, 1))
        orc.relax[:] = r.ravel()
        orc.rFac[:] = k.ravel()
        m.SetRelaxationField(r)
        m.SetRelaxationFactor(k)
    return orc, m


def _cmp(gpu, flat, nq, tol):
    ref = flat.reshape(gpu.shape)
    for f in range(nq):
        err = rel_l2(gpu[:, f], ref[:, f])
        assert err < tol, f"tracer {f}: {err:.3e}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_global_time_derivative_matches_oracle(gpu_lib, case):
    orc, m = _pair(case)
    t = 0.41
    orc.global_time_derivative(t, 1)
    m.GlobalTimeDerivative(t, 1)
    _cmp(m.GetTendency(), orc.tend, case["nq"], 1e-12)
    _cmp(m.GetSolution(), orc.c, case["nq"], 1e-13)   # filtered in place, or not
    m.Trash()


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_rk3_steps_match_oracle(gpu_lib, case):
    orc, m = _pair(case)
    tn, dt = 0.2, 2e-3
    for s in range(3):
        orc.forward_step_rk3(tn + s * dt)
    m.ForwardStepRK3(tn, 3)
    _cmp(m.GetSolution(), orc.c, case["nq"], 1e-12)
    if case.get("filter", True):
        E1, E2, lim = m.GetFilterState()
        nq, E = case["nq"], m.nElems
        assert rel_l2(E1, orc.E1.reshape(E, nq).T) < 1e-12
        assert rel_l2(E2, orc.E2.reshape(E, nq).T) < 1e-11
    m.Trash()


def test_filter_replaces_states_like_the_oracle(gpu_lib):
    """A rough state makes the limiter fire; over four steps the GPU replaces
    as many element states as the oracle's stage-by-stage filter calls."""
    case = dict(N=5, n=(2, 2, 2), nq=1, periodic=(True, True, True), warp=0.03)
    dt = 4e-3
    orc, m = _pair(case, dt=dt, nc=2)
    n1 = 6
    rng = np.random.default_rng(5)
    rough = orc.c + 0.5 * rng.standard_normal(orc.c.size)
    orc.c[:] = rough
    m.SetSolution(rough.reshape(m.nElems, 1, n1, n1, n1))
    m.FilterCount(reset=True)
    # ForwardStepRK3 written out so the oracle's filter count of every stage is seen
    a, b, g = (0.0, -5 / 9, -153 / 128), (0.0, 1 / 3, 3 / 4), (1 / 3, 15 / 16, 8 / 15)
    G = np.zeros_like(orc.c)
    counts = 0
    for s in range(4):
        for st in range(3):
            orc.global_time_derivative(s * dt + b[st] * dt, st + 1)
            counts += orc.nFiltered
            G = a[st] * G + orc.tend
            orc.c[:] = orc.c + g[st] * dt * G
    m.ForwardStepRK3(0.0, 4)
    assert counts > 0
    assert m.FilterCount() == counts
    c = m.GetSolution()
    assert rel_l2(c, orc.c.reshape(c.shape)) < 1e-11
    m.Trash()


def test_tendency_between_steps_follows_the_reference_sequence(gpu_lib):
    """ForwardStepRK3, then GlobalTimeDerivative(t, 2), then ForwardStepRK3:
    the filter runs exactly where the reference runs it."""
    case = dict(N=3, n=(3, 2, 2), nq=1, periodic=(True, True, True), warp=0.04)
    orc, m = _pair(case)
    orc.forward_step_rk3(0.0)
    m.ForwardStepRK3(0.0, 1)
    orc.global_time_derivative(0.1, 2)
    m.GlobalTimeDerivative(0.1, 2)
    _cmp(m.GetTendency(), orc.tend, 1, 1e-12)
    orc.forward_step_rk3(0.2)
    m.ForwardStepRK3(0.2, 1)
    _cmp(m.GetSolution(), orc.c, 1, 1e-12)
    m.Trash()


def test_gaussians_workload_20_steps_matches_oracle(gpu_lib):
    """8 x 8 x 4 periodic elements at N = 7, seeded warp 0.05, u = (1, 0.7, 0.3),
    three seeded Gaussians, filter on (cutoff 3), RK3 at CFL 0.5."""
    from self_amd.workloads import adv3d_gaussians, adv3d_gaussians_fields, adv3d_warp_signs
    m, info = adv3d_gaussians(nX=8, nY=8, nZ=4, N=7)
    nd = po.Nodal3D(7)
    mesh = po.HexMesh(8, 8, 4, 7, nd.nodes, warp=0.05, warpSign=adv3d_warp_signs())
    orc = po.Advection3D(nd, mesh, nTracers=1, dt=info["dt"], nCutoff=3, useFilter=True)
    E = mesh.nElems
    x, y, z = (getattr(mesh, a).reshape(E, 8, 8, 8) for a in ("x", "y", "z"))
    orc.c[:] = adv3d_gaussians_fields(x, y, z).ravel()
    orc.u[:], orc.v[:], orc.w[:] = 1.0, 0.7, 0.3
    for s in range(20):
        orc.forward_step_rk3(s * info["dt"])
    m.ForwardStepRK3(0.0, 20)
    c = m.GetSolution()
    assert rel_l2(c, orc.c.reshape(c.shape)) < 1e-11
    m.Trash()


def test_errors_are_reported(gpu_lib):
    from self_amd import _lib as L
    from self_amd.advection3d import Advection3D, Advection3DParams
    with pytest.raises(L.SelfError):
        Advection3D(Advection3DParams(polyDeg=8))
    m = Advection3D(Advection3DParams(polyDeg=2, nTracers=1))
    with pytest.raises(L.SelfError):
        m.GlobalTimeDerivative(0.0, 4)
    m.Trash()
