"""GPU parity of the advection-diffusion path against the oracle.

Every test calls the HIP kernels through the C ABI (self_amd.advection ->
libself_amd.so) and compares with the FP64 oracle (oracle/orc_adv2d.c) on
identical seeded inputs.  Tolerance: relative L2 1e-12 per tracer (FMA
contraction and summation order differ); 1e-11 after 100 RK3 steps.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2
from adv2d_cases import build_oracle_case

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASES = [
    dict(N=7, nX=4, nY=3, nq=1),
    dict(N=5, nX=3, nY=4, nq=2, periodicX=False, periodicY=False, inflow=True),
    dict(N=3, nX=5, nY=3, nq=2, periodicX=True, periodicY=False, warp=0.04, relax=True, inflow=True),
    dict(N=8, nX=3, nY=3, nq=1, periodicX=False, periodicY=True, warp=0.05, relax=True),
    dict(N=1, nX=4, nY=4, nq=1),
    dict(N=2, nX=4, nY=3, nq=2, quirk=True),
    dict(N=4, nX=3, nY=3, nq=3, periodicX=False, periodicY=False, inflow=True, relax=True),
    dict(N=8, nX=2, nY=3, nq=4, periodicX=False, periodicY=True, warp=0.03, inflow=True),
]


def _ids(c):
    return (f"N{c['N']}-{c['nX']}x{c['nY']}-q{c['nq']}-p{int(c.get('periodicX', 1))}{int(c.get('periodicY', 1))}"
            f"-w{c.get('warp', 0)}{'-in' if c.get('inflow') else ''}{'-rx' if c.get('relax') else ''}"
            f"{'-quirk' if c.get('quirk') else ''}")


def _pair(case, kappa=1e-2, dt=1e-3):
    from self_amd.advection import Advection, AdvectionParams
    kw = dict(periodicX=case.get("periodicX", True), periodicY=case.get("periodicY", True),
              warp=case.get("warp", 0.0), nq=case["nq"], kappa=kappa, dt=dt, inflow=case.get("inflow", False),
              relax=case.get("relax", False), quirk=case.get("quirk", False))
    orc, info = build_oracle_case(case["N"], case["nX"], case["nY"], **kw)
    p = AdvectionParams(polyDeg=case["N"], nXElem=case["nX"], nYElem=case["nY"], nTracers=case["nq"], dt=dt,
                        kappa=kappa, periodicX=kw["periodicX"], periodicY=kw["periodicY"], warp=kw["warp"])
    m = Advection(p, quirk_dmatp_transposed=kw["quirk"], edges=info["edges"])
    m.SetSolution(info["c"])
    m.SetVelocity(info["u"], info["v"])
    if kw["relax"]:
        m.SetRelaxationField(info["relax"])
        m.SetRelaxationFactor(info["rFac"])
    return orc, m, info


def _cmp(gpu, orc_flat, nq, tol):
    ref = orc_flat.reshape(gpu.shape)
    for f in range(nq):
        assert rel_l2(gpu[:, f], ref[:, f]) < tol, f"tracer {f}: {rel_l2(gpu[:, f], ref[:, f]):.3e}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_global_time_derivative_matches_oracle(gpu_lib, case):
    orc, m, info = _pair(case)
    t = 0.37
    orc.global_time_derivative(t)
    m.GlobalTimeDerivative(t)
    gx, gy = m.GetGradient()
    _cmp(gx, orc.dcdx, case["nq"], TOL)
    _cmp(gy, orc.dcdy, case["nq"], TOL)
    _cmp(m.GetTendency(), orc.tend, case["nq"], TOL)
    # the tendency call leaves the state untouched
    assert np.array_equal(m.GetSolution(), info["c"])
    m.Trash()


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_rk3_steps_match_oracle(gpu_lib, case):
    orc, m, info = _pair(case)
    tn, dt = 0.25, 1e-3
    for s in range(3):
        orc.forward_step_rk3(tn + s * dt)
    m.ForwardStepRK3(tn, 3)
    _cmp(m.GetSolution(), orc.c, case["nq"], TOL)
    m.Trash()


def test_tendency_after_steps_uses_current_state(gpu_lib):
    """Set/step/tendency interleavings keep the cached traces consistent."""
    case = CASES[1]
    orc, m, info = _pair(case)
    orc.forward_step_rk3(0.0)
    m.ForwardStepRK3(0.0, 1)
    orc.global_time_derivative(0.5)
    m.GlobalTimeDerivative(0.5)
    _cmp(m.GetTendency(), orc.tend, case["nq"], TOL)
    # a new solution replaces the cached traces
    c2 = info["c"] * 0.5 + 0.1
    m.SetSolution(c2)
    orc.c[:] = c2.ravel()
    orc.forward_step_rk3(1.0)
    m.ForwardStepRK3(1.0, 1)
    _cmp(m.GetSolution(), orc.c, case["nq"], TOL)
    m.Trash()


def test_blob_workload_100_steps_matches_oracle(gpu_lib):
    """10 x 10 periodic elements, N = 7 Gauss, (u, v) = (1, 0.5), kappa = 1e-3,
    Gaussian blob, RK3 dt = 1e-4 for 100 steps."""
    from self_amd.workloads import adv2d_blob, adv2d_blob_fields
    from oracle import pyoracle as po
    m, info = adv2d_blob()
    nodal = po.NodalStorage2D(7)
    mesh = po.QuadMesh(10, 10, 7, nodal.nodes, True, True, 1.0, 1.0)
    orc = po.Advection2D(nodal, mesh, nTracers=1, kappa=1e-3, dt=1e-4)
    E = mesh.nElems
    x, y = mesh.x.reshape(E, 8, 8), mesh.y.reshape(E, 8, 8)
    orc.c[:] = adv2d_blob_fields(x, y).ravel()
    orc.u[:] = 1.0
    orc.v[:] = 0.5
    for s in range(100):
        orc.forward_step_rk3(s * 1e-4)
    m.ForwardStepRK3(0.0, 100)
    c = m.GetSolution()
    ref = orc.c.reshape(c.shape)
    assert rel_l2(c, ref) < 1e-11
    assert np.abs(c - ref).max() < 1e-11
    m.Trash()


def test_errors_are_reported(gpu_lib):
    from self_amd import _lib as L
    from self_amd.advection import Advection, AdvectionParams
    with pytest.raises(L.SelfError):
        Advection(AdvectionParams(polyDeg=9, nXElem=3, nYElem=3))
    with pytest.raises(L.SelfError):
        Advection(AdvectionParams(polyDeg=3, nXElem=3, nYElem=3, nTracers=5))
    m = Advection(AdvectionParams(polyDeg=3, nXElem=3, nYElem=3))
    with pytest.raises(ValueError):
        m.SetSolution(np.zeros((9, 2, 4, 4)))
    m.Trash()


def test_pickup_write_then_read_resumes_the_run(gpu_lib, tmp_path):
    """WritePickup of a model with relaxation set and ReadPickup into a fresh
    one give the same solution, velocity and relaxation fields, and the same
    states after further steps."""
    import os
    from self_amd.workloads import adv2d_blob
    a, info = adv2d_blob(nX=4, N=5, nTracers=2)
    rng = np.random.default_rng(3)
    shape = a.GetSolution().shape
    a.SetRelaxationField(rng.standard_normal(shape))
    a.SetRelaxationFactor(rng.uniform(0, 2, shape))
    a.ForwardStepRK3(0.0, 3)
    path = a.WritePickup(3, directory=str(tmp_path))
    assert os.path.basename(path) == "Advection.0000000003.pickup"
    b, _ = adv2d_blob(nX=4, N=5, nTracers=2)
    b.ReadPickup(3, directory=str(tmp_path))
    assert np.array_equal(b.GetSolution(), a.GetSolution())
    for x, y in zip(b.GetVelocity(), a.GetVelocity()):
        assert np.array_equal(x, y)
    assert np.array_equal(b.GetRelaxationField(), a.GetRelaxationField())
    assert np.array_equal(b.GetRelaxationFactor(), a.GetRelaxationFactor())
    a.ForwardStepRK3(3e-4, 4)
    b.ForwardStepRK3(3e-4, 4)
    assert np.array_equal(b.GetSolution(), a.GetSolution())
    a.Trash()
    b.Trash()
