"""The matrix-core precision variants of the 3-D advection stage
(self_adv3d_set_precision: FP32 and split-BF16 contractions of the volume
weak divergence, FP64 everything else) against the FP64 path, which the
other 3-D tests pin to the oracle and to the compiled reference.

* One time derivative on small meshes (N = 3, 5, 7, periodic and walled,
  warped): relative L2 of the tendency against the FP64 path and the
  oracle: FP32 <= 2e-6, split BF16 <= 2e-4.
* The 3-D advection workload (32 x 32 x 16 elements, N = 7, warp 0.05, filter
  on), 200 RK3 steps: relative L2 of the tracer against the FP64 path:
  FP32 <= 1e-6, split BF16 <= 1e-4, and the tracer integral of each form
  within the same bound of the FP64 form's (the roll-off filter keeps each
  element's reference-space mean, not its physical integral on the warped
  mesh, so the integral drifts in FP64 as well).
"""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL_TEND = {"fp32": 2e-6, "bf16x3": 2e-4}
TOL_RUN = {"fp32": 1e-6, "bf16x3": 1e-4}


def _model(N, n, periodic, warp, nq=1):
    from self_amd.advection3d import Advection3D, Advection3DParams
    p = Advection3DParams(polyDeg=N, nXElem=n[0], nYElem=n[1], nZElem=n[2], nTracers=nq, dt=2e-3,
                          periodic=periodic, warp=warp, warpSign=(1, -1, 1), useFilter=False)
    m = Advection3D(p)
    x, y, z = m.mesh.x, m.mesh.y, m.mesh.z
    c = np.stack([np.sin(2 * np.pi * (x + 0.3 * f)) * np.cos(2 * np.pi * y) + 0.5 * np.cos(2 * np.pi * z)
                  for f in range(nq)], axis=1)
    m.SetSolution(c)
    m.SetVelocity(1.0 + 0.2 * np.sin(2 * np.pi * y), 0.7 + 0.1 * np.cos(2 * np.pi * z),
                  0.3 + 0.1 * np.sin(2 * np.pi * x))
    return m, c


@pytest.mark.parametrize("N,n,periodic,warp,nq", [(7, (2, 2, 2), (True, True, True), 0.05, 1),
                                                  (5, (2, 3, 2), (False, True, False), 0.03, 2),
                                                  (3, (3, 2, 2), (True, True, True), 0.0, 1)])
def test_tendency_variants_against_fp64(gpu_lib, N, n, periodic, warp, nq):
    m, c = _model(N, n, periodic, warp, nq)
    m.GlobalTimeDerivative(0.0, 1)
    ref = m.GetTendency()
    # the oracle agrees with the FP64 path (the folded velocity changes rounding only)
    nd = po.Nodal3D(N)
    mesh = po.HexMesh(*n, N, nd.nodes, periodic=periodic, warp=warp, warpSign=(1, -1, 1))
    o = po.Advection3D(nd, mesh, nTracers=nq, dt=2e-3, useFilter=False)
    o.c[:] = c.ravel()
    x, y, z = m.mesh.x, m.mesh.y, m.mesh.z
    o.u[:] = (1.0 + 0.2 * np.sin(2 * np.pi * y)).ravel()
    o.v[:] = (0.7 + 0.1 * np.cos(2 * np.pi * z)).ravel()
    o.w[:] = (0.3 + 0.1 * np.sin(2 * np.pi * x)).ravel()
    o.global_time_derivative(0.0, 1)
    assert rel_l2(ref.ravel(), o.tend) < 1e-12
    for prec, tol in TOL_TEND.items():
        m.SetSolution(c)
        m.SetPrecision(prec)
        m.GlobalTimeDerivative(0.0, 1)
        t = m.GetTendency()
        err = rel_l2(t.ravel(), ref.ravel())
        print(f"N={N} {prec}: tendency rel L2 vs fp64 {err:.3e}", flush=True)
        assert 0.0 < err < tol, (prec, err)
    m.Trash()


def test_workload_200_steps_variants(gpu_lib):
    from self_amd.workloads import adv3d_gaussians
    w = None
    out = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        m, info = adv3d_gaussians(32, 32, 16, 7)
        m.SetPrecision(prec)
        m.ForwardStepRK3(0.0, 200)
        out[prec] = m.GetSolution()
        if w is None:
            q = m.dgStorage.qWeight
            w = np.einsum("k,j,i->kji", q, q, q) * m.mesh.J
            scale = np.sum(w * np.abs(info["c"][:, 0]))
        out[prec + "_integral"] = np.sum(w * out[prec][:, 0]) / scale
        m.Trash()
    for prec, tol in TOL_RUN.items():
        err = rel_l2(out[prec].ravel(), out["fp64"].ravel())
        dI = abs(out[prec + "_integral"] - out["fp64_integral"])
        print(f"32x32x16 N=7 200 steps {prec}: rel L2 vs fp64 {err:.3e}, integral difference {dI:.3e}", flush=True)
        assert err < tol and dI < tol, (prec, err, dI)
