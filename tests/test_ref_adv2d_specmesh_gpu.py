"""The HIP advection-diffusion path against the reference's own model on
an ISM-V2 mesh file (tests/golden/ref_adv2dspec_annulus_n5.npz: curved,
unstructured, InflowOne / InflowTwo / NO_NORMAL_FLOW sides).  The model is
built on the reference's nodal arrays, edge table and geometry through the
C ABI and must agree to relative L2 1e-12 per tracer, as on the default
mesh (tests/test_ref_adv2d_gpu.py)."""
import os

import numpy as np
import pytest

from test_ref_adv2d_gpu import _cmp, _model

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_adv2dspec_annulus_n5.npz")


def test_global_time_derivative_matches_reference(gpu_lib):
    g = dict(np.load(GOLDEN))
    print("adv2d on the annulus spec mesh: tendency", flush=True)
    m, E4 = _model(g)
    nq = int(g["nTracers"])
    m.GlobalTimeDerivative(float(g["t0"]))
    gx, gy = m.GetGradient()
    _cmp(gx, E4(g["dcdx"]), nq, "dcdx")
    _cmp(gy, E4(g["dcdy"]), nq, "dcdy")
    _cmp(m.GetTendency(), E4(g["tend"]), nq, "tendency")
    m.Trash()


def test_rk3_steps_match_reference(gpu_lib):
    g = dict(np.load(GOLDEN))
    print("adv2d on the annulus spec mesh: RK3", flush=True)
    m, E4 = _model(g)
    m.ForwardStepRK3(float(g["t0"]), int(g["nSteps"]))
    _cmp(m.GetSolution(), E4(g["sol"]), int(g["nTracers"]), "solution")
    m.Trash()
