"""The batched device element operators against the reference's own
routines (tests/golden/ref_nodal.npz: ApplyDerivativeMatrix and
CalculateAtBoundaries of src/nodal/NodalStorage_2D/3D_Class.f90, CoarseToFine
of src/interp/Lagrange_2D/3D_Class.f90, compiled unmodified).

Each fixture field is batched with exact scalings of itself (x2, x-1, x0.5),
whose reference results are the same scalings, through the C ABI with the
reference's own matrices and interpolants; tolerance 1e-13 of the largest
output (the device contracts to FMA).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_nodal.npz"))
CASES = sorted({k.split("/")[0] for k in Z.files})
SCALE = np.array([1.0, 2.0, -1.0, 0.5])


def case(name):
    return {k.split("/", 1)[1]: Z[k] for k in Z.files if k.startswith(name + "/")}


def _close(got, ref):
    assert np.abs(got - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1.0)


@pytest.mark.parametrize("name", CASES)
def test_device_operators_match_reference(gpu_lib, name):
    from self_amd import nodal
    g = case(name)
    s = SCALE
    f2 = s[:, None, None] * g["f2"].T[None]                          # (e, j, i)
    ds, dp = nodal.apply_derivative_2d(g["dMatS"], g["dMatP"], f2)
    _close(ds.reshape(f2.shape), s[:, None, None] * g["dFds2"].T[None])
    _close(dp.reshape(f2.shape), s[:, None, None] * g["dFdp2"].T[None])
    got = nodal.boundary_interp_2d(g["lagWest"], g["lagEast"], g["lagSouth"], g["lagNorth"], f2)
    for o, k in zip(got, ("fWest2", "fEast2", "fSouth2", "fNorth2")):
        _close(o, s[:, None] * g[k][None])

    f3 = s[:, None, None, None] * g["f3"].transpose(2, 1, 0)[None]    # (e, k, j, i)
    outs = nodal.apply_derivative_3d(g["dMat3"], f3)
    for o, k in zip(outs, ("dFds3", "dFdp3", "dFdq3")):
        _close(o.reshape(f3.shape), s[:, None, None, None] * g[k].transpose(2, 1, 0)[None])
    faces = nodal.boundary_interp_3d(g["lagWest"], g["lagEast"], f3).transpose(0, 1, 3, 2)  # (e, face, a, b)
    for i, k in enumerate(("fSouth3", "fEast3", "fNorth3", "fWest3", "fBottom3", "fTop3")):
        _close(faces[:, i], s[:, None, None] * g[k][None])

    T = g["Ts"]
    fine2 = nodal.coarse_to_fine_2d(T, T, f2)
    _close(fine2.transpose(0, 2, 1), s[:, None, None] * g["fine2"][None])
    fine3 = nodal.coarse_to_fine_3d(T, f3)
    _close(fine3.transpose(0, 3, 2, 1), s[:, None, None, None] * g["fine3"][None])
