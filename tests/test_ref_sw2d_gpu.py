"""The HIP shallow-water path against the reference's own conservative
shallow-water class (examples/shallowwater/planewave/localmods/
ConservativeShallowWaterClass.f90, compiled unmodified; fixtures
tests/golden/ref_sw2d_*.npz written by tests/golden/make_ref_sw2d.py).

The model is built through the C ABI on the reference's nodal arrays (its
transposed dMatP included), edge table and geometry, with
self_sw2d_set_variant(1) (that class's wave speed, Source and bathymetry
gradient).  Plain and fused FP64 forms must agree with the reference to
relative L2 1e-12 per field (the GPU contracts to FMA and sums in another
order; the oracle is bitwise equal to the reference); the mixed form is held
to 1e-6, its gate against FP64.
"""
import glob
import os
import types

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_sw2d_*.npz")))


def F(a):
    return np.ascontiguousarray(np.asfortranarray(a).ravel(order="F"))


def _model(g, precision, form):
    from self_amd.shallow_water import ShallowWater, SWParams
    N = int(g["N"])
    nodal = types.SimpleNamespace(N=N, nodes=None, dMatS=g["dMatS"], dMatP=g["dMatP"],
                                  qWeight=np.ascontiguousarray(g["qWeightS"]),
                                  flat=lambda k: F(g[k]),
                                  **{k: np.ascontiguousarray(g[k]) for k in ("lagWest", "lagEast", "lagSouth",
                                                                             "lagNorth")})
    E = int(g["x"].shape[-1])
    mesh = types.SimpleNamespace(nElems=E, nEdges=int(g["edges"].shape[0]), edges=np.ascontiguousarray(g["edges"]),
                                 x=g["x"].transpose(2, 1, 0).copy(), y=g["y"].transpose(2, 1, 0).copy(),
                                 **{k: F(g[k]) for k in ("dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen")})
    p = SWParams(polyDeg=N, nXElem=int(g["nX"]), nYElem=int(g["nY"]), dt=float(g["dt"]), g=float(g["g"]),
                 periodicX=False, periodicY=False)
    m = ShallowWater(p, precision=precision, form=form, mesh=mesh, nodal=nodal)
    m.SetVariant(1)
    E4 = lambda a: np.ascontiguousarray(a.transpose(3, 2, 1, 0))  # (i,j,f,e) -> (e,f,j,i)  # noqa: E731
    m.SetSolution(E4(g["sol"]))
    b = np.zeros((E, 3, N + 1, N + 1))
    b[:, 0] = g["h"].transpose(2, 1, 0)
    m.SetBathymetry(b)
    m.SetPlanetaryVorticity(np.ascontiguousarray(g["f"].transpose(2, 1, 0)))
    m.CalculateBathymetryGradient()
    return m, E4


def _cmp(got, ref, tol, what):
    for f in range(3):
        e = rel_l2(got[:, f], ref[:, f])
        assert e < tol, f"{what} field {f}: {e:.3e}"


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_bathymetry_gradient_and_tendency_match_reference(gpu_lib, path):
    g = dict(np.load(path))
    m, E4 = _model(g, "fp64", "plain")
    b, bref = m.GetBathymetry(), E4(g["bathy"])
    for f in (1, 2):
        assert rel_l2(b[:, f], bref[:, f]) < 1e-12, f"bathymetry gradient {f}"
    assert np.array_equal(b[:, 0], bref[:, 0])
    m.GlobalTimeDerivative(float(g["t0"]))
    _cmp(m.GetTendency(), E4(g["tend"]), 1e-12, "tendency")
    bf = m.GetBoundaryFlux()                      # (E, side, field, node)
    ref = g["bflux"].transpose(3, 2, 1, 0)        # (node, f, side, e) -> (e, side, f, node)
    assert rel_l2(bf, ref) < 1e-12
    m.Trash()


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
@pytest.mark.parametrize("precision,form,tol", [("fp64", "plain", 1e-12), ("fp64", "fused", 1e-12),
                                                ("mixed", "fused", 1e-6)])
def test_rk3_steps_match_reference(gpu_lib, path, precision, form, tol):
    g = dict(np.load(path))
    m, E4 = _model(g, precision, form)
    m.ForwardStepRK3(float(g["t0"]), int(g["nSteps"]))
    _cmp(m.GetSolution(), E4(g["sol_out"]), tol, f"{precision}/{form} RK3")
    m.Trash()
