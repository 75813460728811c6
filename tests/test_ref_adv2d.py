"""The oracle's 2-D advection-diffusion (orc_adv2d.c), its nodal storage and
its mesh against the reference's own model (src/highend/advection-diffusion2d/
AdvectionClass.f90 compiled unmodified, oracle/ref.mk).

Fixture tests read tests/golden/ref_adv2d_*.npz (written by
tests/golden/make_ref_adv2d.py from the reference binary) and run anywhere;
the live test runs the reference binary on a fresh case when it is built.
Inputs are random, so every term of the operator is exercised: the upwind
flux, the BR1 gradient and its boundary values, the inflow states tanh(t)
(InflowOne feeds tracer 1, InflowTwo tracer 2), the relaxation source and
the reference's transposed dMatP.

On the reference's own nodal arrays and mesh the oracle is bitwise equal to
the reference: both are built for x86-64 without fused multiply-add (the
oracle with -ffp-contract=off, the reference with amdflang -O2 and no -march)
and the oracle keeps the reference's order of operations.  The oracle's own
mesh generator agrees with LoadDefaultMesh to round-off (its metric terms
are evaluated in a different order).
"""
import glob
import os
import types

import numpy as np
import pytest

from oracle import pyoracle as po
import ref_run as R

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_adv2d_*.npz")))


def load(path):
    z = np.load(path)
    return {k: z[k] for k in z.files}


def reference_inflow_flags(edges):
    """BuildQuadMesh_Advection (AdvectionClass.f90:304-315): a boundary edge
    whose primary side is west (4) or north (3) gets InflowOne, east (2) or
    south (1) InflowTwo."""
    e = np.array(edges, dtype=np.int32).reshape(-1, 6).copy()
    bnd = e[:, 2] == po.NO_NORMAL_FLOW
    e[bnd & np.isin(e[:, 1], (3, 4)), 2] = po.INFLOW_ONE
    e[bnd & np.isin(e[:, 1], (1, 2)), 2] = po.INFLOW_TWO
    return e


def close(a, b, tol=1e-13):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1.0)
    return np.abs(a - b).max() <= tol * scale


def oracle_on_reference_arrays(g):
    """orc_adv2d on the reference's own nodal arrays and mesh."""
    N, nq = int(g["N"]), int(g["nTracers"])
    F = lambda a: np.asfortranarray(a).ravel(order="F").copy()  # noqa: E731
    nodal = types.SimpleNamespace(N=N, dMatS=g["dMatS"], dMatCorrect=g["dMatP"], qWeight=g["qWeightS"].copy(),
                                  lagWest=g["lagWest"].copy(), lagEast=g["lagEast"].copy(),
                                  lagSouth=g["lagSouth"].copy(), lagNorth=g["lagNorth"].copy())
    mesh = types.SimpleNamespace(nElems=g["x"].shape[-1], nEdges=g["edges"].shape[0], edges=g["edges"],
                                 **{k: F(g[k]) for k in ("dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen")})
    orc = po.Advection2D(nodal, mesh, nTracers=nq, kappa=float(g["kappa"]), dt=float(g["dt"]),
                         dMatP=g["dMatP"], with_relax=True)
    for k in ("c", "u", "v", "relax", "rFac"):
        getattr(orc, k)[:] = F(g[k])
    return orc


def test_fixtures_present():
    assert len(GOLDEN) >= 2


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_nodal_storage_matches_reference(path):
    """orc_nodal2d_build (Gauss, DG) gives the reference's dMatS, its
    transposed dMatP, weights and boundary interpolants."""
    g = load(path)
    ns = po.NodalStorage2D(int(g["N"]), po.GAUSS, po.DG)
    np.testing.assert_array_equal(ns.dMatS, g["dMatS"])
    np.testing.assert_array_equal(ns.dMatP, g["dMatP"])
    assert not close(ns.dMatCorrect, g["dMatP"], 1e-6)  # the reference keeps the transposed one
    np.testing.assert_array_equal(ns.qWeight, g["qWeightS"])
    np.testing.assert_array_equal(ns.qWeight, g["qWeightP"])
    for k in ("lagWest", "lagEast", "lagSouth", "lagNorth"):
        np.testing.assert_array_equal(getattr(ns, k), g[k], err_msg=k)


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_mesh_matches_reference_default_mesh(path):
    """orc_quadmesh_build (walled) + the inflow flags reproduce
    LoadDefaultMesh + ScaleTheMesh: the edge table exactly, geometry to
    round-off."""
    g = load(path)
    N = int(g["N"])
    ns = po.NodalStorage2D(N)
    m = po.QuadMesh(int(g["nX"]), int(g["nY"]), N, ns.nodes, False, False, float(g["xScale"]), float(g["yScale"]))
    np.testing.assert_array_equal(reference_inflow_flags(m.edges), g["edges"])
    F = lambda a: np.asfortranarray(a).ravel(order="F")  # noqa: E731
    for k in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen"):
        assert close(getattr(m, k), F(g[k]), 1e-12), k


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_global_time_derivative_matches_reference(path):
    g = load(path)
    orc = oracle_on_reference_arrays(g)
    orc.global_time_derivative(float(g["t0"]))
    F = lambda a: np.asfortranarray(a).ravel(order="F")  # noqa: E731
    for k in ("dcdx", "dcdy", "tend"):
        np.testing.assert_array_equal(getattr(orc, k), F(g[k]), err_msg=k)


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_rk3_steps_match_reference(path):
    g = load(path)
    orc = oracle_on_reference_arrays(g)
    t0, dt = float(g["t0"]), float(g["dt"])
    for s in range(int(g["nSteps"])):
        orc.forward_step_rk3(t0 + s * dt)
    np.testing.assert_array_equal(orc.c, np.asfortranarray(g["sol"]).ravel(order="F"))


def test_a_mistake_in_the_oracle_input_is_caught():
    """The comparison is sensitive: the correct (untransposed) dMatP, or a
    flipped inflow flag, moves the tendency far outside the tolerance."""
    g = load(GOLDEN[0])
    F = lambda a: np.asfortranarray(a).ravel(order="F")  # noqa: E731
    bad = dict(g)
    bad["dMatP"] = np.ascontiguousarray(g["dMatP"].T)
    orc = oracle_on_reference_arrays(bad)
    orc.global_time_derivative(float(g["t0"]))
    assert not close(orc.tend, F(g["tend"]), 1e-6)
    bad = dict(g)
    e = g["edges"].copy()
    e[e[:, 2] == po.INFLOW_ONE, 2] = po.INFLOW_TWO
    bad["edges"] = e
    orc = oracle_on_reference_arrays(bad)
    orc.global_time_derivative(float(g["t0"]))
    assert not close(orc.tend, F(g["tend"]), 1e-6)


@pytest.mark.skipif(not R.available(), reason="oracle/_ref/ref_adv2d not built (needs the reference sources)")
def test_live_reference_run_matches_oracle():
    N, nX, nY, nq, kappa, dt, t0, nSteps = 4, 5, 2, 2, 3.0e-3, 2.0e-3, 0.05, 5
    inp = R.random_inputs(N, nX, nY, nq, seed=2026)
    g = R.run(N, nX, nY, nq, kappa, dt, t0, nSteps, inp, xScale=1.5, yScale=0.75)
    orc = oracle_on_reference_arrays(g)
    orc.global_time_derivative(t0)
    F = lambda a: np.asfortranarray(a).ravel(order="F")  # noqa: E731
    for k in ("dcdx", "dcdy", "tend"):
        assert close(getattr(orc, k), F(g[k]), 1e-12), k
    orc = oracle_on_reference_arrays(g)
    for s in range(nSteps):
        orc.forward_step_rk3(t0 + s * dt)
    assert close(orc.c, F(g["sol"]), 1e-12)
