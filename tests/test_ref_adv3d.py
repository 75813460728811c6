"""The oracle's 3-D advection with the adaptive roll-off filter (orc_adv3d.c),
its nodal storage, filter matrix and hex mesh against the reference's own
model (src/highend/advection3d/Advection3DClass.f90 compiled unmodified,
oracle/ref.mk).

Fixture tests read tests/golden/ref_adv3d_*.npz (written by
tests/golden/make_ref_adv3d.py from the reference binary) and run anywhere;
the live test runs the reference binary on a fresh case when it is built.
The inputs are smooth fields plus noise of a per-element amplitude, so the
limiter replaces the solution in some elements and keeps it in others;
every boundary face carries the reference's NO_NORMAL_FLOW external state
(the moving Gaussian at the face nodes).

The reference's 3-D SetRelaxationFactor and GetRelaxationFactorAtNode
(Advection3DClass.f90:915-945) act on the relaxation *field* storage, so the
field and the factor are one array and the last Set wins; the driver sets
the field, then the factor, so the reference computes the source
(rFac - c) rFac.  The oracle is given that aliased pair.

On the reference's arrays the oracle's tendency, filtered solution, limiter
state and RK3 results are bitwise equal to the reference's (no FMA on
either side, the reference's order of operations kept).
"""
import glob
import os
import types

import numpy as np
import pytest

from oracle import pyoracle as po
import ref_run as R

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_adv3d_*.npz")))
F = lambda a: np.ascontiguousarray(np.asfortranarray(a).ravel(order="F"))  # noqa: E731


def load(path):
    z = np.load(path)
    return {k: z[k] for k in z.files}


def oracle_on_reference_arrays(g, filterMat=None):
    """orc_adv3d on the reference's nodal arrays, filter and mesh (the
    oracle's own filter matrix when the fixture does not hold one)."""
    N, nq = int(g["N"]), int(g["nTracers"])
    nodal = po.Nodal3D(N)
    nodal.dMat = g["dMatS"]
    nodal.qWeight = np.ascontiguousarray(g["qWeightS"])
    nodal.lagLeft = np.ascontiguousarray(g["lagWest"])
    nodal.lagRight = np.ascontiguousarray(g["lagEast"])
    mesh = types.SimpleNamespace(nElems=int(g["x"].shape[-1]), nFaces=int(g["faces"].shape[0]),
                                 faces=np.ascontiguousarray(g["faces"]),
                                 **{k: F(g[k]) for k in ("cov", "J", "nHat", "nLen", "bpos")})
    orc = po.Advection3D(nodal, mesh, nTracers=nq, dt=float(g["dt"]), nCutoff=int(g["nCutoff"]),
                         useFilter=True, with_relax=True)
    fm = g.get("filterMat") if filterMat is None else filterMat
    if fm is not None:
        orc.F[:] = F(fm)
    for k in ("c", "u", "v", "w"):
        getattr(orc, k)[:] = F(g[k])
    orc.relax[:] = F(g["rFac"])  # the reference's aliased relaxation storage
    orc.rFac[:] = F(g["rFac"])
    return orc


def check_run(g, exact):
    """GlobalTimeDerivative(t0, 1), then nSteps RK3 steps, against the
    reference: tendency, the (possibly filtered) solution, E1, E2, lim."""
    orc = oracle_on_reference_arrays(g)
    t0, dt = float(g["t0"]), float(g["dt"])

    def same(a, b, what):
        if exact:
            np.testing.assert_array_equal(a, b, err_msg=what)
        else:
            assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1.0), what

    orc.global_time_derivative(t0, 1)
    same(orc.tend, F(g["tend1"]), "tendency")
    same(orc.c, F(g["sol1"]), "solution after the stage-1 filter")
    same(orc.E1, np.ascontiguousarray(g["E1_1"]).ravel(), "E1")
    same(orc.E2, np.ascontiguousarray(g["E2_1"]).ravel(), "E2")
    same(orc.lim, F(g["lim1"]), "lim")
    for s in range(int(g["nSteps"])):
        orc.forward_step_rk3(t0 + s * dt)
    same(orc.c, F(g["sol"]), "solution after RK3 steps")
    same(orc.E1, np.ascontiguousarray(g["E1"]).ravel(), "E1 after steps")
    same(orc.E2, np.ascontiguousarray(g["E2"]).ravel(), "E2 after steps")
    same(orc.lim, F(g["lim"]), "lim after steps")


def test_fixtures_present():
    assert len(GOLDEN) >= 2


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_nodal_storage_matches_reference(path):
    """The oracle's 3-D DG storage: one derivative matrix, weights and
    boundary interpolants shared by the three directions, equal to the
    reference's dMatS/P/Q, qWeightS/P/Q and the six interpolant sets."""
    g = load(path)
    ns = po.Nodal3D(int(g["N"]))
    for k in ("dMatS", "dMatP", "dMatQ"):
        assert np.abs(ns.dMat - g[k]).max() <= 1e-13 * np.abs(g[k]).max(), k
    for k in ("qWeightS", "qWeightP", "qWeightQ"):
        np.testing.assert_array_equal(ns.qWeight, g[k], err_msg=k)
    for k in ("lagWest", "lagSouth", "lagBottom"):
        assert np.abs(ns.lagLeft - g[k]).max() <= 1e-14, k
    for k in ("lagEast", "lagNorth", "lagTop"):
        assert np.abs(ns.lagRight - g[k]).max() <= 1e-14, k


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_filter_matrix_matches_reference(path):
    g = load(path)
    if "filterMat" not in g:
        pytest.skip("fixture without the dense filter matrix")
    N = int(g["N"])
    ns = po.Nodal3D(N)
    Fo = po.rolloff_filter3d(N, int(g["nCutoff"]), ns.nodes, ns.qWeight)
    assert np.abs(Fo - g["filterMat"]).max() <= 1e-13


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_mesh_matches_reference_default_mesh(path):
    """orc_hexmesh_build (walled, unwarped) reproduces the reference's default
    hex mesh: the face table exactly (the reference repeats the primary side
    as the secondary side of a boundary face; the oracle stores 0 there and
    never reads it), geometry to round-off."""
    g = load(path)
    N = int(g["N"])
    ns = po.Nodal3D(N)
    m = po.HexMesh(int(g["nX"]), int(g["nY"]), int(g["nZ"]), N, ns.nodes, (False, False, False),
                   (float(g["xScale"]), float(g["yScale"]), float(g["zScale"])))
    ref = g["faces"].copy()
    bnd = ref[:, 2] <= 0
    ref[bnd, 3] = 0
    np.testing.assert_array_equal(m.faces, ref)
    for k in ("x", "y", "z", "J", "cov", "nHat", "nLen", "bpos"):
        assert np.abs(getattr(m, k) - F(g[k])).max() <= 1e-12 * max(np.abs(g[k]).max(), 1.0), k


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_stage_and_steps_match_reference(path):
    """Where the fixture holds no filter matrix the oracle's own is used
    (pinned against the reference's by the N = 3 fixture and the live runs)."""
    check_run(load(path), exact=True)


def test_the_limiter_branches_are_exercised():
    """In the fixtures the stage-1 filter replaces the solution in some
    elements and keeps it in others."""
    for path in GOLDEN:
        g = load(path)
        changed = np.abs(g["sol1"] - g["c"]).max(axis=(0, 1, 2)) > 0
        assert changed.any() and not changed.all(), path


def test_a_mistake_in_the_oracle_input_is_caught():
    """A wrong cutoff (a different filter matrix) or a swapped face
    orientation moves the result far outside the tolerance."""
    g = load(GOLDEN[0])
    N = int(g["N"])
    ns = po.Nodal3D(N)
    wrong = po.rolloff_filter3d(N, int(g["nCutoff"]) + 1, ns.nodes, ns.qWeight)
    orc = oracle_on_reference_arrays(g, filterMat=wrong)
    orc.global_time_derivative(float(g["t0"]), 1)
    assert np.abs(orc.c - F(g["sol1"])).max() > 1e-6
    bad = dict(g)
    f = g["faces"].copy()
    inner = f[:, 2] > 0
    f[inner, 4] = N
    f[inner, 6] = -1
    bad["faces"] = f
    orc = oracle_on_reference_arrays(bad)
    orc.global_time_derivative(float(g["t0"]), 1)
    assert np.abs(orc.tend - F(g["tend1"])).max() > 1e-6


@pytest.mark.skipif(not R.available("adv3d"), reason="oracle/_ref/ref_adv3d not built (needs the reference sources)")
@pytest.mark.parametrize("N,nX,nY,nZ,nq,nc", [(5, 2, 2, 1, 2, 3), (3, 2, 3, 2, 1, 2)])
def test_live_reference_run_matches_oracle(N, nX, nY, nZ, nq, nc):
    inp = R.random_inputs_3d(N, nX, nY, nZ, nq, seed=100 + N)
    g = R.run_3d(N, nX, nY, nZ, nq, nc, 2.0e-3, 0.35, 3, inp, scale=(1.0, 0.8, 1.2))
    ns = po.Nodal3D(N)
    Fo = po.rolloff_filter3d(N, nc, ns.nodes, ns.qWeight)
    assert np.abs(Fo - g["filterMat"]).max() <= 1e-13
    check_run(g, exact=True)
