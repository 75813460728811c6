"""The reference's 2-D advection-diffusion model on an ISM-V2 mesh file.

tests/golden/ref_adv2dspec_annulus_n5.npz holds what the compiled reference
(AdvectionClass.f90 with SpaceManagement % SpecMeshFile: ReadSpecMeshFile,
then ScaleTheMesh, :303-330) built and computed on a curved, unstructured
quarter annulus with InflowOne / InflowTwo / NO_NORMAL_FLOW sides:
* the library's SpecMesh reader + ScaleTheMesh give bitwise the edge table
  and geometry the reference model itself used;
* the oracle on the reference's arrays reproduces its gradients, tendency
  and RK3 solution bitwise;
* live: the same against a fresh reference run on another annulus.
Host-only; the GPU side is tests/test_ref_adv2d_specmesh_gpu.py.
"""
import os

import numpy as np
import pytest

import ref_run as R
import specmesh_gen as G
from test_ref_adv2d import oracle_on_reference_arrays

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_adv2dspec_annulus_n5.npz")


def load(path=GOLDEN):
    z = np.load(path)
    return {k: z[k] for k in z.files}


def reader_matches(g, tmp_path):
    from self_amd import _lib as L
    from self_amd.mesh import SpecMesh
    N = int(g["N"])
    x, w = np.zeros(N + 1), np.zeros(N + 1)
    L.check(L.lib().self_legendre_quadrature(N, L.GAUSS, L.ptr(x), L.ptr(w)))
    path = tmp_path / "m.mesh"
    path.write_bytes(g["mesh"].tobytes())
    m = SpecMesh(path, N, x).ScaleTheMesh(float(g["xScale"]), float(g["yScale"]))
    bad = [] if np.array_equal(m.edges, g["edges"]) else ["edges"]
    for k in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J"):
        if not np.array_equal(getattr(m, k).transpose(2, 1, 0), g[k]):
            bad.append(k)
    if not np.array_equal(m.nHat.transpose(3, 2, 1, 0), g["nHat"]):
        bad.append("nHat")
    if not np.array_equal(m.nLen.transpose(2, 1, 0), g["nLen"]):
        bad.append("nLen")
    return bad


def test_reader_gives_the_reference_models_mesh(tmp_path):
    assert reader_matches(load(), tmp_path) == []


def oracle_matches(g):
    F = lambda a: np.asfortranarray(a).ravel(order="F")  # noqa: E731
    orc = oracle_on_reference_arrays(g)
    orc.global_time_derivative(float(g["t0"]))
    bad = [k for k in ("dcdx", "dcdy", "tend") if not np.array_equal(getattr(orc, k), F(g[k]))]
    orc = oracle_on_reference_arrays(g)
    t0, dt = float(g["t0"]), float(g["dt"])
    for s in range(int(g["nSteps"])):
        orc.forward_step_rk3(t0 + s * dt)
    if not np.array_equal(orc.c, F(g["sol"])):
        bad.append("sol")
    return bad


def test_oracle_matches_the_reference_on_the_spec_mesh():
    g = load()
    assert (g["edges"][:, 3] < 0).any()  # reversed neighbours present
    assert oracle_matches(g) == []


def test_a_flipped_inflow_flag_is_caught():
    g = load()
    e = g["edges"].copy()
    k = np.flatnonzero(e[:, 2] == -103)[0]
    e[k, 2] = -104
    g["edges"] = e
    assert "tend" in oracle_matches(g)


@pytest.mark.skipif(not R.available("adv2d"), reason="oracle/_ref/ref_adv2d not built (needs the reference sources)")
def test_live_reference_on_another_annulus(tmp_path):
    N, nq, E = 4, 2, 2 * 3  # nTracers >= 2: the reference keeps u, v in tracer slots 1, 2
    text = G.quarter_annulus(2, 3, 3, seed=5, names=("InflowTwo", "InflowOne", "NO_NORMAL_FLOW", "InflowOne"))
    inp = R.random_inputs(N, E, 1, nq, 77)
    g = R.run(N, E, 1, nq, 1.0e-2, 2.0e-3, 0.1, 3, inp, xScale=1.5, yScale=1.5, mesh_text=text)
    g.update(xScale=1.5, yScale=1.5, mesh=np.frombuffer(text.encode(), dtype=np.uint8))
    assert reader_matches(g, tmp_path) == []
    assert oracle_matches(g) == []
