"""The request loop (serve.cpp): requests admitted against free HBM and run
concurrently give bitwise the same results as the same requests run
alone, and as the model API run directly (tolerance: exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs():
    from self_amd import serving as S
    specs = []
    for i, (kind, nX, N) in enumerate([(S.REQ_SW2D, 8, 3), (S.REQ_CGSEM2D, 8, 3), (S.REQ_SW2D, 16, 5),
                                       (S.REQ_SW2D, 8, 7), (S.REQ_CGSEM2D, 16, 5), (S.REQ_SW2D, 16, 3),
                                       (S.REQ_CGSEM2D, 8, 7), (S.REQ_SW2D, 32, 3), (S.REQ_SW2D, 8, 5),
                                       (S.REQ_CGSEM2D, 16, 3)]):
        specs.append(S.make_inputs(S.Spec(kind, nX, N, 500 + i, "mixed" if i % 3 == 0 else "fp64", steps=20)))
    return specs


def test_results_are_the_same_alone_and_in_the_mix(gpu_lib):
    from self_amd import serving as S
    specs = _specs()
    srv = S.Server(workers=4)
    mix, summary = S.run_stream(srv, specs)
    srv.close()
    assert summary["ok"] and summary["requests"] == len(specs)
    alone = S.Server(workers=1)
    for (spec, res, out) in mix:
        rid = alone.submit(spec)
        r1, o1 = alone.wait(rid)
        assert r1.status == 0 and res.status == 0
        assert np.array_equal(o1, out), (spec.kind, spec.nX, spec.N)
        if spec.kind == S.REQ_CGSEM2D:
            assert r1.iterations == res.iterations > 0
    alone.close()


def test_served_shallow_water_equals_the_model_api(gpu_lib):
    from self_amd import serving as S
    from self_amd.shallow_water import ShallowWater, SWParams
    spec = S.make_inputs(S.Spec(S.REQ_SW2D, 16, 5, 77, "fp64", steps=15))
    srv = S.Server(workers=2)
    res, out = srv.wait(srv.submit(spec))
    srv.close()
    p = spec.params
    m = ShallowWater(SWParams(polyDeg=5, nXElem=16, nYElem=16, xScale=p["L"], yScale=p["L"], dt=p["dt"], g=p["g"],
                              f0=p["f0"]))
    m.SetBathymetry(spec.arrays["in1"])
    m.SetPlanetaryVorticity(np.full(spec.arrays["in1"].shape[:1] + spec.arrays["in1"].shape[2:], p["f0"]))
    m.CalculateBathymetryGradient()
    m.SetSolution(spec.arrays["in0"])
    m.ForwardStepRK3(0.0, 15)
    assert res.status == 0 and np.array_equal(m.GetSolution(), out)
    assert res.latency_ms >= res.run_ms > 0
    m.Trash()


def test_same_mesh_size_different_domain_side(gpu_lib):
    """Two shallow-water requests with the same nX and N on domains of
    different side L: the server's host-mesh cache must not hand the second
    one the first one's geometry.  Each result equals a direct model run."""
    from self_amd import serving as S
    from self_amd.shallow_water import ShallowWater, SWParams
    specs = [S.make_inputs(S.Spec(S.REQ_SW2D, 8, 3, 41, "fp64", steps=6), Lside=Ls) for Ls in (1.0e6, 2.5e5)]
    srv = S.Server(workers=1)
    results = [srv.wait(srv.submit(sp)) for sp in specs]
    srv.close()
    for spec, (res, out) in zip(specs, results):
        p = spec.params
        m = ShallowWater(SWParams(polyDeg=3, nXElem=8, nYElem=8, xScale=p["L"], yScale=p["L"], dt=p["dt"],
                                  g=p["g"], f0=p["f0"]))
        m.SetBathymetry(spec.arrays["in1"])
        m.SetPlanetaryVorticity(np.full(spec.arrays["in1"].shape[:1] + spec.arrays["in1"].shape[2:], p["f0"]))
        m.CalculateBathymetryGradient()
        m.SetSolution(spec.arrays["in0"])
        m.ForwardStepRK3(0.0, 6)
        assert res.status == 0 and np.array_equal(m.GetSolution(), out), p["L"]
        m.Trash()


def test_admission_with_a_tiny_memory_budget_runs_requests_one_at_a_time(gpu_lib):
    from self_amd import serving as S
    specs = _specs()[:4]
    srv = S.Server(workers=4, memFraction=1e-9)
    out, summary = S.run_stream(srv, specs)
    assert summary["ok"]
    # requests that do not fit run alone: no two runs overlap in time
    spans = sorted((r.queue_ms, r.queue_ms + r.run_ms) for _, r, _ in out)
    srv.close()
    assert len(spans) == 4


def test_bad_requests_are_rejected(gpu_lib):
    from self_amd import _lib as L
    from self_amd import serving as S
    srv = S.Server(workers=1)
    spec = S.make_inputs(S.Spec(S.REQ_SW2D, 8, 3, 1, "fp64", steps=5))
    spec.nX = 2  # periodic shallow water needs nX >= 3
    with pytest.raises(L.SelfError):
        srv.submit(spec)
    srv.close()


def test_batched_shallow_water_requests_equal_alone(gpu_lib):
    """Queued shallow-water requests on the same mesh, degree, precision and
    step count join one ensemble (one fused launch per stage for all); each
    keeps its own state and CFL time step and ends bitwise where it ends
    alone."""
    from self_amd import serving as S
    specs = []
    for i in range(9):
        nX, N, prec = ((16, 5, "mixed"), (8, 3, "fp64"), (16, 5, "mixed"))[i % 3]
        specs.append(S.make_inputs(S.Spec(S.REQ_SW2D, nX, N, 900 + i, prec, steps=12)))
    dts = {s.params["dt"] for s in specs}
    assert len(dts) > 2  # the members step with different time steps
    srv = S.Server(workers=1)
    mix, summary = S.run_stream(srv, specs)
    srv.close()
    assert summary["ok"]
    assert max(r.members for _, r, _ in mix) > 1
    alone = S.Server(workers=1)
    for spec, res, out in mix:
        r1, o1 = alone.wait(alone.submit(spec))
        assert r1.status == 0 and r1.members == 1
        assert np.array_equal(o1, out), (spec.nX, spec.N, spec.precision, res.members)
    alone.close()


def test_grouped_models_of_other_meshes_and_degrees_equal_alone(gpu_lib):
    """Queued shallow-water requests of one precision and step count on
    different meshes and degrees run as several models in the same
    persistent launches (each model an ensemble of the requests that share
    its mesh); each ends bitwise where it ends alone.  A request with another
    step count is never grouped with them."""
    from self_amd import serving as S
    specs = []
    for i, (nX, N) in enumerate([(8, 3), (16, 5), (8, 7), (16, 3), (8, 3), (32, 5), (8, 5), (16, 5), (8, 7)]):
        specs.append(S.make_inputs(S.Spec(S.REQ_SW2D, nX, N, 1300 + i, "fp64", steps=10)))
    odd = S.make_inputs(S.Spec(S.REQ_SW2D, 8, 3, 1399, "fp64", steps=7))
    srv = S.Server(workers=1)
    mix, summary = S.run_stream(srv, specs + [odd])
    srv.close()
    assert summary["ok"]
    assert max(r.models for _, r, _ in mix) > 1
    assert max(r.members for _, r, _ in mix) > 1
    assert mix[-1][1].models == 1 and mix[-1][1].members == 1
    alone = S.Server(workers=1)
    for spec, res, out in mix:
        r1, o1 = alone.wait(alone.submit(spec))
        assert r1.status == 0 and r1.models == 1 and r1.members == 1
        assert np.array_equal(o1, out), (spec.nX, spec.N, res.members, res.models)
    alone.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_request_stream(gpu_lib, precision):
    """The 512-request stream of the serving workload (80 % shallow water on
    8..64 squared meshes, 20 % CGSEM on 16..128 squared, N in {3, 5, 7}),
    with 20 RK3 steps per integration instead of 200 to keep the test
    short: every request completes with status 0 and finite output, and a
    seeded sample of 12 requests rerun alone on a fresh server gives
    bitwise the same output (and the same CG iteration count)."""
    from self_amd import serving as S
    specs = S.request_stream(512, precision=precision, steps=20)
    srv = S.Server(workers=8)
    mix, summary = S.run_stream(srv, specs)
    srv.close()
    assert summary["requests"] == 512
    bad = [(s.kind, s.nX, s.N, r.status) for s, r, _ in mix if r.status != 0]
    assert not bad, bad[:5]
    assert all(np.all(np.isfinite(o)) for _, _, o in mix)
    kinds = {s.kind for s, _, _ in mix}
    assert kinds == {S.REQ_SW2D, S.REQ_CGSEM2D}
    rng = np.random.default_rng(2026)
    sample = rng.choice(len(mix), size=12, replace=False)
    alone = S.Server(workers=1)
    for idx in sample:
        spec, res, out = mix[idx]
        r1, o1 = alone.wait(alone.submit(spec))
        assert r1.status == 0
        assert np.array_equal(o1, out), (idx, spec.kind, spec.nX, spec.N)
        if spec.kind == S.REQ_CGSEM2D:
            assert r1.iterations == res.iterations > 0
    alone.close()
