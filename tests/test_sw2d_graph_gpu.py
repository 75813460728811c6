"""Replaying the fused shallow-water stages from a captured HIP graph
(pairs of RK3 steps) against launching every stage from the host: bitwise
equal states after every call, for FP64 and mixed, single models and
ensembles, odd step counts, and a time-step change between calls (which
must re-capture).  Each replay is followed by host work and a
synchronising read, and every replay is checked."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(precision, members=1):
    from self_amd.workloads import sw2d_geostrophic
    a, info = sw2d_geostrophic(16, 5, precision=precision, form="fused")
    b, _ = sw2d_geostrophic(16, 5, precision=precision, form="fused")
    a.SetGraph(True)
    b.SetGraph(False)
    if members > 1:
        rng = np.random.default_rng(4)
        for m in (a, b):
            m.SetMembers(members)
        for k in range(members):
            s = info["sol"] * (1.0 + 0.01 * k) + 1e-3 * rng.standard_normal(info["sol"].shape) * (k > 0)
            a.SetMemberSolution(k, s)
            b.SetMemberSolution(k, s)
    return a, b, info


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("members", [1, 3])
def test_graph_replay_equals_host_launches(gpu_lib, precision, members):
    a, b, info = _pair(precision, members)
    get = (lambda m: [m.GetMemberSolution(k) for k in range(members)]) if members > 1 else \
        (lambda m: [m.GetSolution()])
    for it, n in enumerate([4, 2, 5, 6, 1, 4]):
        if it == 3:  # a new time step: the captured launches no longer match
            for m in (a, b):
                m.SetTimeStep(0.5 * info["dt"])
        a.ForwardStepRK3(0.0, n)
        b.ForwardStepRK3(0.0, n)
        sa, sb = get(a), get(b)
        for x, y in zip(sa, sb):
            assert np.array_equal(x, y), (precision, members, it)
        _ = np.linalg.norm(sa[0])  # host work between replays
    a.Trash()
    b.Trash()
