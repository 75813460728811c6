"""Shallow-water models of uneven degree stepped together in one persistent
launch per RK stage (self_sw2d_forward_step_rk3_multi) against each model
stepped alone: bitwise equal states, FP64 and mixed, with an ensemble among
the models and an odd number of steps between calls."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODELS = [(12, 3), (10, 5), (8, 7), (6, 4)]


def _build(precision):
    from self_amd.workloads import sw2d_geostrophic
    out = []
    for nX, N in MODELS:
        m, info = sw2d_geostrophic(nX, N, precision=precision, form="fused")
        out.append((m, info))
    return out


def _ensemble(m, info, members=2):
    m.SetMembers(members)
    for k in range(members):
        m.SetMemberSolution(k, info["sol"] * (1.0 + 0.02 * k))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multi_degree_launch_equals_separate_runs(gpu_lib, precision):
    from self_amd.shallow_water import forward_step_rk3_multi
    together, alone = _build(precision), _build(precision)
    for grp in (together, alone):
        _ensemble(*grp[3])
    for n in (5, 4):
        forward_step_rk3_multi([m for m, _ in together], n)
        for m, _ in alone:
            m.ForwardStepRK3(0.0, n)
        for k, ((a, _), (b, _)) in enumerate(zip(together, alone)):
            if k == 3:
                for mb in range(2):
                    assert np.array_equal(a.GetMemberSolution(mb), b.GetMemberSolution(mb)), (precision, k, mb)
            else:
                assert np.array_equal(a.GetSolution(), b.GetSolution()), (precision, k, MODELS[k])
    for m, _ in together + alone:
        m.Trash()


def test_multi_rejects_mixed_precisions(gpu_lib):
    from self_amd import _lib as L
    from self_amd.shallow_water import forward_step_rk3_multi
    from self_amd.workloads import sw2d_geostrophic
    a, _ = sw2d_geostrophic(4, 3, precision="fp64", form="fused")
    b, _ = sw2d_geostrophic(4, 3, precision="mixed", form="fused")
    with pytest.raises(L.SelfError):
        forward_step_rk3_multi([a, b], 1)
    with pytest.raises(L.SelfError):
        forward_step_rk3_multi([a, a], 1)
    a.Trash()
    b.Trash()
