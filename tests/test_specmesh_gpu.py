"""The shallow-water hot path on an unstructured, curved ISM-V2 mesh.

The mesh is a quarter annulus written by tests/specmesh_gen.py (corners
listed from rotated starting corners, curved inner and outer arcs) and read
by the library's SpecMesh reader, whose output the reference's own reader
pins (tests/test_ref_specmesh.py).  The HIP model (self_amd.shallow_water
over the C ABI) and the FP64 oracle run on the same geometry and state; the
reversed neighbours (s2 < 0, start/inc = N-1/-1) and curved metrics
exercise what a structured mesh never does.  Elements are counter-clockwise
as mesh generators write them: the model's flux form assumes a positive
Jacobian (a clockwise element does not even keep a lake at rest in the
oracle).
Tolerances (relative L2 per field): FP64 vs oracle 1e-12, mixed vs FP64 1e-6.
"""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import pyoracle as po
from sw2d_cases import smooth_state, to_oracle_layout
import specmesh_gen as G

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
# walls everywhere (NEUMANN is NO_NORMAL_FLOW, ModelFlags.f90:96); the
# reference's RADIATION state is zero depth, which its flux divides by
NAMES = ("NO_NORMAL_FLOW", "NEUMANN", "NO_NORMAL_FLOW", "NEUMANN")


def setup(tmp_path, N, nr, nt, g, seed, form="fused", precision="fp64", dt=2e-4):
    from self_amd.mesh import SpecMesh
    from self_amd.nodal import NodalStorage2D
    from self_amd.shallow_water import ShallowWater, SWParams
    from self_amd import _lib as L
    path = tmp_path / "annulus.mesh"
    path.write_text(G.quarter_annulus(nr, nt, g, seed=seed, names=NAMES))
    nodal = NodalStorage2D(N, L.GAUSS, L.DG)
    mesh = SpecMesh(path, N, nodal.nodes)
    print(f"specmesh case N={N} {nr}x{nt} g={g} seed={seed} {form} {precision}", flush=True)
    model = ShallowWater(SWParams(polyDeg=N, dt=dt, g=9.81), precision=precision, form=form, mesh=mesh,
                         nodal=nodal)
    onodal = po.NodalStorage2D(N)
    assert np.array_equal(onodal.nodes, nodal.nodes)
    orc = po.ShallowWater2D(onodal, mesh, g=9.81, dt=dt)
    E, n1 = mesh.nElems, N + 1
    U, V, H, h, f = smooth_state(mesh.x, mesh.y, 4.0, 4.0, seed=seed)
    sol = np.stack([U, V, H], axis=1)
    bathy = np.stack([h, np.zeros_like(h), np.zeros_like(h)], axis=1)
    orc.sol[:] = to_oracle_layout(sol)
    orc.bathy[:] = to_oracle_layout(bathy)
    orc.plv[:] = f.ravel()
    orc.bathymetry_gradient()
    model.SetSolution(orc.sol.reshape(E, 3, n1, n1))
    model.SetBathymetry(orc.bathy.reshape(E, 3, n1, n1))
    model.SetPlanetaryVorticity(orc.plv.reshape(E, n1, n1))
    return model, orc, mesh


CASES = [
    dict(N=5, nr=3, nt=4, g=4, seed=1),
    dict(N=7, nr=4, nt=6, g=6, seed=2),
    dict(N=3, nr=2, nt=3, g=3, seed=3),
]


def _ids(c):
    return f"N{c['N']}-{c['nr']}x{c['nt']}-g{c['g']}-s{c['seed']}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("form", ["plain", "fused"])
def test_time_derivative_and_rk3_match_oracle(gpu_lib, tmp_path, case, form):
    m, orc, mesh = setup(tmp_path, form=form, **case)
    assert (mesh.edges[:, 3] < 0).any()
    E, n1 = mesh.nElems, case["N"] + 1
    orc.global_time_derivative(0.0)
    m.GlobalTimeDerivative(0.0)
    tend, ref = m.GetTendency(), orc.tend.reshape(E, 3, n1, n1)
    for f in range(3):
        assert rel_l2(tend[:, f], ref[:, f]) < TOL64, (f, rel_l2(tend[:, f], ref[:, f]))
    nsteps = 4
    for it in range(nsteps):
        orc.forward_step_rk3(it * 2e-4)
    m.ForwardStepRK3(0.0, nsteps)
    sol, ref = m.GetSolution(), orc.sol.reshape(E, 3, n1, n1)
    for f in range(3):
        assert rel_l2(sol[:, f], ref[:, f]) < TOL64, (f, rel_l2(sol[:, f], ref[:, f]))
    m.Trash()


def test_mixed_precision_on_curved_mesh(gpu_lib, tmp_path):
    """Curved elements have non-zero metric deviations in the mixed form."""
    case = CASES[1]
    m64, _, mesh = setup(tmp_path, precision="fp64", **case)
    mmx, _, _ = setup(tmp_path, precision="mixed", **case)
    m64.ForwardStepRK3(0.0, 20)
    mmx.ForwardStepRK3(0.0, 20)
    a, b = mmx.GetSolution(), m64.GetSolution()
    for f in range(3):
        assert rel_l2(a[:, f], b[:, f]) < 1e-6, (f, rel_l2(a[:, f], b[:, f]))
    assert not np.array_equal(a, b)
    m64.Trash()
    mmx.Trash()
