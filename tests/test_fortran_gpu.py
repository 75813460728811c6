"""The Fortran iso_c_binding layer (self_amd/fortran) driven the way the
reference's driver programs drive ShallowWater: per-element Set/Get around
ForwardStepRK3(tn).  The driver's result must be bitwise identical to the
Python host path over the same C ABI (same kernels, same inputs)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "self_amd", "fortran", "bin", "sw_driver")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,form", [(0, 1), (1, 1), (0, 0)])
def test_fortran_driver_matches_python_host_path(gpu_lib, tmp_path, precision, form):
    from self_amd.shallow_water import ShallowWater, SWParams
    assert os.path.exists(DRIVER), "Fortran driver not built (run __graft_entry__.build())"
    out = tmp_path / "sw.bin"
    nSteps = 7
    r = subprocess.run([DRIVER, str(out), str(precision), str(form), str(nSteps)], capture_output=True, text=True,
                       timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    N, nX, nY = 5, 6, 5
    E, n1 = nX * nY, N + 1
    raw = np.fromfile(out, dtype=np.float64)
    assert raw.size == 5 * E * 3 * n1 * n1
    sol0, bathy, final, after, resumed = (a.reshape(E, 3, n1, n1) for a in np.split(raw, 5))
    # WritePickup from Fortran: the reference's file, read back by the library
    from self_amd import pickup as P
    s, h, f = P.read_sw2d(tmp_path / P.filename("ShallowWater", nSteps), N, E)
    assert np.array_equal(s, final) and np.array_equal(h[:, 0], bathy[:, 0]) and np.all(f == 0.1)
    # ReadPickup into a second model resumes the run exactly
    assert np.array_equal(resumed, after) and not np.array_equal(after, final)
    p = SWParams(polyDeg=N, nXElem=nX, nYElem=nY, xScale=2.0, yScale=3.0, dt=1e-3, g=9.81, f0=0.1, warp=0.03)
    m = ShallowWater(p, precision=("fp64", "mixed")[precision], form=("plain", "fused")[form])
    b = bathy.copy()
    b[:, 1:] = 0.0
    m.SetBathymetry(b)
    m.CalculateBathymetryGradient()
    assert np.array_equal(m.GetBathymetry(), bathy)
    m.SetSolution(sol0)
    m.ForwardStepRK3(0.0, nSteps)
    assert np.array_equal(m.GetSolution(), final)
    assert np.abs(final - sol0).max() > 1e-6  # the state moved
    m.Trash()


@pytest.mark.parametrize("precision", [0, 1])
def test_fortran_driver_on_a_spec_mesh_file(gpu_lib, tmp_path, precision):
    """SWParamsAMD % SpecMeshFile: the drop-in type reads an ISM-V2 mesh and
    scales it (BuildQuadMesh, ConservativeShallowWaterClass.f90:289-304); the
    result is bitwise the Python path on SpecMesh(...).ScaleTheMesh(2, 3)."""
    import specmesh_gen as G
    from self_amd.mesh import SpecMesh
    from self_amd.nodal import NodalStorage2D
    from self_amd.shallow_water import ShallowWater, SWParams
    from self_amd import _lib as L
    assert os.path.exists(DRIVER), "Fortran driver not built (run __graft_entry__.build())"
    mesh_file = tmp_path / "annulus.mesh"
    mesh_file.write_text(G.quarter_annulus(3, 4, 5, seed=4, names=("NO_NORMAL_FLOW",) * 4))
    out = tmp_path / "sw.bin"
    nSteps = 5
    print(f"fortran driver on {mesh_file.name}, precision {precision}", flush=True)
    r = subprocess.run([DRIVER, str(out), str(precision), "1", str(nSteps), str(mesh_file)], capture_output=True,
                       text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    N = 5
    nodal = NodalStorage2D(N, L.GAUSS, L.DG)
    mesh = SpecMesh(mesh_file, N, nodal.nodes).ScaleTheMesh(2.0, 3.0)
    E, n1 = mesh.nElems, N + 1
    raw = np.fromfile(out, dtype=np.float64)
    assert raw.size == 5 * E * 3 * n1 * n1
    sol0, bathy, final, after, resumed = (a.reshape(E, 3, n1, n1) for a in np.split(raw, 5))
    assert np.array_equal(resumed, after)
    p = SWParams(polyDeg=N, xScale=2.0, yScale=3.0, dt=1e-3, g=9.81, f0=0.1)
    m = ShallowWater(p, precision=("fp64", "mixed")[precision], form="fused", mesh=mesh, nodal=nodal)
    b = bathy.copy()
    b[:, 1:] = 0.0
    m.SetBathymetry(b)
    m.CalculateBathymetryGradient()
    assert np.array_equal(m.GetBathymetry(), bathy)
    m.SetSolution(sol0)
    m.ForwardStepRK3(0.0, nSteps)
    assert np.array_equal(m.GetSolution(), final)
    assert np.abs(final - sol0).max() > 1e-6
    m.Trash()
