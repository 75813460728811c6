"""The reference-named shallow-water drop-in (self_amd/fortran/sw_dropin.f90:
modules ModelPrecision, ConstantsDictionary, SWParamsClass, QuadMeshClass,
ConservativeShallowWaterClass, TYPE ShallowWater) run the way the
reference's drivers run it: Build reads ./shallowwater.params and the pickup
at iterInit, GlobalTimeDerivative + WriteTecplot, nTimeSteps ForwardStepRK3
with dumps every dumpFreq, WritePickup, mesh % WriteTecplot.  The final
pickup equals the Python path's state bitwise (same kernels from the same
initial pickup); the Tecplot files hold the reference's layout with the
state interpolated to the plotting points (checked against an independent
numpy interpolation); the mesh file holds the mesh's geometry.  That the
reference's own driver programs compile unchanged against these modules is
checked on the host (tests/test_fortran_bindings.py)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "self_amd", "fortran", "bin", "sw_dropin_driver")

pytestmark = pytest.mark.gpu

PARAMS = """&MODEL_FORM
ModelFormulation = 'Conservative',
/
&TIME_MANAGEMENT
dt = 20.0,
iterInit = 0,
nTimeSteps = 6,
dumpFreq = 3,
/
&SPACE_MANAGEMENT
polyDeg = 5,
nXElem = 6,
nYElem = 4,
nPlot = 7,
xScale = 1000000.0,
yScale = 600000.0,
/
&PHYSICAL_PARAMETERS
g = 9.81,
f0 = 0.0001,
betaX = 0.0,
betaY = 1.0e-11,
linearDrag = 0.0,
/
&DIPOLE_PARAMETERS
/
"""


def _model():
    from self_amd.shallow_water import ShallowWater, SWParams
    p = SWParams(polyDeg=5, nXElem=6, nYElem=4, xScale=1.0e6, yScale=6.0e5, dt=20.0, g=9.81, f0=1e-4, betaY=1e-11,
                 periodicX=False, periodicY=False)
    return ShallowWater(p, quirk_dmatp_transposed=True)  # the reference's nodal storage, as the drop-in


def _lagrange_matrix(nodes, z):
    """T[a, i] = l_i(z_a), independent of the library (product form)."""
    T = np.ones((len(z), len(nodes)))
    for i, xi in enumerate(nodes):
        for k, xk in enumerate(nodes):
            if k != i:
                T[:, i] *= (z - xk) / (xi - xk)
    return T


def _read_tec(path, ncol):
    """Zones of a list-directed Tecplot file (a record may wrap over lines)."""
    zones, cur = [], None
    with open(path) as f:
        head = f.readline()
        for line in f:
            if "ZONE" in line:
                cur = []
                zones.append(cur)
            else:
                cur.extend(float(v) for v in line.split())
    return head, [np.array(z).reshape(-1, ncol) for z in zones]


def test_reference_named_driver_run(gpu_lib, tmp_path):
    assert os.path.exists(DRIVER), "drop-in driver not built (run __graft_entry__.build())"
    (tmp_path / "shallowwater.params").write_text(PARAMS)
    m = _model()
    x, y = m.mesh.x, m.mesh.y
    E, n1 = m.nElems, 6
    h = 1000.0 + 50.0 * np.sin(2 * np.pi * x / 1.0e6) * np.cos(np.pi * y / 6.0e5)
    bathy = np.stack([h, np.zeros_like(h), np.zeros_like(h)], axis=1)
    eta = 0.5 * np.exp(-((x - 5e5) ** 2 + (y - 3e5) ** 2) / (1.2e5) ** 2)
    sol = np.stack([0.02 * (h + eta) * np.sin(np.pi * y / 6e5), -0.01 * (h + eta) * np.sin(np.pi * x / 1e6),
                    h + eta], axis=1)
    m.SetBathymetry(bathy)
    m.SetPlanetaryVorticity(1e-4 + 1e-11 * y)
    m.SetSolution(sol)
    m.CalculateBathymetryGradient()
    m.WritePickup(0, str(tmp_path))
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    # the final pickup: the Python path from the same initial pickup
    ref = _model()
    ref.ReadPickup(0, str(tmp_path))
    ref.CalculateBathymetryGradient()
    ref.ForwardStepRK3(0.0, 6)
    out = _model()
    out.ReadPickup(6, str(tmp_path))
    final = out.GetSolution()
    assert np.array_equal(final, ref.GetSolution())
    assert np.abs(final - sol).max() > 1e-6
    # Tecplot dumps at 0, 3, 6 in the reference's layout
    for it in (0, 3, 6):
        assert (tmp_path / f"ShallowWater.{it:010d}.tec").exists(), it
    head, zones = _read_tec(tmp_path / "ShallowWater.0000000006.tec", 6)
    assert head.strip().startswith("VARIABLES") and len(zones) == E
    nPlot = 7
    z = -1.0 + 2.0 * np.arange(nPlot + 1) / nPlot
    T = _lagrange_matrix(m.dgStorage.nodes, z)
    hb = out.GetBathymetry()[:, 0]
    for e in (0, 7, E - 1):
        assert zones[e].shape == ((nPlot + 1) ** 2, 6)

        def plot(F):  # F[j(p), i(s)] at the Gauss nodes -> G[a(s), b(p)]
            return T @ F.T @ T.T

        X, Y, D = plot(x[e]), plot(y[e]), plot(hb[e])
        U, V, H = plot(final[e, 0]), plot(final[e, 1]), plot(final[e, 2])
        # rows: iP outer, iS inner of the reference's transposed CoarseToFine output
        exp = np.array([[X[iP, iS], Y[iP, iS], D[iP, iS], U[iP, iS] / H[iP, iS], V[iP, iS] / H[iP, iS],
                         H[iP, iS] - D[iP, iS]] for iP in range(nPlot + 1) for iS in range(nPlot + 1)])
        scale = np.abs(exp).max(axis=0)
        assert np.all(np.abs(zones[e] - exp).max(axis=0) <= 1e-12 * scale + 1e-12), e
    # the mesh file
    head, mz = _read_tec(tmp_path / "mesh.tec", 7)
    assert len(mz) == E and mz[0].shape == (n1 * n1, 7)
    exp = np.stack([m.mesh.x[0].ravel(), m.mesh.y[0].ravel(), m.mesh.J[0].ravel()], axis=1)
    assert np.allclose(mz[0][:, :3], exp, rtol=1e-14, atol=0)
    for mm in (m, ref, out):
        mm.Trash()
