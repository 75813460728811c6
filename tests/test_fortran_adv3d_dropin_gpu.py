"""The reference's own 3-D advection programs
(src/highend/advection3d/Advection3D_InitialConditions.f90 and
Advection3D_Driver.f90), compiled unchanged against the drop-in modules
(self_amd/fortran/adv3d_dropin.f90 with sw_dropin.f90) by oracle/ref.mk
into oracle/_ref/dropin_adv3d_ic and dropin_adv3d_driver, run on the GPU,
against the same programs compiled against the reference's modules (fixture
tests/golden/refdriver_adv3d.npz, made by
tests/golden/make_ref_adv3d_driver.py; 2 tracers, N = 5, roll-off cutoff 3,
a 3 x 2 x 2 walled mesh scaled to 1 x 1.2 x 0.8, the Gaussian
exp(-|x|^2/0.04) carried by the velocity (1, 1, 1), boundary states the
moving Gaussian, 20 RK3 steps of 2e-3 with the adaptive roll-off filter at
every stage):
 * the initial-condition program: its pickup (tracers to 1e-13 of their
   largest value, relaxation and velocity bitwise), Advection.init.tec and
   mesh.tec;
 * the driver, started from the reference's initial pickup: the final
   pickup (tracers to 1e-12 of their largest value, relaxation and velocity
   bitwise), the Tecplot dumps at iterations 0, 10 and 20 (every column to
   1e-11 of its largest value), mesh.tec (to 1e-13 of the largest value) and
   the timing table Timing.stats (File-I/O measured twice per dump, as the
   reference's StopThisTimer + AccumulateTimings count it)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cgsem_outputs import numeric_rows

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
GOLDEN = os.path.join(ROOT, "tests", "golden", "refdriver_adv3d.npz")
PARAMS = os.path.join(ROOT, "tests", "golden", "adv3d_driver_advection.params")
N, NE, NQ = 5, 12, 2


def _pickup(raw):
    """Advection3DClass.f90:1905-1972: per element nEq tracers, nEq
    relaxation fields, nEq relaxation factors, u, v, w, each a big-endian
    (0:N)^3 record."""
    n1 = N + 1
    a = np.frombuffer(bytes(raw), dtype=">f8").reshape(NE, 3 * NQ + 3, n1, n1, n1)
    return {"c": a[:, :NQ], "relax": a[:, NQ:2 * NQ], "rFac": a[:, 2 * NQ:3 * NQ], "vel": a[:, 3 * NQ:]}


def _cols(a, b):
    scale = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
    return np.max(np.abs(a - b), axis=0) / scale


def _run(exe, cwd):
    if not os.path.exists(exe):
        pytest.skip(f"{os.path.basename(exe)} not built (make -f oracle/ref.mk)")
    print(f"running {os.path.basename(exe)}", flush=True)
    r = subprocess.run([exe], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])


def _same_pickup(got, want, tol):
    g, w = _pickup(got), _pickup(want)
    e = float(np.max(np.abs(g["c"] - w["c"])) / np.max(np.abs(w["c"])))
    print(f"tracers: max difference / max {e:.3e}", flush=True)
    assert 0 < np.max(np.abs(w["c"])) and e < tol
    for k in ("relax", "rFac", "vel"):
        assert np.array_equal(g[k], w[k]), k


def test_reference_adv3d_initial_conditions_on_the_drop_in(gpu_lib, tmp_path):
    ref = dict(np.load(GOLDEN))
    shutil.copy(PARAMS, tmp_path / "advection.params")
    _run(os.path.join(REF, "dropin_adv3d_ic"), tmp_path)
    _same_pickup(np.fromfile(tmp_path / "Advection.0000000000.pickup", np.uint8), ref["pickup0"], 1e-13)
    g = numeric_rows(str(tmp_path / "Advection.init.tec"), 8)
    assert g.shape == ref["tecinit"].shape
    assert _cols(g, ref["tecinit"]).max() < 1e-12
    m = numeric_rows(str(tmp_path / "mesh.tec"), 13)
    assert np.max(np.abs(m - ref["mesh"])) < 1e-13 * np.max(np.abs(ref["mesh"]))


def test_reference_adv3d_driver_on_the_drop_in(gpu_lib, tmp_path):
    ref = dict(np.load(GOLDEN))
    shutil.copy(PARAMS, tmp_path / "advection.params")
    (tmp_path / "Advection.0000000000.pickup").write_bytes(ref["pickup0"].tobytes())
    _run(os.path.join(REF, "dropin_adv3d_driver"), tmp_path)
    got = np.fromfile(tmp_path / "Advection.0000000020.pickup", np.uint8)
    _same_pickup(got, ref["pickup20"], 1e-12)
    # the tracers moved: the final state is not the initial one
    assert np.max(np.abs(_pickup(got)["c"] - _pickup(ref["pickup0"])["c"])) > 1e-3
    for it in (0, 10, 20):
        g = numeric_rows(str(tmp_path / f"Advection.{it:010d}.tec"), 8)
        w = ref[f"tec{it}"]
        assert g.shape == w.shape, it
        c = _cols(g, w)
        print(f"Advection.{it:010d}.tec: worst column {c.max():.3e}", flush=True)
        assert c.max() < 1e-11, (it, c)
    m = numeric_rows(str(tmp_path / "mesh.tec"), 13)
    assert m.shape == ref["mesh"].shape
    assert np.max(np.abs(m - ref["mesh"])) < 1e-13 * np.max(np.abs(ref["mesh"]))
    stats = (tmp_path / "Timing.stats").read_text().split("\n")
    k = next(i for i, line in enumerate(stats) if "File-I/O" in line)
    assert float(stats[k + 1].split(":")[1]) == 6.0
