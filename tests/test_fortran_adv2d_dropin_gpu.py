"""The reference's own 2-D advection-diffusion driver
(src/highend/advection-diffusion2d/Advection_Driver.f90), compiled unchanged
against the drop-in modules (self_amd/fortran/adv_dropin.f90 with
sw_dropin.f90) by oracle/ref.mk into oracle/_ref/dropin_adv2d_driver, run on
the GPU from the initial pickup the reference's own initial-condition
program wrote, against the same driver compiled against the reference's
modules (fixture tests/golden/refdriver_adv2d.npz, made by
tests/golden/make_ref_adv2d_driver.py; 2 tracers, kappa 1e-3, N = 5 on a
4 x 3 mesh scaled to 2 x 1.5, inflow boundaries tanh(t), 20 RK3 steps):
 * the final pickup: tracers to 1e-12 of their largest value, velocity,
   relaxation field and factor bitwise (they do not change);
 * the Tecplot dumps at iterations 0, 10, 20 (positions, velocity, tracers
   and their BR1 gradients at the plotting points): every column to 1e-11
   of its largest value;
 * mesh.tec: positions, Jacobian and metric terms to 1e-13 of the largest
   value (dx/dp and dy/ds vanish on this mesh)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cgsem_outputs import numeric_rows

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_adv2d_driver")
GOLDEN = os.path.join(ROOT, "tests", "golden", "refdriver_adv2d.npz")


def _cols(a, b):
    scale = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
    return np.max(np.abs(a - b), axis=0) / scale


def test_reference_adv2d_driver_on_the_drop_in(gpu_lib, tmp_path):
    from self_amd.pickup import read_adv2d
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/dropin_adv2d_driver not built (make -f oracle/ref.mk)")
    ref = dict(np.load(GOLDEN))
    shutil.copy(os.path.join(ROOT, "tests", "golden", "adv2d_driver_advection.params"), tmp_path / "advection.params")
    (tmp_path / "Advection.0000000000.pickup").write_bytes(ref["pickup0"].tobytes())
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    (tmp_path / "ref.pickup").write_bytes(ref["pickup20"].tobytes())
    got = read_adv2d(tmp_path / "Advection.0000000020.pickup", 5, 12, 2)
    want = read_adv2d(tmp_path / "ref.pickup", 5, 12, 2)
    e = float(np.max(np.abs(got["c"] - want["c"])) / np.max(np.abs(want["c"])))
    print(f"final tracers: max difference / max {e:.3e}", flush=True)
    assert 0 < np.max(np.abs(want["c"])) and e < 1e-12
    for k in ("vel", "relax", "rFac"):
        assert np.array_equal(got[k], want[k]), k
    for it in (0, 10, 20):
        g = numeric_rows(str(tmp_path / f"Advection.{it:010d}.tec"), 10)
        w = ref[f"tec{it}"]
        assert g.shape == w.shape, it
        c = _cols(g, w)
        print(f"Advection.{it:010d}.tec: worst column {c.max():.3e}", flush=True)
        assert c.max() < 1e-11, (it, c)
    m = numeric_rows(str(tmp_path / "mesh.tec"), 7)
    assert m.shape == ref["mesh"].shape
    assert np.max(np.abs(m - ref["mesh"])) < 1e-13 * np.max(np.abs(ref["mesh"]))
