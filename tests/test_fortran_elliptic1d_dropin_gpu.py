"""The reference's own 1-D elliptic driver
(examples/elliptic1d/parabola/localmods/Elliptic1D_Driver.f90), compiled
unchanged against the drop-in modules (self_amd/fortran/elliptic1d_dropin.f90
with sw_dropin.f90) by oracle/ref.mk into oracle/_ref/dropin_elliptic1d_driver,
run on the GPU with the example's runtime.params (N = 5, 10 elements,
tolerance 1e-8, at most 1000 CG iterations), against the outputs the example
ships under run/ (tests/golden/elliptic1d_{residual,elliptic}.curve):
 * residual.curve: 1001 rows '(I4,1x,E17.8)'; the iterations whose residual
   is O(1) (0..24) to the 8 printed digits, the cancellation step 25 below
   1e-6 on both sides, then round-off until the loop stops and zeros after;
 * elliptic.curve: node positions to 1e-15 and the unmasked solution to
   1e-13 under '#solution', the source (1 everywhere) under '#source'."""
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_elliptic1d_driver")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _curve(path):
    out, cur = {}, None
    for line in open(path):
        t = line.strip()
        if t.startswith("#"):
            cur = t[1:]
            out[cur] = []
        elif t:
            out[cur].append([float(v) for v in t.split()])
    return {k: np.array(v) for k, v in out.items()}


def test_reference_elliptic1d_driver_on_the_drop_in(gpu_lib, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("dropin_elliptic1d_driver not built (make -f oracle/ref.mk)")
    shutil.copy(os.path.join(GOLDEN, "elliptic1d_runtime.params"), tmp_path / "runtime.params")
    print("running dropin_elliptic1d_driver", flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    assert "SOLVERCRITERIA" in r.stdout.upper() and "failed to converge" not in r.stdout

    got = _curve(tmp_path / "residual.curve")["residual"]
    ref = _curve(os.path.join(GOLDEN, "elliptic1d_residual.curve"))["residual"]
    assert got.shape == ref.shape == (1001, 2)
    assert np.array_equal(got[:, 0], np.arange(1001))
    above = ref[:, 1] > 1e-3
    assert above.sum() == 25
    rel = np.abs(got[above, 1] - ref[above, 1]) / ref[above, 1]
    print(f"residual.curve: worst relative difference over the O(1) iterations {rel.max():.2e}", flush=True)
    assert rel.max() < 1e-7
    assert got[25, 1] < 1e-6 and ref[25, 1] < 1e-6
    last = int(np.max(np.nonzero(got[:, 1])))
    assert 60 <= last <= 120 and got[last, 1] < 3e-16
    assert np.all(got[last + 1:, 1] == 0.0)

    g = _curve(tmp_path / "elliptic.curve")
    w = _curve(os.path.join(GOLDEN, "elliptic1d_elliptic.curve"))
    for k in ("solution", "source"):
        assert g[k].shape == w[k].shape == (60, 2), k
        assert np.abs(g[k][:, 0] - w[k][:, 0]).max() < 1e-15, k
    assert np.abs(g["solution"][:, 1] - w["solution"][:, 1]).max() < 1e-13
    assert np.array_equal(g["source"][:, 1], w["source"][:, 1])
