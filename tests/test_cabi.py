"""The C ABI on the CPU: libself_amd.so loads, exports every function that
include/*.h declares, its host set-up routines agree with the oracle, and
bad arguments come back as error codes with a message (no GPU needed)."""
import ctypes as C
import glob
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(self_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from self_amd import _lib as L
    lib = L.lib()
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (self_\w+)", out))
    assert set(names) <= exported
    # every declared function has a ctypes signature in the Python binding
    assert set(names) <= set(L.exported_symbols())


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7, 10])
@pytest.mark.parametrize("kind", [1, -1])
def test_setup_quadrature_and_nodal_bitwise_equal_oracle(N, kind):
    from self_amd import nodal
    x, w = nodal.legendre_quadrature(N, kind)
    xo, wo = po.quadrature(N, kind)
    assert np.array_equal(x, xo) and np.array_equal(w, wo)
    ns = nodal.NodalStorage2D(N, kind, 2000)
    no = po.NodalStorage2D(N, kind, po.DG)
    assert np.array_equal(ns.dMatS, no.dMatS)
    assert np.array_equal(ns.dMatP, no.dMatCorrect)
    assert np.array_equal(nodal.NodalStorage2D(N, kind, 2000, True).dMatP, no.dMatP)
    for a in ("lagWest", "lagEast", "lagSouth", "lagNorth"):
        assert np.array_equal(getattr(ns, a), getattr(no, a))
    bw = nodal.barycentric_weights(x)
    assert np.array_equal(bw, po.barycentric_weights(x))
    assert np.array_equal(nodal.derivative_matrix(x, bw), po.derivative_matrix(x, bw))
    z = np.linspace(-1, 1, 7)
    assert np.array_equal(nodal.interpolation_matrix(x, bw, z), po.interpolation_matrix(x, bw, z))


@pytest.mark.parametrize("args", [(4, 3, True, True, 0.0), (5, 4, False, True, 0.05), (3, 3, False, False, 0.0),
                                  (6, 5, True, False, 0.03)])
def test_setup_quadmesh_matches_oracle(args):
    from self_amd import mesh, nodal
    nx, ny, px, py, warp = args
    N = 5
    ns = nodal.NodalStorage2D(N)
    m = mesh.QuadMesh(nx, ny, N, ns.nodes, px, py, 2.0, 3.0, warp)
    mo = po.QuadMesh(nx, ny, N, ns.nodes, px, py, 2.0, 3.0, warp)
    assert np.array_equal(m.edges, mo.edges)
    for a in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J"):
        ref = getattr(mo, a)
        assert np.abs(getattr(m, a).ravel() - ref).max() <= 4e-16 * max(1.0, np.abs(ref).max()), a
    assert np.abs(m.nHat.ravel() - mo.nHat).max() < 1e-15
    assert np.abs(m.nLen.ravel() - mo.nLen).max() < 1e-14


def test_errors_are_reported_not_raised_by_the_library():
    from self_amd import _lib as L
    lib = L.lib()
    x = np.zeros(4)
    w = np.zeros(4)
    rc = lib.self_legendre_quadrature(3, 7, L.ptr(x), L.ptr(w))
    assert rc == -5
    assert b"invalid form" in lib.self_last_error()
    nE, nEd = C.c_int(), C.c_int()
    assert lib.self_quadmesh_counts(0, 3, 0, 0, C.byref(nE), C.byref(nEd)) == -5
    # periodic direction with two elements is rejected (ambiguous edge discovery)
    e = np.zeros(6 * 16, dtype=np.int32)
    arrs = [np.zeros(64 * 4) for _ in range(7)] + [np.zeros(2 * 8 * 4 * 4), np.zeros(8 * 4 * 4)]
    rc = lib.self_quadmesh_structured(2, 2, 7, 1, 1, 1.0, 1.0, 0.0, L.ptr(np.zeros(8)), L.ptr(e, L.ip),
                                      *[L.ptr(a) for a in arrs])
    assert rc == -5 and b"periodic" in lib.self_last_error()
    with pytest.raises(L.SelfError):
        L.check(lib.self_nodal2d_build(3, 1, 1234, 0, *[L.ptr(np.zeros(16)) for _ in range(8)]))


def test_noalloc_coarse_to_fine_rejects_oversized_calls_before_touching_the_device():
    """self_coarse_to_fine_2d_noalloc checks its 2^21-double limit first, so
    an oversized call returns -5 with a message and no HIP call (CPU)."""
    from self_amd import _lib as L
    Ts = np.zeros(11 * 8)
    f = np.zeros(64)
    out = np.zeros(121)
    nE = (1 << 21) // (64 + 121) + 1
    rc = L.lib().self_coarse_to_fine_2d_noalloc(7, 7, nE, L.ptr(Ts), L.ptr(Ts), 10, 10, L.ptr(f), L.ptr(out))
    assert rc == -5 and b"scratch" in L.lib().self_last_error()
    assert L.lib().self_coarse_to_fine_2d_noalloc(7, 7, 1, None, None, 10, 10, None, None) == -5


def test_tuning_table_file_round_trip_on_the_host(tmp_path):
    """The autotuning table's text format: load, list, save, reload (no
    device work)."""
    from self_amd import shallow_water as SW
    p = tmp_path / "t.txt"
    p.write_text("sw2d 2 7 1 12 1 0 0 0 0.0123\nsw2d_ensemble 2 5 0 9 2 1 0 0 0.5\n")
    SW.clear_tuning_table()
    SW.load_tuning_table(str(p))
    t = SW.tuning_table()
    assert [(e["kind"], e["N"], e["dtype"], e["sizeClass"], e["config"][:2]) for e in t] == [
        ("sw2d", 7, 1, 12, [1, 0]), ("sw2d_ensemble", 5, 0, 9, [2, 1])]
    q = tmp_path / "u.txt"
    SW.save_tuning_table(str(q))
    SW.clear_tuning_table()
    SW.load_tuning_table(str(q))
    assert SW.tuning_table() == t
    bad = tmp_path / "bad.txt"
    bad.write_text("sw2d 2 7\n")
    from self_amd import _lib as L
    with pytest.raises(L.SelfError):
        SW.load_tuning_table(str(bad))
    SW.clear_tuning_table()
