"""Pickup (restart) files against the reference's own WritePickup/ReadPickup.

The 2-D advection model's routines (AdvectionClass.f90:2040-2171) are
compiled unmodified (oracle/ref.mk -> oracle/_ref/ref_pickup).  The
library's writer produces the reference's file byte for byte, its reader
reads the reference's file exactly, and the reference reads the library's
file back exactly (tests/golden/ref_pickup_adv2d_n3_3x2.npz holds one file
the reference wrote; the live tests run the reference when it is built).

The shallow-water model does not compile (WriteTecplot, :1992-2009), so its
record order (U, V, H, h, f per element, ConservativeShallowWaterClass.f90:
2013-2125) is pinned against a numpy statement of the same format, whose
conventions (direct access, big endian, one (0:N,0:N) record per element
field, no record markers) the advection files above pin against the
reference.  These are host-only file operations; no GPU is used.
"""
import glob
import os

import numpy as np
import pytest

import ref_run as R
from self_amd import _lib as L
from self_amd import pickup as P

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_pickup_adv2d_n3_3x2.npz")


def em(a):
    """(i, j, f, e) element-last reference layout -> (e, f, j, i)."""
    return np.ascontiguousarray(a.transpose(3, 2, 1, 0))


def em2(a):
    """(i, j, e) -> (e, j, i)."""
    return np.ascontiguousarray(a.transpose(2, 1, 0))


def ours(inp):
    vel = np.ascontiguousarray(np.stack([em2(inp["u"]), em2(inp["v"])], axis=1))
    return em(inp["c"]), em(inp["relax"]), em(inp["rFac"]), vel


def test_filename_is_the_reference_pattern():
    assert P.filename("ShallowWater", 0) == "ShallowWater.0000000000.pickup"
    assert P.filename("Advection", 1234567) == "Advection.0001234567.pickup"
    with pytest.raises(L.SelfError):
        P.filename("Advection", -1)


def test_writer_matches_the_reference_file_and_reader_reads_it(tmp_path):
    g = np.load(GOLDEN)
    inp = {k: g[k] for k in ("c", "u", "v", "relax", "rFac")}
    N, E, q = int(g["N"]), int(g["nX"]) * int(g["nY"]), int(g["nTracers"])
    ref_bytes = g["pickup"].tobytes()
    assert len(ref_bytes) == E * (3 * q + 2) * (N + 1) ** 2 * 8  # no record markers
    c, rl, rf, vel = ours(inp)
    path = tmp_path / P.filename("Advection", int(g["iter"]))
    P.write_adv2d(path, c, rl, rf, vel)
    assert path.read_bytes() == ref_bytes
    path.write_bytes(ref_bytes)
    back = P.read_adv2d(path, N, E, q)
    for k, v in (("c", c), ("relax", rl), ("rFac", rf), ("vel", vel)):
        assert np.array_equal(back[k], v), k


def test_a_transposed_record_is_caught(tmp_path):
    """Writing (j, i) instead of (i, j) records changes the bytes."""
    g = np.load(GOLDEN)
    inp = {k: g[k] for k in ("c", "u", "v", "relax", "rFac")}
    c, rl, rf, vel = ours(inp)
    path = tmp_path / "x.pickup"
    P.write_adv2d(path, np.ascontiguousarray(c.transpose(0, 1, 3, 2)), rl, rf, vel)
    assert path.read_bytes() != g["pickup"].tobytes()


@pytest.mark.skipif(not R.available("pickup"), reason="oracle/_ref/ref_pickup not built (needs the reference sources)")
@pytest.mark.parametrize("N,nX,nY,q", [(1, 2, 2, 1), (5, 3, 4, 3), (7, 2, 3, 4)])
def test_live_reference_round_trips(tmp_path, N, nX, nY, q):
    inp = R.random_inputs(N, nX, nY, q, 100 + N)
    ref_bytes = R.run_pickup_write(N, nX, nY, q, 42, inp)
    c, rl, rf, vel = ours(inp)
    path = tmp_path / P.filename("Advection", 42)
    P.write_adv2d(path, c, rl, rf, vel)
    assert path.read_bytes() == ref_bytes
    back = R.run_pickup_read(N, nX, nY, q, 42, path.read_bytes())
    for k in inp:
        assert np.array_equal(back[k], inp[k]), k


def _sw_bytes(sol, bathy, plv):
    """ConservativeShallowWaterClass.f90:2044-2061 stated with numpy: per
    element the records U, V, H, h, f, big-endian float64, (0:N,0:N)
    column-major (our (j, i) C-order arrays are exactly that)."""
    parts = []
    for e in range(sol.shape[0]):
        for f in range(3):
            parts.append(sol[e, f].astype(">f8").tobytes())
        parts.append(bathy[e, 0].astype(">f8").tobytes())
        parts.append(plv[e].astype(">f8").tobytes())
    return b"".join(parts)


@pytest.mark.parametrize("N,E", [(3, 6), (7, 5)])
def test_shallow_water_pickup_format_and_round_trip(tmp_path, N, E):
    rng = np.random.default_rng(N)
    n1 = N + 1
    sol = rng.standard_normal((E, 3, n1, n1))
    bathy = rng.standard_normal((E, 3, n1, n1))
    plv = rng.standard_normal((E, n1, n1))
    path = tmp_path / P.filename("ShallowWater", 3)
    P.write_sw2d(path, sol, bathy, plv)
    assert path.read_bytes() == _sw_bytes(sol, bathy, plv)
    s2, b2, p2 = P.read_sw2d(path, N, E)
    assert np.array_equal(s2, sol) and np.array_equal(p2, plv)
    assert np.array_equal(b2[:, 0], bathy[:, 0])
    assert not b2[:, 1:].any()  # ReadPickup: bathy = ZERO, only h is read (:2088, :2110)


def test_size_and_open_errors(tmp_path):
    rng = np.random.default_rng(0)
    sol = rng.standard_normal((4, 3, 4, 4))
    path = tmp_path / "s.pickup"
    P.write_sw2d(path, sol, sol, sol[:, 0])
    with pytest.raises(L.SelfError, match="holds"):
        P.read_sw2d(path, 3, 5)  # one element too many
    with pytest.raises(L.SelfError, match="holds"):
        P.read_sw2d(path, 4, 4)  # another degree
    with pytest.raises(L.SelfError, match="cannot open"):
        P.read_sw2d(tmp_path / "missing.pickup", 3, 4)
    path.write_bytes(path.read_bytes()[:-8])
    with pytest.raises(L.SelfError, match="holds"):
        P.read_sw2d(path, 3, 4)


def test_fixture_present():
    assert glob.glob(GOLDEN)
