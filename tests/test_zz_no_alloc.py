"""self_coarse_to_fine_2d_noalloc (no device allocation, staging in the
library's static device array) against self_coarse_to_fine_2d, called
through the C ABI: bitwise the same output on two calls in a row, the
second with a smaller input (so stale contents of the shared array from the
first call would show), and CoarseToFine_2D's values checked on the oracle's
statement of it."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(nS, nP, nNewS, nNewP, nE, seed):
    rng = np.random.default_rng(seed)
    Ts = np.asfortranarray(rng.uniform(-1, 1, (nNewS + 1, nS + 1)))
    Tp = np.asfortranarray(rng.uniform(-1, 1, (nNewP + 1, nP + 1)))
    f = rng.uniform(-1, 1, nE * (nS + 1) * (nP + 1))
    return Ts, Tp, f


def _call(name, nS, nP, nNewS, nNewP, nE, Ts, Tp, f):
    from self_amd import _lib as L
    out = np.full(nE * (nNewS + 1) * (nNewP + 1), np.nan)
    L.check(getattr(L.lib(), name)(nS, nP, nE, L.ptr(np.ascontiguousarray(Ts.ravel(order="F"))),
                                   L.ptr(np.ascontiguousarray(Tp.ravel(order="F"))), nNewS, nNewP, L.ptr(f),
                                   L.ptr(out)))
    return out


def test_noalloc_equals_the_allocating_entry_point_twice(gpu_lib):
    from oracle import pyoracle as po
    cases = [(7, 7, 10, 10, 3000, 1), (5, 4, 6, 8, 37, 2)]  # second call smaller than the first
    for nS, nP, nNewS, nNewP, nE, seed in cases:
        Ts, Tp, f = _inputs(nS, nP, nNewS, nNewP, nE, seed)
        a = _call("self_coarse_to_fine_2d_noalloc", nS, nP, nNewS, nNewP, nE, Ts, Tp, f)
        b = _call("self_coarse_to_fine_2d", nS, nP, nNewS, nNewP, nE, Ts, Tp, f)
        assert np.all(np.isfinite(a)) and np.array_equal(a, b), (nS, nP, nE)
        if nS != nP or nNewS != nNewP:
            continue  # the oracle's statement takes square elements
        # the first three elements against the oracle's CoarseToFine_2D
        n_in, n_out = (nS + 1) * (nP + 1), (nNewS + 1) * (nNewP + 1)
        for e in range(3):
            fe = f[e * n_in:(e + 1) * n_in].reshape(nP + 1, nS + 1).T
            ref = po.coarse_to_fine_2d(fe, Ts, Tp)
            got = a[e * n_out:(e + 1) * n_out].reshape(nNewS + 1, nNewP + 1).T
            assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_noalloc_rejects_calls_beyond_its_array(gpu_lib):
    from self_amd import _lib as L
    Ts, Tp, f = _inputs(7, 7, 10, 10, 1, 3)
    out = np.zeros(121)
    nE = (1 << 21) // (64 + 121) + 1
    rc = L.lib().self_coarse_to_fine_2d_noalloc(7, 7, nE, L.ptr(np.ascontiguousarray(Ts.ravel(order="F"))),
                                                L.ptr(np.ascontiguousarray(Tp.ravel(order="F"))), 10, 10,
                                                L.ptr(f), L.ptr(out))
    assert rc == -5
    assert b"scratch" in L.lib().self_last_error()
