"""The reference's pickup (restart) files, read and written by the library.

WritePickup/ReadPickup of the shallow-water and advection-diffusion models
(src/highend/shallowwater/ConservativeShallowWaterClass.f90:2013-2125,
src/highend/advection-diffusion2d/AdvectionClass.f90:2040-2171): an
unformatted direct-access file, big-endian, RECL = 8 (N+1)^2 bytes, one
record per (0:N,0:N) element field, elements in order -- shallow water:
U, V, H, h, f; advection: c(1..nTracers), relaxation field(1..nTracers),
relaxation factor(1..nTracers), u, v.  Arrays here are element-major
(nElems, nf, N+1, N+1) as everywhere in this package; the C library does
the I/O.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def filename(model: str, it: int) -> str:
    """'<model>.<it as 10 digits>.pickup', e.g. filename("ShallowWater", 0)."""
    buf = C.create_string_buffer(len(model) + 32)
    L.check(L.lib().self_pickup_filename(model.encode(), int(it), buf, len(buf)))
    return buf.value.decode()


def write_sw2d(path, sol, bathy, plv):
    """sol (nElems, 3, N+1, N+1) U, V, H; bathy (nElems, 3, N+1, N+1) (only
    h is stored); plv (nElems, N+1, N+1)."""
    sol, bathy, plv = L.f64(sol), L.f64(bathy), L.f64(plv)
    E, N = sol.shape[0], sol.shape[-1] - 1
    L.check(L.lib().self_sw2d_pickup_write(str(path).encode(), N, E, L.ptr(sol), L.ptr(bathy), L.ptr(plv)))


def read_sw2d(path, N, nElems):
    """Returns sol, bathy (h; gradient zero, as ReadPickup leaves it), plv."""
    n1 = N + 1
    sol = np.zeros((nElems, 3, n1, n1))
    bathy = np.zeros((nElems, 3, n1, n1))
    plv = np.zeros((nElems, n1, n1))
    L.check(L.lib().self_sw2d_pickup_read(str(path).encode(), N, nElems, L.ptr(sol), L.ptr(bathy), L.ptr(plv)))
    return sol, bathy, plv


def write_adv2d(path, c, relax, rFac, vel):
    """c, relax, rFac (nElems, nTracers, N+1, N+1); vel (nElems, 2, N+1, N+1)."""
    c, relax, rFac, vel = (L.f64(a) for a in (c, relax, rFac, vel))
    E, q, N = c.shape[0], c.shape[1], c.shape[-1] - 1
    L.check(L.lib().self_adv2d_pickup_write(str(path).encode(), N, E, q, L.ptr(c), L.ptr(relax), L.ptr(rFac),
                                            L.ptr(vel)))


def read_adv2d(path, N, nElems, nTracers):
    """Returns dict c, relax, rFac (nElems, nTracers, N+1, N+1), vel (nElems, 2, N+1, N+1)."""
    n1 = N + 1
    out = {k: np.zeros((nElems, nTracers, n1, n1)) for k in ("c", "relax", "rFac")}
    out["vel"] = np.zeros((nElems, 2, n1, n1))
    L.check(L.lib().self_adv2d_pickup_read(str(path).encode(), N, nElems, nTracers, L.ptr(out["c"]),
                                           L.ptr(out["relax"]), L.ptr(out["rFac"]), L.ptr(out["vel"])))
    return out
