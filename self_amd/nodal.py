"""Host-side interpolation and nodal storage (set-up), over the C ABI.

Mirrors src/interp/Legendre.f90, src/interp/Lagrange_1D_Class.f90 and
src/nodal/NodalStorage_2D_Class.f90 (BuildNodalStorage_2D :101-230).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def legendre_quadrature(N: int, quadType: int = L.GAUSS):
    """GenerateLegendreQuadrature (Legendre.f90:180-212): nodes, weights."""
    x = np.zeros(N + 1)
    w = np.zeros(N + 1)
    L.check(L.lib().self_legendre_quadrature(N, quadType, L.ptr(x), L.ptr(w)))
    return x, w


def barycentric_weights(x):
    x = L.f64(x)
    bw = np.zeros_like(x)
    L.check(L.lib().self_barycentric_weights(len(x) - 1, L.ptr(x), L.ptr(bw)))
    return bw


def lagrange_polynomials(x, bw, xEval):
    x = L.f64(x)
    out = np.zeros_like(x)
    L.check(L.lib().self_lagrange_polynomials(len(x) - 1, L.ptr(x), L.ptr(L.f64(bw)), float(xEval), L.ptr(out)))
    return out


def derivative_matrix(x, bw):
    x = L.f64(x)
    n1 = len(x)
    D = np.zeros(n1 * n1)
    L.check(L.lib().self_derivative_matrix(n1 - 1, L.ptr(x), L.ptr(L.f64(bw)), L.ptr(D)))
    return D.reshape(n1, n1, order="F")


def interpolation_matrix(x, bw, znew):
    x = L.f64(x)
    znew = L.f64(znew)
    T = np.zeros(len(znew) * len(x))
    L.check(L.lib().self_interpolation_matrix(len(x) - 1, L.ptr(x), L.ptr(L.f64(bw)), len(znew) - 1,
                                              L.ptr(znew), L.ptr(T)))
    return T.reshape(len(znew), len(x), order="F")


class NodalStorage2D:
    """NodalStorage_2D: quadrature, derivative matrices, boundary interpolants.

    quirk_dmatp_transposed reproduces the reference's DG dMatP (the transpose
    of the weak operator, NodalStorage_2D_Class.f90:209-218); by default both
    directions use the weak operator as the reference's 1-D/3-D storages do.
    Matrices are returned as (N+1, N+1) numpy arrays, element (row, col).
    """

    def __init__(self, N: int, quadrature: int = L.GAUSS, form: int = L.DG,
                 quirk_dmatp_transposed: bool = False):
        n1 = N + 1
        self.N = N
        self.quadrature = quadrature
        self.form = form
        self.nodes = np.zeros(n1)
        self.qWeight = np.zeros(n1)
        dms = np.zeros(n1 * n1)
        dmp = np.zeros(n1 * n1)
        self.lagWest, self.lagEast = np.zeros(n1), np.zeros(n1)
        self.lagSouth, self.lagNorth = np.zeros(n1), np.zeros(n1)
        L.check(L.lib().self_nodal2d_build(N, quadrature, form, int(quirk_dmatp_transposed), L.ptr(self.nodes),
                                           L.ptr(self.qWeight), L.ptr(dms), L.ptr(dmp), L.ptr(self.lagWest),
                                           L.ptr(self.lagEast), L.ptr(self.lagSouth), L.ptr(self.lagNorth)))
        self.dMatS = dms.reshape(n1, n1, order="F")
        self.dMatP = dmp.reshape(n1, n1, order="F")

    def flat(self, name):
        """Column-major flat copy of a matrix, as the C ABI takes it."""
        return np.ascontiguousarray(getattr(self, name).ravel(order="F"))


class NodalStorage3D:
    """NodalStorage_3D (src/nodal/NodalStorage_3D_Class.f90:107-264), equal
    degree in s, p, q: nodes, qWeight, dMat (weak DG operator for L.DG),
    lagLeft (at -1) and lagRight (at +1)."""

    def __init__(self, N: int, quadrature: int = L.GAUSS, form: int = L.DG):
        n1 = N + 1
        self.N = N
        self.nodes, self.qWeight = np.zeros(n1), np.zeros(n1)
        dm = np.zeros(n1 * n1)
        self.lagLeft, self.lagRight = np.zeros(n1), np.zeros(n1)
        L.check(L.lib().self_nodal3d_build(N, quadrature, form, L.ptr(self.nodes), L.ptr(self.qWeight), L.ptr(dm),
                                           L.ptr(self.lagLeft), L.ptr(self.lagRight)))
        self.dMat = dm.reshape(n1, n1, order="F")


def rolloff_filter1d(N, nCutoff, nodes, weights):
    """Direction factor F1 of RollOffFilter3D ((N+1) x (N+1)); the dense
    filter is kron(F1, kron(F1, F1))."""
    F1 = np.zeros((N + 1) * (N + 1))
    L.check(L.lib().self_rolloff_filter1d(N, nCutoff, L.ptr(L.f64(nodes)), L.ptr(L.f64(weights)), L.ptr(F1)))
    return F1.reshape(N + 1, N + 1, order="F")


# ---------------------------------------------------------------- batched library operators
def _elems(f, per):
    f = L.f64(f)
    if f.size % per:
        raise ValueError(f"array size {f.size} is not a multiple of the element size {per}")
    return f, f.size // per


def apply_derivative_2d(dMatS, dMatP, f):
    """ApplyDerivativeMatrix_NodalStorage_2D for every element of f
    (nElems, N+1, N+1) indexed [e, j(p), i(s)]: returns (dFds, dFdp)."""
    n1 = np.asarray(dMatS).shape[0]
    f, nE = _elems(f, n1 * n1)
    ds, dp_ = np.zeros_like(f), np.zeros_like(f)
    L.check(L.lib().self_nodal2d_apply_derivative(n1 - 1, nE, L.ptr(L.f64(np.asarray(dMatS).ravel(order="F"))),
                                                  L.ptr(L.f64(np.asarray(dMatP).ravel(order="F"))), L.ptr(f),
                                                  L.ptr(ds), L.ptr(dp_)))
    return ds, dp_


def boundary_interp_2d(lW, lE, lS, lN, f):
    """CalculateAtBoundaries_NodalStorage_2D: (fWest, fEast, fSouth, fNorth), each (nElems, N+1)."""
    n1 = len(lW)
    f, nE = _elems(f, n1 * n1)
    out = [np.zeros((nE, n1)) for _ in range(4)]
    L.check(L.lib().self_nodal2d_boundary_interp(n1 - 1, nE, *[L.ptr(L.f64(a)) for a in (lW, lE, lS, lN)], L.ptr(f),
                                                 *[L.ptr(o) for o in out]))
    return tuple(out)


def apply_derivative_3d(dMat, f):
    """ApplyDerivativeMatrix_NodalStorage_3D: (dFds, dFdp, dFdq) shaped like f."""
    n1 = np.asarray(dMat).shape[0]
    f, nE = _elems(f, n1 ** 3)
    out = [np.zeros_like(f) for _ in range(3)]
    L.check(L.lib().self_nodal3d_apply_derivative(n1 - 1, nE, L.ptr(L.f64(np.asarray(dMat).ravel(order="F"))),
                                                  L.ptr(f), *[L.ptr(o) for o in out]))
    return tuple(out)


def boundary_interp_3d(lagLeft, lagRight, f):
    """CalculateAtBoundaries_NodalStorage_3D: (nElems, 6, N+1, N+1), faces
    south, east, north, west, bottom, top."""
    n1 = len(lagLeft)
    f, nE = _elems(f, n1 ** 3)
    out = np.zeros((nE, 6, n1, n1))
    L.check(L.lib().self_nodal3d_boundary_interp(n1 - 1, nE, L.ptr(L.f64(lagLeft)), L.ptr(L.f64(lagRight)), L.ptr(f),
                                                 L.ptr(out)))
    return out


def coarse_to_fine_2d(Ts, Tp, f):
    """CoarseToFine_2D: f (nElems, nP+1, nS+1) [e, j, i] -> (nElems, nNewS+1, nNewP+1),
    i.e. the reference's fNew(0:nNewP, 0:nNewS) per element."""
    Ts, Tp = np.asarray(Ts), np.asarray(Tp)
    ms, s1 = Ts.shape
    mp, p1 = Tp.shape
    f, nE = _elems(f, s1 * p1)
    out = np.zeros((nE, ms, mp))
    L.check(L.lib().self_coarse_to_fine_2d(s1 - 1, p1 - 1, nE, L.ptr(L.f64(Ts.ravel(order="F"))),
                                           L.ptr(L.f64(Tp.ravel(order="F"))), ms - 1, mp - 1, L.ptr(f), L.ptr(out)))
    return out


def coarse_to_fine_3d(T, f):
    """CoarseToFine_3D: f (nElems, N+1, N+1, N+1) -> (nElems, nNew+1, nNew+1, nNew+1)."""
    T = np.asarray(T)
    m1, n1 = T.shape
    f, nE = _elems(f, n1 ** 3)
    out = np.zeros((nE, m1, m1, m1))
    L.check(L.lib().self_coarse_to_fine_3d(n1 - 1, nE, L.ptr(L.f64(T.ravel(order="F"))), m1 - 1, L.ptr(f),
                                           L.ptr(out)))
    return out
