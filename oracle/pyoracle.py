"""ctypes front end of the oracle library (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  Never imported by self_amd/.

Each wrapper returns numpy arrays in the reference's Fortran layouts (see
oracle.h); the restated reference routine is cited in oracle.h beside each C
entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

GAUSS, GAUSS_LOBATTO, DG, CG = 1, -1, 2000, 2001
NO_NORMAL_FLOW, RADIATION = -100, -101

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build() -> str:
    """Compile the oracle with its Makefile (gcc, plain C99)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


def _declare(L):
    L.orc_generate_legendre_quadrature.argtypes = [C.c_int, _dp, _dp, C.c_int]
    L.orc_generate_legendre_quadrature.restype = C.c_int
    L.orc_legendre_polynomial.argtypes = [C.c_int, C.c_double, _dp, _dp]
    L.orc_barycentric_weights.argtypes = [C.c_int, _dp, _dp]
    L.orc_lagrange_polynomials_at.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp]
    L.orc_lagrange_interpolate.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp]
    L.orc_lagrange_interpolate.restype = C.c_double
    L.orc_lagrange_derivative.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp]
    L.orc_lagrange_derivative.restype = C.c_double
    L.orc_derivative_matrix.argtypes = [C.c_int, _dp, _dp, _dp]
    L.orc_interpolation_matrix.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    L.orc_coarse_to_fine_2d.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, _dp]
    L.orc_coarse_to_fine_3d.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp]
    L.orc_nodal2d_build.argtypes = [C.POINTER(Nodal2D)]
    L.orc_nodal2d_build.restype = C.c_int
    L.orc_apply_derivative_2d.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.orc_boundary_interp_2d.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_apply_derivative_3d.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.orc_boundary_interp_3d.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    L.orc_quadmesh_counts.argtypes = [C.POINTER(QuadMeshDesc)]
    L.orc_quadmesh_build.argtypes = [C.POINTER(QuadMeshDesc), _dp, _ip, _ip] + [_dp] * 9
    L.orc_quadmesh_build.restype = C.c_int
    L.orc_sw2d_riemann.argtypes = [_dp, _dp, _dp, C.c_double, _dp]
    L.orc_sw2d_riemann_form.argtypes = [_dp, _dp, _dp, C.c_double, C.c_int, _dp]
    L.orc_sw2d_external_state.argtypes = [C.c_int, _dp, _dp, _dp]
    for name in ("orc_sw2d_solution_at_boundaries", "orc_sw2d_bathymetry_gradient"):
        getattr(L, name).argtypes = [C.POINTER(SW2D)]
    for name in ("orc_sw2d_edge_flux", "orc_sw2d_mapped_time_derivative",
                 "orc_sw2d_global_time_derivative", "orc_sw2d_forward_step_rk3"):
        getattr(L, name).argtypes = [C.POINTER(SW2D), C.c_double]
    L.orc_sw2d_integrate.argtypes = [C.POINTER(SW2D), C.c_double, C.c_int, C.c_int]
    for name in ("orc_solve_cg", "orc_solve_pcg", "orc_solve_gmres", "orc_solve_pcgmres"):
        getattr(L, name).argtypes = [C.POINTER(Krylov)]
        getattr(L, name).restype = C.c_int
    L.orc_elliptic1d_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _ip]
    L.orc_elliptic1d_solve.restype = C.c_int
    for extra in ("_declare_extra_adv2d", "_declare_extra_adv3d", "_declare_extra_cgsem2d", "_declare_extra_hpe"):
        fn = globals().get(extra)
        if fn is not None:
            fn(L)


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


class Nodal2D(C.Structure):
    _fields_ = [("N", C.c_int), ("quadrature", C.c_int), ("form", C.c_int)] + [
        (n, _dp) for n in ("nodes", "qWeight", "bWeights", "dMatS", "dMatP", "dMatCorrect",
                           "lagWest", "lagEast", "lagSouth", "lagNorth")]


class QuadMeshDesc(C.Structure):
    _fields_ = [("nX", C.c_int), ("nY", C.c_int), ("N", C.c_int), ("periodicX", C.c_int),
                ("periodicY", C.c_int), ("xScale", C.c_double), ("yScale", C.c_double),
                ("warp", C.c_double), ("nElems", C.c_int), ("nEdges", C.c_int), ("nNodes", C.c_int)]


class SW2D(C.Structure):
    _fields_ = [("N", C.c_int), ("nElems", C.c_int), ("nEdges", C.c_int),
                ("g", C.c_double), ("dt", C.c_double), ("linearDrag", C.c_double)] + [
        (n, _dp) for n in ("dMatS", "dMatP", "qWeightS", "qWeightP", "lagWest", "lagEast",
                           "lagSouth", "lagNorth")] + [("edges", _ip)] + [
        (n, _dp) for n in ("dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen", "sol", "tend",
                           "bsol", "bflux", "bathy", "bbathy", "relax", "plv", "G")] + [("form", C.c_int)]


VecOp = C.CFUNCTYPE(None, C.c_void_p, _dp, _dp)


class Krylov(C.Structure):
    _fields_ = [("nDOF", C.c_int), ("maxIters", C.c_int), ("mInnerIters", C.c_int),
                ("tolerance", C.c_double), ("ctx", C.c_void_p), ("matrixAction", VecOp),
                ("residual", VecOp), ("precondition", VecOp), ("sol", _dp), ("resi", _dp),
                ("resiLen", C.c_int), ("iterations", C.c_int)]


# --------------------------------------------------------------------------- L1/L2
def quadrature(N: int, kind: int = GAUSS):
    x = np.zeros(N + 1)
    w = np.zeros(N + 1)
    rc = lib().orc_generate_legendre_quadrature(N, _p(x), _p(w), kind)
    if rc != 0:
        raise ValueError("invalid quadrature type")
    return x, w


def legendre(nP: int, x: float):
    l, d = C.c_double(), C.c_double()
    lib().orc_legendre_polynomial(nP, x, C.byref(l), C.byref(d))
    return l.value, d.value


def barycentric_weights(nodes):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    bw = np.zeros_like(nodes)
    lib().orc_barycentric_weights(len(nodes) - 1, _p(nodes), _p(bw))
    return bw


def lagrange_at(nodes, bw, x):
    out = np.zeros(len(nodes))
    lib().orc_lagrange_polynomials_at(len(nodes) - 1, _p(nodes), _p(bw), float(x), _p(out))
    return out


def lagrange_derivative(nodes, bw, x, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().orc_lagrange_derivative(len(nodes) - 1, _p(nodes), _p(bw), float(x), _p(f))


def derivative_matrix(nodes, bw):
    n1 = len(nodes)
    D = np.zeros(n1 * n1)
    lib().orc_derivative_matrix(n1 - 1, _p(nodes), _p(bw), _p(D))
    return D.reshape(n1, n1, order="F")


def interpolation_matrix(nodes, bw, znew):
    znew = np.ascontiguousarray(znew, dtype=np.float64)
    T = np.zeros(len(znew) * len(nodes))
    lib().orc_interpolation_matrix(len(nodes) - 1, _p(nodes), _p(bw), len(znew) - 1, _p(znew), _p(T))
    return T.reshape(len(znew), len(nodes), order="F")


def coarse_to_fine_2d(fold, Ts, Tp):
    fold = np.asfortranarray(fold, dtype=np.float64)
    n = fold.shape[0] - 1
    m = Ts.shape[0] - 1
    Ts = np.asfortranarray(Ts)
    Tp = np.asfortranarray(Tp)
    out = np.zeros((m + 1) * (m + 1))
    lib().orc_coarse_to_fine_2d(n, n, _p(fold.ravel(order="F")), m, m,
                                _p(Ts.ravel(order="F")), _p(Tp.ravel(order="F")), _p(out))
    return out.reshape(m + 1, m + 1, order="F")


def coarse_to_fine_3d(fold, T):
    fold = np.asfortranarray(fold, dtype=np.float64)
    n = fold.shape[0] - 1
    m = T.shape[0] - 1
    T = np.asfortranarray(T)
    out = np.zeros((m + 1) ** 3)
    lib().orc_coarse_to_fine_3d(n, _p(fold.ravel(order="F")), m, _p(T.ravel(order="F")), _p(out))
    return out.reshape(m + 1, m + 1, m + 1, order="F")


class NodalStorage2D:
    """NodalStorage_2D (src/nodal/NodalStorage_2D_Class.f90:41-230)."""

    def __init__(self, N: int, quadrature: int = GAUSS, form: int = DG):
        n1 = N + 1
        self.N = N
        self.nodes = np.zeros(n1)
        self.qWeight = np.zeros(n1)
        self.bWeights = np.zeros(n1)
        self.dMatS = np.zeros(n1 * n1)
        self.dMatP = np.zeros(n1 * n1)
        self.dMatCorrect = np.zeros(n1 * n1)
        self.lagWest = np.zeros(n1)
        self.lagEast = np.zeros(n1)
        self.lagSouth = np.zeros(n1)
        self.lagNorth = np.zeros(n1)
        s = Nodal2D(N, quadrature, form, *[_p(getattr(self, n)) for n in (
            "nodes", "qWeight", "bWeights", "dMatS", "dMatP", "dMatCorrect",
            "lagWest", "lagEast", "lagSouth", "lagNorth")])
        if lib().orc_nodal2d_build(C.byref(s)) != 0:
            raise ValueError("invalid nodal storage arguments")
        for n in ("dMatS", "dMatP", "dMatCorrect"):
            setattr(self, n, getattr(self, n).reshape(n1, n1, order="F"))


def apply_derivative_2d(N, dMatS, dMatP, f):
    n1 = N + 1
    f = np.asfortranarray(f).ravel(order="F")
    a = np.zeros(n1 * n1)
    b = np.zeros(n1 * n1)
    lib().orc_apply_derivative_2d(N, _p(np.asfortranarray(dMatS).ravel(order="F")),
                                  _p(np.asfortranarray(dMatP).ravel(order="F")), _p(f), _p(a), _p(b))
    return a.reshape(n1, n1, order="F"), b.reshape(n1, n1, order="F")


def boundary_interp_2d(N, lW, lE, lS, lN, f):
    n1 = N + 1
    f = np.asfortranarray(f).ravel(order="F")
    outs = [np.zeros(n1) for _ in range(4)]
    lib().orc_boundary_interp_2d(N, _p(lW), _p(lE), _p(lS), _p(lN), _p(f), *[_p(o) for o in outs])
    return outs  # west, east, south, north


def apply_derivative_3d(N, dMat, f):
    n1 = N + 1
    f = np.asfortranarray(f).ravel(order="F")
    outs = [np.zeros(n1 ** 3) for _ in range(3)]
    lib().orc_apply_derivative_3d(N, _p(np.asfortranarray(dMat).ravel(order="F")), _p(f),
                                  *[_p(o) for o in outs])
    return [o.reshape(n1, n1, n1, order="F") for o in outs]


def boundary_interp_3d(N, lL, lR, f):
    n1 = N + 1
    f = np.asfortranarray(f).ravel(order="F")
    out = np.zeros(6 * n1 * n1)
    lib().orc_boundary_interp_3d(N, _p(lL), _p(lR), _p(f), _p(out))
    return out.reshape(6, n1, n1).transpose(0, 2, 1)  # [face][a][b] with a fastest in C


# --------------------------------------------------------------------------- mesh
class QuadMesh:
    """Structured quad mesh with reference edge conventions (oracle.h)."""

    def __init__(self, nX, nY, N, sNodes, periodicX=True, periodicY=True,
                 xScale=1.0, yScale=1.0, warp=0.0):
        d = QuadMeshDesc(nX, nY, N, int(periodicX), int(periodicY), xScale, yScale, warp, 0, 0, 0)
        lib().orc_quadmesh_counts(C.byref(d))
        n1 = N + 1
        self.N, self.nElems, self.nEdges = N, d.nElems, d.nEdges
        E = d.nElems
        self.elemNodes = np.zeros(4 * E, dtype=np.int32)
        self.edges = np.zeros(6 * d.nEdges, dtype=np.int32)
        for n in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J"):
            setattr(self, n, np.zeros(n1 * n1 * E))
        self.nHat = np.zeros(2 * n1 * 4 * E)
        self.nLen = np.zeros(n1 * 4 * E)
        sn = np.ascontiguousarray(sNodes, dtype=np.float64)
        rc = lib().orc_quadmesh_build(C.byref(d), _p(sn), _pi(self.elemNodes), _pi(self.edges),
                                      *[_p(getattr(self, n)) for n in (
                                          "x", "y", "dxds", "dxdp", "dyds", "dydp", "J",
                                          "nHat", "nLen")])
        if rc != 0:
            raise ValueError(f"orc_quadmesh_build failed ({rc})")
        self.edges = self.edges.reshape(-1, 6)


# --------------------------------------------------------------------------- SW2D
class ShallowWater2D:
    """Conservative shallow water (ConservativeShallowWaterClass.f90) on the oracle.

    Arrays are flat numpy buffers in the per-element Fortran layouts of oracle.h.
    """

    def __init__(self, nodal: NodalStorage2D, mesh: QuadMesh, g=9.81, dt=1.0,
                 linearDrag=0.0, dMatP=None, with_relax=False, form=0):
        N, E = nodal.N, mesh.nElems
        n1 = N + 1
        self.N, self.nElems = N, E
        self._keep = []
        dms = np.asfortranarray(nodal.dMatS).ravel(order="F").copy()
        dmp = np.asfortranarray(nodal.dMatCorrect if dMatP is None else dMatP).ravel(order="F").copy()
        self._keep += [dms, dmp]
        self.sol = np.zeros(n1 * n1 * 3 * E)
        self.tend = np.zeros_like(self.sol)
        self.G = np.zeros_like(self.sol)
        self.bsol = np.zeros(n1 * 3 * 4 * E)
        self.bflux = np.zeros_like(self.bsol)
        self.bathy = np.zeros(n1 * n1 * 3 * E)
        self.bbathy = np.zeros(n1 * 3 * 4 * E)
        self.relax = np.zeros(n1 * n1 * 4 * E) if with_relax else None
        self.plv = np.zeros(n1 * n1 * E)
        self.mesh = mesh
        self.nodal = nodal
        self.s = SW2D(N, E, mesh.nEdges, g, dt, linearDrag,
                      _p(dms), _p(dmp), _p(nodal.qWeight), _p(nodal.qWeight),
                      _p(nodal.lagWest), _p(nodal.lagEast), _p(nodal.lagSouth), _p(nodal.lagNorth),
                      _pi(np.ascontiguousarray(mesh.edges.ravel())),
                      *[_p(getattr(mesh, n)) for n in ("dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen")],
                      _p(self.sol), _p(self.tend), _p(self.bsol), _p(self.bflux), _p(self.bathy),
                      _p(self.bbathy), _p(self.relax) if self.relax is not None else None,
                      _p(self.plv), _p(self.G), form)
        self._edges = np.ascontiguousarray(mesh.edges.ravel())
        self.s.edges = _pi(self._edges)

    def set_dt(self, dt):
        self.s.dt = dt

    def global_time_derivative(self, t=0.0):
        lib().orc_sw2d_global_time_derivative(C.byref(self.s), t)

    def forward_step_rk3(self, tn=0.0):
        lib().orc_sw2d_forward_step_rk3(C.byref(self.s), tn)

    def integrate(self, t0, nsteps, nthreads=1):
        lib().orc_sw2d_integrate(C.byref(self.s), t0, nsteps, nthreads)

    def bathymetry_gradient(self):
        lib().orc_sw2d_bathymetry_gradient(C.byref(self.s))


def riemann(inS, outS, nHat, g, form=0):
    out = np.zeros(3)
    a = np.ascontiguousarray(inS, dtype=np.float64)
    b = np.ascontiguousarray(outS, dtype=np.float64)
    n = np.ascontiguousarray(nHat, dtype=np.float64)
    lib().orc_sw2d_riemann_form(_p(a), _p(b), _p(n), g, form, _p(out))
    return out


# --------------------------------------------------------------------------- Krylov
class DenseCtx(C.Structure):
    _fields_ = [("n", C.c_int), ("A", _dp), ("b", _dp)]


def _dense_solver(fn_name, A, b, x0=None, maxIters=200, m=30, tol=1e-12, precond=False, native=False):
    """native=True applies the operator with the oracle's C helpers
    (orc_dense_matvec/residual/jacobi: column sums in ascending order, b - As,
    s / diag A) instead of numpy."""
    L = lib()
    n = A.shape[0]
    A = np.asfortranarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    Af = A.ravel(order="F").copy()

    def mv(ctx, i, o):
        v = np.ctypeslib.as_array(i, shape=(n,))
        np.ctypeslib.as_array(o, shape=(n,))[:] = A @ v

    def res(ctx, i, o):
        v = np.ctypeslib.as_array(i, shape=(n,))
        np.ctypeslib.as_array(o, shape=(n,))[:] = b - A @ v

    def pc(ctx, i, o):
        v = np.ctypeslib.as_array(i, shape=(n,))
        np.ctypeslib.as_array(o, shape=(n,))[:] = v / np.diag(A)

    ctx = None
    if native:
        ctx = DenseCtx(n, _p(Af), _p(b))
        cb = tuple(C.cast(getattr(L, f), VecOp) for f in ("orc_dense_matvec", "orc_dense_residual",
                                                          "orc_dense_jacobi"))
    else:
        cb = (VecOp(mv), VecOp(res), VecOp(pc))
    sol = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    resiLen = maxIters * (m if "gmres" in fn_name else 1) + 1
    resi = np.zeros(resiLen)
    k = Krylov(n, maxIters, m, tol, C.cast(C.pointer(ctx), C.c_void_p) if native else None, cb[0], cb[1], cb[2],
               _p(sol), _p(resi), resiLen, 0)
    ioerr = getattr(L, fn_name)(C.byref(k))
    del Af
    return sol, resi, ioerr, k.iterations


def solve_cg(A, b, **kw):
    return _dense_solver("orc_solve_cg", A, b, **kw)


def solve_pcg(A, b, **kw):
    return _dense_solver("orc_solve_pcg", A, b, **kw)


def solve_gmres(A, b, **kw):
    return _dense_solver("orc_solve_gmres", A, b, **kw)


def solve_pcgmres(A, b, **kw):
    return _dense_solver("orc_solve_pcgmres", A, b, **kw)


def elliptic1d(N=5, nElems=10, maxIters=1000, tol=1e-8):
    n = (N + 1) * nElems
    x = np.zeros(n)
    u = np.zeros(n)
    resi = np.zeros(maxIters + 1)
    ioerr = C.c_int()
    iters = lib().orc_elliptic1d_solve(N, nElems, maxIters, tol, _p(x), _p(u), _p(resi), C.byref(ioerr))
    return x, u, resi, ioerr.value, iters


# --------------------------------------------------------------------------- CGSEM 2-D
class CGSEM2DStruct(C.Structure):
    _fields_ = [("N", C.c_int), ("nElems", C.c_int), ("nEdges", C.c_int), ("nNodes", C.c_int),
                ("edges", _ip), ("nodeElemStart", _ip), ("nodeElem", _ip), ("nodeLocal", _ip),
                ("nodeType", _ip), ("D", _dp), ("w", _dp), ("dxds", _dp), ("dxdp", _dp), ("dyds", _dp),
                ("dydp", _dp), ("J", _dp), ("fluxCoeff", _dp), ("solution", _dp), ("source", _dp),
                ("dirichletValue", C.c_double), ("tolerance", C.c_double), ("maxIters", C.c_int),
                ("resi", _dp), ("resiLen", C.c_int), ("q6AsBuilt", C.c_int)]


def _declare_extra_cgsem2d(L):
    P = C.POINTER(CGSEM2DStruct)
    L.orc_cgsem2d_connectivity.argtypes = [P, _ip, C.c_int]
    for n in ("orc_cgsem2d_mask", "orc_cgsem2d_unmask", "orc_cgsem2d_global_sum", "orc_cgsem2d_residual",
              "orc_cgsem2d_diagonal"):
        getattr(L, n).argtypes = [P, _dp]
    L.orc_cgsem2d_set_dirichlet.argtypes = [P]
    L.orc_cgsem2d_matrix_action.argtypes = [P, _dp, _dp]
    L.orc_cgsem2d_dot.argtypes = [P, _dp, _dp]
    L.orc_cgsem2d_dot.restype = C.c_double
    L.orc_cgsem2d_solve_pcg.argtypes = [P, C.c_int, _ip]
    L.orc_cgsem2d_solve_pcg.restype = C.c_int
    L.orc_cgsem2d_pc_size.restype = C.c_size_t
    L.orc_cgsem2d_pc_build.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int]
    L.orc_cgsem2d_pc_free.argtypes = [C.c_void_p]
    L.orc_cgsem2d_pc_solve.argtypes = [C.c_void_p, _dp, _dp, _ip]
    L.orc_cgsem2d_solve_pcg_fem.argtypes = [P, C.c_void_p, _ip, _ip]


class CGSEM2D:
    """CGsemElliptic_2D on a structured walled mesh (all boundaries Dirichlet),
    GLL nodes, CG derivative matrix.  Arrays element-local (nElems, N+1, N+1)
    indexed [e, j(p), i(s)].  q6_as_built: Mask leaves the secondary sides of
    interior edges, as the reference compiled with amdflang -O2 does."""

    def __init__(self, N, nX, nY, Lx=1.0, Ly=1.0, warp=0.0, tolerance=1e-10, maxIters=2000, q6_as_built=False):
        self.N = N
        n1 = N + 1
        self.nodes, self.w = quadrature(N, GAUSS_LOBATTO)
        bw = barycentric_weights(self.nodes)
        self.D = np.ascontiguousarray(derivative_matrix(self.nodes, bw).ravel(order="F"))
        d = QuadMeshDesc(nX, nY, N, 0, 0, Lx, Ly, warp, 0, 0, 0)
        lib().orc_quadmesh_counts(C.byref(d))
        self.mesh = QuadMesh(nX, nY, N, self.nodes, False, False, Lx, Ly, warp)
        E = self.mesh.nElems
        self.nElems = E
        self.fluxCoeff = np.ones(n1 * n1 * E)
        self.solution = np.zeros(n1 * n1 * E)
        self.source = np.zeros(n1 * n1 * E)
        self.resi = np.zeros(maxIters + 1)
        nN = d.nNodes
        self.edges = np.ascontiguousarray(self.mesh.edges.ravel().astype(np.int32))
        self.nodeElemStart = np.zeros(nN + 1, dtype=np.int32)
        self.nodeElem = np.zeros(4 * E, dtype=np.int32)
        self.nodeLocal = np.zeros(4 * E, dtype=np.int32)
        self.nodeType = np.zeros(nN, dtype=np.int32)
        m = self.mesh
        self.s = CGSEM2DStruct(N, E, m.nEdges, nN, _pi(self.edges), _pi(self.nodeElemStart), _pi(self.nodeElem),
                               _pi(self.nodeLocal), _pi(self.nodeType), _p(self.D), _p(self.w), _p(m.dxds),
                               _p(m.dxdp), _p(m.dyds), _p(m.dydp), _p(m.J), _p(self.fluxCoeff),
                               _p(self.solution), _p(self.source), 0.0, tolerance, maxIters, _p(self.resi),
                               maxIters + 1, int(q6_as_built))
        lib().orc_cgsem2d_connectivity(C.byref(self.s), _pi(np.ascontiguousarray(m.elemNodes)), nN)

    def shape(self, a):
        n1 = self.N + 1
        return a.reshape(self.nElems, n1, n1)

    def matrix_action(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).ravel().copy()
        out = np.zeros_like(u)
        lib().orc_cgsem2d_matrix_action(C.byref(self.s), _p(u), _p(out))
        return out, u

    def residual(self):
        r = np.zeros_like(self.solution)
        lib().orc_cgsem2d_residual(C.byref(self.s), _p(r))
        return r

    def mask(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).ravel().copy()
        lib().orc_cgsem2d_mask(C.byref(self.s), _p(u))
        return u

    def unmask(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).ravel().copy()
        lib().orc_cgsem2d_unmask(C.byref(self.s), _p(u))
        return u

    def global_sum(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).ravel().copy()
        lib().orc_cgsem2d_global_sum(C.byref(self.s), _p(u))
        return u

    def diagonal(self):
        d = np.zeros_like(self.solution)
        lib().orc_cgsem2d_diagonal(C.byref(self.s), _p(d))
        return d

    def solve(self, precond=0):
        it = C.c_int()
        ioerr = lib().orc_cgsem2d_solve_pcg(C.byref(self.s), precond, C.byref(it))
        return ioerr, it.value

    def build_preconditioner(self, x, y, fluxCoeff=None, pcTolerance=1e-4, pcIterates=50):
        """CGsemPreconditioner_2D on the linear mesh through the node
        positions x, y (element-local layout); fluxCoeff defaults to ONE, the
        value the reference maps at Build."""
        self._pc = C.create_string_buffer(lib().orc_cgsem2d_pc_size())
        self._pcx = np.ascontiguousarray(x, dtype=np.float64).ravel().copy()
        self._pcy = np.ascontiguousarray(y, dtype=np.float64).ravel().copy()
        k = np.ones_like(self._pcx) if fluxCoeff is None else np.ascontiguousarray(fluxCoeff, dtype=np.float64).ravel()
        self._pck = k.copy()
        lib().orc_cgsem2d_pc_build(self._pc, self.N, self.nElems, _p(self._pcx), _p(self._pcy), _p(self._pck),
                                   pcTolerance, pcIterates)

    def precondition(self, b):
        x = np.zeros_like(self.solution)
        bb = np.ascontiguousarray(b, dtype=np.float64).ravel().copy()
        it = C.c_int()
        io = lib().orc_cgsem2d_pc_solve(self._pc, _p(x), _p(bb), C.byref(it))
        return x, io, it.value

    def solve_fem(self):
        """Solve_PCConjugateGradient with the FEM preconditioner (pcIterates > 0)."""
        it, pit = C.c_int(), C.c_int()
        ioerr = lib().orc_cgsem2d_solve_pcg_fem(C.byref(self.s), self._pc, C.byref(it), C.byref(pit))
        return ioerr, it.value, pit.value

    def __del__(self):
        if getattr(self, "_pc", None) is not None:
            try:
                lib().orc_cgsem2d_pc_free(self._pc)
            except Exception:
                pass


# --------------------------------------------------------------------------- advection-diffusion 2-D
INFLOW_ONE, INFLOW_TWO = -103, -104


class ADV2DStruct(C.Structure):
    _fields_ = [("N", C.c_int), ("nq", C.c_int), ("nElems", C.c_int), ("nEdges", C.c_int),
                ("kappa", C.c_double), ("dt", C.c_double)] + [
        (n, _dp) for n in ("dMatS", "dMatP", "qWeight", "lagWest", "lagEast", "lagSouth", "lagNorth")] + [
        ("edges", _ip)] + [
        (n, _dp) for n in ("dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen", "c", "tend", "G", "dcdx",
                           "dcdy", "u", "v", "relax", "rFac", "bc", "bdx", "bdy", "fc", "fgx", "fgy", "bu", "bv")]


def _declare_extra_adv2d(L):
    P = C.POINTER(ADV2DStruct)
    L.orc_adv2d_global_time_derivative.argtypes = [P, C.c_double]
    L.orc_adv2d_forward_step_rk3.argtypes = [P, C.c_double]


class Advection2D:
    """Advection-diffusion (advection-diffusion2d/AdvectionClass.f90) on the oracle.

    c, tend, dcdx, dcdy, relax, rFac: flat (nElems, nTracers, N+1, N+1) with
    s fastest; u, v: flat (nElems, N+1, N+1).  dMatP defaults to the weak
    operator (dMatCorrect); pass nodal.dMatP for the reference's transposed one.
    """

    def __init__(self, nodal: NodalStorage2D, mesh: QuadMesh, nTracers=1, kappa=0.0, dt=1.0, dMatP=None,
                 with_relax=False, edges=None):
        N, E = nodal.N, mesh.nElems
        n1, nq = N + 1, nTracers
        self.N, self.nElems, self.nq = N, E, nq
        self.dms = np.asfortranarray(nodal.dMatS).ravel(order="F").copy()
        self.dmp = np.asfortranarray(nodal.dMatCorrect if dMatP is None else dMatP).ravel(order="F").copy()
        self.edges = np.ascontiguousarray((mesh.edges if edges is None else edges).ravel().astype(np.int32))
        vol, tr = n1 * n1 * nq * E, n1 * nq * 4 * E
        for n in ("c", "tend", "G", "dcdx", "dcdy"):
            setattr(self, n, np.zeros(vol))
        self.relax = np.zeros(vol) if with_relax else None
        self.rFac = np.zeros(vol) if with_relax else None
        self.u = np.zeros(n1 * n1 * E)
        self.v = np.zeros(n1 * n1 * E)
        for n in ("bc", "bdx", "bdy", "fc", "fgx", "fgy"):
            setattr(self, n, np.zeros(tr))
        self.bu = np.zeros(n1 * 4 * E)
        self.bv = np.zeros(n1 * 4 * E)
        self.mesh, self.nodal = mesh, nodal
        self.s = ADV2DStruct(N, nq, E, mesh.nEdges, kappa, dt, _p(self.dms), _p(self.dmp), _p(nodal.qWeight),
                             _p(nodal.lagWest), _p(nodal.lagEast), _p(nodal.lagSouth), _p(nodal.lagNorth),
                             _pi(self.edges),
                             *[_p(getattr(mesh, n)) for n in ("dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen")],
                             _p(self.c), _p(self.tend), _p(self.G), _p(self.dcdx), _p(self.dcdy), _p(self.u),
                             _p(self.v), _p(self.relax), _p(self.rFac), _p(self.bc), _p(self.bdx), _p(self.bdy),
                             _p(self.fc), _p(self.fgx), _p(self.fgy), _p(self.bu), _p(self.bv))

    def global_time_derivative(self, t=0.0):
        lib().orc_adv2d_global_time_derivative(C.byref(self.s), t)

    def forward_step_rk3(self, tn=0.0):
        lib().orc_adv2d_forward_step_rk3(C.byref(self.s), tn)


# --------------------------------------------------------------------------- 3-D: mesh, filter, advection
class HexMeshDesc(C.Structure):
    _fields_ = [("nX", C.c_int), ("nY", C.c_int), ("nZ", C.c_int), ("N", C.c_int), ("periodicX", C.c_int),
                ("periodicY", C.c_int), ("periodicZ", C.c_int), ("xScale", C.c_double), ("yScale", C.c_double),
                ("zScale", C.c_double), ("warp", C.c_double), ("warpSign", C.c_int * 3), ("nElems", C.c_int),
                ("nFaces", C.c_int)]


class ADV3DStruct(C.Structure):
    _fields_ = [("N", C.c_int), ("nq", C.c_int), ("nElems", C.c_int), ("nFaces", C.c_int), ("useFilter", C.c_int),
                ("dt", C.c_double)] + [(n, _dp) for n in ("dMat", "qWeight", "lagLeft", "lagRight", "F")] + [
        ("faces", _ip)] + [(n, _dp) for n in ("cov", "J", "nHat", "nLen", "bpos", "c", "tend", "G", "u", "v", "w",
                                             "relax", "rFac", "bc", "fc", "bu", "bv", "bw", "E1", "E2", "lim")] + [
        ("nFiltered", C.c_int)]


def _declare_extra_adv3d(L):
    L.orc_hexmesh_counts.argtypes = [C.POINTER(HexMeshDesc)]
    L.orc_hexmesh_build.argtypes = [C.POINTER(HexMeshDesc), _dp, _ip] + [_dp] * 8
    L.orc_hexmesh_build.restype = C.c_int
    L.orc_rolloff_filter3d_build.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
    L.orc_rolloff_filter3d_apply.argtypes = [C.c_int, _dp, _dp, _dp]
    P = C.POINTER(ADV3DStruct)
    L.orc_adv3d_adaptive_filter.argtypes = [P, C.c_int]
    L.orc_adv3d_global_time_derivative.argtypes = [P, C.c_double, C.c_int]
    L.orc_adv3d_forward_step_rk3.argtypes = [P, C.c_double]


class HexMesh:
    """Structured hex mesh of the unit cube (orc_adv3d.h).  Volume arrays
    flat per element (i fastest); faces (nFaces, 9)."""

    def __init__(self, nX, nY, nZ, N, nodes, periodic=(True, True, True), scale=(1.0, 1.0, 1.0), warp=0.0,
                 warpSign=(1, 1, 1)):
        d = HexMeshDesc(nX, nY, nZ, N, int(periodic[0]), int(periodic[1]), int(periodic[2]), scale[0], scale[1],
                        scale[2], warp, (C.c_int * 3)(*warpSign), 0, 0)
        lib().orc_hexmesh_counts(C.byref(d))
        n1 = N + 1
        E, nF = d.nElems, d.nFaces
        self.N, self.nElems, self.nFaces = N, E, nF
        self.faces = np.zeros(9 * nF, dtype=np.int32)
        for n in ("x", "y", "z", "J"):
            setattr(self, n, np.zeros(n1 ** 3 * E))
        self.cov = np.zeros(9 * n1 ** 3 * E)
        self.nHat = np.zeros(3 * n1 * n1 * 6 * E)
        self.nLen = np.zeros(n1 * n1 * 6 * E)
        self.bpos = np.zeros(3 * n1 * n1 * 6 * E)
        rc = lib().orc_hexmesh_build(C.byref(d), _p(np.ascontiguousarray(nodes, dtype=np.float64)),
                                     _pi(self.faces), _p(self.x), _p(self.y), _p(self.z), _p(self.cov), _p(self.J),
                                     _p(self.nHat), _p(self.nLen), _p(self.bpos))
        if rc != 0:
            raise ValueError(f"orc_hexmesh_build failed ({rc})")
        self.faces = self.faces.reshape(-1, 9)


def rolloff_filter3d(N, nc, nodes, w):
    nD = (N + 1) ** 3
    F = np.zeros(nD * nD)
    lib().orc_rolloff_filter3d_build(N, nc, _p(np.ascontiguousarray(nodes, dtype=np.float64)),
                                     _p(np.ascontiguousarray(w, dtype=np.float64)), _p(F))
    return F.reshape(nD, nD, order="F")


class Nodal3D:
    """NodalStorage_3D DG (src/nodal/NodalStorage_3D_Class.f90:107-264) on Gauss nodes."""

    def __init__(self, N):
        self.N = N
        self.nodes, self.qWeight = quadrature(N, GAUSS)
        bw = barycentric_weights(self.nodes)
        D = derivative_matrix(self.nodes, bw)
        n1 = N + 1
        self.dMat = np.zeros((n1, n1))
        for j in range(n1):
            for i in range(n1):
                self.dMat[i, j] = -D[j, i] * self.qWeight[j] / self.qWeight[i]
        self.lagLeft = lagrange_at(self.nodes, bw, -1.0)
        self.lagRight = lagrange_at(self.nodes, bw, 1.0)


class Advection3D:
    """Advection3D (advection3d/Advection3DClass.f90) on the oracle.

    c, tend, relax, rFac: flat (nElems, nTracers, N+1, N+1, N+1) i fastest;
    u, v, w: flat (nElems, N+1, N+1, N+1)."""

    def __init__(self, nodal: Nodal3D, mesh: HexMesh, nTracers=1, dt=1e-3, nCutoff=3, useFilter=True,
                 with_relax=False, faces=None):
        N, E = nodal.N, mesh.nElems
        n1, nq = N + 1, nTracers
        self.N, self.nElems, self.nq = N, E, nq
        self.dm = np.ascontiguousarray(nodal.dMat.ravel(order="F"))
        self.F = np.ascontiguousarray(rolloff_filter3d(N, nCutoff, nodal.nodes, nodal.qWeight).ravel(order="F"))
        self.faces = np.ascontiguousarray((mesh.faces if faces is None else faces).ravel().astype(np.int32))
        vol, tr = n1 ** 3 * nq * E, n1 * n1 * nq * 6 * E
        for n in ("c", "tend", "G"):
            setattr(self, n, np.zeros(vol))
        self.relax = np.zeros(vol) if with_relax else None
        self.rFac = np.zeros(vol) if with_relax else None
        for n in ("u", "v", "w"):
            setattr(self, n, np.zeros(n1 ** 3 * E))
        self.bc = np.zeros(tr)
        self.fc = np.zeros(tr)
        for n in ("bu", "bv", "bw"):
            setattr(self, n, np.zeros(n1 * n1 * 6 * E))
        self.E1 = np.zeros(nq * E)
        self.E2 = np.zeros(nq * E)
        self.lim = np.zeros(3 * nq * E)
        self.mesh, self.nodal = mesh, nodal
        self.s = ADV3DStruct(N, nq, E, mesh.nFaces, int(useFilter), dt, _p(self.dm), _p(nodal.qWeight),
                             _p(nodal.lagLeft), _p(nodal.lagRight), _p(self.F), _pi(self.faces), _p(mesh.cov),
                             _p(mesh.J), _p(mesh.nHat), _p(mesh.nLen), _p(mesh.bpos), _p(self.c), _p(self.tend),
                             _p(self.G), _p(self.u), _p(self.v), _p(self.w), _p(self.relax), _p(self.rFac),
                             _p(self.bc), _p(self.fc), _p(self.bu), _p(self.bv), _p(self.bw), _p(self.E1),
                             _p(self.E2), _p(self.lim), 0)

    @property
    def nFiltered(self):
        return self.s.nFiltered

    def adaptive_filter(self, stage):
        lib().orc_adv3d_adaptive_filter(C.byref(self.s), stage)

    def global_time_derivative(self, t=0.0, stage=1):
        lib().orc_adv3d_global_time_derivative(C.byref(self.s), t, stage)

    def forward_step_rk3(self, tn=0.0):
        lib().orc_adv3d_forward_step_rk3(C.byref(self.s), tn)


class HPEStruct(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("N", "M", "nHElems", "nLayers", "nElems", "nFaces", "diagnoseP")] + [
        ("dt", C.c_double)] + [(n, _dp) for n in ("nodes", "bWeights", "dMat", "qWeight", "lagLeft", "lagRight",
                                                   "intNodes", "intWeights")] + [("faces", _ip)] + [
        (n, _dp) for n in ("cov", "J", "nHat", "nLen", "bpos", "sol", "bg", "bsol", "bflux", "tend", "G", "divU",
                           "bottomW")]


def _declare_extra_hpe(L):
    P = C.POINTER(HPEStruct)
    for n in ("traces", "divu", "diagnose_p", "diagnose_w", "edge_flux", "mapped", "global_time_derivative",
              "forward_step_rk3", "forward_step_ab2"):
        getattr(L, "orc_hpe_" + n).argtypes = [P]


class HydrostaticPrimitive:
    """The hydrostatic primitive equations (hydrostaticprimitives/
    HydrostaticPrimitivesClass.f90) on the oracle, on a stacked structured hex
    mesh of nX x nY columns and nLayers layers (orc_hpe.h).

    sol, bg: (nElems, 5, N+1, N+1, N+1) i fastest, variables u v b w p;
    tend, G: (nElems, 3, ...); divU (nElems, ...); bottomW (nHElems, N+1, N+1)."""

    def __init__(self, N, nX, nY, nLayers, periodic=(False, False), scale=(1.0, 1.0, 1.0), warp=0.0,
                 warpSign=(1, 1, 1), dt=1e-3, diagnoseP=True):
        nodal = Nodal3D(N)
        mesh = HexMesh(nX, nY, nLayers, N, nodal.nodes, periodic=(periodic[0], periodic[1], False), scale=scale,
                       warp=warp, warpSign=warpSign)
        n1, E = N + 1, mesh.nElems
        self.N, self.nElems, self.nHElems, self.nLayers = N, E, nX * nY, nLayers
        self.M = (N + 1) // 2 + 1
        self.intNodes, self.intWeights = quadrature(self.M, GAUSS_LOBATTO)
        self.bWeights = barycentric_weights(nodal.nodes)
        self.dm = np.ascontiguousarray(nodal.dMat.ravel(order="F"))
        self.faces = np.ascontiguousarray(mesh.faces.ravel().astype(np.int32))
        self.sol = np.zeros(5 * n1 ** 3 * E)
        self.bg = np.zeros(5 * n1 ** 3 * E)
        self.bsol = np.zeros(5 * n1 * n1 * 6 * E)
        self.bflux = np.zeros(3 * n1 * n1 * 6 * E)
        self.tend = np.zeros(3 * n1 ** 3 * E)
        self.G = np.zeros(3 * n1 ** 3 * E)
        self.divU = np.zeros(n1 ** 3 * E)
        self.bottomW = np.zeros(n1 * n1 * self.nHElems)
        self.mesh, self.nodal = mesh, nodal
        self.s = HPEStruct(N, self.M, self.nHElems, nLayers, E, mesh.nFaces, int(diagnoseP), dt, _p(nodal.nodes),
                           _p(self.bWeights), _p(self.dm), _p(nodal.qWeight), _p(nodal.lagLeft),
                           _p(nodal.lagRight), _p(self.intNodes), _p(self.intWeights), _pi(self.faces),
                           _p(mesh.cov), _p(mesh.J), _p(mesh.nHat), _p(mesh.nLen), _p(mesh.bpos), _p(self.sol),
                           _p(self.bg), _p(self.bsol), _p(self.bflux), _p(self.tend), _p(self.G), _p(self.divU),
                           _p(self.bottomW))

    def view(self, a, nv):
        n1 = self.N + 1
        return a.reshape(self.nElems, nv, n1, n1, n1)

    def call(self, name):
        getattr(lib(), "orc_hpe_" + name)(C.byref(self.s))
