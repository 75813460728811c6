"""Krylov solvers and the CG spectral-element elliptic solver on the GPU.

LinOp mirrors the reference's IterativeSolver plug-in point
(src/iterativesolve/IterativeSolvers_Class.f90:55-186): an operator object
whose MatrixAction / Residual / PreConditionInvert the solver calls.  Here the
operator lives on the device and Solve runs CG, PCG, GMRES(m) or FGMRES(m)
with the reference's control flow and stop tests entirely on the device.

CGsemElliptic2D mirrors CGsemElliptic_2D (src/cgsem/CGsemElliptic_2D_Class.f90).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .mesh import QuadMesh
from .nodal import legendre_quadrature

CG, PCG, GMRES, PCGMRES, CGSEM_PCG = 0, 1, 2, 3, 4
_METHODS = {"cg": CG, "pcg": PCG, "gmres": GMRES, "pcgmres": PCGMRES, "cgsem_pcg": CGSEM_PCG}
MGS, CGS2 = 0, 1


class LinOp:
    """A device operator for the generic solvers."""

    def __init__(self, handle):
        self._h = handle
        n = C.c_int()
        L.check(L.lib().self_linop_size(self._h, C.byref(n)))
        self.nDOF = n.value

    @classmethod
    def dense(cls, A, b):
        A = np.asfortranarray(A, dtype=np.float64)
        n = A.shape[0]
        h = C.c_void_p()
        L.check(L.lib().self_linop_dense_create(n, L.ptr(np.ascontiguousarray(A.ravel(order="F"))),
                                                L.ptr(L.f64(b)), C.byref(h)))
        return cls(h)

    @classmethod
    def elliptic1d(cls, N=5, nElems=10, xScale=1.0):
        """Elliptic1D (examples/elliptic1d/parabola/localmods/Elliptic1D_Class.f90)."""
        h = C.c_void_p()
        L.check(L.lib().self_linop_elliptic1d_create(N, nElems, xScale, C.byref(h)))
        op = cls(h)
        op.N, op.nElems = N, nElems
        return op

    def Solve(self, method="cg", tolerance=1e-8, maxIters=1000, mInner=30, ortho=MGS, sol=None):
        """IterativeSolver%Solve: returns (sol, resi, iterations, ioerr)."""
        m = _METHODS[method] if isinstance(method, str) else method
        x = np.zeros(self.nDOF) if sol is None else np.array(sol, dtype=np.float64)
        rl = maxIters * (mInner if m in (GMRES, PCGMRES) else 1) + 1
        resi = np.zeros(rl)
        it, io = C.c_int(), C.c_int()
        L.check(L.lib().self_krylov_solve(self._h, m, tolerance, maxIters, mInner, ortho, L.ptr(x), L.ptr(resi), rl,
                                          C.byref(it), C.byref(io)))
        return x, resi, it.value, io.value

    def UnMaskSolution(self, sol):
        u = np.zeros(self.nDOF)
        x = np.zeros(self.nDOF)
        L.check(L.lib().self_linop_elliptic1d_unmask(self._h, L.ptr(L.f64(sol)), L.ptr(u), L.ptr(x)))
        return x, u

    def Trash(self):
        if getattr(self, "_h", None):
            L.check(L.lib().self_linop_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.Trash()
        except Exception:
            pass


@dataclass
class CGsemParams:
    """Subset of CGsemParams (src/cgsem/CGsemParams_Class.f90:40-144)."""
    polyDeg: int = 7
    nXElem: int = 4
    nYElem: int = 4
    xScale: float = 1.0
    yScale: float = 1.0
    tolerance: float = 1e-10
    maximumIterates: int = 5000
    warp: float = 0.0


_PRECOND = {"none": 0, "jacobi": 1, "fem": 2}


class CGsemElliptic2D:
    """div(k grad u) = s with Dirichlet boundaries, GLL nodes (CG form).

    Arrays are element-local (nElems, N+1, N+1) indexed [e, j(p), i(s)],
    the reference's solution(0:N,0:N,1:nElems)."""

    def __init__(self, params: CGsemParams, mesh=None):
        """mesh: a SpecMesh read on the GLL nodes of degree polyDeg (the
        reference's SpecMeshFile, BuildQuadMesh :169-179); default the
        structured nXElem x nYElem mesh (LoadDefaultMesh)."""
        self.params = params
        N = params.polyDeg
        self.nS = self.nP = N
        self.nodes, self.qWeight = legendre_quadrature(N, L.GAUSS_LOBATTO)
        if mesh is not None:
            if mesh.N != N or not np.array_equal(mesh._sNodes, self.nodes):
                raise L.SelfError(L.SELF_ERR_ARG, "CGsemElliptic2D: the mesh must be read on the GLL nodes of degree N")
            self.mesh = mesh
            self.nElems = mesh.nElems
            self.elemNodes = np.ascontiguousarray(mesh.elemNodes.ravel())
            nNodes = mesh.nNodes
        else:
            self.mesh = QuadMesh(params.nXElem, params.nYElem, N, self.nodes, False, False, params.xScale,
                                 params.yScale, params.warp)
            self.nElems = self.mesh.nElems
            nX, nY = params.nXElem, params.nYElem
            # corner node ids of LoadDefaultMesh (walled): SW, SE, NE, NW
            en = np.zeros((self.nElems, 4), dtype=np.int32)
            for iY in range(1, nY + 1):
                for iX in range(1, nX + 1):
                    e = iX + (iY - 1) * nX - 1
                    n0 = iX + (iY - 1) * (nX + 1)
                    en[e] = (n0, n0 + 1, n0 + 1 + nX + 1, n0 + nX + 1)
            self.elemNodes = np.ascontiguousarray(en.ravel())
            nNodes = (nX + 1) * (nY + 1)
        m = self.mesh
        h = C.c_void_p()
        L.check(L.lib().self_cgsem2d_create(N, m.nElems, m.nEdges, nNodes, L.ptr(m.edges, L.ip),
                                            L.ptr(self.elemNodes, L.ip), L.ptr(m.dxds), L.ptr(m.dxdp),
                                            L.ptr(m.dyds), L.ptr(m.dydp), L.ptr(m.J), C.byref(h)))
        self._h = h
        g = C.c_int()
        L.check(L.lib().self_cgsem2d_global_dofs(self._h, C.byref(g)))
        self.nGlobal = g.value
        self.resi = None

    def _arr(self, a):
        a = L.f64(a)
        if a.size != self.nElems * (self.nS + 1) ** 2:
            raise ValueError("expected an element-local array (nElems, N+1, N+1)")
        return a

    def _new(self):
        return np.zeros((self.nElems, self.nS + 1, self.nS + 1))

    def SetFluxCoefficient(self, k):
        L.check(L.lib().self_cgsem2d_set_flux_coefficient(self._h, L.ptr(self._arr(k))))

    def SetSource(self, s):
        L.check(L.lib().self_cgsem2d_set_source(self._h, L.ptr(self._arr(s))))

    def SetDirichletBoundaryCondition(self, values):
        L.check(L.lib().self_cgsem2d_set_dirichlet(self._h, L.ptr(self._arr(values))))

    def SetSolution(self, u):
        L.check(L.lib().self_cgsem2d_set_solution(self._h, L.ptr(self._arr(u))))

    def GetSolution(self):
        out = self._new()
        L.check(L.lib().self_cgsem2d_get_solution(self._h, L.ptr(out)))
        return out

    def MatrixAction(self, u):
        out = self._new()
        L.check(L.lib().self_cgsem2d_matrix_action(self._h, L.ptr(self._arr(u)), L.ptr(out)))
        return out

    def Residual(self):
        out = self._new()
        L.check(L.lib().self_cgsem2d_residual(self._h, L.ptr(out)))
        return out

    def Solve(self, method="cgsem_pcg", precond="none", tolerance=None, maxIters=None, mInner=30, ortho=MGS):
        """Returns (ioerr, iterations, elapsed_ms); the residual history is self.resi."""
        m = _METHODS[method] if isinstance(method, str) else method
        tol = self.params.tolerance if tolerance is None else tolerance
        mi = self.params.maximumIterates if maxIters is None else maxIters
        rl = mi * (mInner if m in (GMRES, PCGMRES) else 1) + 1
        self.resi = np.zeros(rl)
        io, it, ms = C.c_int(), C.c_int(), C.c_double()
        pc = _PRECOND[precond]
        L.check(L.lib().self_cgsem2d_solve(self._h, m, pc, tol, mi, mInner, ortho, C.byref(io), C.byref(it),
                                           L.ptr(self.resi), rl, C.byref(ms)))
        return io.value, it.value, ms.value

    def SetPreconditioner(self, pcIterates=50, pcTolerance=1.0e-4, coefficient="one"):
        """CGsemPreconditioner_2D % Build (CGsemPreconditioner_2D_Class.f90:88-126):
        the operator at N = 1 on the linear mesh through the Gauss-Lobatto
        nodes, inverted by pcIterates CG iterations to pcTolerance.
        coefficient "one" is what the reference maps at Build (fluxCoeff =
        ONE there); "flux" uses the coefficient set by SetFluxCoefficient.
        Select it with precond="fem" in Solve and SolveGMRESMixed."""
        x = L.f64(self.mesh.x)
        y = L.f64(self.mesh.y)
        L.check(L.lib().self_cgsem2d_set_preconditioner(self._h, L.ptr(x), L.ptr(y), int(pcIterates),
                                                        float(pcTolerance), {"one": 0, "flux": 1}[coefficient]))

    def ApplyPreconditioner(self, b):
        """CGsemPreconditioner_2D % Solve: returns (x, inner iterations)."""
        out = self._new()
        it = C.c_int()
        L.check(L.lib().self_cgsem2d_apply_preconditioner(self._h, L.ptr(self._arr(b)), L.ptr(out), C.byref(it)))
        return out, it.value

    def SolveGMRESMixed(self, tolerance=1e-10, maxCycles=50, mInner=30, precond="fem", ortho=MGS):
        """Restarted FGMRES(mInner) with FP32 operator and preconditioner
        applications and FP64 basis, orthogonalisation and restart
        residuals: (ioerr, iterations, relResidual, ms)."""
        io, it, rel, ms = C.c_int(), C.c_int(), C.c_double(), C.c_double()
        L.check(L.lib().self_cgsem2d_solve_gmres_mixed(self._h, _PRECOND[precond], tolerance, maxCycles, mInner,
                                                       ortho, C.byref(io), C.byref(it), C.byref(rel), C.byref(ms)))
        return io.value, it.value, rel.value, ms.value

    def SolveMixed(self, tolerance=1e-10, innerTolerance=1e-5, maxInner=20000, maxOuter=20, precond="jacobi"):
        """FP32 inner PCG + FP64 iterative refinement: (ioerr, outer, inner, relResidual, ms)."""
        io, outer, inner, rel, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double(), C.c_double()
        pc = {"none": 0, "jacobi": 1}[precond]
        L.check(L.lib().self_cgsem2d_solve_mixed(self._h, pc, tolerance, innerTolerance, maxInner, maxOuter,
                                                 C.byref(io), C.byref(outer), C.byref(inner), C.byref(rel),
                                                 C.byref(ms)))
        return io.value, outer.value, inner.value, rel.value, ms.value

    def SolveReliable(self, tolerance=1e-10, delta=0.2, maxIters=40000, maxUpdates=400, precond="jacobi",
                      inner="fp32"):
        """PCG with reliable updates (one never-restarted sequence, FP64
        solution and residual refreshed whenever the recursive residual has
        dropped by delta): inner "fp32" runs the sequence in FP32, "fp64" in
        FP64 (then the solve stops on the true residual).  Returns (ioerr,
        updates, iterations, relResidual, ms)."""
        io, up, it, rel, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double(), C.c_double()
        pc = {"none": 0, "jacobi": 1}[precond]
        fn = {"fp32": L.lib().self_cgsem2d_solve_reliable, "fp64": L.lib().self_cgsem2d_solve_reliable_fp64}[inner]
        L.check(fn(self._h, pc, tolerance, delta, maxIters, maxUpdates, C.byref(io), C.byref(up), C.byref(it),
                   C.byref(rel), C.byref(ms)))
        return io.value, up.value, it.value, rel.value, ms.value

    def Trash(self):
        if getattr(self, "_h", None):
            L.check(L.lib().self_cgsem2d_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.Trash()
        except Exception:
            pass


def SolveBatched(models, precond="jacobi", tolerance=1e-10, maxIters=20000):
    """The CGSEM PCG (CGsemElliptic_2D_Class.f90:1114-1207) for many
    CGsemElliptic2D models at once (any mesh sizes and degrees): every CG
    iteration is two launches for all models still iterating.  Each model's
    solution, ioerr and iteration count are bitwise those of its own
    Solve("cgsem_pcg", precond, tolerance, maxIters).  Returns
    (ioerr list, iterations list, elapsed_ms of the whole batch)."""
    n = len(models)
    if n == 0:
        return [], [], 0.0
    hs = (C.c_void_p * n)(*[m._h for m in models])
    io = (C.c_int * n)()
    it = (C.c_int * n)()
    ms = C.c_double()
    pc = {"none": 0, "jacobi": 1}[precond]
    L.check(L.lib().self_cgsem2d_solve_batched(hs, n, pc, tolerance, maxIters, io, it, C.byref(ms)))
    return list(io), list(it), ms.value
