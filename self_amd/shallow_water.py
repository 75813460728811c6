"""Conservative 2-D shallow water on the GPU: the ShallowWater type of
src/highend/shallowwater/ConservativeShallowWaterClass.f90 behind the C ABI.

Method names are the reference's type-bound procedures (Build, Trash,
ForwardStepRK3, GlobalTimeDerivative, CalculateSolutionAtBoundaries,
EdgeFlux, MappedTimeDerivative, CalculateBathymetryGradient, Get*/Set*).
Field arrays are numpy arrays in the reference's per-element layout with
the element index first: solution (nElems, 3, N+1, N+1) indexed
[e, field, j(p), i(s)], i.e. sol(iEl)%solution(i, j, field) stacked.
Device state stays resident between calls; Get* copies back.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import os

import numpy as np

from . import _lib as L
from .mesh import QuadMesh
from .nodal import NodalStorage2D


@dataclass
class SWParams:
    """Subset of SWParams (src/highend/shallowwater/SWParamsClass.f90:41-82)
    used by the conservative model's hot path, plus the mesh periodicity
    this implementation adds."""
    polyDeg: int = 7
    nXElem: int = 4
    nYElem: int = 4
    xScale: float = 1.0
    yScale: float = 1.0
    dt: float = 1.0
    g: float = 9.81
    f0: float = 0.0
    betaX: float = 0.0
    betaY: float = 0.0
    linearDrag: float = 0.0
    periodicX: bool = True
    periodicY: bool = True
    warp: float = 0.0
    # SpaceManagement % SpecMeshFile: an ISM-V2 mesh file read (and scaled by
    # xScale, yScale) instead of the structured nXElem x nYElem mesh
    SpecMeshFile: str = ""


_FIELDS = {"solution": (0, 3), "tendency": (1, 3), "bathymetry": (2, 3), "vorticity": (3, 1),
           "relax": (4, 4), "boundarySolution": (5, None), "boundaryFlux": (6, None)}


class ShallowWater:
    """Build_ShallowWater (:163-245) with the nodal storage, mesh and fields
    held on the device.

    precision: "fp64" or "mixed"; form: "fused" (one kernel per RK stage) or
    "plain" (one kernel per reference phase).
    mixedGuard: for precision "mixed", the first ForwardStepRK3 of a single
    state first runs CheckMixed(mixedGuard): where the mixed stage's momentum
    tendency differs from the FP64 one by more than this (relative L2), the
    model continues in FP64.  0 or None: no check.
    """

    MIXED_GUARD = 2e-3

    def __init__(self, params: SWParams, precision: str = "fp64", form: str = "fused",
                 quirk_dmatp_transposed: bool = False, mesh: QuadMesh | None = None,
                 nodal: NodalStorage2D | None = None, mixedGuard: float | None = MIXED_GUARD):
        self.params = params
        self.mixedGuard = mixedGuard if precision == "mixed" else None
        self.mixedGuardResult = None  # (relErr, switched) once the check has run
        self.members = 1
        N = params.polyDeg
        self.nS = self.nP = N
        self.nEq = 3
        self.dgStorage = nodal or NodalStorage2D(N, L.GAUSS, L.DG, quirk_dmatp_transposed)
        if mesh is None and params.SpecMeshFile:
            # BuildQuadMesh (ConservativeShallowWaterClass.f90:289-304): read, then scale
            from .mesh import SpecMesh
            mesh = SpecMesh(params.SpecMeshFile, N, self.dgStorage.nodes).ScaleTheMesh(params.xScale,
                                                                                       params.yScale)
        self.mesh = mesh or QuadMesh(params.nXElem, params.nYElem, N, self.dgStorage.nodes,
                                     params.periodicX, params.periodicY, params.xScale, params.yScale,
                                     params.warp)
        self.nElems = self.mesh.nElems
        prec = {"fp64": L.FP64, "mixed": L.MIXED}[precision]
        frm = {"plain": L.PLAIN, "fused": L.FUSED}[form]
        self.precision, self.form = precision, form
        desc = L.SWDesc(N, self.mesh.nElems, self.mesh.nEdges, prec, frm, params.g, params.dt, params.linearDrag)
        h = C.c_void_p()
        L.check(L.lib().self_sw2d_create(C.byref(desc), C.byref(h)))
        self._h = h
        ns = self.dgStorage
        L.check(L.lib().self_sw2d_set_nodal(self._h, L.ptr(ns.flat("dMatS")), L.ptr(ns.flat("dMatP")),
                                            L.ptr(ns.qWeight), L.ptr(ns.qWeight), L.ptr(ns.lagWest),
                                            L.ptr(ns.lagEast), L.ptr(ns.lagSouth), L.ptr(ns.lagNorth)))
        m = self.mesh
        L.check(L.lib().self_sw2d_set_geometry(self._h, L.ptr(m.edges, L.ip), L.ptr(m.dxds), L.ptr(m.dxdp),
                                               L.ptr(m.dyds), L.ptr(m.dydp), L.ptr(m.J), L.ptr(m.nHat),
                                               L.ptr(m.nLen)))
        # Build_ShallowWater defaults: depth hDefault = 1, zero gradient (:187-190);
        # planetary vorticity f0 + betaX x + betaY y
        n1 = N + 1
        bathy = np.zeros((self.nElems, 3, n1, n1))
        bathy[:, 0] = 1.0
        self.SetBathymetry(bathy)
        self.SetPlanetaryVorticity(params.f0 + params.betaX * m.x + params.betaY * m.y)

    # ------------------------------------------------------------- lifetime
    def Trash(self):
        if getattr(self, "_h", None):
            L.check(L.lib().self_sw2d_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.Trash()
        except Exception:
            pass

    # ------------------------------------------------------------- fields
    def _shape(self, nf):
        n1 = self.nS + 1
        return (self.nElems, nf, n1, n1) if nf > 1 else (self.nElems, n1, n1)

    def _set(self, field, arr, nf):
        a = L.f64(arr)
        if a.shape != self._shape(nf):
            raise ValueError(f"expected shape {self._shape(nf)}, got {a.shape}")
        L.check(L.lib().self_sw2d_set_field(self._h, field, L.ptr(a)))

    def _get(self, field, nf):
        out = np.zeros(self._shape(nf))
        L.check(L.lib().self_sw2d_get_field(self._h, field, L.ptr(out)))
        return out

    def SetSolution(self, sol):
        self._set(0, sol, 3)

    def GetSolution(self):
        return self._get(0, 3)

    def GetTendency(self):
        return self._get(1, 3)

    def SetBathymetry(self, b):
        self._set(2, b, 3)

    def GetBathymetry(self):
        return self._get(2, 3)

    def SetPlanetaryVorticity(self, f):
        self._set(3, f, 1)

    def GetPlanetaryVorticity(self):
        return self._get(3, 1)

    def SetRelaxationField(self, r):
        """relax vars 1..3 (U_r, V_r, H_r) and var 4 (1/timescale), (nElems, 4, N+1, N+1)."""
        self._set(4, r, 4)

    def GetRelaxationField(self):
        return self._get(4, 4)

    def GetBoundarySolution(self):
        """(nElems, 4 sides, 3 fields, N+1) -- sol%boundarySolution(0:N,1:3,1:4)."""
        n1 = self.nS + 1
        out = np.zeros((self.nElems, 4, 3, n1))
        L.check(L.lib().self_sw2d_get_field(self._h, 5, L.ptr(out)))
        return out

    def GetBoundaryFlux(self):
        n1 = self.nS + 1
        out = np.zeros((self.nElems, 4, 3, n1))
        L.check(L.lib().self_sw2d_get_field(self._h, 6, L.ptr(out)))
        return out

    def SetTimeStep(self, dt):
        self.params.dt = dt
        L.check(L.lib().self_sw2d_set_params(self._h, self.params.g, dt, self.params.linearDrag))

    # ------------------------------------------------------------- hot path
    def CalculateSolutionAtBoundaries(self):
        L.check(L.lib().self_sw2d_calculate_solution_at_boundaries(self._h))

    def EdgeFlux(self, t=0.0):
        L.check(L.lib().self_sw2d_edge_flux(self._h, t))

    def MappedTimeDerivative(self, t=0.0):
        L.check(L.lib().self_sw2d_mapped_time_derivative(self._h, t))

    def GlobalTimeDerivative(self, t=0.0):
        L.check(L.lib().self_sw2d_global_time_derivative(self._h, t))

    def CheckMixed(self, tolerance=1e-5):
        """Mixed form only: relative L2 difference of the momentum tendency of
        the current state between the mixed stage and the FP64 path; above
        tolerance the model switches to FP64.  Returns (relErr, switched)."""
        e, s = C.c_double(), C.c_int()
        L.check(L.lib().self_sw2d_check_mixed(self._h, tolerance, C.byref(e), C.byref(s)))
        if s.value:
            self.precision = "fp64"
        return e.value, bool(s.value)

    def _guard(self):
        if self.mixedGuard and self.mixedGuardResult is None and self.members == 1:
            self.mixedGuardResult = self.CheckMixed(self.mixedGuard)

    def ForwardStepRK3(self, tn=0.0, nSteps=1):
        self._guard()
        L.check(L.lib().self_sw2d_forward_step_rk3(self._h, tn, nSteps))

    def ForwardStepRK3Timed(self, tn=0.0, nSteps=1):
        """Returns device milliseconds for nSteps steps (HIP events on the model's stream)."""
        self._guard()
        ms = C.c_double()
        L.check(L.lib().self_sw2d_forward_step_rk3_timed(self._h, tn, nSteps, C.byref(ms)))
        return ms.value

    def KernelTimes(self, tn=0.0, nSteps=10):
        """Mean device ms per launch of each stage kernel (fused: [stage];
        plain: [traces, edge flux, mapped tendency, RK update])."""
        out = np.zeros(4)
        nk = C.c_int()
        L.check(L.lib().self_sw2d_kernel_times(self._h, tn, nSteps, L.ptr(out), C.byref(nk)))
        return out[:nk.value].tolist()

    def CalculateBathymetryGradient(self):
        L.check(L.lib().self_sw2d_calculate_bathymetry_gradient(self._h))

    def SetVariant(self, variant):
        """0: the src class; 1: the class as the examples ship it
        (examples/shallowwater/planewave/localmods/ConservativeShallowWaterClass.f90:
        its wave speed, a Source without drag and relaxation, grad(h^2/2)/h)."""
        L.check(L.lib().self_sw2d_set_variant(self._h, int(variant)))

    def SetGroups(self, groups):
        """Element groups per workgroup of the fused kernel (0 = heuristic)."""
        L.check(L.lib().self_sw2d_set_groups(self._h, groups))

    def SetGraph(self, enable=True):
        """Fused form: replay pairs of RK3 steps from one captured HIP graph
        (default) or launch every stage from the host."""
        L.check(L.lib().self_sw2d_set_graph(self._h, int(bool(enable))))

    def WritePickup(self, iter, directory="."):
        """WritePickup (ConservativeShallowWaterClass.f90:2013-2065): state,
        bathymetry and vorticity to <directory>/ShallowWater.<iter>.pickup."""
        from .pickup import filename
        path = os.path.join(directory, filename("ShallowWater", iter))
        L.check(L.lib().self_sw2d_write_pickup(self._h, path.encode()))
        return path

    def ReadPickup(self, iter, directory="."):
        """ReadPickup (:2069-2125) and the bathymetry gradient Build computes
        after it (:224-228)."""
        from .pickup import filename
        path = os.path.join(directory, filename("ShallowWater", iter))
        L.check(L.lib().self_sw2d_read_pickup(self._h, path.encode()))
        return path

    def SetMembers(self, members):
        """Ensemble of `members` states advanced together (fused form): one
        launch per RK stage for all of them.  States are reset to zero;
        Set/GetSolution address member 0."""
        L.check(L.lib().self_sw2d_set_members(self._h, int(members)))
        self.members = int(members)

    def SetMemberSolution(self, member, sol):
        a = L.f64(sol)
        if a.size != self.nElems * 3 * (self.nS + 1) ** 2:
            raise ValueError("expected (nElems, 3, N+1, N+1)")
        L.check(L.lib().self_sw2d_set_member_solution(self._h, int(member), L.ptr(a)))

    def SetMemberTimeStep(self, member, dt):
        L.check(L.lib().self_sw2d_set_member_time_step(self._h, int(member), float(dt)))

    def GetMemberSolution(self, member):
        out = np.zeros(self._shape(3))
        L.check(L.lib().self_sw2d_get_member_solution(self._h, int(member), L.ptr(out)))
        return out

    def Autotune(self, nSteps=50):
        """Element groups per workgroup x graph replay: the process-wide
        table's choice for this degree, arithmetic and size class, or the
        fastest setting timed on a copy of the state (then recorded).
        Returns (groups, ms_per_step)."""
        g, ms = C.c_int(), C.c_double()
        L.check(L.lib().self_sw2d_autotune(self._h, nSteps, C.byref(g), C.byref(ms)))
        return g.value, ms.value

    def Synchronize(self):
        L.check(L.lib().self_sw2d_synchronize(self._h))

    def DeviceBytes(self):
        b = C.c_size_t()
        L.check(L.lib().self_sw2d_device_bytes(self._h, C.byref(b)))
        return b.value


def forward_step_rk3_multi(models, nSteps=1, timed=False):
    """ForwardStepRK3 nSteps times for several fused-form ShallowWater models
    of any degrees (one precision), each RK stage of all of them in one
    persistent launch.  Returns the device milliseconds when timed."""
    import ctypes as C
    arr = (C.c_void_p * len(models))(*[m._h for m in models])
    ms = C.c_double()
    L.check(L.lib().self_sw2d_forward_step_rk3_multi(arr, len(models), nSteps, C.byref(ms) if timed else None))
    return ms.value if timed else None


def tuning_table():
    """The process-wide autotuning table: a list of dicts (kind, dim, N,
    dtype, sizeClass, config, ms)."""
    n = C.c_int()
    L.check(L.lib().self_tune_count(C.byref(n)))
    out = []
    for i in range(n.value):
        kind = C.create_string_buffer(64)
        key = (C.c_int * 4)()
        cfg = (C.c_int * 4)()
        ms = C.c_double()
        L.check(L.lib().self_tune_entry(i, kind, 64, key, cfg, C.byref(ms)))
        out.append({"kind": kind.value.decode(), "dim": key[0], "N": key[1], "dtype": key[2], "sizeClass": key[3],
                    "config": list(cfg), "ms": ms.value})
    return out


def save_tuning_table(path):
    L.check(L.lib().self_tune_save(os.fsencode(path)))


def load_tuning_table(path):
    L.check(L.lib().self_tune_load(os.fsencode(path)))


def clear_tuning_table():
    L.check(L.lib().self_tune_clear())
