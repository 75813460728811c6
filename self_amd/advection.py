"""2-D advection-diffusion on the GPU: the Advection type of
src/highend/advection-diffusion2d/AdvectionClass.f90 behind the C ABI.

Method names follow the reference's type-bound procedures (Build, Trash,
SetSolution, SetVelocity, SetRelaxationField, SetRelaxationFactor,
GlobalTimeDerivative, ForwardStepRK3, GetTendency, GetGradient).  Arrays are
numpy arrays in the reference's per-element layout with the element index
first: solution (nElems, nTracers, N+1, N+1) indexed [e, tracer, j(p), i(s)];
velocity (nElems, 2, N+1, N+1) with u, v in slots 0, 1.
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
class AdvectionParams:
    """Subset of AdvectionParams (src/highend/advection-diffusion2d/
    AdvectionParamsClass.f90) used by the hot path, plus mesh periodicity."""
    polyDeg: int = 7
    nXElem: int = 4
    nYElem: int = 4
    nTracers: int = 1
    xScale: float = 1.0
    yScale: float = 1.0
    dt: float = 1e-4
    kappa: float = 0.0
    periodicX: bool = True
    periodicY: bool = True
    warp: float = 0.0
    # SpaceManagement % SpecMeshFile: an ISM-V2 mesh file read (and scaled by
    # xScale, yScale) instead of the structured nXElem x nYElem mesh
    SpecMeshFile: str = ""


class Advection:
    """Build_Advection (:176-252): nodal storage, mesh and state on the device.

    edges: optional (nEdges, 6) edge table replacing the mesh's (for example
    to flag boundary edges SELF_INFLOW_ONE / SELF_INFLOW_TWO).
    """

    def __init__(self, params: AdvectionParams, quirk_dmatp_transposed: bool = False,
                 mesh: QuadMesh | None = None, nodal: NodalStorage2D | None = None, edges=None):
        self.params = params
        N = params.polyDeg
        self.nS = self.nP = N
        self.nEq = params.nTracers
        self.dgStorage = nodal or NodalStorage2D(N, L.GAUSS, L.DG, quirk_dmatp_transposed)
        if mesh is None and params.SpecMeshFile:
            # BuildQuadMesh (AdvectionClass.f90:303-330): read, then scale
            from .mesh import SpecMesh
            mesh = SpecMesh(params.SpecMeshFile, N, self.dgStorage.nodes).ScaleTheMesh(params.xScale,
                                                                                       params.yScale)
        self.mesh = mesh or QuadMesh(params.nXElem, params.nYElem, N, self.dgStorage.nodes, params.periodicX,
                                     params.periodicY, params.xScale, params.yScale, params.warp)
        self.nElems = self.mesh.nElems
        m, ns = self.mesh, self.dgStorage
        self.edges = L.i32(m.edges if edges is None else edges)
        h = C.c_void_p()
        L.check(L.lib().self_adv2d_create(
            N, params.nTracers, m.nElems, m.nEdges, L.ptr(self.edges, L.ip),
            *[L.ptr(getattr(m, n)) for n in ("dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen")],
            L.ptr(L.f64(ns.dMatS.ravel(order="F"))), L.ptr(L.f64(ns.dMatP.ravel(order="F"))), L.ptr(ns.qWeight),
            L.ptr(ns.lagWest), L.ptr(ns.lagEast), L.ptr(ns.lagSouth), L.ptr(ns.lagNorth),
            float(params.kappa), float(params.dt), C.byref(h)))
        self._h = h

    def Trash(self):
        if getattr(self, "_h", None):
            L.check(L.lib().self_adv2d_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.Trash()
        except Exception:
            pass

    def _shape(self, nf):
        n1 = self.nS + 1
        return (self.nElems, nf, n1, n1)

    def _set(self, field, arr, nf):
        a = L.f64(arr)
        if a.shape != self._shape(nf):
            raise ValueError(f"expected shape {self._shape(nf)}, got {a.shape}")
        L.check(L.lib().self_adv2d_set_field(self._h, field, L.ptr(a)))

    def _get(self, field):
        out = np.zeros(self._shape(self.nEq))
        L.check(L.lib().self_adv2d_get_field(self._h, field, L.ptr(out)))
        return out

    def SetSolution(self, c):
        self._set(0, c, self.nEq)

    def GetSolution(self):
        return self._get(0)

    def SetVelocity(self, u, v):
        """u, v: (nElems, N+1, N+1)."""
        self._set(1, np.stack([L.f64(u), L.f64(v)], axis=1), 2)

    def SetRelaxationField(self, r):
        self._set(2, r, self.nEq)

    def SetRelaxationFactor(self, f):
        self._set(3, f, self.nEq)

    def GetVelocity(self):
        """(u, v), each (nElems, N+1, N+1)."""
        out = np.zeros(self._shape(2))
        L.check(L.lib().self_adv2d_get_field(self._h, 4, L.ptr(out)))
        return out[:, 0].copy(), out[:, 1].copy()

    def GetRelaxationField(self):
        return self._get(5)

    def GetRelaxationFactor(self):
        return self._get(6)

    def WritePickup(self, iter, directory="."):
        """WritePickup (AdvectionClass.f90:2040-2103) to
        <directory>/Advection.<iter>.pickup."""
        from .pickup import filename
        path = os.path.join(directory, filename("Advection", iter))
        L.check(L.lib().self_adv2d_write_pickup(self._h, path.encode()))
        return path

    def ReadPickup(self, iter, directory="."):
        """ReadPickup (:2107-2171): solution, relaxation field and factor, velocity."""
        from .pickup import filename
        path = os.path.join(directory, filename("Advection", iter))
        L.check(L.lib().self_adv2d_read_pickup(self._h, path.encode()))
        return path

    def GetTendency(self):
        return self._get(1)

    def GetGradient(self):
        """(dc/dx, dc/dy) from the last CalculateConcentrationGradient."""
        return self._get(2), self._get(3)

    def GlobalTimeDerivative(self, t=0.0):
        L.check(L.lib().self_adv2d_global_time_derivative(self._h, t))

    def ForwardStepRK3(self, tn=0.0, nSteps=1):
        L.check(L.lib().self_adv2d_forward_step_rk3(self._h, tn, nSteps, None))

    def ForwardStepRK3Timed(self, tn=0.0, nSteps=1):
        ms = C.c_double()
        L.check(L.lib().self_adv2d_forward_step_rk3(self._h, tn, nSteps, C.byref(ms)))
        return ms.value

    def Synchronize(self):
        L.check(L.lib().self_adv2d_synchronize(self._h))
