"""The hydrostatic primitive equations on the GPU: the HydrostaticPrimitive
type of src/highend/hydrostaticprimitives/HydrostaticPrimitivesClass.f90
behind the C ABI.

Method names follow the reference (SetSolution, SetBackgroundState,
SetPreviousTendency, GetTendency, GlobalTimeDerivative, ForwardStepAB2,
ForwardStepRK3 of the implicit variant).  Arrays have the element index
first: solution and background (nElems, 5, N+1, N+1, N+1) indexed
[e, var, k(q), j(p), i(s)] with variables u, v, b, w, p; tendency
(nElems, 3, ...).  Element e = hE + nHElems * layer, layer 0 at the sea
floor.  The mesh spans [0, xScale] x [0, yScale] x [0, zScale]; only
differences of z enter the equations (Source is zero), so z = 0 stands for
the reference's sea floor at -H.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .mesh import HexMesh
from .nodal import NodalStorage3D


@dataclass
class HydrostaticParams:
    """Subset of HydrostaticParams (HydrostaticParams_Class.f90:40-66):
    polyDeg, nXElem, nYElem, nZElem (the layers), dt, x/y/zScale, plus the
    lateral periodicity and the warp of this build's structured mesh and
    the diagnoseP switch."""
    polyDeg: int = 5
    nXElem: int = 4
    nYElem: int = 4
    nZElem: int = 4
    dt: float = 1e-3
    xScale: float = 1.0
    yScale: float = 1.0
    zScale: float = 1.0
    periodic: tuple = (False, False)
    warp: float = 0.0
    warpSign: tuple = (1, 1, 1)
    diagnoseP: bool = True


class HydrostaticPrimitive:
    """Build (:219-317) with the stacked hex mesh, nodal storage and state on the device."""

    def __init__(self, params: HydrostaticParams):
        self.params = params
        N = params.polyDeg
        self.nS = self.nP = self.nQ = N
        self.dgStorage = NodalStorage3D(N)
        ns = self.dgStorage
        self.mesh = HexMesh(params.nXElem, params.nYElem, params.nZElem, N, ns.nodes,
                            (params.periodic[0], params.periodic[1], False),
                            (params.xScale, params.yScale, params.zScale), params.warp, params.warpSign)
        m = self.mesh
        self.nElems = m.nElems
        self.nHElems = params.nXElem * params.nYElem
        self.nLayers = params.nZElem
        h = C.c_void_p()
        L.check(L.lib().self_hpe_create(
            N, self.nHElems, self.nLayers, m.nFaces, L.ptr(L.i32(m.faces), L.ip), L.ptr(m.cov), L.ptr(m.J),
            L.ptr(m.nHat), L.ptr(m.nLen), L.ptr(m.bpos), L.ptr(ns.nodes), L.ptr(ns.qWeight),
            L.ptr(L.f64(ns.dMat.ravel(order="F"))), L.ptr(ns.lagLeft), L.ptr(ns.lagRight), int(params.diagnoseP),
            float(params.dt), C.byref(h)))
        self._h = h

    def Trash(self):
        if getattr(self, "_h", None):
            L.check(L.lib().self_hpe_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.Trash()
        except Exception:
            pass

    def _shape(self, nf):
        n1 = self.nS + 1
        return (self.nElems, nf, n1, n1, n1)

    def _set(self, field, arr, nf):
        a = L.f64(arr)
        if a.shape != self._shape(nf):
            raise ValueError(f"expected shape {self._shape(nf)}, got {a.shape}")
        L.check(L.lib().self_hpe_set_field(self._h, field, L.ptr(a)))

    def _get(self, field, nf):
        out = np.zeros(self._shape(nf))
        L.check(L.lib().self_hpe_get_field(self._h, field, L.ptr(out)))
        return out

    def SetSolution(self, sol):
        self._set(0, sol, 5)

    def GetSolution(self):
        return self._get(0, 5)

    def SetBackgroundState(self, bg):
        self._set(1, bg, 5)

    def SetPreviousTendency(self, g):
        self._set(2, g, 3)

    def GetTendency(self):
        return self._get(1, 3)

    def GetDivU(self):
        return self._get(2, 1)[:, 0]

    def GetFaceValues(self):
        """Face values of b, w, p of the last time derivative, (nElems, 6, 3, N+1, N+1)."""
        n1 = self.nS + 1
        out = np.zeros((self.nElems, 6, 3, n1, n1))
        L.check(L.lib().self_hpe_get_field(self._h, 3, L.ptr(out)))
        return out

    def GlobalTimeDerivative(self):
        L.check(L.lib().self_hpe_global_time_derivative(self._h))

    def ForwardStepRK3(self, nSteps=1):
        L.check(L.lib().self_hpe_forward_step(self._h, 0, nSteps, None))

    def ForwardStepAB2(self, nSteps=1):
        L.check(L.lib().self_hpe_forward_step(self._h, 1, nSteps, None))

    def ForwardStepRK3Timed(self, nSteps=1):
        ms = C.c_double()
        L.check(L.lib().self_hpe_forward_step(self._h, 0, nSteps, C.byref(ms)))
        return ms.value

    def KernelTimes(self, nSteps=2):
        """Average ms of (column kernel, stage kernel) per RK3 stage."""
        out = np.zeros(2)
        L.check(L.lib().self_hpe_kernel_times(self._h, nSteps, L.ptr(out)))
        return float(out[0]), float(out[1])

    def Synchronize(self):
        L.check(L.lib().self_hpe_synchronize(self._h))
