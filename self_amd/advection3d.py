"""3-D linear advection on the GPU: the Advection type of
src/highend/advection3d/Advection3DClass.f90 behind the C ABI.

Method names follow the reference (Build, Trash, SetSolution, SetVelocity,
SetRelaxationField, SetRelaxationFactor, GlobalTimeDerivative(t, m),
ForwardStepRK3, GetTendency).  Arrays have the element index first:
solution (nElems, nTracers, N+1, N+1, N+1) indexed [e, tracer, k(q), j(p), i(s)],
velocity components (nElems, N+1, N+1, N+1).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .mesh import HexMesh
from .nodal import NodalStorage3D, rolloff_filter1d


@dataclass
class Advection3DParams:
    """Subset of AdvectionParams (src/highend/advection3d/Advection3DParamsClass.f90),
    whose defaults are nTracers = 2, polyDeg = 5, nCutoff = 3, plus mesh
    periodicity and warp."""
    polyDeg: int = 5
    nXElem: int = 2
    nYElem: int = 2
    nZElem: int = 2
    nTracers: int = 2
    nCutoff: int = 3
    dt: float = 1e-3
    xScale: float = 1.0
    yScale: float = 1.0
    zScale: float = 1.0
    periodic: tuple = (True, True, True)
    warp: float = 0.0
    warpSign: tuple = (1, 1, 1)
    useFilter: bool = True


class Advection3D:
    """Build_Advection (:167-272) with the hex mesh, nodal storage, filter
    and state on the device.  faces: optional (nFaces, 9) face table
    replacing the mesh's (for boundary flags).

    quirk_relaxation_alias: the reference's SetRelaxationFactor and
    GetRelaxationFactorAtNode (:915-945) use the relaxation-field storage,
    so field and factor are one array and the last Set wins.  False (the
    default) keeps them separate; True reproduces the reference (each Set
    writes both device arrays)."""

    def __init__(self, params: Advection3DParams, mesh: HexMesh | None = None, faces=None,
                 quirk_relaxation_alias: bool = False):
        self.params = params
        self.quirk_relaxation_alias = quirk_relaxation_alias
        N = params.polyDeg
        self.nS = self.nP = self.nQ = N
        self.nEq = params.nTracers
        self.dgStorage = NodalStorage3D(N)
        ns = self.dgStorage
        self.mesh = mesh or HexMesh(params.nXElem, params.nYElem, params.nZElem, N, ns.nodes, params.periodic,
                                    (params.xScale, params.yScale, params.zScale), params.warp, params.warpSign)
        m = self.mesh
        self.nElems = m.nElems
        self.faces = L.i32(m.faces if faces is None else faces)
        self.F1 = rolloff_filter1d(N, params.nCutoff, ns.nodes, ns.qWeight)
        h = C.c_void_p()
        L.check(L.lib().self_adv3d_create(
            N, params.nTracers, m.nElems, m.nFaces, L.ptr(self.faces, L.ip), L.ptr(m.cov), L.ptr(m.J),
            L.ptr(m.nHat), L.ptr(m.nLen), L.ptr(m.bpos), L.ptr(L.f64(ns.dMat.ravel(order="F"))), L.ptr(ns.qWeight),
            L.ptr(ns.lagLeft), L.ptr(ns.lagRight), L.ptr(L.f64(self.F1.ravel(order="F"))), int(params.useFilter),
            float(params.dt), C.byref(h)))
        self._h = h

    def Trash(self):
        if getattr(self, "_h", None):
            L.check(L.lib().self_adv3d_destroy(self._h))
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
        L.check(L.lib().self_adv3d_set_field(self._h, field, L.ptr(a)))

    def _get(self, field):
        out = np.zeros(self._shape(self.nEq))
        L.check(L.lib().self_adv3d_get_field(self._h, field, L.ptr(out)))
        return out

    def SetSolution(self, c):
        self._set(0, c, self.nEq)

    def GetSolution(self):
        return self._get(0)

    def GetTendency(self):
        return self._get(1)

    def SetVelocity(self, u, v, w):
        self._set(1, np.stack([L.f64(u), L.f64(v), L.f64(w)], axis=1), 3)

    def SetRelaxationField(self, r):
        self._set(2, r, self.nEq)
        if self.quirk_relaxation_alias:
            self._set(3, r, self.nEq)

    def SetRelaxationFactor(self, f):
        self._set(3, f, self.nEq)
        if self.quirk_relaxation_alias:
            self._set(2, f, self.nEq)

    def GetFilterState(self):
        """(E1, E2, lim): E1, E2 (nTracers, nElems); lim (3, nTracers, nElems)."""
        nq, E = self.nEq, self.nElems
        out = np.zeros(5 * nq * E)
        L.check(L.lib().self_adv3d_get_field(self._h, 2, L.ptr(out)))
        return out[:nq * E].reshape(nq, E), out[nq * E:2 * nq * E].reshape(nq, E), out[2 * nq * E:].reshape(3, nq, E)

    def FilterCount(self, reset=True):
        n = C.c_uint()
        L.check(L.lib().self_adv3d_filter_count(self._h, int(reset), C.byref(n)))
        return n.value

    def GlobalTimeDerivative(self, t=0.0, m=1):
        L.check(L.lib().self_adv3d_global_time_derivative(self._h, t, m))

    def ForwardStepRK3(self, tn=0.0, nSteps=1):
        L.check(L.lib().self_adv3d_forward_step_rk3(self._h, tn, nSteps, None))

    def ForwardStepRK3Timed(self, tn=0.0, nSteps=1):
        ms = C.c_double()
        L.check(L.lib().self_adv3d_forward_step_rk3(self._h, tn, nSteps, C.byref(ms)))
        return ms.value

    def SetPrecision(self, precision="fp64"):
        """Arithmetic of the volume contractions: "fp64" (the reference's),
        "fp32" (matrix-core FP32) or "bf16x3" (matrix-core split BF16)."""
        L.check(L.lib().self_adv3d_set_precision(self._h, {"fp64": 0, "fp32": 1, "bf16x3": 2}[precision]))

    def Synchronize(self):
        L.check(L.lib().self_adv3d_synchronize(self._h))
