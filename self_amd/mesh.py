"""Structured quad meshes (host set-up) over the C ABI.

Mirrors QuadMesh%LoadDefaultMesh / ConstructEdges (src/geom/QuadMeshClass.f90
:1489-1627, :1718-1859) and MappedGeometry_2D metrics
(src/geom/MappedGeometryClass_2D.f90:611-710), with optional periodic pairing.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L


class QuadMesh:
    """Arrays per element (element slowest):
    x, y, dxds, dxdp, dyds, dydp, J : (nElems, N+1, N+1) indexed [e, j, i]
    nHat : (nElems, 4, N+1, 2) [e, side, node, component]; nLen (nElems, 4, N+1)
    edges: (nEdges, 6) int32 = e1, s1, e2, s2, start, inc (1-based, reference convention)
    """

    def __init__(self, nXElem: int, nYElem: int, N: int, sNodes, periodicX=True, periodicY=True,
                 xScale=1.0, yScale=1.0, warp=0.0):
        lib = L.lib()
        import ctypes as C
        nE, nEd = C.c_int(), C.c_int()
        L.check(lib.self_quadmesh_counts(nXElem, nYElem, int(periodicX), int(periodicY), C.byref(nE), C.byref(nEd)))
        self.nX, self.nY, self.N = nXElem, nYElem, N
        self.nElems, self.nEdges = nE.value, nEd.value
        self.periodicX, self.periodicY = bool(periodicX), bool(periodicY)
        self.xScale, self.yScale, self.warp = xScale, yScale, warp
        n1 = N + 1
        E = self.nElems
        self.edges = np.zeros((self.nEdges, 6), dtype=np.int32)
        shp = (E, n1, n1)
        for name in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J"):
            setattr(self, name, np.zeros(shp))
        self.nHat = np.zeros((E, 4, n1, 2))
        self.nLen = np.zeros((E, 4, n1))
        L.check(lib.self_quadmesh_structured(
            nXElem, nYElem, N, int(periodicX), int(periodicY), xScale, yScale, warp,
            L.ptr(L.f64(sNodes)), L.ptr(self.edges, L.ip),
            *[L.ptr(getattr(self, n)) for n in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen")]))


class SpecMesh:
    """An ISM-V2 SpecMesh file (QuadMesh%ReadSpecMeshFile, src/geom/
    QuadMeshClass.f90:2100-2465): unstructured quads with curved boundaries,
    geometry at the interpolation nodes sNodes by transfinite interpolation
    of the side curves.  Same arrays as QuadMesh, plus elemNodes
    (nElems, 4), the corner node types nodeType and positions nodeX, nodeY
    (nNodes,).  Raises SelfError on a file that is not ISM-V2."""

    def __init__(self, path, N: int, sNodes):
        import ctypes as C
        lib = L.lib()
        nN, nEd, nE, g = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        p = str(path).encode()
        L.check(lib.self_specmesh_counts(p, C.byref(nN), C.byref(nEd), C.byref(nE), C.byref(g)))
        self.N, self.nNodes, self.nEdges, self.nElems, self.gPolyDeg = N, nN.value, nEd.value, nE.value, g.value
        n1, E = N + 1, self.nElems
        self.edges = np.zeros((self.nEdges, 6), dtype=np.int32)
        self.elemNodes = np.zeros((E, 4), dtype=np.int32)
        self.nodeType = np.zeros(self.nNodes, dtype=np.int32)
        self.nodeX = np.zeros(self.nNodes)
        self.nodeY = np.zeros(self.nNodes)
        for name in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J"):
            setattr(self, name, np.zeros((E, n1, n1)))
        self.nHat = np.zeros((E, 4, n1, 2))
        self.nLen = np.zeros((E, 4, n1))
        L.check(lib.self_specmesh_read(
            p, N, L.ptr(L.f64(sNodes)), L.ptr(self.edges, L.ip), L.ptr(self.elemNodes, L.ip),
            L.ptr(self.nodeType, L.ip), L.ptr(self.nodeX), L.ptr(self.nodeY),
            *[L.ptr(getattr(self, n)) for n in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen")]))
        self._sNodes = L.f64(sNodes)

    def ScaleTheMesh(self, xScale: float, yScale: float):
        """QuadMesh%ScaleTheMesh (QuadMeshClass.f90:1691-1714), in place; the
        shallow-water model applies it after reading a mesh file
        (ConservativeShallowWaterClass.f90:302-304)."""
        L.check(L.lib().self_quadmesh_scale(
            self.N, L.ptr(self._sNodes), self.nElems, self.nNodes, float(xScale), float(yScale),
            L.ptr(self.nodeX), L.ptr(self.nodeY),
            *[L.ptr(getattr(self, n)) for n in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J", "nHat", "nLen")]))
        return self


class HexMesh:
    """Structured hex mesh of the unit cube (self_hexmesh_structured).

    x, y, z, J: (nElems, N+1, N+1, N+1) indexed [e, k(q), j(p), i(s)];
    cov: (nElems, N+1, N+1, N+1, 9); nHat, bpos: (nElems, 6, N+1, N+1, 3)
    indexed [e, side, b, a, component]; nLen: (nElems, 6, N+1, N+1);
    faces: (nFaces, 9) int32 = e1, s1, e2, s2, iStart, jStart, iInc, jInc, swap.
    """

    def __init__(self, nX, nY, nZ, N, nodes, periodic=(True, True, True), scale=(1.0, 1.0, 1.0), warp=0.0,
                 warpSign=(1, 1, 1)):
        import ctypes as C
        lib = L.lib()
        nE, nF = C.c_int(), C.c_int()
        L.check(lib.self_hexmesh_counts(nX, nY, nZ, int(periodic[0]), int(periodic[1]), int(periodic[2]),
                                        C.byref(nE), C.byref(nF)))
        self.nX, self.nY, self.nZ, self.N = nX, nY, nZ, N
        self.nElems, self.nFaces = nE.value, nF.value
        n1, E = N + 1, self.nElems
        self.faces = np.zeros((self.nFaces, 9), dtype=np.int32)
        for name in ("x", "y", "z", "J"):
            setattr(self, name, np.zeros((E, n1, n1, n1)))
        self.cov = np.zeros((E, n1, n1, n1, 9))
        self.nHat = np.zeros((E, 6, n1, n1, 3))
        self.nLen = np.zeros((E, 6, n1, n1))
        self.bpos = np.zeros((E, 6, n1, n1, 3))
        sign = L.i32(warpSign)
        L.check(lib.self_hexmesh_structured(
            nX, nY, nZ, N, int(periodic[0]), int(periodic[1]), int(periodic[2]), scale[0], scale[1], scale[2],
            warp, L.ptr(sign, L.ip), L.ptr(L.f64(nodes)), L.ptr(self.faces, L.ip),
            *[L.ptr(getattr(self, n)) for n in ("x", "y", "z", "cov", "J", "nHat", "nLen", "bpos")]))
