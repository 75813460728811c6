"""Advection-diffusion test cases shared by the CPU (oracle) and GPU tests.

A case is a mesh, nodal storage, an edge table (optionally with inflow
flags on boundary edges) and a smooth seeded state loaded identically into
the oracle (oracle/pyoracle.py) and, for GPU tests, into the HIP model
(self_amd.advection).
"""
import math

import numpy as np

from oracle import pyoracle as po


def flag_inflow(edges, west=po.INFLOW_ONE, south=po.INFLOW_TWO):
    """Copy of an (nEdges, 6) edge table whose boundary edges on west
    (side 4) and south (side 1) element sides carry inflow flags."""
    e = np.array(edges, dtype=np.int32).reshape(-1, 6).copy()
    bnd = e[:, 2] <= 0
    e[bnd & (e[:, 1] == 4), 2] = west
    e[bnd & (e[:, 1] == 1), 2] = south
    return e


def smooth_tracers(x, y, Lx, Ly, nq, seed=3):
    rng = np.random.default_rng(seed)
    kx, ky = 2 * math.pi / Lx, 2 * math.pi / Ly
    out = []
    for f in range(nq):
        a = rng.uniform(0.5, 1.0, 4)
        out.append(a[0] * np.sin(kx * x + 0.4 * f) * np.cos(ky * y) + a[1] * np.cos(2 * kx * x - ky * y)
                   + a[2] + 0.2 * a[3] * np.sin(ky * y + 0.1))
    return np.stack(out, axis=1)


def smooth_velocity(x, y, Lx, Ly, seed=4):
    rng = np.random.default_rng(seed)
    kx, ky = 2 * math.pi / Lx, 2 * math.pi / Ly
    a = rng.uniform(0.5, 1.0, 4)
    u = 1.0 + a[0] * 0.5 * np.cos(ky * y) + 0.2 * np.sin(kx * x)
    v = 0.5 + a[1] * 0.5 * np.sin(kx * x + 0.3) - 0.1 * np.cos(ky * y)
    return u, v


def build_oracle_case(N, nX, nY, periodicX=True, periodicY=True, Lx=1.0, Ly=1.0, warp=0.0, nq=1,
                      kappa=1e-2, dt=1e-3, inflow=False, relax=False, quirk=False, seed=3):
    nodal = po.NodalStorage2D(N)
    mesh = po.QuadMesh(nX, nY, N, nodal.nodes, periodicX, periodicY, Lx, Ly, warp)
    edges = flag_inflow(mesh.edges) if inflow else np.array(mesh.edges, dtype=np.int32)
    orc = po.Advection2D(nodal, mesh, nTracers=nq, kappa=kappa, dt=dt, dMatP=nodal.dMatP if quirk else None,
                         with_relax=relax, edges=edges)
    n1, E = N + 1, mesh.nElems
    x = mesh.x.reshape(E, n1, n1)
    y = mesh.y.reshape(E, n1, n1)
    c = smooth_tracers(x, y, Lx, Ly, nq, seed)
    u, v = smooth_velocity(x, y, Lx, Ly, seed + 1)
    orc.c[:] = c.ravel()
    orc.u[:] = u.ravel()
    orc.v[:] = v.ravel()
    info = dict(c=c, u=u, v=v, nodal=nodal, mesh=mesh, edges=edges, x=x, y=y)
    if relax:
        rng = np.random.default_rng(seed + 2)
        r = np.cos(2 * math.pi * x / Lx)[:, None] * rng.uniform(0.5, 1.0, nq)[None, :, None, None]
        k = (0.5 + 0.25 * np.sin(2 * math.pi * y / Ly))[:, None] * np.ones((1, nq, 1, 1))
        orc.relax[:] = r.ravel()
        orc.rFac[:] = k.ravel()
        info.update(relax=r, rFac=k)
    return orc, info
