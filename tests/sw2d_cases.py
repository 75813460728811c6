"""Shallow-water test cases shared by the CPU (oracle) and GPU tests.

A case is a mesh, nodal storage and a smooth, non-trivial state that is
loaded identically into the oracle (oracle/pyoracle.py) and, for GPU tests,
into the HIP model (self_amd.shallow_water).
"""
import math

import numpy as np

from oracle import pyoracle as po


def smooth_state(x, y, Lx, Ly, H0=10.0, seed=1):
    """A smooth, unbalanced (U, V, H) with O(1) tendencies; periodic on
    [0,Lx]x[0,Ly]."""
    rng = np.random.default_rng(seed)
    kx, ky = 2 * math.pi / Lx, 2 * math.pi / Ly
    a = rng.uniform(0.5, 1.0, 6)
    H = H0 + a[0] * np.sin(kx * x + 0.3) * np.cos(ky * y) + a[1] * np.cos(2 * kx * x) * 0.3
    U = a[2] * np.cos(kx * x) * np.sin(ky * y + 0.2) * H
    V = a[3] * np.sin(kx * x - 0.1) * np.cos(2 * ky * y) * H
    h = H0 + a[4] * np.cos(kx * x + ky * y)
    f = 0.1 + 0.05 * np.sin(ky * y)
    return U, V, H, h, f


def to_oracle_layout(arr_e_f_j_i):
    """(E, nf, n1, n1) [e,f,j,i] -> flat per-element Fortran layout (i fastest)."""
    return np.ascontiguousarray(arr_e_f_j_i).ravel()


def build_oracle_case(N, nX, nY, periodicX=True, periodicY=True, Lx=2.0, Ly=3.0, warp=0.0, g=9.81,
                      dt=1e-3, seed=1, quirk=False, drag=0.0, relax=None):
    nodal = po.NodalStorage2D(N)
    mesh = po.QuadMesh(nX, nY, N, nodal.nodes, periodicX, periodicY, Lx, Ly, warp)
    orc = po.ShallowWater2D(nodal, mesh, g=g, dt=dt, linearDrag=drag,
                            dMatP=nodal.dMatP if quirk else None, with_relax=relax is not None)
    n1 = N + 1
    E = mesh.nElems
    x = mesh.x.reshape(E, n1, n1)
    y = mesh.y.reshape(E, n1, n1)
    U, V, H, h, f = smooth_state(x, y, Lx, Ly, seed=seed)
    sol = np.stack([U, V, H], axis=1)
    bathy = np.stack([h, np.zeros_like(h), np.zeros_like(h)], axis=1)
    orc.sol[:] = to_oracle_layout(sol)
    orc.bathy[:] = to_oracle_layout(bathy)
    orc.plv[:] = f.ravel()
    if relax is not None:
        orc.relax[:] = to_oracle_layout(relax)
    return orc, dict(sol=sol, bathy=bathy, f=f, nodal=nodal, mesh=mesh, x=x, y=y)
