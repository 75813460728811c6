"""Seeded synthetic workloads shared by the tests, the smoke check and the
benchmark.

Random numbers come from splitmix64 (a counter-based generator with a
published constant set, Steele/Lea/Flood 2014), mapped to U[0,1) by the top
53 bits, so the same inputs can be regenerated anywhere.

sw2d_geostrophic(): 2-D nonlinear shallow water on a doubly periodic
[0, L]^2 f-plane.  Depth h = H0 + sum_{k=1..8} a_k cos(2 pi (m_k x + n_k y)/L + phi_k)
with a_k ~ U[0, amp/8] (so |h - H0| <= amp), (m_k, n_k) uniform on
{-4..4}^2 minus the origin, phi_k ~ U[0, 2 pi).  Free surface
eta = eta0 exp(-r^2 / width^2) about the domain centre, geostrophic
velocity u = -(g/f) d(eta)/dy, v = (g/f) d(eta)/dx evaluated analytically,
H = h + eta, U = H u, V = H v.  The time step follows the nodal-DG CFL
estimate dt = cfl * dx_elem / ((N+1)^2 max(|u| + sqrt(g H))); with
cfl = 0.5 this is 1.22 s for 64x64 elements at N = 7 (a (2N+1) denominator
instead gives 5.2 s, which is above the RK3 stability limit of this
Gauss-node DGSEM and blows up within 20 steps).
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


DEFAULT_SEED = 20261202


def bathymetry_modes(seed=DEFAULT_SEED, nmodes=8, amp=50.0, kmax=4):
    rng = SplitMix64(seed)
    modes = []
    for _ in range(nmodes):
        a = rng.uniform() * amp / nmodes
        while True:
            m = rng.randint(-kmax, kmax)
            n = rng.randint(-kmax, kmax)
            if m != 0 or n != 0:
                break
        phi = 2.0 * math.pi * rng.uniform()
        modes.append((a, m, n, phi))
    return modes


def sw2d_geostrophic_fields(x, y, L=1.0e6, H0=1000.0, g=9.81, f0=1.0e-4, eta0=1.0, width=1.0e5,
                            seed=DEFAULT_SEED, nmodes=8, amp=50.0, centre=(0.5, 0.5)):
    """Fields at node positions x, y (any matching shapes): returns
    h (depth), U, V, H.  The anomaly is centred at centre * L."""
    h = np.full_like(x, H0)
    for a, m, n, phi in bathymetry_modes(seed, nmodes, amp):
        h = h + a * np.cos(2.0 * math.pi * (m * x + n * y) / L + phi)
    xc, yc = centre[0] * L, centre[1] * L
    eta = eta0 * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / width ** 2)
    detadx = -2.0 * (x - xc) / width ** 2 * eta
    detady = -2.0 * (y - yc) / width ** 2 * eta
    u = -(g / f0) * detady
    v = (g / f0) * detadx
    H = h + eta
    return h, H * u, H * v, H


def sw2d_cfl_dt(U, V, H, g, dx_elem, N, cfl=0.5):
    u = np.abs(U / H)
    v = np.abs(V / H)
    lam = np.max(np.maximum(u, v) + np.sqrt(g * H))
    return cfl * dx_elem / ((N + 1) ** 2 * lam)


def sw2d_geostrophic(nX=64, N=7, precision="fp64", form="fused", L=1.0e6, seed=DEFAULT_SEED,
                     cfl=0.5, g=9.81, f0=1.0e-4, quirk_dmatp_transposed=False, compute_gradient=True,
                     mixedGuard="default"):
    """Build a ShallowWater model with the geostrophic case loaded.

    Returns (model, info) with info holding the host arrays that were set.
    mixedGuard: ShallowWater's (the model's default unless given).
    """
    from .shallow_water import ShallowWater, SWParams

    p = SWParams(polyDeg=N, nXElem=nX, nYElem=nX, xScale=L, yScale=L, g=g, f0=f0,
                 periodicX=True, periodicY=True)
    guard = ShallowWater.MIXED_GUARD if mixedGuard == "default" else mixedGuard
    model = ShallowWater(p, precision=precision, form=form, quirk_dmatp_transposed=quirk_dmatp_transposed,
                         mixedGuard=guard)
    m = model.mesh
    h, U, V, H = sw2d_geostrophic_fields(m.x, m.y, L=L, g=g, f0=f0, seed=seed)
    sol = np.stack([U, V, H], axis=1)
    bathy = np.stack([h, np.zeros_like(h), np.zeros_like(h)], axis=1)
    model.SetBathymetry(bathy)
    model.SetPlanetaryVorticity(np.full_like(h, f0))
    model.SetSolution(sol)
    if compute_gradient:
        model.CalculateBathymetryGradient()
    dt = sw2d_cfl_dt(U, V, H, g, L / nX, N, cfl)
    model.SetTimeStep(dt)
    return model, {"sol": sol, "bathy": bathy, "dt": dt, "f0": f0, "g": g, "L": L}


def sw2d_dipole_fields(x, y, g=9.81, f0=1.0e-4, vMax=1.75, x0=2.0e5, y0=2.05e5, x1=2.0e5, y1=1.95e5, L=1.0e4,
                      hMin=200.0, hMax=3000.0, Lshelf=2.5e5):
    """The dipole run's state (examples/shallowwater/dipole/run/
    shallowwater.params; localmods/ShallowWater_InitialConditions.f90):
    shelf bathymetry h = hMin + (hMax - hMin) tanh(x / Lshelf), a free-surface
    dipole dn (exp(-c0 / 2L^2) - exp(-c1 / 2L^2)) with dn = f0 vMax L / g, and
    the geostrophic velocity of that surface, here from its analytic gradient
    (the run takes it from a first GlobalTimeDerivative).  Returns h, U, V, H
    in the conservative variables."""
    dn = f0 * vMax * L / g
    e0 = np.exp(-0.5 * ((x - x0) ** 2 + (y - y0) ** 2) / (L * L))
    e1 = np.exp(-0.5 * ((x - x1) ** 2 + (y - y1) ** 2) / (L * L))
    eta = dn * (e0 - e1)
    detadx = dn * (-(x - x0) * e0 + (x - x1) * e1) / (L * L)
    detady = dn * (-(y - y0) * e0 + (y - y1) * e1) / (L * L)
    h = hMin + (hMax - hMin) * np.tanh(x / Lshelf)
    u = -(g / f0) * detady
    v = (g / f0) * detadx
    H = h + eta
    return h, H * u, H * v, H


def sw2d_dipole_box(mesh_path, N=10, precision="fp64", form="fused", dt=0.1, g=9.81, f0=1.0e-4,
                    scale=4.0e5, mixedGuard="default"):
    """The shallow-water model on the dipole run's unstructured mesh file
    (448 elements, read with the SpecMesh reader and scaled by 4e5 as the
    run's parameters ask), degree N, its time step and state
    (sw2d_dipole_fields).  Returns (model, info).  mixedGuard: ShallowWater's
    (the model's default unless given; 0 runs the mixed form unchecked)."""
    from .mesh import SpecMesh
    from .nodal import NodalStorage2D
    from .shallow_water import ShallowWater, SWParams
    from . import _lib as L

    nodal = NodalStorage2D(N, L.GAUSS, L.DG)
    mesh = SpecMesh(mesh_path, N, nodal.nodes).ScaleTheMesh(scale, scale)
    p = SWParams(polyDeg=N, dt=dt, g=g, f0=f0)
    guard = ShallowWater.MIXED_GUARD if mixedGuard == "default" else mixedGuard
    model = ShallowWater(p, precision=precision, form=form, mesh=mesh, nodal=nodal, mixedGuard=guard)
    h, U, V, H = sw2d_dipole_fields(mesh.x, mesh.y, g=g, f0=f0)
    sol = np.stack([U, V, H], axis=1)
    bathy = np.stack([h, np.zeros_like(h), np.zeros_like(h)], axis=1)
    model.SetBathymetry(bathy)
    model.SetPlanetaryVorticity(np.full_like(h, f0))
    model.SetSolution(sol)
    model.CalculateBathymetryGradient()
    return model, {"sol": sol, "bathy": bathy, "dt": dt, "mesh": mesh}


def adv2d_blob_fields(x, y, x0=0.5, y0=0.5, width=0.1, amp=1.0):
    """Gaussian tracer blob amp * exp(-r^2 / (2 width^2)), r measured from
    (x0, y0) (the 2 sigma^2 convention of the commented-out initial condition
    at advection-diffusion2d/Advection_InitialConditions.f90:91)."""
    return amp * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2.0 * width ** 2))


def adv2d_blob(nX=10, N=7, u=1.0, v=0.5, kappa=1.0e-3, dt=1.0e-4, nTracers=1, quirk_dmatp_transposed=False):
    """Advection-diffusion of a Gaussian blob on the doubly periodic unit
    square: nX x nX elements of degree N on Legendre-Gauss nodes, constant
    velocity (u, v), diffusivity kappa, RK3 step dt.  Returns (model, info)."""
    from .advection import Advection, AdvectionParams

    p = AdvectionParams(polyDeg=N, nXElem=nX, nYElem=nX, nTracers=nTracers, dt=dt, kappa=kappa)
    model = Advection(p, quirk_dmatp_transposed=quirk_dmatp_transposed)
    m = model.mesh
    c = np.zeros((m.nElems, nTracers, N + 1, N + 1))
    c[:, 0] = adv2d_blob_fields(m.x, m.y)
    model.SetSolution(c)
    model.SetVelocity(np.full_like(m.x, u), np.full_like(m.x, v))
    return model, {"c": c, "u": u, "v": v, "kappa": kappa, "dt": dt}


def cgsem_rhs_modes(seed=DEFAULT_SEED, nmodes=16, kmax=8):
    """16 seeded Fourier modes (m, n) in [1..kmax]^2 with a ~ N(0, 1)
    (Box-Muller on splitmix64 uniforms), amplitude a / (m^2 + n^2)."""
    rng = SplitMix64(seed + 4)
    modes = []
    for _ in range(nmodes):
        m = rng.randint(1, kmax)
        n = rng.randint(1, kmax)
        u1 = max(rng.uniform(), 1e-300)
        u2 = rng.uniform()
        a = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
        modes.append((a / (m * m + n * n), m, n))
    return modes


def cgsem_poisson_fields(x, y, seed=DEFAULT_SEED):
    """k = 1 + 0.5 sin(2 pi x) sin(2 pi y) and f = sum a_mn sin(m pi x) sin(n pi y)
    on [0,1]^2, for -div(k grad u) = f."""
    k = 1.0 + 0.5 * np.sin(2 * math.pi * x) * np.sin(2 * math.pi * y)
    f = np.zeros_like(x)
    sx, sy = {}, {}  # each sine evaluated once (same values as evaluating it per term)
    for a, m, n in cgsem_rhs_modes(seed):
        if m not in sx:
            sx[m] = np.sin(m * math.pi * x)
        if n not in sy:
            sy[n] = np.sin(n * math.pi * y)
        f += a * sx[m] * sy[n]
    return k, f


def cgsem_helmholtz(nX=128, N=7, seed=DEFAULT_SEED, tolerance=1e-10, maxIters=20000):
    """CGsemElliptic_2D with the variable-coefficient Poisson problem
    -div(k grad u) = f, u = 0 on the boundary of [0,1]^2, nX x nX elements of
    degree N on Gauss-Lobatto nodes.  The reference's operator is
    div(k grad u) = s, so the source is s = -f.  Returns (model, info)."""
    from .krylov import CGsemElliptic2D, CGsemParams

    p = CGsemParams(polyDeg=N, nXElem=nX, nYElem=nX, tolerance=tolerance, maximumIterates=maxIters)
    model = CGsemElliptic2D(p)
    m = model.mesh
    k, f = cgsem_poisson_fields(m.x, m.y, seed)
    model.SetFluxCoefficient(k)
    model.SetSource(-f)
    model.SetDirichletBoundaryCondition(np.zeros_like(k))
    model.SetSolution(np.zeros_like(k))
    return model, {"k": k, "f": f}


def cgsem_batch(n=64, seed=DEFAULT_SEED, sizes=(16, 32, 64, 128), degrees=(3, 5, 7), tolerance=1e-10,
                maxIters=20000):
    """The batched CGSEM case: n problems, problem i on an nX x nX mesh
    (nX seeded from sizes) of degree N (seeded from degrees), each the
    variable-coefficient Poisson problem of cgsem_helmholtz with its own
    seeded right-hand side (seed + 100 + i).  Returns (models, [(nX, N,
    seed_i)])."""
    rng = SplitMix64(seed + 41)
    models, info = [], []
    for i in range(n):
        nX = sizes[rng.randint(0, len(sizes) - 1)]
        N = degrees[rng.randint(0, len(degrees) - 1)]
        m, _ = cgsem_helmholtz(nX, N, seed=seed + 100 + i, tolerance=tolerance, maxIters=maxIters)
        models.append(m)
        info.append((nX, N, seed + 100 + i))
    return models, info


def adv3d_gaussians_fields(x, y, z, seed=DEFAULT_SEED, n=3):
    """Sum of n seeded Gaussians: centres ~ U[0.2, 0.8]^3, widths ~ U[0.05, 0.1],
    exp(-r^2 / (2 sigma^2))."""
    rng = SplitMix64(seed + 3)
    c = np.zeros_like(x)
    for _ in range(n):
        x0, y0, z0 = (0.2 + 0.6 * rng.uniform() for _ in range(3))
        sig = 0.05 + 0.05 * rng.uniform()
        c += np.exp(-((x - x0) ** 2 + (y - y0) ** 2 + (z - z0) ** 2) / (2.0 * sig * sig))
    return c


def adv3d_warp_signs(seed=DEFAULT_SEED):
    rng = SplitMix64(seed + 5)
    return tuple(1 if rng.uniform() < 0.5 else -1 for _ in range(3))


def adv3d_cfl_dt(N, nX, nY, nZ, u=(1.0, 0.7, 0.3), cfl=0.5):
    """dt = cfl / ((N+1)^2 sum_d |u_d| / h_d) on the unit cube."""
    rate = abs(u[0]) * nX + abs(u[1]) * nY + abs(u[2]) * nZ
    return cfl / ((N + 1) ** 2 * rate)


def adv3d_gaussians(nX=32, nY=32, nZ=16, N=7, warp=0.05, u=(1.0, 0.7, 0.3), nTracers=1, useFilter=True,
                    seed=DEFAULT_SEED, cfl=0.5, nCutoff=3):
    """3-D linear advection on the periodic unit cube: nX x nY x nZ elements
    of degree N, seeded warp of amplitude `warp`, uniform velocity u, tracer
    = sum of 3 seeded Gaussians, RK3 dt from the CFL number.  Returns (model, info)."""
    from .advection3d import Advection3D, Advection3DParams
    dt = adv3d_cfl_dt(N, nX, nY, nZ, u, cfl)
    p = Advection3DParams(polyDeg=N, nXElem=nX, nYElem=nY, nZElem=nZ, nTracers=nTracers, nCutoff=nCutoff, dt=dt,
                          warp=warp, warpSign=adv3d_warp_signs(seed), useFilter=useFilter)
    model = Advection3D(p)
    m = model.mesh
    c = np.zeros((m.nElems, nTracers, N + 1, N + 1, N + 1))
    c[:, 0] = adv3d_gaussians_fields(m.x, m.y, m.z, seed)
    model.SetSolution(c)
    model.SetVelocity(np.full_like(m.x, u[0]), np.full_like(m.x, u[1]), np.full_like(m.x, u[2]))
    return model, {"c": c, "dt": dt, "u": u}


def hpe_buoyancy_anomaly_fields(x, y, z, H, amp=0.01, width=0.1):
    """Gaussian buoyancy anomaly of the first vertical mode:
    b = amp exp(-((x-0.5)^2 + (y-0.5)^2) / (2 width^2)) cos(pi z / H), z in [-H, 0]."""
    return amp * np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / (2.0 * width * width)) * np.cos(np.pi * z / H)


def hpe_buoyancy_anomaly(nX=32, nY=32, nLayers=16, N=7, H=0.1, B=1.0, dt=2.0e-3, diagnoseP=True):
    """Hydrostatic primitive equations in the walled unit square of depth H
    (flat sea floor, NO_NORMAL_FLOW walls and lid): fluid at rest with a
    Gaussian buoyancy anomaly of the first vertical mode on a constant
    background buoyancy B.  The fastest wave speed is about sqrt(B H / pi),
    0.18 for the defaults, so dt = 2e-3 is well inside the RK3 limit at
    N = 7.  Returns (model, info)."""
    from .hydrostatic import HydrostaticParams, HydrostaticPrimitive
    p = HydrostaticParams(polyDeg=N, nXElem=nX, nYElem=nY, nZElem=nLayers, dt=dt, zScale=H, diagnoseP=diagnoseP)
    model = HydrostaticPrimitive(p)
    m = model.mesh
    sol = np.zeros((m.nElems, 5, N + 1, N + 1, N + 1))
    sol[:, 2] = hpe_buoyancy_anomaly_fields(m.x, m.y, m.z - H, H)
    bg = np.zeros_like(sol)
    bg[:, 2] = B
    model.SetSolution(sol)
    model.SetBackgroundState(bg)
    return model, {"sol": sol, "dt": dt, "H": H, "B": B}
