"""Host side of the request loop (serve.cpp) and the mixed request stream.

Server wraps self_serve_*: submit() hands a request with its input arrays
to the native scheduler (the arrays are kept alive here until wait()),
wait() returns its timings and output.  run_stream() replays a list of
requests with Poisson arrivals and reports throughput and latency
percentiles.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .workloads import DEFAULT_SEED, SplitMix64, cgsem_poisson_fields, sw2d_cfl_dt, sw2d_geostrophic_fields

REQ_SW2D, REQ_CGSEM2D = 1, 2


Request, Result = L.Request, L.Result


@dataclass
class Spec:
    """One request of the stream: kind, mesh, degree and its inputs."""
    kind: int
    nX: int
    N: int
    seed: int
    precision: str = "fp64"
    steps: int = 200
    arrays: dict = field(default_factory=dict)
    params: dict = field(default_factory=dict)

    @property
    def dofs(self):
        n = self.nX * self.nX * (self.N + 1) ** 2
        return 3 * n if self.kind == REQ_SW2D else n


class Server:
    def __init__(self, workers=8, memFraction=0.9):
        h = C.c_void_p()
        L.check(L.lib().self_serve_create(workers, memFraction, C.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, spec: Spec):
        a = spec.arrays
        out = np.zeros_like(a["in0"])
        prec = {"fp64": L.FP64, "mixed": L.MIXED}[spec.precision]
        p = spec.params
        r = Request(spec.kind, spec.nX, spec.N, prec, spec.steps, p.get("precond", 1), p.get("maxIters", 40000),
                    p.get("L", 1.0e6), p.get("g", 9.81), p.get("f0", 1.0e-4), p.get("dt", 1.0), p.get("tol", 1e-10),
                    L.ptr(a["in0"]), L.ptr(a["in1"]), L.ptr(a["in2"]) if a.get("in2") is not None else None,
                    L.ptr(out))
        rid = C.c_longlong()
        L.check(L.lib().self_serve_submit(self._h, C.byref(r), C.byref(rid)))
        self._keep[rid.value] = (a, out)
        return rid.value

    def wait(self, rid):
        res = Result()
        L.check(L.lib().self_serve_wait(self._h, rid, C.byref(res)))
        _, out = self._keep.pop(rid)
        return res, out

    def stats(self):
        q, r, rb, bb = C.c_int(), C.c_int(), C.c_size_t(), C.c_size_t()
        L.check(L.lib().self_serve_stats(self._h, C.byref(q), C.byref(r), C.byref(rb), C.byref(bb)))
        return {"queued": q.value, "running": r.value, "reserved": rb.value, "budget": bb.value}

    def close(self):
        if getattr(self, "_h", None):
            L.check(L.lib().self_serve_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_MESH = {}


def _positions(kind, nX, N, Lside):
    """Node positions of the request's mesh (cached per kind, size, degree)."""
    key = (kind, nX, N)
    if key not in _MESH:
        from .mesh import QuadMesh
        from .nodal import NodalStorage2D
        if kind == REQ_SW2D:
            ns = NodalStorage2D(N, L.GAUSS, L.DG)
            m = QuadMesh(nX, nX, N, ns.nodes, True, True, Lside, Lside)
        else:
            ns = NodalStorage2D(N, L.GAUSS_LOBATTO, L.CG)
            m = QuadMesh(nX, nX, N, ns.nodes, False, False, 1.0, 1.0)
        _MESH[key] = (m.x.copy(), m.y.copy())
    return _MESH[key]


def make_inputs(spec: Spec, Lside=1.0e6, g=9.81, f0=1.0e-4):
    """Seeded inputs of a request: SW gets a geostrophic Gaussian anomaly at a
    seeded centre over the shallow-water workload's bathymetry (8 Fourier
    modes of the default seed, the same for every request on a mesh, so such
    requests can share an ensemble), dt from CFL 0.5; CGSEM gets the
    variable-coefficient Poisson problem with a seeded 16-mode source."""
    if spec.arrays:
        return spec
    x, y = _positions(spec.kind, spec.nX, spec.N, Lside)
    if spec.kind == REQ_SW2D:
        rng = SplitMix64(spec.seed)
        centre = (0.3 + 0.4 * rng.uniform(), 0.3 + 0.4 * rng.uniform())
        h, U, V, H = sw2d_geostrophic_fields(x, y, L=Lside, g=g, f0=f0, seed=DEFAULT_SEED, centre=centre)
        sol = np.ascontiguousarray(np.stack([U, V, H], axis=1))
        bathy = np.ascontiguousarray(np.stack([h, 0 * h, 0 * h], axis=1))
        spec.arrays = {"in0": sol, "in1": bathy}
        spec.params = {"L": Lside, "g": g, "f0": f0, "dt": sw2d_cfl_dt(U, V, H, g, Lside / spec.nX, spec.N)}
    else:
        k, f = cgsem_poisson_fields(x, y, seed=spec.seed)
        spec.arrays = {"in0": np.ascontiguousarray(k), "in1": np.ascontiguousarray(-f)}
        spec.params = {"tol": 1e-10, "precond": 1, "maxIters": 40000}
    return spec


def request_stream(n=512, seed=DEFAULT_SEED, frac_sw=0.8, precision="fp64", steps=200,
                   sw_sizes=(8, 16, 32, 64), cg_sizes=(16, 32, 64, 128), degrees=(3, 5, 7)):
    """The mixed stream: 80 % shallow-water integrations (mesh from sw_sizes,
    N from degrees, `steps` RK3 steps), 20 % CGSEM solves (mesh from cg_sizes,
    N from degrees), each with its own seed."""
    rng = SplitMix64(seed + 17)
    specs = []
    for i in range(n):
        sw = rng.uniform() < frac_sw
        sizes = sw_sizes if sw else cg_sizes
        nX = sizes[rng.randint(0, len(sizes) - 1)]
        N = degrees[rng.randint(0, len(degrees) - 1)]
        specs.append(Spec(REQ_SW2D if sw else REQ_CGSEM2D, nX, N, seed + 1000 + i, precision, steps))
    return specs


def run_stream(server: Server, specs, rate=None, seed=DEFAULT_SEED):
    """Submit specs with Poisson arrivals at `rate` requests/s (None: all at
    once), wait for all; returns per-request results and a summary."""
    for s in specs:
        make_inputs(s)
    rng = SplitMix64(seed + 29)
    t0 = time.perf_counter()
    tnext = 0.0
    ids = []
    for s in specs:
        if rate:
            tnext += -math.log(max(rng.uniform(), 1e-300)) / rate
            delay = tnext - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
        ids.append((s, server.submit(s)))
    out = []
    for s, rid in ids:
        res, arr = server.wait(rid)
        out.append((s, res, arr))
    wall = time.perf_counter() - t0
    lat = np.array([r.latency_ms for _, r, _ in out])
    ok = all(r.status == 0 for _, r, _ in out)
    dof_work = sum(s.dofs * (3 * s.steps if s.kind == REQ_SW2D else 1) for s, _, _ in out)
    summary = {"requests": len(out), "ok": ok, "wall_s": wall, "throughput_req_s": len(out) / wall,
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
               "sw_dof_stage_per_s": sum(s.dofs * 3 * s.steps for s, _, _ in out if s.kind == REQ_SW2D) / wall,
               "sw_mean_ensemble": float(np.mean([r.members for s, r, _ in out if s.kind == REQ_SW2D] or [0])),
               "sw_mean_models_per_launch": float(np.mean([r.models for s, r, _ in out if s.kind == REQ_SW2D]
                                                          or [0])),
               "arrival_rate_req_s": rate}
    del dof_work
    return out, summary
