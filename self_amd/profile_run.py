"""Run one workload for a fixed number of steps, for use under a profiler.

    python -m self_amd.profile_run copy
    python -m self_amd.profile_run sw2d --precision mixed --form fused --steps 10
    python -m self_amd.profile_run sw2d-dipole --mesh box.mesh --N 10 --form plain --precision fp64
    python -m self_amd.profile_run adv2d --steps 10
    python -m self_amd.profile_run adv3d --steps 5 [--contraction fp32|bf16x3]
    python -m self_amd.profile_run cgsem --nx 128 --precond 1
    python -m self_amd.profile_run cgsem --precision mixed --steps 2000
    python -m self_amd.profile_run hpe --steps 5
    python -m self_amd.profile_run cgsem-batched --nx 64 [--iters 64]

`sw2d-dipole` runs the shallow-water model on an ISM-V2 mesh file (the
dipole run's box.mesh, scaled by 4e5) with the dipole state.
`copy` runs the STREAM-style copy kernel on 24 MiB (read and written: about
the shallow-water stage's 49.5 MB of traffic at 64x64, N = 7; 21
launches), 64 MiB (21 launches) and 1 GiB (11 launches): the calibration for hardware byte counters (FETCH_SIZE /
WRITE_SIZE of a copy of known size).  Every other mode builds the named
workload with its seeded inputs and runs the HIP path only.  For cgsem,
--precision fp64 runs the reference's PCG loop and mixed the FP32 PCG with
FP64 reliable updates, each for at most --steps iterations.  cgsem-batched
solves the --nx problems of the batched workload (meshes 16^2..128^2, N 3, 5,
7; Jacobi-PCG to 1e-10) in one launch sequence.
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys

from . import _lib as L


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m self_amd.profile_run")
    ap.add_argument("mode", choices=["copy", "sw2d", "sw2d-dipole", "adv2d", "adv3d", "cgsem", "hpe",
                                     "cgsem-batched"])
    ap.add_argument("--mesh", default=None, help="ISM-V2 mesh file (sw2d-dipole)")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--form", default="fused", choices=["fused", "plain"])
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--N", type=int, default=7)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--precond", type=int, default=1)
    ap.add_argument("--iters", type=int, default=20000, help="iteration limit (cgsem-batched)")
    ap.add_argument("--contraction", default="fp64", choices=["fp64", "fp32", "bf16x3"],
                    help="arithmetic of the 3-D volume contractions (adv3d)")
    a = ap.parse_args(argv)
    L.check(L.lib().self_init(0))
    if a.mode == "copy":
        g, ms = C.c_double(), C.c_double()
        for size, iters in ((24 << 20, 20), (64 << 20, 20), (1 << 30, 10)):
            L.check(L.lib().self_stream_copy(size, iters, C.byref(g), C.byref(ms)))
            print(f"copy {size} B: {g.value:.1f} GB/s, {ms.value * 1e3:.2f} us per copy", flush=True)
    elif a.mode == "sw2d":
        from .workloads import sw2d_geostrophic
        m, _ = sw2d_geostrophic(a.nx or 64, a.N, precision=a.precision, form=a.form)
        m.ForwardStepRK3(0.0, a.steps)
        m.Synchronize()
        print(f"sw2d {a.precision} {a.form}: {a.steps} RK3 steps", flush=True)
    elif a.mode == "sw2d-dipole":
        from .workloads import sw2d_dipole_box
        if not a.mesh:
            ap.error("sw2d-dipole needs --mesh")
        m, _ = sw2d_dipole_box(a.mesh, a.N, precision=a.precision, form=a.form)
        m.ForwardStepRK3(0.0, a.steps)
        m.Synchronize()
        print(f"sw2d-dipole N={a.N} {a.precision} {a.form}: {a.steps} RK3 steps", flush=True)
    elif a.mode == "hpe":
        from .workloads import hpe_buoyancy_anomaly
        m, _ = hpe_buoyancy_anomaly(a.nx or 32, a.nx or 32, 16, a.N)
        m.ForwardStepRK3(a.steps)
        m.Synchronize()
        print(f"hpe {a.nx or 32}x{a.nx or 32}x16 N={a.N}: {a.steps} RK3 steps", flush=True)
    elif a.mode == "adv2d":
        from .workloads import adv2d_blob
        m, _ = adv2d_blob(nX=a.nx or 10, N=a.N)
        m.ForwardStepRK3(0.0, a.steps)
        m.Synchronize()
        print(f"adv2d: {a.steps} RK3 steps", flush=True)
    elif a.mode == "adv3d":
        from .workloads import adv3d_gaussians
        m, _ = adv3d_gaussians(nX=a.nx or 32, nY=a.nx or 32, nZ=(a.nx or 32) // 2, N=a.N)
        m.SetPrecision(a.contraction)
        m.ForwardStepRK3(0.0, a.steps)
        m.Synchronize()
        print(f"adv3d {a.contraction}: {a.steps} RK3 steps", flush=True)
    elif a.mode == "cgsem-batched":
        from .krylov import SolveBatched
        from .workloads import cgsem_batch
        models, _ = cgsem_batch(a.nx or 64)
        io, its, ms = SolveBatched(models, precond="jacobi", tolerance=1e-10, maxIters=a.iters)
        print(f"cgsem batched: {len(models)} problems, {ms:.1f} ms, iterations {min(its)}..{max(its)}", flush=True)
        for m in models:
            m.Trash()
    elif a.mode == "cgsem":
        from .workloads import cgsem_helmholtz
        m, _ = cgsem_helmholtz(nX=a.nx or 128, N=a.N)
        pc = ("none", "jacobi")[a.precond]
        if a.precision == "mixed":
            r = m.SolveReliable(tolerance=1e-10, delta=0.2, maxIters=a.steps, precond=pc)
        else:
            r = m.Solve("cgsem_pcg", precond=pc, tolerance=1e-10, maxIters=a.steps)
        print(f"cgsem {a.precision}: {r}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
