"""Benchmark of the shallow-water DGSEM hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 2-D nonlinear
shallow water, 64x64 quad elements on a doubly periodic 1e6 m x 1e6 m
f-plane (f = 1e-4 1/s, g = 9.81 m/s^2), N = 7 Gauss nodes (262144 nodes x 3
fields), seeded bathymetry (8 Fourier modes, |h - 1000 m| <= 50 m) and a
geostrophically balanced 1 m Gaussian, RK3 at dt = 1.22 s (CFL 0.5 on the
(N+1)^2 scaling; the scheme diverges above about 2.5-3 s).  One "step" is one
low-storage RK3 step (3 stages) over the whole state; inputs are resident in
HBM when the timed region starts.

value = DOF-stage/s of the FP64 fused path (DOF = node x field, one RK3
stage), the reference's own arithmetic, whole job on one GPU.  The mixed
precision fused path (FP32 volume contractions, FP64 state) and the FP64
plain drop-in (one kernel per reference phase) are timed beside it under
`variants`; the mixed result is compared with the FP64 one after the same
steps, and again at dt = 2.5 s.

roofline: algorithmic bytes per stage launch of the dominant kernel (the
FP64 fused stage kernel, 11 520 B per element at N = 7) over its mean launch
time measured with HIP events on the model's stream; `traffic` is the HBM
bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and
stored under profiles/ (named in `traffic_source`), not re-measured here.

cpu_baseline: the reference's own conservative shallow-water class (Fortran,
compiled from the reference's sources into oracle/_ref, serial as the
reference builds it) on a bounded sample of the same workload (its default
walled mesh of the same size and state); the oracle (C restatement) is timed
beside it at 1 thread and at all host cores, and on samples of the other
configurations.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DOF-stage/s, SW2D N=7 FP64 vs mixed @1e-6 rel L2; p50/p99 batched-solve latency"
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def sw_bytes_per_element(N, s_q, s_G, s_m, s_b, s_n):
    """Compulsory bytes per element per RK stage: q and G read+write,
    5 metric terms, 4 bathymetry/Coriolis fields per node, and
    (n_x, n_y, |n|) per side node."""
    np_ = (N + 1) ** 2
    return np_ * (2 * 3 * s_q + 2 * 3 * s_G + 5 * s_m + 4 * s_b) + 4 * (N + 1) * 3 * s_n


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable_cores": usable, "cpu_model": model}


def _c2_oracle(nX, N, periodic=True):
    import numpy as np
    from oracle import pyoracle as po
    from self_amd.workloads import sw2d_cfl_dt, sw2d_geostrophic_fields
    L = 1.0e6
    nodal = po.NodalStorage2D(N)
    mesh = po.QuadMesh(nX, nX, N, nodal.nodes, periodic, periodic, L, L)
    n1 = N + 1
    x = mesh.x.reshape(-1, n1, n1)
    y = mesh.y.reshape(-1, n1, n1)
    h, U, V, H = sw2d_geostrophic_fields(x, y, L=L, seed=SEED)
    dt = sw2d_cfl_dt(U, V, H, 9.81, L / nX, N)
    orc = po.ShallowWater2D(nodal, mesh, g=9.81, dt=dt)
    orc.sol[:] = np.stack([U, V, H], axis=1).ravel()
    orc.bathy[:] = np.stack([h, 0 * h, 0 * h], axis=1).ravel()
    orc.plv[:] = 1.0e-4
    orc.bathymetry_gradient()
    return orc, (h, U, V, H), dt


def _time_oracle_sw(nX, N, threads, target_s):
    orc, _, _ = _c2_oracle(nX, N)
    orc.integrate(0.0, 2, threads)  # warm the threads and caches
    steps, el = 0, 0.0
    t0 = time.perf_counter()
    while el < target_s and steps < 5000:  # chunks of 10 steps until the sample is long enough
        orc.integrate(0.0, 10, threads)
        steps += 10
        el = time.perf_counter() - t0
    dof = 3 * nX * nX * (N + 1) ** 2
    return {"value": dof * 3 * steps / el, "unit": "DOF-stage/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 C restatement, {nX}x{nX} N={N} geostrophic case, {steps} RK3 steps, "
                      f"{threads} OpenMP thread(s), {el:.2f} s"}


def _time_reference_sw(nX, N, steps):
    """The reference's own class (oracle/_ref/ref_sw2d, serial) on its
    default walled nX x nX mesh scaled to 1e6 m, with the geostrophic state."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ref_run as R
    if not R.available("sw2d"):
        return None
    orc, (h, U, V, H), dt = _c2_oracle(nX, N, periodic=False)
    T = lambda a: np.ascontiguousarray(a.transpose(2, 1, 0))  # noqa: E731  (e, j, i) -> (i, j, e)
    inp = {"sol": np.stack([T(U), T(V), T(H)], axis=2), "h": T(h), "f": np.full_like(T(h), 1.0e-4)}
    sec = R.time_sw2d(N, nX, nX, 9.81, dt, steps, inp, xScale=1.0e6, yScale=1.0e6)
    dof = 3 * nX * nX * (N + 1) ** 2
    return {"value": dof * 3 * steps / sec, "unit": "DOF-stage/s", "cores": 1, "kind": "reference",
            "sample": f"reference ConservativeShallowWaterClass (planewave example's copy, Fortran, amdflang -O2, "
                      f"serial as the reference builds it), default walled {nX}x{nX} mesh scaled to 1e6 m, N={N}, "
                      f"geostrophic state, {steps} ForwardStepRK3 calls, {sec:.2f} s"}


def _time_oracle_others():
    """1-thread oracle samples of the other configurations (the C
    restatements of the advection, 3-D advection and CGSEM loops are serial)."""
    import numpy as np
    from oracle import pyoracle as po
    from self_amd.workloads import (adv2d_blob_fields, adv3d_gaussians_fields, cgsem_poisson_fields)
    out = {}
    # C1: advection-diffusion blob, 10x10 N=7 periodic, 100 RK3 steps
    N = 7
    nodal = po.NodalStorage2D(N)
    mesh = po.QuadMesh(10, 10, N, nodal.nodes, True, True, 1.0, 1.0)
    orc = po.Advection2D(nodal, mesh, nTracers=1, kappa=1e-3, dt=1e-4)
    orc.c[:] = adv2d_blob_fields(mesh.x, mesh.y)
    orc.u[:] = 1.0
    orc.v[:] = 0.5
    t0 = time.perf_counter()
    for k in range(100):
        orc.forward_step_rk3(k * 1e-4)
    el = time.perf_counter() - t0
    out["C1_adv2d"] = {"value": 6400 * 3 * 100 / el, "unit": "DOF-stage/s", "cores": 1, "kind": "port",
                       "ms_per_step": 1e3 * el / 100, "sample": f"10x10 N=7, 100 RK3 steps, {el:.2f} s"}
    # C3: 3-D advection with the dense roll-off filter, sampled on 8x8x4 (per element
    # the work does not depend on the mesh size)
    nd = po.Nodal3D(N)
    hm = po.HexMesh(8, 8, 4, N, nd.nodes, warp=0.05)
    o3 = po.Advection3D(nd, hm, nTracers=1, dt=1e-4, nCutoff=3, useFilter=True)
    o3.c[:] = adv3d_gaussians_fields(hm.x, hm.y, hm.z)
    o3.u[:], o3.v[:], o3.w[:] = 1.0, 0.7, 0.3
    t0 = time.perf_counter()
    for k in range(2):
        o3.forward_step_rk3(k * 1e-4)
    el = time.perf_counter() - t0
    dof = hm.nElems * 512
    out["C3_adv3d"] = {"value": dof * 3 * 2 / el, "unit": "DOF-stage/s", "cores": 1, "kind": "port",
                       "sample": f"8x8x4 N=7 (1/64 of 32x32x16), dense filter on, 2 RK3 steps, {el:.2f} s"}
    # C4: CGSEM PCG iterations at full size, 128x128 N=7, 30 iterations
    cg = po.CGSEM2D(N, 128, 128, tolerance=1e-30, maxIters=30)
    n1 = N + 1
    x = cg.mesh.x.reshape(-1, n1, n1)
    y = cg.mesh.y.reshape(-1, n1, n1)
    k, f = cgsem_poisson_fields(x, y)
    cg.fluxCoeff[:] = k.ravel()
    cg.source[:] = -f.ravel()
    t0 = time.perf_counter()
    io, it = cg.solve(precond=0)
    el = time.perf_counter() - t0
    out["C4_cgsem_cg"] = {"value": it / el, "unit": "Krylov iterations/s", "cores": 1, "kind": "port",
                          "sample": f"128x128 N=7 GLL, {it} CG iterations of the reference's CGSEM loop, {el:.2f} s"}
    return out


def cpu_baseline(nX, N, target_s=12.0, others=True):
    """The reference's own Fortran on a bounded sample of the workload (kind
    "reference"; the oracle when the reference build is absent), with the
    oracle at 1 thread and at all usable cores beside it."""
    info = host_info()
    allc = max(1, min(16, info["usable_cores"]))  # the GPU box gives one GPU 16 host cores
    port1 = _time_oracle_sw(nX, N, 1, target_s)
    portN = _time_oracle_sw(nX, N, allc, target_s) if allc > 1 else port1
    ref = None
    try:
        ref = _time_reference_sw(nX, N, 60)
    except Exception as exc:  # the reference build is optional; say why it is absent
        info["reference_error"] = repr(exc)[:200]
    out = dict(ref if ref else port1)
    out["host"] = info
    out["port_1_thread"] = port1
    out["port_all_cores"] = portN
    if others:
        out["other_configs"] = _time_oracle_others()
    return out


def bench_adv2d(steps=100, warm=10):
    """Advection-diffusion blob (10x10 periodic, N=7 Gauss, (u,v)=(1,0.5),
    kappa=1e-3, dt=1e-4): device time per RK3 step; two launches per stage,
    0.6 MB per stage, so launch latency rather than HBM bounds it."""
    from self_amd.workloads import adv2d_blob
    m, info = adv2d_blob()
    m.ForwardStepRK3(0.0, warm)
    ms = m.ForwardStepRK3Timed(warm * info["dt"], steps)
    m.Trash()
    dof = 100 * 64
    return {"workload": "adv-diff 2D blob 10x10 N=7 periodic, kappa=1e-3, dt=1e-4", "steps": steps,
            "ms_per_step": ms / steps, "us_per_stage": 1e3 * ms / steps / 3.0,
            "dof_stage_per_s": dof * 3 * steps / (ms * 1e-3)}


def bench_adv3d(steps=20, warm=2, nX=32, nY=32, nZ=16, N=7, acc_steps=200):
    """3-D linear advection (32x32x16 periodic, N=7, seeded warp 0.05,
    u=(1,0.7,0.3), three seeded Gaussians, adaptive roll-off filter on,
    RK3 at CFL 0.5), FP64, and the matrix-core variants of the volume
    contractions (FP32, split BF16) with their relative L2 against FP64
    after acc_steps steps.  Bytes per DOF-stage: 2 s_c + 2 s_G + 4 s_m (the
    static velocity folded into the three cofactor rows, and J) plus
    6 (N+1)^2 4 s_n per element = 64 B + 12 B at N=7."""
    import numpy as np
    from self_amd.workloads import adv3d_gaussians
    out = {}
    finals = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        m, info = adv3d_gaussians(nX, nY, nZ, N)
        m.SetPrecision(prec)
        m.ForwardStepRK3(0.0, warm)
        m.FilterCount(reset=True)
        ms = m.ForwardStepRK3Timed(warm * info["dt"], steps)
        filtered = m.FilterCount(reset=True)
        E, n1 = m.nElems, N + 1
        dof = E * n1 ** 3
        bytes_stage = E * (n1 ** 3 * (2 * 8 + 2 * 8 + 4 * 8) + 6 * n1 * n1 * 4 * 8)
        stage_ms = ms / steps / 3.0
        m.SetSolution(info["c"])
        m.ForwardStepRK3(0.0, acc_steps)
        finals[prec] = m.GetSolution()
        m.Trash()
        out[prec] = {"ms_per_step": ms / steps, "stage_ms": stage_ms, "dof_stage_per_s": dof * 3 * steps / (ms * 1e-3),
                     "algorithmic_bytes_per_stage": bytes_stage,
                     "achieved_GBs": round(bytes_stage / (stage_ms * 1e-3) / 1e9, 1),
                     "filtered_element_stages": filtered}
    for prec in ("fp32", "bf16x3"):
        d = finals[prec] - finals["fp64"]
        out[prec][f"rel_l2_vs_fp64_after_{acc_steps}_steps"] = float(np.linalg.norm(d) / np.linalg.norm(finals["fp64"]))
    head = out.pop("fp64")
    head.update({"workload": f"3D advection {nX}x{nY}x{nZ} N={N} periodic, warp 0.05, filter on (cutoff 3)",
                 "steps": steps, "dt": info["dt"], "variants": out})
    return head


def bench_sw2d_uneven(nX=16, degrees=(3, 5, 7, 3, 5, 7, 3, 5), steps=50, warm=4):
    """Eight small shallow-water models (16x16 meshes, N = 3, 5, 7, the
    request mix of the serving workload; geostrophic state, FP64 fused):
    stepped one after another with their own launches (24 launches per
    step), against all eight in one persistent launch per RK stage (3 per
    step).  Small models cannot fill the GPU one at a time."""
    from self_amd.shallow_water import forward_step_rk3_multi
    from self_amd.workloads import sw2d_geostrophic
    models = [sw2d_geostrophic(nX, N, precision="fp64", form="fused")[0] for N in degrees]
    dof = sum(m.nElems * (N + 1) ** 2 for m, N in zip(models, degrees))
    for m in models:
        m.SetGraph(False)
        m.ForwardStepRK3(0.0, warm)
    for m in models:
        m.Synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for m in models:
            m.ForwardStepRK3(0.0, 1)
    for m in models:
        m.Synchronize()
    sep = (time.perf_counter() - t0) * 1e3
    forward_step_rk3_multi(models, warm)
    t0 = time.perf_counter()
    forward_step_rk3_multi(models, steps)
    one = (time.perf_counter() - t0) * 1e3
    for m in models:
        m.Trash()
    return {"workload": f"{len(degrees)} shallow-water models {nX}x{nX} at N = {', '.join(map(str, degrees))}, "
                        f"FP64 fused, {steps} steps, wall clock", "dof": dof,
            "separate_ms_per_step": sep / steps, "one_launch_ms_per_step": one / steps,
            "separate_dof_stage_per_s": dof * 3 * steps / (sep * 1e-3),
            "one_launch_dof_stage_per_s": dof * 3 * steps / (one * 1e-3), "speedup": sep / one}


def hpe_bytes_per_element_stage(N):
    """Algorithmic bytes of one stage per element (FP64): the column kernel
    reads u, v, b, six cofactor terms and J at every node, the column
    height, and the four normal values on six faces, and writes p, w and
    the face values of b, w, p; the stage kernel reads p, w, u, v, b, the
    background buoyancy, nine cofactor terms and J, its own and the
    neighbour's face values, the face data (normal, |n|, background) and
    the RK register, and writes u, v, b and the register."""
    n1 = N + 1
    np_, n2 = n1 ** 3, n1 * n1
    col = 12 * np_ + 43 * n2
    stage = 25 * np_ + 72 * n2
    return 8 * col, 8 * stage


def bench_hpe(steps=10, warm=2, nX=32, nY=32, nL=16, N=7):
    """Hydrostatic primitive equations, 32x32 columns x 16 layers at N=7
    (16384 elements, as the 3-D advection workload), walled box of depth
    0.1 with a buoyancy anomaly at rest, FP64, RK3.  Two kernels per stage;
    each kernel's time comes from events around its launches."""
    import numpy as np
    from self_amd.workloads import hpe_buoyancy_anomaly
    m, info = hpe_buoyancy_anomaly(nX, nY, nL, N)
    m.ForwardStepRK3(warm)
    ms = m.ForwardStepRK3Timed(steps)
    col_ms, st_ms = m.KernelTimes(3)
    ok = bool(np.all(np.isfinite(m.GetSolution())))
    E = m.nElems
    dof = E * (N + 1) ** 3
    bcol, bst = hpe_bytes_per_element_stage(N)
    m.Trash()
    stage_ms = ms / steps / 3.0
    return {"workload": f"hydrostatic primitive equations {nX}x{nY} columns x {nL} layers N={N}, walled, "
                        f"buoyancy anomaly at rest, dt={info['dt']}", "steps": steps,
            "ms_per_step": ms / steps, "stage_ms": stage_ms, "dof_stage_per_s": dof * 3 * steps / (ms * 1e-3),
            "kernel_ms": {"k_hpe_columns": col_ms, "k_hpe_stage": st_ms},
            "algorithmic_bytes_per_stage": {"k_hpe_columns": bcol * E, "k_hpe_stage": bst * E},
            "achieved_GBs": {"k_hpe_columns": round(bcol * E / (col_ms * 1e-3) / 1e9, 1),
                             "k_hpe_stage": round(bst * E / (st_ms * 1e-3) / 1e9, 1),
                             "stage": round((bcol + bst) * E / (stage_ms * 1e-3) / 1e9, 1)},
            "finite": ok}


def bench_sw2d_ensemble(members=8, nX=64, N=7, steps=100, warm=10):
    """Ensemble of `members` states of the 64x64, N=7 geostrophic case (each
    with its own 0.1 m depth offset) advanced together: one fused launch per
    RK stage for all members.  Bytes: per member and element the state and
    RK register read+write (Np 6 (s_q + s_G)); shared by the members the
    metrics, bathymetry/Coriolis and normals (Np (5 s_m + 4 s_b) +
    4 (N+1) 3 s_n), read once per launch."""
    from self_amd.workloads import sw2d_geostrophic
    out = {"members": members, "workload": f"{members} x SW2D geostrophic {nX}x{nX} N={N}"}
    E, np_ = nX * nX, (N + 1) ** 2
    for prec, s_b in (("mixed", 4), ("fp64", 8)):
        m, info = sw2d_geostrophic(nX, N, precision=prec, form="fused", seed=SEED)
        m.ForwardStepRK3(0.0, warm)
        single = m.ForwardStepRK3Timed(0.0, steps) / steps / 3.0
        m.SetMembers(members)
        for k in range(members):
            sol = info["sol"].copy()
            sol[:, 2] += 0.1 * k
            m.SetMemberSolution(k, sol)
        m.ForwardStepRK3(0.0, warm)
        stage_ms = m.ForwardStepRK3Timed(0.0, steps) / steps / 3.0
        m.Trash()
        per_member = np_ * 6 * (8 + 8)
        shared = np_ * (5 * 8 + 4 * s_b) + 4 * (N + 1) * 3 * s_b
        bytes_launch = E * (members * per_member + shared)
        out[prec] = {"stage_ms": stage_ms, "member_stage_us": 1e3 * stage_ms / members,
                     "dof_stage_per_s": members * 3 * E * np_ / (stage_ms * 1e-3),
                     "single_member_stage_us": 1e3 * single, "speedup_vs_single": single * members / stage_ms,
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "achieved_GBs": round(bytes_launch / (stage_ms * 1e-3) / 1e9, 1)}
    return out


def bench_sw2d_dipole(N=10, steps=200, warm=10):
    """The reference's GlobalTimeDerivative profile case
    (doc/shallowwater/ShallowWater-Projects.tex:316-329): the dipole run's
    448-element unstructured mesh (ISM-V2 file, scaled by 4e5), N = 10, its
    dt = 0.1 s and state.  The plain form's per-kernel device times give the
    share of GlobalTimeDerivative spent in the MappedTimeDerivative kernel,
    the figure the reference quotes (85.68 % on its CPU); the fused stage
    times give the path's throughput on this mesh."""
    import tempfile
    import numpy as np
    from self_amd.workloads import sw2d_dipole_box
    src = os.path.join(ROOT, "tests", "golden", "dipole_box_mesh.npz")
    if not os.path.exists(src):
        return {"skipped": "mesh file absent"}
    out = {"workload": "SW2D dipole run mesh, 448 unstructured elements, N=10, dt=0.1 s",
           "reference_mapped_share_of_gtd_pct": 85.68}
    sols = {}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "box.mesh")
        with open(path, "wb") as f:
            f.write(np.load(src)["mesh"].tobytes())
        for name, prec, form in (("fp64_plain", "fp64", "plain"), ("fp64_fused", "fp64", "fused"),
                                 ("mixed_fused", "mixed", "fused")):
            # the mixed form unchecked, to time it and measure its error here
            m, info = sw2d_dipole_box(path, N, precision=prec, form=form, mixedGuard=0)
            E = m.nElems
            m.ForwardStepRK3(0.0, warm)
            stage_ms = m.ForwardStepRK3Timed(0.0, steps) / steps / 3.0
            r = {"stage_us": 1e3 * stage_ms,
                 "dof_stage_per_s": 3 * E * (N + 1) ** 2 / (stage_ms * 1e-3)}
            if form == "plain":
                kt = m.KernelTimes(0.0, 20)
                gtd = kt[0] + kt[1] + kt[2]
                r["kernel_ms_traces_edgeflux_mapped_rk"] = kt
                r["mapped_share_of_gtd_pct"] = round(100.0 * kt[2] / gtd, 2)
            else:
                s_b = 4 if prec == "mixed" else 8
                b = sw_bytes_per_element(N, 8, 8, 8, s_b, s_b) * E
                r["algorithmic_bytes_per_launch"] = b
                r["achieved_GBs"] = round(b / (stage_ms * 1e-3) / 1e9, 1)
                sols[name] = m.GetSolution()
                if prec == "fp64":
                    # eight members of this case in one launch per stage fill the CUs
                    # the single 448-element state leaves idle
                    m.SetMembers(8)
                    for k in range(8):
                        sol = info["sol"].copy()
                        sol[:, 2] += 0.01 * k
                        m.SetMemberSolution(k, sol)
                    m.ForwardStepRK3(0.0, warm)
                    ens_ms = m.ForwardStepRK3Timed(0.0, steps) / steps / 3.0
                    r["members8_member_stage_us"] = 1e3 * ens_ms / 8
                    r["members8_speedup_vs_single"] = 8 * stage_ms / ens_ms
            out[name] = r
            m.Trash()
        # what the model does by default with precision "mixed": the guard
        # compares the mixed and FP64 momentum tendencies and selects FP64
        m, _ = sw2d_dipole_box(path, N, precision="mixed", form="fused")
        err, switched = m.CheckMixed(m.mixedGuard)
        out["mixed_guard"] = {"tendency_rel_l2_mixed_vs_fp64": err, "threshold": m.mixedGuard,
                              "switched_to_fp64": switched}
        m.Trash()
    a, b = sols["mixed_fused"], sols["fp64_fused"]
    out["mixed_unchecked_vs_fp64_rel_l2_UVH"] = [float(np.linalg.norm(a[:, f] - b[:, f]) / np.linalg.norm(b[:, f]))
                                                 for f in range(3)]
    out["steps_compared"] = warm + steps
    return out


def bench_cgsem(nX=128, N=7, tol=1e-10):
    """CGSEM 2-D variable-coefficient Poisson (128x128, N=7 GLL, Dirichlet):
    Jacobi-PCG in FP64 (the reference's CGSEM PCG loop), FP32-inner PCG with
    FP64 iterative refinement (restarted inner solves, and one FP32 sequence
    with FP64 reliable updates), and the reference's FEM preconditioner
    (inner CG on the linear sub-mesh through the Gauss-Lobatto nodes) in the
    PCG loop and as the right preconditioner of FGMRES(30), FP64 and mixed."""
    import numpy as np
    from self_amd.workloads import cgsem_helmholtz
    m, info = cgsem_helmholtz(nX, N, tolerance=tol)
    nodes = m.nElems * (N + 1) ** 2
    zero = np.zeros((m.nElems, N + 1, N + 1))
    r0 = float(np.linalg.norm(m.Residual()))
    out = {"workload": f"CGSEM Poisson {nX}x{nX} N={N} GLL, k=1+0.5 sin sin, 16-mode RHS", "global_dofs": m.nGlobal,
           "element_local_nodes": nodes}
    io, it, ms = m.Solve("cgsem_pcg", precond="jacobi", tolerance=tol, maxIters=40000)
    rel = float(np.linalg.norm(m.Residual())) / r0
    bytes_per_iter = 198.0 * nodes
    out["pcg_fp64"] = {"ioerr": io, "iterations": it, "ms": ms, "iters_per_s": it / (ms * 1e-3),
                       "true_rel_residual": rel,
                       "achieved_GBs": round(bytes_per_iter * it / (ms * 1e-3) / 1e9, 1)}
    m.SetSolution(zero)
    io, outer, inner, relm, ms = m.SolveMixed(tolerance=tol, innerTolerance=1e-5, maxInner=40000, maxOuter=30)
    out["pcg_fp32_inner_fp64_refinement"] = {"ioerr": io, "outer": outer, "inner_iterations": inner, "ms": ms,
                                             "rel_residual": relm,
                                             "true_rel_residual": float(np.linalg.norm(m.Residual())) / r0}
    m.SetSolution(zero)
    io, updates, its, relr, ms = m.SolveReliable(tolerance=tol, delta=0.2, maxIters=40000)
    out["pcg_fp32_reliable_updates_fp64"] = {"ioerr": io, "updates": updates, "iterations": its, "ms": ms,
                                             "iters_per_s": its / (ms * 1e-3), "rel_residual": relr,
                                             "true_rel_residual": float(np.linalg.norm(m.Residual())) / r0,
                                             "speedup_vs_pcg_fp64": out["pcg_fp64"]["ms"] / ms}
    m.SetSolution(zero)
    # the reference's loop stops on its recursive residual (true residual
    # above tol); the FP64 sequence with reliable updates stops on the true one
    io, updates, its, relr, ms = m.SolveReliable(tolerance=tol, delta=1e-2, maxIters=40000, inner="fp64")
    out["pcg_fp64_reliable_updates"] = {"ioerr": io, "updates": updates, "iterations": its, "ms": ms,
                                        "iters_per_s": its / (ms * 1e-3), "rel_residual": relr,
                                        "true_rel_residual": float(np.linalg.norm(m.Residual())) / r0}
    m.SetSolution(zero)
    # the reference's own preconditioner (inner CG on the linear sub-mesh)
    # in its PCG loop, and as the right preconditioner of FGMRES(30) in FP64
    # and in the mixed form (FP32 operator and preconditioner, FP64 basis
    # and restart residuals)
    t0 = time.perf_counter()
    m.SetPreconditioner(pcIterates=100, pcTolerance=1e-4, coefficient="flux")
    out["fem_preconditioner_build_s"] = time.perf_counter() - t0
    io, it, ms = m.Solve("cgsem_pcg", precond="fem", tolerance=tol, maxIters=2000)
    out["pcg_fem_fp64"] = {"ioerr": io, "iterations": it, "ms": ms, "pcIterates": 100, "pcTolerance": 1e-4,
                           "true_rel_residual": float(np.linalg.norm(m.Residual())) / r0}
    m.SetSolution(zero)
    io, it, ms = m.Solve("pcgmres", precond="fem", tolerance=tol, maxIters=200, mInner=30)
    out["fgmres30_fem_fp64"] = {"ioerr": io, "iterations": it, "ms": ms, "pcIterates": 100,
                                "true_rel_residual": float(np.linalg.norm(m.Residual())) / r0}
    m.SetSolution(zero)
    io, it, relg, ms = m.SolveGMRESMixed(tolerance=tol, maxCycles=200, mInner=30, precond="fem")
    out["fgmres30_fem_mixed"] = {"ioerr": io, "iterations": it, "ms": ms, "pcIterates": 100, "rel_residual": relg,
                                 "true_rel_residual": float(np.linalg.norm(m.Residual())) / r0,
                                 "speedup_vs_fgmres30_fem_fp64": out["fgmres30_fem_fp64"]["ms"] / ms}
    m.Trash()
    return out


def bench_cgsem_batched(n=64, tol=1e-10):
    """Batched CGSEM: 64 Jacobi-PCG solves to 1e-10 (meshes 16^2..128^2,
    N in {3,5,7}, seeded right-hand sides) in one launch sequence, against
    the same 64 problems solved one after the other with the single solve
    (same results bit for bit; checked here).  Bytes: 198 per element-local
    node per iteration, as for the single solve."""
    import numpy as np
    from self_amd.krylov import SolveBatched
    from self_amd.workloads import cgsem_batch
    t0 = time.perf_counter()
    models, info = cgsem_batch(n, tolerance=tol)
    setup_s = time.perf_counter() - t0
    zeros = [np.zeros((m.nElems, N + 1, N + 1)) for m, (_, N, _) in zip(models, info)]
    io, its, ms_b = SolveBatched(models, precond="jacobi", tolerance=tol, maxIters=20000)
    ub = [m.GetSolution() for m in models]
    ms_s, same, io_s = 0.0, True, []
    for m, z, u, it in zip(models, zeros, ub, its):
        m.SetSolution(z)
        io1, it1, ms1 = m.Solve("cgsem_pcg", precond="jacobi", tolerance=tol, maxIters=20000)
        ms_s += ms1
        io_s.append(io1)
        same = same and it1 == it and np.array_equal(m.GetSolution(), u)
    nodes = [m.nElems * (N + 1) ** 2 for m, (_, N, _) in zip(models, info)]
    work = sum(198.0 * nd * it for nd, it in zip(nodes, its))
    dof_it = sum(m.nGlobal * it for m, it in zip(models, its))
    for m in models:
        m.Trash()
    mix = {}
    for nX, N, _ in info:
        mix[f"{nX}^2 N={N}"] = mix.get(f"{nX}^2 N={N}", 0) + 1
    return {"workload": f"{n} CGSEM Poisson problems, Jacobi-PCG to {tol:g}", "mix": mix,
            "element_local_nodes": sum(nodes), "iterations_min_max_total": [min(its), max(its), sum(its)],
            "converged": all(e == -2 for e in io), "batched_ms": ms_b, "sequential_single_ms": ms_s,
            "speedup_vs_sequential": ms_s / ms_b, "bitwise_equal_to_single_solves": bool(same),
            "batched_global_dof_iterations_per_s": dof_it / (ms_b * 1e-3),
            "batched_achieved_GBs": round(work / (ms_b * 1e-3) / 1e9, 1),
            "sequential_achieved_GBs": round(work / (ms_s * 1e-3) / 1e9, 1), "host_setup_s": round(setup_s, 1)}


def bench_serving(n=512, workers=8, steps=200):
    """Request stream: 80 % shallow-water integrations (8..64 squared
    elements, N in {3,5,7}, 200 RK3 steps), 20 % CGSEM solves (16..128
    squared, N in {3,5,7}, Jacobi-PCG to 1e-10), for the FP64 and the mixed
    forms.  A saturation pass (all requests at once) measures capacity; the
    reported latencies come from Poisson arrivals at 80 % of it."""
    from self_amd import serving as S
    out = {}
    for prec in ("fp64", "mixed"):
        specs = S.request_stream(n, precision=prec, steps=steps)
        srv = S.Server(workers=workers)
        _, sat = S.run_stream(srv, specs)
        rate = 0.8 * sat["throughput_req_s"]
        specs = S.request_stream(n, precision=prec, steps=steps)
        _, poi = S.run_stream(srv, specs, rate=rate)
        srv.close()
        out[prec] = {"capacity_req_s": sat["throughput_req_s"], "arrival_rate_req_s": rate,
                     "throughput_req_s": poi["throughput_req_s"], "p50_ms": poi["p50_ms"], "p99_ms": poi["p99_ms"],
                     "mean_ms": poi["mean_ms"], "sw_dof_stage_per_s": poi["sw_dof_stage_per_s"],
                     "sw_mean_ensemble": poi["sw_mean_ensemble"], "capacity_sw_mean_ensemble": sat["sw_mean_ensemble"],
                     "sw_mean_models_per_launch": poi["sw_mean_models_per_launch"],
                     "capacity_sw_mean_models_per_launch": sat["sw_mean_models_per_launch"], "ok": poi["ok"]}
    out["stream"] = f"{n} requests, 80% SW (8..64^2, N 3/5/7, {steps} RK3 steps) / 20% CGSEM (16..128^2, N 3/5/7)"
    out["workers"] = workers
    return out


def progress(msg):
    """One line per finished part on stderr, so a long run shows it is alive."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true", help="skip the advection and CGSEM lines")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--serve-requests", type=int, default=512)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the shallow-water path "
              "shards by element blocks with a face-trace halo, not run here.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from self_amd import _lib as L
    from self_amd.workloads import sw2d_geostrophic

    L.check(L.lib().self_init(0))
    nX, N = 64, 7
    K, W = args.steps, args.warmup
    E = nX * nX
    dof = 3 * E * (N + 1) ** 2
    results, sols = {}, {}
    for name, prec, form in (("mixed_fused", "mixed", "fused"), ("fp64_fused", "fp64", "fused"),
                             ("fp64_plain", "fp64", "plain")):
        m, info = sw2d_geostrophic(nX, N, precision=prec, form=form, seed=SEED)
        guard = None
        if prec == "mixed":  # the model's first-step check, here before the timed region
            m.mixedGuardResult = guard = m.CheckMixed(m.mixedGuard)
        if W:
            m.ForwardStepRK3(0.0, W)
        m.Synchronize()
        t0 = time.perf_counter()
        ms = m.ForwardStepRK3Timed(0.0, K)
        wall = time.perf_counter() - t0
        sols[name] = m.GetSolution()
        kt = m.KernelTimes(0.0, 20)
        results[name] = {"ms_per_step": ms / K, "wall_ms_per_step": 1e3 * wall / K,
                         "dof_stage_per_s": dof * 3 * K / (ms * 1e-3), "kernel_ms": kt, "dt": info["dt"]}
        if guard:
            results[name]["mixed_guard"] = {"tendency_rel_l2_mixed_vs_fp64": guard[0], "threshold": m.mixedGuard,
                                            "switched_to_fp64": guard[1]}
        m.Trash()
        progress(f"{name}: {ms / K:.4f} ms per step")

    def rel(a, b):
        return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))

    acc = [rel(sols["mixed_fused"][:, f], sols["fp64_fused"][:, f]) for f in range(3)]
    acc_plain = [rel(sols["fp64_fused"][:, f], sols["fp64_plain"][:, f]) for f in range(3)]
    # the same gate near the scheme's stability limit: 500 steps at dt = 2.5 s
    # (1250 s of model time; the scheme diverges between 2.5 and 3 s)
    long_sols = {}
    for prec in ("fp64", "mixed"):
        m, info = sw2d_geostrophic(nX, N, precision=prec, form="fused", seed=SEED)
        m.SetTimeStep(2.5)
        m.ForwardStepRK3(0.0, 500)
        long_sols[prec] = m.GetSolution()
        m.Trash()
    acc_long = [rel(long_sols["mixed"][:, f], long_sols["fp64"][:, f]) for f in range(3)]

    head = results["fp64_fused"]
    # algorithmic (compulsory) bytes per stage launch of the FP64 fused kernel:
    # q and G read + written, 5 metric terms, (h, h_x, h_y, f) per node and
    # (n_x, n_y, |n|) per side node, all FP64: 11 520 B per element at N = 7
    b_fp64 = sw_bytes_per_element(N, 8, 8, 8, 8, 8) * E
    # the mixed kernel's own arrays: q FP64 r/w, G FP32 r/w, metric deviations
    # and h_x, h_y, f FP32 per node, 5 FP64 metric values per element, FP32
    # normal deviations per side node and 3 FP64 values per side (h is read
    # only with drag, absent here)
    np_ = (N + 1) ** 2
    b_mixed = E * (np_ * (2 * 3 * 8 + 2 * 3 * 4 + 5 * 4 + 3 * 4) + 5 * 8 + 4 * (N + 1) * 3 * 4 + 4 * 3 * 8)
    # mean stage-kernel duration: the timed region is 3K back-to-back launches
    # of the fused stage kernel on the model's stream
    stage_ms = head["ms_per_step"] / 3.0
    achieved = b_fp64 / (stage_ms * 1e-3) / 1e9
    tname = "sw2d_stage_fused_fp64_64x64_N7_traffic.json"
    traffic = load_traffic(tname)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic.get("bytes_per_launch") if traffic else None,
                "traffic_source": (f"profiles/{tname}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this kernel "
                                   "at this size, calibrated on a copy kernel; stored, not measured in this run")
                if traffic else None,
                "kernel": "k_sw_stage_fused<7,fp64>", "algorithmic_bytes_per_launch": b_fp64,
                "mean_launch_ms": stage_ms, "event_pair_ms": head["kernel_ms"][0]}
    mixed = results["mixed_fused"]
    mixed_stage_ms = mixed["ms_per_step"] / 3.0
    out = {
        "metric": METRIC,
        "value": head["dof_stage_per_s"],
        "unit": "DOF-stage/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "SW2D geostrophic f-plane 64x64 elements N=7 periodic, RK3 dt 1.22 s",
                   "elements": E, "nodes": E * (N + 1) ** 2, "fields": 3, "N": N, "dt_s": head["dt"],
                   "seed": SEED, "parallelism": "single GPU"},
        "roofline": roofline,
        "variants": {
            "mixed_fused": {"dtype": "f64 state and RK accumulation + f32 volume contractions and tendency register",
                            "dof_stage_per_s": mixed["dof_stage_per_s"], "ms_per_step": mixed["ms_per_step"],
                            "stage_kernel_ms": mixed_stage_ms, "bytes_per_launch": b_mixed,
                            "achieved_GBs": round(b_mixed / (mixed_stage_ms * 1e-3) / 1e9, 1),
                            "frac": round(b_mixed / (mixed_stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "speedup_vs_fp64_fused": mixed["dof_stage_per_s"] / head["dof_stage_per_s"],
                            "rel_l2_vs_fp64_UVH": acc, "mixed_guard": mixed.get("mixed_guard")},
            "fp64_plain": {"dof_stage_per_s": results["fp64_plain"]["dof_stage_per_s"],
                           "ms_per_step": results["fp64_plain"]["ms_per_step"],
                           "kernel_ms_traces_edgeflux_mapped_rk": results["fp64_plain"]["kernel_ms"]},
        },
        "speedup_fp64_fused_vs_fp64_plain": head["dof_stage_per_s"] / results["fp64_plain"]["dof_stage_per_s"],
        "speedup_mixed_vs_fp64_plain": mixed["dof_stage_per_s"] / results["fp64_plain"]["dof_stage_per_s"],
        "accuracy": {"mixed_vs_fp64_rel_l2_UVH": acc, "fp64_fused_vs_plain_rel_l2_UVH": acc_plain,
                     "steps_compared": K + W, "dt_s": head["dt"],
                     "mixed_vs_fp64_rel_l2_UVH_dt2p5_500_steps": acc_long,
                     "gate": 1e-6, "pass": max(acc + acc_long) <= 1e-6},
    }
    progress("accuracy gates done")
    if not args.no_extras:
        for key, fn in (("sw2d_ensemble", bench_sw2d_ensemble), ("sw2d_uneven_degree", bench_sw2d_uneven),
                        ("sw2d_dipole_N10", bench_sw2d_dipole),
                        ("adv2d", bench_adv2d), ("adv3d", bench_adv3d), ("hpe", bench_hpe),
                        ("cgsem", bench_cgsem), ("cgsem_batched_64", bench_cgsem_batched)):
            out[key] = fn()
            progress(f"{key} done")
    if not args.no_serve:
        out["serving"] = bench_serving(args.serve_requests)
        progress("serving done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nX, N, others=not args.no_extras)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
