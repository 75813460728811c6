"""Runs the reference's own models, built by oracle/ref.mk from the
reference's sources into oracle/_ref, on one case and reads back what they
computed:
  ref_adv2d  src/highend/advection-diffusion2d/AdvectionClass.f90
  ref_adv3d  src/highend/advection3d/Advection3DClass.f90
  ref_cgsem2d  src/cgsem/CGsemElliptic_2D_Class.f90
The file formats are the ones documented at the top of oracle/ref_adv2d.f90,
oracle/ref_adv3d.f90 and oracle/ref_cgsem2d.f90.  Inputs are seeded random fields: the comparison
is of the discrete operator, which needs no smoothness.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def available(model="adv2d") -> bool:
    return os.access(os.path.join(REF_DIR, f"ref_{model}"), os.X_OK)


def _execute(model, params_text, input_arrays, header, timeout, params_file="advection.params", place=None):
    """Writes the namelist file (params_file), ref_in.bin and the files of
    `place` ({name: text}) into a fresh directory, runs oracle/_ref/ref_<model>
    there and returns the bytes of ref_out.bin."""
    exe = os.path.join(REF_DIR, f"ref_{model}")
    if not available(model):
        raise FileNotFoundError(exe)
    with tempfile.TemporaryDirectory() as d:
        for name, text in (place or {}).items():
            with open(os.path.join(d, name), "w") as f:
                f.write(text)
        with open(os.path.join(d, params_file), "w") as f:
            f.write(params_text)
        with open(os.path.join(d, "ref_in.bin"), "wb") as f:
            f.write(header)
            for a in input_arrays:
                f.write(np.asfortranarray(a, dtype=np.float64).tobytes(order="F"))
        p = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"ref_{model} failed ({p.returncode}): {p.stderr[-2000:]}")
        with open(os.path.join(d, "ref_out.bin"), "rb") as f:
            return f.read()


class _Reader:
    def __init__(self, raw):
        self.raw, self.pos = raw, 0

    def take(self, dtype, shape):
        cnt = int(np.prod(shape))
        a = np.frombuffer(self.raw, dtype=dtype, count=cnt, offset=self.pos)
        self.pos += cnt * np.dtype(dtype).itemsize
        return a.reshape(shape, order="F").copy()

    def done(self):
        assert self.pos == len(self.raw), (self.pos, len(self.raw))


def random_inputs(N, nX, nY, nTracers, seed):
    """c ~ U[-1, 1], u, v ~ U[-1, 1] + (0.6, -0.3), relax ~ U[-1, 1],
    rFac ~ U[0, 2]; arrays element-last (Fortran order of the reference)."""
    rng = np.random.default_rng(seed)
    n1, E = N + 1, nX * nY
    c = rng.uniform(-1, 1, (n1, n1, nTracers, E))
    u = rng.uniform(-1, 1, (n1, n1, E)) + 0.6
    v = rng.uniform(-1, 1, (n1, n1, E)) - 0.3
    rl = rng.uniform(-1, 1, (n1, n1, nTracers, E))
    rf = rng.uniform(0, 2, (n1, n1, nTracers, E))
    return {"c": c, "u": u, "v": v, "relax": rl, "rFac": rf}


def _params_text(N, nX, nY, nTracers, kappa, dt, xScale, yScale, spec_mesh=None):
    mesh = f" SpecMeshFile = '{spec_mesh}',\n" if spec_mesh else ""
    return ("&TracerSetup\n"
            f" nTracers = {nTracers},\n kappa = {kappa!r},\n/\n"
            "&TimeManagement\n"
            f" dt = {dt!r},\n iterInit = 0,\n nTimeSteps = 0,\n dumpFreq = 1,\n/\n"
            "&SpaceManagement\n" + mesh +
            f" polyDeg = {N},\n nXElem = {nX},\n nYElem = {nY},\n nPlot = 4,\n"
            f" xScale = {xScale!r},\n yScale = {yScale!r},\n/\n")


def run(N, nX, nY, nTracers, kappa, dt, t0, nSteps, inputs, xScale=1.0, yScale=1.0, timeout=120, mesh_text=None):
    """Returns a dict: the inputs, the reference's nodal arrays (dMatS,
    dMatP (n1, n1), qWeightS, qWeightP, lagWest/East/South/North), its mesh
    (edges (nEdges, 6) int32, x, y, dxds, dxdp, dyds, dydp, J (n1, n1, E),
    nHat (2, n1, 4, E), nLen (n1, 4, E)), the tendency and gradients after
    GlobalTimeDerivative(t0) and the solution after nSteps RK3 steps.
    With mesh_text the model reads that ISM-V2 file (SpecMeshFile) instead
    of its default nX x nY mesh; the inputs then have nX * nY = its element
    count."""
    n1, E = N + 1, nX * nY
    place = {"m.mesh": mesh_text} if mesh_text else None
    raw = _execute("adv2d", _params_text(N, nX, nY, nTracers, kappa, dt, xScale, yScale,
                                         "m.mesh" if mesh_text else None),
                   [inputs[k] for k in ("c", "u", "v", "relax", "rFac")],
                   np.int32(nSteps).tobytes() + np.float64(t0).tobytes(), timeout, place=place)
    r = _Reader(raw)
    take = r.take
    out = dict(inputs)
    nE, nEd = take(np.int32, (2,))
    assert nE == E
    out["dMatS"] = take(np.float64, (n1, n1))
    out["dMatP"] = take(np.float64, (n1, n1))
    for k in ("qWeightS", "qWeightP", "lagWest", "lagEast", "lagSouth", "lagNorth"):
        out[k] = take(np.float64, (n1,))
    out["edges"] = take(np.int32, (6, int(nEd))).T.copy()
    for k in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J"):
        out[k] = take(np.float64, (n1, n1, E))
    out["nHat"] = take(np.float64, (2, n1, 4, E))
    out["nLen"] = take(np.float64, (n1, 4, E))
    shp = (n1, n1, nTracers, E)
    out["tend"] = take(np.float64, shp)
    out["dcdx"] = take(np.float64, shp)
    out["dcdy"] = take(np.float64, shp)
    out["sol"] = take(np.float64, shp)
    r.done()
    out.update(N=N, nX=nX, nY=nY, nTracers=nTracers, kappa=kappa, dt=dt, t0=t0, nSteps=nSteps)
    return out


def random_inputs_3d(N, nX, nY, nZ, nTracers, seed, noise=1.0):
    """c: a smooth function of the node indices plus noise whose amplitude
    (noise x U[0,1]^3) varies per element (so the adaptive filter fires in some elements and
    not in others); u, v, w ~ U[-0.5, 0.5] + (1, 0.7, 0.3); relax ~ U[-1, 1],
    rFac ~ U[0, 2].  Arrays element-last (Fortran order)."""
    rng = np.random.default_rng(seed)
    n1, E = N + 1, nX * nY * nZ
    i = np.arange(n1)
    I, J, K = np.meshgrid(i, i, i, indexing="ij")
    c = np.empty((n1, n1, n1, nTracers, E))
    for e in range(E):
        for f in range(nTracers):
            ph = rng.uniform(0, 2 * np.pi, 3)
            smooth = np.cos(0.4 * I + ph[0]) * np.cos(0.3 * J + ph[1]) * np.cos(0.35 * K + ph[2])
            c[..., f, e] = smooth + noise * rng.uniform(0, 1) ** 3 * rng.uniform(-1, 1, (n1, n1, n1))
    vel = [rng.uniform(-0.5, 0.5, (n1, n1, n1, E)) + m for m in (1.0, 0.7, 0.3)]
    rl = rng.uniform(-1, 1, (n1, n1, n1, nTracers, E))
    rf = rng.uniform(0, 2, (n1, n1, n1, nTracers, E))
    return {"c": c, "u": vel[0], "v": vel[1], "w": vel[2], "relax": rl, "rFac": rf}


def _params_text_3d(N, nX, nY, nZ, nTracers, nCutoff, dt, scale):
    return ("&TracerSetup\n"
            f" nTracers = {nTracers},\n/\n"
            "&TimeManagement\n"
            f" dt = {dt!r},\n iterInit = 0,\n nTimeSteps = 0,\n dumpFreq = 1,\n/\n"
            "&SpaceManagement\n"
            f" polyDeg = {N},\n nCutoff = {nCutoff},\n nXElem = {nX},\n nYElem = {nY},\n nZElem = {nZ},\n"
            f" nPlot = 4,\n xScale = {scale[0]!r},\n yScale = {scale[1]!r},\n zScale = {scale[2]!r},\n/\n")


def run_3d(N, nX, nY, nZ, nTracers, nCutoff, dt, t0, nSteps, inputs, scale=(1.0, 1.0, 1.0), timeout=300):
    """Returns a dict: the inputs; the reference's nodal arrays (dMatS,
    dMatP, dMatQ (n1, n1), qWeightS/P/Q, lagWest/East/South/North/Bottom/Top,
    filterMat (n1^3, n1^3)); its mesh (faces (nFaces, 9) int32, x, y, z, J
    (n1, n1, n1, E), cov (9, n1, n1, n1, E), nHat (3, n1, n1, 6, E), nLen
    (n1, n1, 6, E), bpos (3, n1, n1, 6, E)); after GlobalTimeDerivative(t0, 1):
    tend1, sol1, E1_1, E2_1 (E, nTracers), lim1 (nTracers, E, 3); after
    nSteps RK3 steps that follow: sol, E1, E2, lim."""
    n1, E = N + 1, nX * nY * nZ
    raw = _execute("adv3d", _params_text_3d(N, nX, nY, nZ, nTracers, nCutoff, dt, scale),
                   [inputs[k] for k in ("c", "u", "v", "w", "relax", "rFac")],
                   np.int32(nSteps).tobytes() + np.float64(t0).tobytes(), timeout)
    r = _Reader(raw)
    take = r.take
    out = dict(inputs)
    nE, nF = take(np.int32, (2,))
    assert nE == E
    for k in ("dMatS", "dMatP", "dMatQ"):
        out[k] = take(np.float64, (n1, n1))
    for k in ("qWeightS", "qWeightP", "qWeightQ", "lagWest", "lagEast", "lagSouth", "lagNorth", "lagBottom",
              "lagTop"):
        out[k] = take(np.float64, (n1,))
    out["filterMat"] = take(np.float64, (n1 ** 3, n1 ** 3))
    out["faces"] = take(np.int32, (9, int(nF))).T.copy()
    for k in ("x", "y", "z", "J"):
        out[k] = take(np.float64, (n1, n1, n1, E))
    out["cov"] = take(np.float64, (9, n1, n1, n1, E))
    out["nHat"] = take(np.float64, (3, n1, n1, 6, E))
    out["nLen"] = take(np.float64, (n1, n1, 6, E))
    out["bpos"] = take(np.float64, (3, n1, n1, 6, E))
    shp = (n1, n1, n1, nTracers, E)
    out["tend1"] = take(np.float64, shp)
    out["sol1"] = take(np.float64, shp)
    out["E1_1"] = take(np.float64, (E, nTracers))
    out["E2_1"] = take(np.float64, (E, nTracers))
    out["lim1"] = take(np.float64, (nTracers, E, 3))
    out["sol"] = take(np.float64, shp)
    out["E1"] = take(np.float64, (E, nTracers))
    out["E2"] = take(np.float64, (E, nTracers))
    out["lim"] = take(np.float64, (nTracers, E, 3))
    r.done()
    out.update(N=N, nX=nX, nY=nY, nZ=nZ, nTracers=nTracers, nCutoff=nCutoff, dt=dt, t0=t0, nSteps=nSteps,
               xScale=scale[0], yScale=scale[1], zScale=scale[2])
    return out


def random_inputs_cgsem2d(N, nX, nY, seed):
    """fluxCoeff ~ U[0.5, 1.5], source ~ U[-1, 1], solution ~ U[-0.5, 0.5],
    test vectors u, v ~ U[-1, 1] (element-last, Fortran order)."""
    rng = np.random.default_rng(seed)
    shp = (N + 1, N + 1, nX * nY)
    return {"fluxCoeff": rng.uniform(0.5, 1.5, shp), "source": rng.uniform(-1, 1, shp),
            "solution": rng.uniform(-0.5, 0.5, shp), "u": rng.uniform(-1, 1, shp), "v": rng.uniform(-1, 1, shp)}


def run_cgsem2d(N, nX, nY, maxIters, tolerance, dirValue, inputs, timeout=300, mesh_text=None, pcIterates=0,
                pcTolerance=1.0e-4):
    """The reference's CGsemElliptic_2D, by default with pcIterates = 0
    (identity preconditioner; pcIterates > 0 runs its inner-CG FEM
    preconditioner).  Returns a dict: the inputs; dMat, qWeight; edges
    (nEdges, 6) int32, nodeType (nNodes,) int32; x, y, dxds, dxdp, dyds,
    dydp, J (n1, n1, E); mask_v, unmask_mask_v, gsum_v; Au, u_masked (after
    MatrixAction(u)); dot_uv; r, sol_after_residual; ioerr, resi
    (maxIters + 1,), sol_after_solve."""
    n1, E = N + 1, nX * nY
    params = ("&SolverCriteria\n"
              f" MaximumIterates = {maxIters},\n pcIterates = {int(pcIterates)},\n tolerance = {float(tolerance)!r},\n"
              f" pcTolerance = {float(pcTolerance)!r},\n pcDegree = 1,\n/\n"
              "&SpaceManagement\n" + (" SpecMeshFile = 'm.mesh',\n" if mesh_text else "") +
              f" polyDeg = {N},\n nXElem = {nX},\n nYElem = {nY},\n nPlot = 4,\n xScale = 1.0,\n/\n")
    raw = _execute("cgsem2d", params, [inputs[k] for k in ("fluxCoeff", "source", "solution", "u", "v")],
                   np.float64(dirValue).tobytes(), timeout, params_file="runtime.params",
                   place={"m.mesh": mesh_text} if mesh_text else None)
    r = _Reader(raw)
    take = r.take
    out = dict(inputs)
    nE, nEd, nNo = take(np.int32, (3,))
    assert nE == E
    out["dMat"] = take(np.float64, (n1, n1))
    out["qWeight"] = take(np.float64, (n1,))
    out["edges"] = take(np.int32, (6, int(nEd))).T.copy()
    out["nodeType"] = take(np.int32, (int(nNo),))
    shp = (n1, n1, E)
    for k in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J", "mask_v", "unmask_mask_v", "gsum_v", "Au",
              "u_masked"):
        out[k] = take(np.float64, shp)
    out["dot_uv"] = take(np.float64, (1,))[0]
    out["r"] = take(np.float64, shp)
    out["sol_after_residual"] = take(np.float64, shp)
    out["ioerr"] = int(take(np.int32, (1,))[0])
    out["resi"] = take(np.float64, (maxIters + 1,))
    out["sol_after_solve"] = take(np.float64, shp)
    r.done()
    out.update(N=N, nX=nX, nY=nY, maxIters=maxIters, tolerance=tolerance, dirValue=dirValue, pcIterates=pcIterates,
               pcTolerance=pcTolerance)
    return out


KRYLOV_METHODS = {"cg": 1, "pcg": 2, "gmres": 3, "pcgmres": 4}


def run_krylov(A, b, x0, method, maxIters, mInner, tolerance, timeout=120):
    """The reference's Solve_<method> of IterativeSolvers_Class on the dense
    system A x = b from x0 (oracle/ref_krylov.f90).  Returns ioerr, resi (the
    whole history array) and x."""
    n = A.shape[0]
    header = (np.array([n, KRYLOV_METHODS[method], maxIters, mInner], dtype=np.int32).tobytes()
              + np.float64(tolerance).tobytes())
    raw = _execute("krylov", "", [A, b, x0], header, timeout, params_file="unused.params")
    r = _Reader(raw)
    ioerr, resiLen = (int(v) for v in r.take(np.int32, (2,)))
    resi = r.take(np.float64, (resiLen,))
    x = r.take(np.float64, (n,))
    r.done()
    return {"ioerr": ioerr, "resi": resi, "x": x}


def run_nodal(N, quadrature, form, nNew, f2, f3, timeout=60):
    """The reference's quadrature, 2-D/3-D nodal storage, ApplyDerivativeMatrix,
    CalculateAtBoundaries, interpolation matrix and CoarseToFine for one
    degree (oracle/ref_nodal.f90).  f2 (N+1, N+1) and f3 (N+1, N+1, N+1)
    indexed [i, j(, k)]."""
    n1, m1 = N + 1, nNew + 1
    header = np.array([N, quadrature, form, nNew], dtype=np.int32).tobytes()
    raw = _execute("nodal", "", [f2, f3], header, timeout, params_file="unused.params")
    r = _Reader(raw)
    t = r.take
    out = {"N": N, "quadrature": quadrature, "form": form, "nNew": nNew, "f2": f2, "f3": f3}
    out["nodes"], out["weights"] = t(np.float64, (n1,)), t(np.float64, (n1,))
    out["dMatS"], out["dMatP"] = t(np.float64, (n1, n1)), t(np.float64, (n1, n1))
    out["qWeightS"] = t(np.float64, (n1,))
    for k in ("lagWest", "lagEast", "lagSouth", "lagNorth"):
        out[k] = t(np.float64, (n1,))
    out["dFds2"], out["dFdp2"] = t(np.float64, (n1, n1)), t(np.float64, (n1, n1))
    for k in ("fWest2", "fEast2", "fSouth2", "fNorth2"):
        out[k] = t(np.float64, (n1,))
    out["dMat3"] = t(np.float64, (n1, n1))
    for k in ("dFds3", "dFdp3", "dFdq3"):
        out[k] = t(np.float64, (n1, n1, n1))
    for k in ("fWest3", "fEast3", "fSouth3", "fNorth3", "fBottom3", "fTop3"):
        out[k] = t(np.float64, (n1, n1))
    out["Ts"] = t(np.float64, (m1, n1))
    out["fine2"] = t(np.float64, (m1, m1))
    out["fine3"] = t(np.float64, (m1, m1, m1))
    r.done()
    return out


def _execute_files(model, params_text, header, input_arrays, want, timeout=60, place=None):
    """As _execute, with files placed in the directory beforehand (place:
    name -> bytes) and the named output files returned (want: names)."""
    exe = os.path.join(REF_DIR, f"ref_{model}")
    if not available(model):
        raise FileNotFoundError(exe)
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "advection.params"), "w") as f:
            f.write(params_text)
        with open(os.path.join(d, "ref_in.bin"), "wb") as f:
            f.write(header)
            for a in input_arrays:
                f.write(np.asfortranarray(a, dtype=np.float64).tobytes(order="F"))
        for name, data in (place or {}).items():
            with open(os.path.join(d, name), "wb") as f:
                f.write(data)
        p = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"ref_{model} failed ({p.returncode}): {p.stderr[-2000:]}")
        out = {}
        for name in want:
            with open(os.path.join(d, name), "rb") as f:
                out[name] = f.read()
        return out


def pickup_name(model, it):
    """'<model>.<it, 10 digits>.pickup' (WRITE(iterChar,'(I10.10)'))."""
    return f"{model}.{it:010d}.pickup"


def run_pickup_write(N, nX, nY, nTracers, it, inputs, timeout=60):
    """The reference's WritePickup_Advection of the given fields: returns the
    bytes of Advection.<it>.pickup (inputs as random_inputs)."""
    params = _params_text(N, nX, nY, nTracers, 1e-3, 1e-4, 1.0, 1.0)
    header = np.array([1, it], dtype=np.int32).tobytes()
    arrays = [inputs[k] for k in ("c", "u", "v", "relax", "rFac")]
    name = pickup_name("Advection", it)
    return _execute_files("pickup", params, header, arrays, [name], timeout)[name]


def run_pickup_read(N, nX, nY, nTracers, it, data, timeout=60):
    """The reference's ReadPickup_Advection of the given file bytes; returns
    the fields read back (c, u, v, relax, rFac; element-last shapes)."""
    params = _params_text(N, nX, nY, nTracers, 1e-3, 1e-4, 1.0, 1.0)
    header = np.array([2, it], dtype=np.int32).tobytes()
    raw = _execute_files("pickup", params, header, [], ["ref_out.bin"], timeout,
                         place={pickup_name("Advection", it): data})["ref_out.bin"]
    n1, E = N + 1, nX * nY
    r = _Reader(raw)
    out = {"c": r.take("<f8", (n1, n1, nTracers, E)), "u": r.take("<f8", (n1, n1, E)),
           "v": r.take("<f8", (n1, n1, E)), "relax": r.take("<f8", (n1, n1, nTracers, E)),
           "rFac": r.take("<f8", (n1, n1, nTracers, E))}
    r.done()
    return out


def run_specmesh(path, N, quadrature=1, scale=None, timeout=120):
    """The reference's ReadSpecMeshFile_QuadMesh on a mesh file with the DG
    nodal storage of degree N (quadrature 1 Gauss, -1 Gauss-Lobatto), then,
    with scale = (xScale, yScale), its ScaleTheMesh: edges, element corners,
    node types and positions, and the element geometry (element-last
    shapes)."""
    name = os.path.abspath(str(path)).encode()
    header = np.array([N, quadrature, len(name)], dtype=np.int32).tobytes() + name
    header += np.array([0 if scale is None else 1], dtype=np.int32).tobytes()
    header += np.array(scale or (1.0, 1.0), dtype=np.float64).tobytes()
    exe = os.path.join(REF_DIR, "ref_specmesh")
    if not available("specmesh"):
        raise FileNotFoundError(exe)
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "ref_in.bin"), "wb") as f:
            f.write(header)
        p = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"ref_specmesh failed ({p.returncode}): {p.stderr[-2000:]}")
        with open(os.path.join(d, "ref_out.bin"), "rb") as f:
            raw = f.read()
    r = _Reader(raw)
    nE, nEd, nNd = (int(v) for v in r.take("<i4", (3,)))
    n1 = N + 1
    out = {"nElems": nE, "nEdges": nEd, "nNodes": nNd, "edges": r.take("<i4", (6, nEd)),
           "elemNodes": r.take("<i4", (4, nE)), "nodeType": r.take("<i4", (nNd,)),
           "nodeX": r.take("<f8", (nNd,)), "nodeY": r.take("<f8", (nNd,))}
    for k in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J"):
        out[k] = r.take("<f8", (n1, n1, nE))
    out["nHat"] = r.take("<f8", (2, n1, 4, nE))
    out["nLen"] = r.take("<f8", (n1, 4, nE))
    r.done()
    return out


def random_inputs_sw2d(N, nX, nY, seed):
    """Shallow-water state per node: H ~ U[1, 2], U, V ~ U[-0.5, 0.5] H,
    bathymetry h ~ U[0.8, 1.2] (smooth in nothing: the comparison is of the
    discrete operator), Coriolis f ~ U[-1, 1]; arrays element-last."""
    rng = np.random.default_rng(seed)
    n1, E = N + 1, nX * nY
    H = rng.uniform(1.0, 2.0, (n1, n1, E))
    sol = np.stack([rng.uniform(-0.5, 0.5, (n1, n1, E)) * H, rng.uniform(-0.5, 0.5, (n1, n1, E)) * H, H], axis=2)
    h = rng.uniform(0.8, 1.2, (n1, n1, E))
    f = rng.uniform(-1.0, 1.0, (n1, n1, E))
    return {"sol": sol, "h": h, "f": f}


def _params_text_sw2d(N, nX, nY, g, dt, xScale, yScale):
    g, dt, xScale, yScale = float(g), float(dt), float(xScale), float(yScale)
    return ("&MODEL_FORM\n/\n"
            "&TIME_MANAGEMENT\n"
            f" dt = {dt!r},\n iterInit = 0,\n nTimeSteps = 0,\n dumpFreq = 1,\n/\n"
            "&SPACE_MANAGEMENT\n"
            f" polyDeg = {N},\n nXElem = {nX},\n nYElem = {nY},\n nPlot = 4,\n"
            f" xScale = {xScale!r},\n yScale = {yScale!r},\n/\n"
            "&PHYSICAL_PARAMETERS\n"
            f" g = {g!r},\n f0 = 0.0,\n betaX = 0.0,\n betaY = 0.0,\n linearDrag = 0.0,\n/\n"
            "&DIPOLE_PARAMETERS\n/\n")


def run_sw2d(N, nX, nY, g, dt, t0, nSteps, inputs, xScale=1.0, yScale=1.0, timeout=120):
    """The reference's conservative shallow-water class (planewave example,
    oracle/ref_sw2d.f90).  Returns a dict: the inputs; the reference's nodal
    arrays and mesh as run(); bathy (n1, n1, 3, E) after
    CalculateBathymetryGradient; tend (n1, n1, 3, E) and bflux (n1, 3, 4, E)
    after GlobalTimeDerivative(t0); sol_out (n1, n1, 3, E) after nSteps RK3
    steps."""
    n1, E = N + 1, nX * nY
    raw = _execute("sw2d", _params_text_sw2d(N, nX, nY, g, dt, xScale, yScale),
                   [inputs[k] for k in ("sol", "h", "f")],
                   np.int32(nSteps).tobytes() + np.float64(t0).tobytes(), timeout,
                   params_file="shallowwater.params")
    r = _Reader(raw)
    take = r.take
    out = dict(inputs)
    nE, nEd = take(np.int32, (2,))
    assert nE == E
    out["dMatS"] = take(np.float64, (n1, n1))
    out["dMatP"] = take(np.float64, (n1, n1))
    for k in ("qWeightS", "qWeightP", "lagWest", "lagEast", "lagSouth", "lagNorth"):
        out[k] = take(np.float64, (n1,))
    out["edges"] = take(np.int32, (6, int(nEd))).T.copy()
    for k in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J"):
        out[k] = take(np.float64, (n1, n1, E))
    out["nHat"] = take(np.float64, (2, n1, 4, E))
    out["nLen"] = take(np.float64, (n1, 4, E))
    out["bathy"] = take(np.float64, (n1, n1, 3, E))
    out["tend"] = take(np.float64, (n1, n1, 3, E))
    out["bflux"] = take(np.float64, (n1, 3, 4, E))
    out["sol_out"] = take(np.float64, (n1, n1, 3, E))
    r.done()
    out.update(N=N, nX=nX, nY=nY, g=g, dt=dt, t0=t0, nSteps=nSteps, xScale=xScale, yScale=yScale)
    return out


def time_sw2d(N, nX, nY, g, dt, nSteps, inputs, xScale=1.0, yScale=1.0, timeout=600):
    """Wall seconds of nSteps ForwardStepRK3 calls of the reference's
    shallow-water class (oracle/ref_sw2d timing mode) on its default walled
    nX x nY mesh, from the given state; the reference is serial."""
    import resource
    try:  # the class keeps its RK3 register on the stack
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    except (ValueError, OSError):
        pass
    raw = _execute("sw2d", _params_text_sw2d(N, nX, nY, g, dt, xScale, yScale),
                   [inputs[k] for k in ("sol", "h", "f")],
                   np.int32(-nSteps).tobytes() + np.float64(0.0).tobytes(), timeout,
                   params_file="shallowwater.params")
    r = _Reader(raw)
    steps = int(r.take(np.int32, (1,))[0])
    sec = float(r.take(np.float64, (1,))[0])
    r.done()
    assert steps == nSteps
    return sec
