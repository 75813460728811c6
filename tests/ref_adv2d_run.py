"""Runs the reference's own 2-D advection-diffusion model (oracle/_ref/ref_adv2d,
built by oracle/ref.mk from src/highend/advection-diffusion2d/AdvectionClass.f90
and its modules) on one case and reads back what it computed.

The file formats are the ones documented at the top of oracle/ref_adv2d.f90.
Inputs are seeded random fields: the comparison is of the discrete operator,
which needs no smoothness.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_adv2d")


def available() -> bool:
    return os.access(REF_BIN, os.X_OK)


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


def _params_text(N, nX, nY, nTracers, kappa, dt, xScale, yScale):
    return ("&TracerSetup\n"
            f" nTracers = {nTracers},\n kappa = {kappa!r},\n/\n"
            "&TimeManagement\n"
            f" dt = {dt!r},\n iterInit = 0,\n nTimeSteps = 0,\n dumpFreq = 1,\n/\n"
            "&SpaceManagement\n"
            f" polyDeg = {N},\n nXElem = {nX},\n nYElem = {nY},\n nPlot = 4,\n"
            f" xScale = {xScale!r},\n yScale = {yScale!r},\n/\n")


def run(N, nX, nY, nTracers, kappa, dt, t0, nSteps, inputs, xScale=1.0, yScale=1.0, timeout=120):
    """Returns a dict: the inputs, the reference's nodal arrays (dMatS,
    dMatP (n1, n1), qWeightS, qWeightP, lagWest/East/South/North), its mesh
    (edges (nEdges, 6) int32, x, y, dxds, dxdp, dyds, dydp, J (n1, n1, E),
    nHat (2, n1, 4, E), nLen (n1, 4, E)), the tendency and gradients after
    GlobalTimeDerivative(t0) and the solution after nSteps RK3 steps."""
    if not available():
        raise FileNotFoundError(REF_BIN)
    n1, E = N + 1, nX * nY
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "advection.params"), "w") as f:
            f.write(_params_text(N, nX, nY, nTracers, kappa, dt, xScale, yScale))
        with open(os.path.join(d, "ref_in.bin"), "wb") as f:
            f.write(np.int32(nSteps).tobytes())
            f.write(np.float64(t0).tobytes())
            for k in ("c", "u", "v", "relax", "rFac"):
                f.write(np.asfortranarray(inputs[k], dtype=np.float64).tobytes(order="F"))
        p = subprocess.run([REF_BIN], cwd=d, capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"ref_adv2d failed ({p.returncode}): {p.stderr[-2000:]}")
        raw = open(os.path.join(d, "ref_out.bin"), "rb").read()

    pos = 0

    def take(dtype, shape):
        nonlocal pos
        cnt = int(np.prod(shape))
        a = np.frombuffer(raw, dtype=dtype, count=cnt, offset=pos)
        pos += cnt * np.dtype(dtype).itemsize
        return a.reshape(shape, order="F").copy()

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
    assert pos == len(raw), (pos, len(raw))
    out.update(N=N, nX=nX, nY=nY, nTracers=nTracers, kappa=kappa, dt=dt, t0=t0, nSteps=nSteps)
    return out
