"""Writes tests/golden/ref_adv2d_<name>.npz: inputs and outputs of the
reference's own 2-D advection-diffusion model (src/highend/advection-
diffusion2d/AdvectionClass.f90, compiled unmodified by oracle/ref.mk into
oracle/_ref/ref_adv2d) on small seeded cases.

Each file holds the random inputs (c, u, v, relax, rFac), the reference's
nodal arrays and mesh (default walled mesh with its inflow flags), the
tendency and gradients of GlobalTimeDerivative(t0) and the solution after
nSteps ForwardStepRK3 calls; layouts are those of tests/ref_adv2d_run.py.

    make -f oracle/ref.mk && python tests/golden/make_ref_adv2d.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_run as R  # noqa: E402

# name: (N, nX, nY, nTracers, kappa, dt, t0, nSteps, xScale, yScale, seed)
CASES = {
    "n5_4x3": (5, 4, 3, 2, 1.0e-2, 1.0e-3, 0.3, 4, 1.0, 1.0, 11),
    "n7_3x2": (7, 3, 2, 3, 0.0, 5.0e-4, 1.2, 3, 2.0, 0.5, 12),
}


def main():
    for name, (N, nX, nY, nq, kappa, dt, t0, nSteps, xs, ys, seed) in CASES.items():
        inp = R.random_inputs(N, nX, nY, nq, seed)
        out = R.run(N, nX, nY, nq, kappa, dt, t0, nSteps, inp, xScale=xs, yScale=ys)
        out.update(xScale=xs, yScale=ys, seed=seed)
        path = os.path.join(HERE, f"ref_adv2d_{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in out.items()})
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
