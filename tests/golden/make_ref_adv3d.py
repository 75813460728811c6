"""Writes tests/golden/ref_adv3d_<name>.npz: inputs and outputs of the
reference's own 3-D advection model with the adaptive roll-off filter
(src/highend/advection3d/Advection3DClass.f90, compiled unmodified by
oracle/ref.mk into oracle/_ref/ref_adv3d) on small seeded cases.

Each file holds the inputs (c, u, v, w, relax, rFac), the reference's nodal
arrays and default mesh, the tendency, filtered solution and limiter state
after GlobalTimeDerivative(t0, 1), and the solution and limiter state after
nSteps ForwardStepRK3 calls that follow; layouts are those of
tests/ref_run.py.  The dense filter matrix is kept only where it is small
(N = 3: 64 x 64).

    make -f oracle/ref.mk && python tests/golden/make_ref_adv3d.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_run as R  # noqa: E402

# name: (N, nX, nY, nZ, nTracers, nCutoff, dt, t0, nSteps, scale, seed, noise, keep filterMat)
CASES = {
    "n3_3x2x2": (3, 3, 2, 2, 2, 2, 2.0e-3, 0.3, 3, (1.0, 1.0, 1.0), 21, 1.0, True),
    "n5_2x2x2": (5, 2, 2, 2, 1, 3, 1.0e-3, 0.4, 2, (1.0, 0.75, 1.25), 22, 3.0, False),
}


def main():
    for name, (N, nX, nY, nZ, nq, nc, dt, t0, nSteps, scale, seed, noise, keepF) in CASES.items():
        inp = R.random_inputs_3d(N, nX, nY, nZ, nq, seed, noise)
        out = R.run_3d(N, nX, nY, nZ, nq, nc, dt, t0, nSteps, inp, scale=scale)
        out.update(seed=seed, noise=noise)
        if not keepF:
            del out["filterMat"]
        path = os.path.join(HERE, f"ref_adv3d_{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in out.items()})
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
