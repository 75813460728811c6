"""Writes tests/golden/ref_sw2d_<name>.npz: inputs and outputs of the
reference's conservative shallow-water class as the planewave example ships
it (examples/shallowwater/planewave/localmods/ConservativeShallowWaterClass.f90,
compiled unmodified by oracle/ref.mk into oracle/_ref/ref_sw2d) on small
seeded cases on its default walled mesh.

Each file holds the random inputs (sol, h, f), the reference's nodal arrays
and mesh, the bathymetry after CalculateBathymetryGradient, the tendency and
boundary fluxes of GlobalTimeDerivative(t0) and the solution after nSteps
ForwardStepRK3 calls; layouts are those of tests/ref_run.py run_sw2d.

    make -f oracle/ref.mk && python tests/golden/make_ref_sw2d.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_run as R  # noqa: E402

# name: (N, nX, nY, g, dt, t0, nSteps, xScale, yScale, seed)
CASES = {
    "n5_4x3": (5, 4, 3, 9.81, 1.0e-4, 0.5, 4, 2.0, 1.5, 21),
    "n7_3x3": (7, 3, 3, 9.81, 5.0e-5, 1.0, 3, 1.0, 1.0, 22),
    "n3_5x2": (3, 5, 2, 1.0, 5.0e-4, 0.0, 5, 3.0, 1.0, 23),
}


def main():
    for name, (N, nX, nY, g, dt, t0, nSteps, xs, ys, seed) in CASES.items():
        inp = R.random_inputs_sw2d(N, nX, nY, seed)
        out = R.run_sw2d(N, nX, nY, g, dt, t0, nSteps, inp, xScale=xs, yScale=ys)
        out.update(seed=seed)
        path = os.path.join(HERE, f"ref_sw2d_{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in out.items()})
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
