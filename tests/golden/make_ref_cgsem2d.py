"""Writes tests/golden/ref_cgsem2d_<name>.npz: inputs and outputs of the
reference's own CG spectral-element elliptic solver (src/cgsem/
CGsemElliptic_2D_Class.f90, compiled unmodified by oracle/ref.mk into
oracle/_ref/ref_cgsem2d) with pcIterates = 0, on small seeded cases, and
tests/golden/ref_cgsempc_<name>.npz: the same with its FEM preconditioner
(pcIterates > 0, zero initial guess).

Each file holds the inputs (fluxCoeff, source, solution, u, v, the Dirichlet
value), the reference's GLL derivative matrix and weights, its default mesh
(edges, node types, geometry), Mask/UnMask/GlobalSum of v, MatrixAction of u,
DotProduct(u, v), Residual, and the ioerr, residual history and solution of
Solve; layouts are those of tests/ref_run.py.

    make -f oracle/ref.mk && python tests/golden/make_ref_cgsem2d.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_run as R  # noqa: E402

# name: (N, nX, nY, maxIters, tolerance, dirValue, seed, zero start)
# A zero start also zeroes the test vector u on the boundary, so that its
# MatrixAction is the one a Krylov search direction sees (Dirichlet nodes 0).
CASES = {
    "n4_3x2_random": (4, 3, 2, 200, 1.0e-8, 0.3, 7, False),
    "n5_3x3_zero": (5, 3, 3, 2000, 1.0e-12, 0.0, 8, True),
}


# with the reference's FEM preconditioner (pcIterates > 0):
# name: (N, nX, nY, maxIters, tol, seed, pcIterates, pcTolerance)
PC_CASES = {
    "n5_3x3_pc3": (5, 3, 3, 300, 1.0e-10, 8, 3, 1.0e-4),
    "n4_4x3_pc50": (4, 4, 3, 300, 1.0e-10, 9, 50, 1.0e-4),
}


def main_pc():
    for name, (N, nX, nY, maxIters, tol, seed, pcIt, pcTol) in PC_CASES.items():
        inp = R.random_inputs_cgsem2d(N, nX, nY, seed)
        inp["solution"][:] = 0.0
        out = R.run_cgsem2d(N, nX, nY, maxIters, tol, 0.0, inp, pcIterates=pcIt, pcTolerance=pcTol)
        out.update(seed=seed)
        path = os.path.join(HERE, f"ref_cgsempc_{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in out.items()})
        print(path, os.path.getsize(path))


def main():
    for name, (N, nX, nY, maxIters, tol, dirValue, seed, zero) in CASES.items():
        inp = R.random_inputs_cgsem2d(N, nX, nY, seed)
        if zero:
            inp["solution"][:] = 0.0
            u = inp["u"].reshape(N + 1, N + 1, nY, nX)  # element e = iX + nX iY
            u[0, :, :, 0] = 0.0
            u[N, :, :, nX - 1] = 0.0
            u[:, 0, 0, :] = 0.0
            u[:, N, nY - 1, :] = 0.0
        out = R.run_cgsem2d(N, nX, nY, maxIters, tol, dirValue, inp)
        out.update(seed=seed)
        path = os.path.join(HERE, f"ref_cgsem2d_{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in out.items()})
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
    main_pc()
