"""Writes tests/golden/ref_krylov_<name>.npz: dense systems and what the
reference's own Krylov loops (src/iterativesolve/IterativeSolvers_Class.f90
Solve_ConjugateGradient, Solve_PCConjugateGradient, Solve_GMRES,
Solve_PCGMRES, compiled unmodified by oracle/ref.mk into
oracle/_ref/ref_krylov) computed on them: ioerr, the residual history and
the solution.  The operator callbacks are those of oracle/ref_krylov.f90.

    make -f oracle/ref.mk && python tests/golden/make_ref_krylov.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_run as R  # noqa: E402


def spd(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * np.linspace(1.0, 60.0, n)) @ Q.T
    A = 0.5 * (A + A.T) + np.diag(rng.uniform(0.0, 5.0, n))  # unequal diagonal: Jacobi is not a scaling of I
    return A, rng.standard_normal(n), 0.1 * rng.standard_normal(n)


def nonsym(n, seed):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) * 0.4 + np.diag(rng.uniform(0.0, 2.0, n))
    return A, rng.standard_normal(n), 0.1 * rng.standard_normal(n)


# name: (system, n, seed, method, maxIters, mInner, tolerance, zero initial guess)
CASES = {
    "cg_spd40": (spd, 40, 31, "cg", 200, 1, 1.0e-7, False),
    "pcg_spd40": (spd, 40, 31, "pcg", 200, 1, 1.0e-7, False),
    "cg_spd40_x0": (spd, 40, 32, "cg", 200, 1, 1.0e-7, True),
    "gmres_ns40_m8": (nonsym, 40, 33, "gmres", 10, 8, 1.0e-11, False),
    "pcgmres_ns40_m8": (nonsym, 40, 33, "pcgmres", 10, 8, 1.0e-11, False),
    "gmres_ns30_notconv": (nonsym, 30, 34, "gmres", 2, 5, 1.0e-30, True),
}


def main():
    for name, (system, n, seed, method, maxIters, mInner, tol, zero) in CASES.items():
        A, b, x0 = system(n, seed)
        if zero:
            x0 = np.zeros(n)
        out = R.run_krylov(A, b, x0, method, maxIters, mInner, tol)
        out.update(A=A, b=b, x0=x0, method=method, maxIters=maxIters, mInner=mInner, tolerance=tol, seed=seed)
        path = os.path.join(HERE, f"ref_krylov_{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in out.items()})
        print(path, os.path.getsize(path), "ioerr", out["ioerr"], "its", np.count_nonzero(out["resi"]) - 1)


if __name__ == "__main__":
    main()
