"""Writes tests/golden/ref_nodal.npz: for a set of degrees, quadratures and
forms, what the reference's own quadrature, nodal-storage and interpolation
routines (src/interp/Legendre.f90, Lagrange_2D/3D_Class.f90,
src/nodal/NodalStorage_2D/3D_Class.f90, compiled unmodified by oracle/ref.mk
into oracle/_ref/ref_nodal) computed on seeded random element fields.
Keys are prefixed "<case>/" with case = "N<N>_<gauss|lobatto>_<dg|cg>".

    make -f oracle/ref.mk && python tests/golden/make_ref_nodal.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_run as R  # noqa: E402

GAUSS, GAUSS_LOBATTO, DG, CG = 1, -1, 2000, 2001  # src/common/ModelFlags.f90
CASES = [(1, GAUSS, DG), (3, GAUSS, DG), (4, GAUSS, DG), (7, GAUSS, DG), (8, GAUSS, DG),
         (2, GAUSS_LOBATTO, CG), (5, GAUSS_LOBATTO, CG), (7, GAUSS_LOBATTO, CG), (6, GAUSS, CG),
         (5, GAUSS_LOBATTO, DG)]
NNEW = 9


def case_name(N, q, f):
    return f"N{N}_{'gauss' if q == GAUSS else 'lobatto'}_{'dg' if f == DG else 'cg'}"


def main():
    rng = np.random.default_rng(41)
    store = {}
    for N, q, f in CASES:
        n1 = N + 1
        out = R.run_nodal(N, q, f, NNEW, rng.standard_normal((n1, n1)), rng.standard_normal((n1, n1, n1)))
        for k, v in out.items():
            store[f"{case_name(N, q, f)}/{k}"] = np.asarray(v)
    path = os.path.join(HERE, "ref_nodal.npz")
    np.savez_compressed(path, **store)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
