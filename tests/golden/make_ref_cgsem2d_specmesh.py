"""Writes tests/golden/ref_cgsem2dspec_annulus_n4.npz: the reference's own
CGsemElliptic_2D (src/cgsem/CGsemElliptic_2D_Class.f90, compiled unmodified
by oracle/ref.mk into oracle/_ref/ref_cgsem2d) on an ISM-V2 mesh file
(SpaceManagement % SpecMeshFile; BuildQuadMesh :169-179 reads it, boundary
edges become DIRICHLET): a curved quarter annulus from tests/specmesh_gen.py,
zero start and zero Dirichlet values, identity preconditioner.

Same keys as tests/golden/ref_cgsem2d_*.npz (nX = element count, nY = 1),
plus the mesh file's text (uint8).

    make -f oracle/ref.mk && python tests/golden/make_ref_cgsem2d_specmesh.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_run as R  # noqa: E402
import specmesh_gen as G  # noqa: E402

N, NR, NT, GDEG, SEED, MAXIT, TOL = 4, 3, 4, 4, 31, 2000, 1.0e-12
NAMES = ("DIRICHLET", "NO_NORMAL_FLOW", "DIRICHLET", "NO_NORMAL_FLOW")


def mesh_text():
    return G.quarter_annulus(NR, NT, GDEG, seed=SEED, names=NAMES)


def main():
    text = mesh_text()
    E = NR * NT
    inp = R.random_inputs_cgsem2d(N, E, 1, SEED)
    inp["solution"][:] = 0.0
    out = R.run_cgsem2d(N, E, 1, MAXIT, TOL, 0.0, inp, mesh_text=text)
    out.update(seed=SEED, mesh=np.frombuffer(text.encode(), dtype=np.uint8))
    path = os.path.join(HERE, "ref_cgsem2dspec_annulus_n4.npz")
    np.savez_compressed(path, **{k: np.asarray(v) for k, v in out.items()})
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
