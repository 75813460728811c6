"""Writes tests/golden/ref_adv2dspec_annulus_n5.npz: the reference's own
2-D advection-diffusion model (src/highend/advection-diffusion2d/
AdvectionClass.f90, compiled unmodified by oracle/ref.mk into
oracle/_ref/ref_adv2d) run on an ISM-V2 mesh file (SpaceManagement %
SpecMeshFile; BuildQuadMesh reads it and applies ScaleTheMesh, :303-330)
instead of its default mesh: a curved quarter annulus from
tests/specmesh_gen.py with InflowOne on the inner arc, InflowTwo on the
outer and NO_NORMAL_FLOW on the straight sides.

Same keys as tests/golden/ref_adv2d_*.npz (nX = element count, nY = 1),
plus the mesh file's text (uint8).

    make -f oracle/ref.mk && python tests/golden/make_ref_adv2d_specmesh.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_run as R  # noqa: E402
import specmesh_gen as G  # noqa: E402

N, NR, NT, GDEG, SEED = 5, 3, 4, 4, 21
NQ, KAPPA, DT, T0, NSTEPS, XS, YS = 2, 5.0e-3, 1.0e-3, 0.4, 4, 2.0, 0.5
NAMES = ("InflowOne", "InflowTwo", "NO_NORMAL_FLOW", "NO_NORMAL_FLOW")


def mesh_text():
    return G.quarter_annulus(NR, NT, GDEG, seed=SEED, names=NAMES)


def main():
    text = mesh_text()
    E = NR * NT
    inp = R.random_inputs(N, E, 1, NQ, SEED)
    out = R.run(N, E, 1, NQ, KAPPA, DT, T0, NSTEPS, inp, xScale=XS, yScale=YS, mesh_text=text)
    out.update(xScale=XS, yScale=YS, seed=SEED, mesh=np.frombuffer(text.encode(), dtype=np.uint8))
    path = os.path.join(HERE, "ref_adv2dspec_annulus_n5.npz")
    np.savez_compressed(path, **{k: np.asarray(v) for k, v in out.items()})
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
