"""Writes tests/golden/ref_pickup_adv2d_n3_3x2.npz: fields and the pickup
file the reference's own WritePickup_Advection (src/highend/
advection-diffusion2d/AdvectionClass.f90:2040-2103, compiled unmodified by
oracle/ref.mk into oracle/_ref/ref_pickup) wrote for them.

The fixture holds the inputs (c, u, v, relax, rFac, element-last as the
reference stores them) and the file's bytes (uint8).

    make -f oracle/ref.mk && python tests/golden/make_ref_pickup.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_run as R  # noqa: E402

N, NX, NY, NQ, ITER, SEED = 3, 3, 2, 2, 7, 11


def main():
    inp = R.random_inputs(N, NX, NY, NQ, SEED)
    data = R.run_pickup_write(N, NX, NY, NQ, ITER, inp)
    path = os.path.join(HERE, "ref_pickup_adv2d_n3_3x2.npz")
    np.savez_compressed(path, N=N, nX=NX, nY=NY, nTracers=NQ, iter=ITER, seed=SEED,
                        pickup=np.frombuffer(data, dtype=np.uint8), **inp)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
