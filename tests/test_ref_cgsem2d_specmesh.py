"""The reference's CGSEM-2D on an ISM-V2 mesh file: the mesh it builds.

tests/golden/ref_cgsem2dspec_annulus_n4.npz holds what the compiled
reference CGsemElliptic_2D (SpecMeshFile, BuildQuadMesh :169-179: read, no
scaling, boundary edges NO_NORMAL_FLOW -> DIRICHLET) built and solved on a
curved quarter annulus.  The library's reader on the GLL nodes gives the
same edge table (after that remap) and bitwise the same geometry; the GPU
solve is tests/test_ref_cgsem2d_specmesh_gpu.py.
"""
import os

import numpy as np
import pytest

import ref_run as R
import specmesh_gen as G

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_cgsem2dspec_annulus_n4.npz")


def read_on_gll(g, tmp_path):
    from self_amd import _lib as L
    from self_amd.mesh import SpecMesh
    N = int(g["N"])
    x, w = np.zeros(N + 1), np.zeros(N + 1)
    L.check(L.lib().self_legendre_quadrature(N, L.GAUSS_LOBATTO, L.ptr(x), L.ptr(w)))
    path = tmp_path / "m.mesh"
    path.write_bytes(g["mesh"].tobytes())
    return SpecMesh(path, N, x)


def mismatches(m, g):
    e = m.edges.copy()
    e[e[:, 2] == -100, 2] = -200  # BuildQuadMesh: NO_NORMAL_FLOW -> DIRICHLET
    bad = [] if np.array_equal(e, g["edges"]) else ["edges"]
    for k in ("x", "y", "dxds", "dxdp", "dyds", "dydp", "J"):
        if not np.array_equal(getattr(m, k).transpose(2, 1, 0), g[k]):
            bad.append(k)
    return bad


def test_reader_gives_the_reference_solvers_mesh(tmp_path):
    g = dict(np.load(GOLDEN))
    m = read_on_gll(g, tmp_path)
    assert mismatches(m, g) == []
    assert (m.edges[:, 3] < 0).any()


@pytest.mark.skipif(not R.available("cgsem2d"), reason="oracle/_ref/ref_cgsem2d not built (needs the reference sources)")
def test_live_reference_on_another_annulus(tmp_path):
    N, E = 3, 2 * 3
    text = G.quarter_annulus(2, 3, 2, seed=9, names=("NO_NORMAL_FLOW",) * 4)
    inp = R.random_inputs_cgsem2d(N, E, 1, 5)
    g = R.run_cgsem2d(N, E, 1, 300, 1e-10, 0.0, inp, mesh_text=text)
    g["mesh"] = np.frombuffer(text.encode(), dtype=np.uint8)
    assert mismatches(read_on_gll(g, tmp_path), g) == []
