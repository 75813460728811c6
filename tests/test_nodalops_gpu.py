"""GPU parity of the batched nodal library operators (nodalops.hip)
against the oracle's single-element restatements (oracle/orc_interp.c),
applied element by element.  Tolerance: 1e-13 relative to the largest
output (summation order differs)."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _close(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() < 1e-13 * scale, np.abs(a - b).max() / scale


@pytest.mark.parametrize("N,nE,quirk", [(7, 100, False), (3, 17, True), (1, 5, False), (11, 9, False)])
def test_apply_derivative_and_boundaries_2d(gpu_lib, N, nE, quirk):
    from self_amd import nodal
    ns = nodal.NodalStorage2D(N, quirk_dmatp_transposed=quirk)
    f = np.random.default_rng(N).standard_normal((nE, N + 1, N + 1))
    ds, dp = nodal.apply_derivative_2d(ns.dMatS, ns.dMatP, f)
    w, e, s, n = nodal.boundary_interp_2d(ns.lagWest, ns.lagEast, ns.lagSouth, ns.lagNorth, f)
    for k in range(nE):
        rs, rp = po.apply_derivative_2d(N, ns.dMatS, ns.dMatP, f[k].T)
        _close(ds.reshape(f.shape)[k].T, rs)
        _close(dp.reshape(f.shape)[k].T, rp)
        ref = po.boundary_interp_2d(N, ns.lagWest, ns.lagEast, ns.lagSouth, ns.lagNorth, f[k].T)
        for got, r in zip((w, e, s, n), ref):
            _close(got[k], r)


@pytest.mark.parametrize("N,nE", [(7, 20), (4, 33), (2, 3)])
def test_apply_derivative_and_boundaries_3d(gpu_lib, N, nE):
    from self_amd import nodal
    ns = nodal.NodalStorage3D(N)
    f = np.random.default_rng(N + 1).standard_normal((nE, N + 1, N + 1, N + 1))
    outs = nodal.apply_derivative_3d(ns.dMat, f)
    faces = nodal.boundary_interp_3d(ns.lagLeft, ns.lagRight, f)
    for k in range(nE):
        fe = f[k].transpose(2, 1, 0)  # (i, j, k) Fortran order
        ref = po.apply_derivative_3d(N, ns.dMat, fe)
        for got, r in zip(outs, ref):
            _close(got.reshape(f.shape)[k].transpose(2, 1, 0), r)
        rf = po.boundary_interp_3d(N, ns.lagLeft, ns.lagRight, fe)   # [face][a][b]
        _close(faces[k].transpose(0, 2, 1), rf)


@pytest.mark.parametrize("N,M,nE", [(7, 10, 12), (3, 5, 4)])
def test_coarse_to_fine_2d_and_3d(gpu_lib, N, M, nE):
    from self_amd import nodal
    ns = nodal.NodalStorage2D(N)
    bw = po.barycentric_weights(ns.nodes)
    T = po.interpolation_matrix(ns.nodes, bw, np.linspace(-1, 1, M + 1))
    f2 = np.random.default_rng(3).standard_normal((nE, N + 1, N + 1))
    g2 = nodal.coarse_to_fine_2d(T, T, f2)
    for k in range(nE):
        _close(g2[k].T, po.coarse_to_fine_2d(f2[k].T, T, T))
    f3 = np.random.default_rng(4).standard_normal((nE, N + 1, N + 1, N + 1))
    g3 = nodal.coarse_to_fine_3d(T, f3)
    for k in range(nE):
        _close(g3[k].transpose(2, 1, 0), po.coarse_to_fine_3d(f3[k].transpose(2, 1, 0), T))


def test_empty_batch_and_errors(gpu_lib):
    from self_amd import nodal
    from self_amd import _lib as L
    ns = nodal.NodalStorage2D(3)
    ds, dp = nodal.apply_derivative_2d(ns.dMatS, ns.dMatP, np.zeros((0, 4, 4)))
    assert ds.size == 0 and dp.size == 0
    with pytest.raises(ValueError):
        nodal.apply_derivative_2d(ns.dMatS, ns.dMatP, np.zeros(17))
    assert L.lib().self_nodal2d_apply_derivative(-1, 1, None, None, None, None, None) == L.SELF_ERR_ARG
