"""Host set-up of the 3-D path (CPU only, no device calls): nodal storage,
the roll-off filter factor and the structured hex mesh of the C ABI against
the oracle."""
import numpy as np
import pytest

from oracle import pyoracle as po


@pytest.mark.parametrize("N", [1, 3, 5, 7])
def test_nodal3d_matches_oracle(N):
    from self_amd.nodal import NodalStorage3D
    ns, o = NodalStorage3D(N), po.Nodal3D(N)
    assert np.array_equal(ns.nodes, o.nodes) and np.array_equal(ns.qWeight, o.qWeight)
    assert np.array_equal(ns.lagLeft, o.lagLeft) and np.array_equal(ns.lagRight, o.lagRight)
    assert np.abs(ns.dMat - o.dMat).max() < 1e-13 * np.abs(o.dMat).max()


@pytest.mark.parametrize("N,nc", [(3, 2), (5, 3), (7, 3)])
def test_filter_factor_reproduces_the_dense_reference_filter(N, nc):
    from self_amd.nodal import NodalStorage3D, rolloff_filter1d
    ns = NodalStorage3D(N)
    F1 = rolloff_filter1d(N, nc, ns.nodes, ns.qWeight)
    F = po.rolloff_filter3d(N, nc, ns.nodes, ns.qWeight)
    assert np.abs(np.kron(F1, np.kron(F1, F1)) - F).max() < 1e-14


@pytest.mark.parametrize("periodic", [(True, True, True), (True, False, True), (False, False, False)])
def test_hexmesh_matches_oracle_bitwise(periodic):
    from self_amd.mesh import HexMesh
    from self_amd.nodal import NodalStorage3D
    N = 4
    ns = NodalStorage3D(N)
    m = HexMesh(3, 2, 2, N, ns.nodes, periodic=periodic, warp=0.05, warpSign=(1, -1, 1))
    om = po.HexMesh(3, 2, 2, N, ns.nodes, periodic=periodic, warp=0.05, warpSign=(1, -1, 1))
    assert np.array_equal(m.faces, om.faces)
    for a in ("x", "y", "z", "cov", "J", "nHat", "nLen", "bpos"):
        assert np.array_equal(getattr(m, a).ravel(), getattr(om, a)), a
    # every element side is on exactly one face
    sides = {(e1, s1) for e1, s1, *_ in m.faces} | {(e2, s2) for _, _, e2, s2, *_ in m.faces if e2 > 0}
    assert len(sides) == 6 * m.nElems
