"""Quadrature, nodal storage and the element operators against the
reference's own routines (src/interp/Legendre.f90, Lagrange_2D/3D_Class.f90,
src/nodal/NodalStorage_2D/3D_Class.f90, compiled unmodified, oracle/ref.mk),
fixture tests/golden/ref_nodal.npz (tests/golden/make_ref_nodal.py).

Checked on the CPU: the oracle (orc_interp.c) and the library's host set-up
(self_legendre_quadrature, self_nodal2d_build with the reference's
transposed dMatP, self_nodal3d_build, self_interpolation_matrix) for Gauss
and Gauss-Lobatto nodes, DG and CG forms, N = 1..8; the oracle's element
operators (ApplyDerivativeMatrix, CalculateAtBoundaries, CoarseToFine in
2-D and 3-D).  All of it is bitwise equal to the reference (both sides
built without fused multiply-add, the reference's order of operations kept).
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

Z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_nodal.npz"))
CASES = sorted({k.split("/")[0] for k in Z.files})


def case(name):
    return {k.split("/", 1)[1]: Z[k] for k in Z.files if k.startswith(name + "/")}


def close(a, b, tol=0.0):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1.0)


def test_fixture_covers_both_quadratures_and_forms():
    assert {(int(case(c)["quadrature"]), int(case(c)["form"])) for c in CASES} == {
        (1, 2000), (-1, 2001), (1, 2001), (-1, 2000)}


@pytest.mark.parametrize("name", CASES)
def test_oracle_quadrature_and_2d_storage_are_the_reference(name):
    g = case(name)
    N, q, f = int(g["N"]), int(g["quadrature"]), int(g["form"])
    x, w = po.quadrature(N, q)
    np.testing.assert_array_equal(x, g["nodes"])
    np.testing.assert_array_equal(w, g["weights"])
    ns = po.NodalStorage2D(N, q, f)
    np.testing.assert_array_equal(ns.dMatS, g["dMatS"])
    np.testing.assert_array_equal(ns.dMatP, g["dMatP"])
    np.testing.assert_array_equal(ns.qWeight, g["qWeightS"])
    for k in ("lagWest", "lagEast", "lagSouth", "lagNorth"):
        np.testing.assert_array_equal(getattr(ns, k), g[k], err_msg=k)


@pytest.mark.parametrize("name", CASES)
def test_library_host_setup_is_the_reference(name):
    """The C ABI's host set-up, called on the CPU (no device needed)."""
    from self_amd import nodal
    g = case(name)
    N, q, f = int(g["N"]), int(g["quadrature"]), int(g["form"])
    x, w = nodal.legendre_quadrature(N, q)
    np.testing.assert_array_equal(x, g["nodes"])
    np.testing.assert_array_equal(w, g["weights"])
    ns = nodal.NodalStorage2D(N, q, f, quirk_dmatp_transposed=True)
    np.testing.assert_array_equal(ns.dMatS, g["dMatS"])
    np.testing.assert_array_equal(ns.dMatP, g["dMatP"])
    for k in ("lagWest", "lagEast", "lagSouth", "lagNorth"):
        np.testing.assert_array_equal(getattr(ns, k), g[k], err_msg=k)
    n3 = nodal.NodalStorage3D(N, q, f)
    assert close(n3.dMat, g["dMat3"])
    bw = nodal.barycentric_weights(x)
    assert close(nodal.interpolation_matrix(x, bw, np.linspace(-1.0, 1.0, int(g["nNew"]) + 1)), g["Ts"])


@pytest.mark.parametrize("name", CASES)
def test_oracle_element_operators_are_the_reference(name):
    g = case(name)
    N, q, f = int(g["N"]), int(g["quadrature"]), int(g["form"])
    ns = po.NodalStorage2D(N, q, f)
    a, b = po.apply_derivative_2d(N, g["dMatS"], g["dMatP"], g["f2"])
    assert close(a, g["dFds2"]) and close(b, g["dFdp2"])
    for got, k in zip(po.boundary_interp_2d(N, ns.lagWest, ns.lagEast, ns.lagSouth, ns.lagNorth, g["f2"]),
                      ("fWest2", "fEast2", "fSouth2", "fNorth2")):
        assert close(got, g[k]), k
    d3 = po.apply_derivative_3d(N, g["dMat3"], g["f3"])
    for got, k in zip(d3, ("dFds3", "dFdp3", "dFdq3")):
        assert close(got, g[k]), k
    faces = po.boundary_interp_3d(N, ns.lagWest, ns.lagEast, g["f3"])  # [south, east, north, west, bottom, top]
    for i, k in enumerate(("fSouth3", "fEast3", "fNorth3", "fWest3", "fBottom3", "fTop3")):
        assert close(faces[i], g[k]), k
    x, _ = po.quadrature(N, q)
    T = po.interpolation_matrix(x, po.barycentric_weights(x), np.linspace(-1.0, 1.0, int(g["nNew"]) + 1))
    assert close(T, g["Ts"])
    assert close(po.coarse_to_fine_2d(g["f2"], g["Ts"], g["Ts"]), g["fine2"])
    assert close(po.coarse_to_fine_3d(g["f3"], g["Ts"]), g["fine3"])


def test_a_mistake_in_the_oracle_is_caught():
    """The other quadrature, or the untransposed DG dMatP, differs from the
    reference by far more than round-off."""
    g = case("N4_gauss_dg")
    x, _ = po.quadrature(4, po.GAUSS_LOBATTO)
    assert not close(x, g["nodes"], 1e-6)
    ns = po.NodalStorage2D(4)
    assert not close(ns.dMatCorrect, g["dMatP"], 1e-6)
