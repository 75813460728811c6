"""Pins of the oracle's quadrature, Lagrange interpolation and nodal storage
(oracle/orc_interp.c) against closed forms and numpy's independent
implementations -- chosen so that a dropped term, a transposed operand or a
wrong weight fails."""
import math

import numpy as np
import pytest
from numpy.polynomial import legendre as Leg

from oracle import pyoracle as po


@pytest.mark.parametrize("N", range(0, 16))
def test_gauss_nodes_weights_match_numpy(N):
    x, w = po.quadrature(N, po.GAUSS)
    xr, wr = Leg.leggauss(N + 1)
    order = np.argsort(x)
    assert np.allclose(x[order], xr, rtol=0, atol=2e-15)
    assert np.allclose(w[order], wr, rtol=0, atol=2e-15)


def test_gauss_n3_node_order_quirk():
    """For nP = 3 the reference's root loop (Legendre.f90:255, jX = 0..(nP+1)/2)
    revisits the middle root from the far side: the set is right, the order
    is +x, -x in the interior."""
    x, w = po.quadrature(3, po.GAUSS)
    a = math.sqrt(3.0 / 7.0 - 2.0 / 7.0 * math.sqrt(6.0 / 5.0))
    b = math.sqrt(3.0 / 7.0 + 2.0 / 7.0 * math.sqrt(6.0 / 5.0))
    assert np.allclose(x, [-b, a, -a, b], atol=1e-15)
    assert np.allclose(np.sort(x), [-b, -a, a, b], atol=1e-15)
    assert np.isclose(w.sum(), 2.0, atol=1e-15)


@pytest.mark.parametrize("N", range(1, 16))
def test_gauss_lobatto_closed_form(N):
    x, w = po.quadrature(N, po.GAUSS_LOBATTO)
    c = np.zeros(N + 1)
    c[-1] = 1.0
    roots = np.sort(Leg.legroots(Leg.legder(c))) if N > 1 else np.array([])
    xr = np.concatenate([[-1.0], roots, [1.0]])
    assert np.allclose(x, xr, atol=3e-15)
    wr = 2.0 / (N * (N + 1) * Leg.legval(xr, c) ** 2)
    assert np.allclose(w, wr, atol=3e-15)
    # exact for polynomials of degree 2N-1
    for k in range(2 * N):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-13


@pytest.mark.parametrize("n", [0, 1, 2, 5, 9])
def test_legendre_polynomial_matches_numpy(n):
    c = np.zeros(n + 1)
    c[-1] = 1.0
    for x in (-0.9, -0.3, 0.0, 0.45, 1.0):
        l, d = po.legendre(n, x)
        assert abs(l - Leg.legval(x, c)) < 1e-14
        assert abs(d - Leg.legval(x, Leg.legder(c))) < 1e-12


def _nodes(N, kind=po.GAUSS):
    return po.quadrature(N, kind)[0]


@pytest.mark.parametrize("N", [1, 3, 5, 7, 10])
def test_barycentric_weights_closed_form(N):
    x = _nodes(N)
    bw = po.barycentric_weights(x)
    ref = np.array([1.0 / np.prod([x[j] - x[k] for k in range(N + 1) if k != j]) for j in range(N + 1)])
    assert np.allclose(bw, ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("N", [2, 5, 7])
def test_lagrange_polynomials_match_product_form(N):
    x = _nodes(N)
    bw = po.barycentric_weights(x)
    for xe in (-1.0, -0.77, 0.123, 1.0):
        l = po.lagrange_at(x, bw, xe)
        ref = np.array([np.prod([(xe - x[k]) / (x[j] - x[k]) for k in range(N + 1) if k != j])
                        for j in range(N + 1)])
        assert np.allclose(l, ref, atol=1e-13)
    # at a node it is the unit vector
    assert np.array_equal(po.lagrange_at(x, bw, x[1]), np.eye(N + 1)[1])


@pytest.mark.parametrize("kind", [po.GAUSS, po.GAUSS_LOBATTO])
@pytest.mark.parametrize("N", [1, 3, 5, 7, 11])
def test_derivative_matrix_exact_on_polynomials(N, kind):
    x = _nodes(N, kind)
    bw = po.barycentric_weights(x)
    D = po.derivative_matrix(x, bw)
    assert np.abs(D @ np.ones(N + 1)).max() < 1e-13          # D*1 = 0 (negative-sum trick)
    for k in range(1, N + 1):
        assert np.abs(D @ x ** k - k * x ** (k - 1)).max() < 1e-11 * (N + 1) ** 2
    # entries equal l_j'(x_k) from numpy's polynomial fit
    for j in range(N + 1):
        e = np.zeros(N + 1)
        e[j] = 1.0
        p = np.polynomial.Polynomial.fit(x, e, N).convert()
        assert np.allclose(D[:, j], p.deriv()(x), atol=1e-9)
    # the reference's off-node EvaluateDerivative agrees
    f = np.sin(x)
    assert abs(po.lagrange_derivative(x, bw, 0.3141, f) -
               np.polynomial.Polynomial.fit(x, f, N).convert().deriv()(0.3141)) < 1e-9


@pytest.mark.parametrize("N", [3, 7])
def test_interpolation_matrix_exact_on_polynomials(N):
    x = _nodes(N)
    bw = po.barycentric_weights(x)
    z = np.linspace(-1, 1, 11)
    T = po.interpolation_matrix(x, bw, z)
    for k in range(N + 1):
        assert np.abs(T @ x ** k - z ** k).max() < 1e-13
    assert np.allclose(T.sum(axis=1), 1.0, atol=1e-14)


def test_coarse_to_fine_2d_is_transposed_tensor_interpolation():
    """CoarseToFine_2D (Lagrange_2D_Class.f90:508-510) returns fNew(a, b) =
    f(s_new(b), p_new(a)): the output is transposed relative to fOld."""
    N, M = 5, 9
    x = _nodes(N)
    bw = po.barycentric_weights(x)
    z = np.linspace(-1, 1, M + 1)
    T = po.interpolation_matrix(x, bw, z)
    f = lambda s, p: s ** 2 * p - 3 * p ** 3 + 0.5 * s ** 4
    S, P = np.meshgrid(x, x, indexing="ij")
    fnew = po.coarse_to_fine_2d(f(S, P), T, T)
    Sn, Pn = np.meshgrid(z, z, indexing="ij")
    assert np.abs(fnew - f(Sn, Pn).T).max() < 1e-12


def test_coarse_to_fine_3d_tensor_interpolation():
    N, M = 4, 6
    x = _nodes(N)
    bw = po.barycentric_weights(x)
    z = np.linspace(-1, 1, M + 1)
    T = po.interpolation_matrix(x, bw, z)
    f = lambda s, p, q: s * p ** 2 - q ** 3 + s * q + 1.0
    S, P, Q = np.meshgrid(x, x, x, indexing="ij")
    fnew = po.coarse_to_fine_3d(f(S, P, Q), T)
    Sn, Pn, Qn = np.meshgrid(z, z, z, indexing="ij")
    assert np.abs(fnew - f(Sn, Pn, Qn)).max() < 1e-12


@pytest.mark.parametrize("N", [1, 3, 5, 7, 10])
def test_nodal2d_dg_weak_operator_and_lift(N):
    """dMatS = -W^-1 D^T W, and Dhat f + lift(f(+-1)) = f' for degree <= N."""
    ns = po.NodalStorage2D(N, po.GAUSS, po.DG)
    x, w = ns.nodes, ns.qWeight
    D = po.derivative_matrix(x, po.barycentric_weights(x))
    ref = -(D.T * w[None, :]) / w[:, None]
    assert np.allclose(ns.dMatS, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert np.array_equal(ns.dMatP, ns.dMatS.T)        # quirk Q1 (NodalStorage_2D_Class.f90:209-218)
    assert np.array_equal(ns.dMatCorrect, ns.dMatS)
    for k in range(N + 1):
        f = x ** k
        lift = (1.0 ** k * ns.lagEast - (-1.0) ** k * ns.lagWest) / w
        assert np.abs(ns.dMatS @ f + lift - k * x ** max(k - 1, 0) * (k > 0)).max() < 1e-10 * (N + 1) ** 2
    # the transposed operator is wrong in general (the quirk is observable)
    if N >= 2:
        f = x ** 2
        lift = (ns.lagEast - ns.lagWest) / w
        assert np.abs(ns.dMatP @ f + lift - 2 * x).max() > 1e-3
    # boundary interpolants reproduce polynomial end values
    for k in range(N + 1):
        assert abs(ns.lagEast @ x ** k - 1.0) < 1e-13
        assert abs(ns.lagWest @ x ** k - (-1.0) ** k) < 1e-13


def test_nodal2d_cg_stores_plain_derivative():
    ns = po.NodalStorage2D(5, po.GAUSS_LOBATTO, po.CG)
    D = po.derivative_matrix(ns.nodes, po.barycentric_weights(ns.nodes))
    assert np.array_equal(ns.dMatS, D) and np.array_equal(ns.dMatP, D)


@pytest.mark.parametrize("N", [3, 7])
def test_apply_derivative_and_boundaries_2d_exact(N):
    ns = po.NodalStorage2D(N, po.GAUSS, po.CG)
    x = ns.nodes
    S, P = np.meshgrid(x, x, indexing="ij")
    f = S ** 2 * P + P ** 3 - 2 * S
    ds, dp = po.apply_derivative_2d(N, ns.dMatS, ns.dMatP, f)
    assert np.abs(ds - (2 * S * P - 2)).max() < 1e-11
    assert np.abs(dp - (S ** 2 + 3 * P ** 2)).max() < 1e-11
    fw, fe, fs, fn = po.boundary_interp_2d(N, ns.lagWest, ns.lagEast, ns.lagSouth, ns.lagNorth, f)
    assert np.abs(fw - (1 * x + x ** 3 + 2)).max() < 1e-12
    assert np.abs(fe - (x + x ** 3 - 2)).max() < 1e-12
    assert np.abs(fs - (-(x ** 2) - 1 - 2 * x)).max() < 1e-12
    assert np.abs(fn - (x ** 2 + 1 - 2 * x)).max() < 1e-12


def test_apply_derivative_and_boundaries_3d_exact():
    N = 5
    x = _nodes(N)
    D = po.derivative_matrix(x, po.barycentric_weights(x))
    S, P, Q = np.meshgrid(x, x, x, indexing="ij")
    f = S * P ** 2 + Q ** 3 - S * Q
    a, b, c = po.apply_derivative_3d(N, D, f)
    assert np.abs(a - (P ** 2 - Q)).max() < 1e-11
    assert np.abs(b - 2 * S * P).max() < 1e-11
    assert np.abs(c - (3 * Q ** 2 - S)).max() < 1e-11
    bw = po.barycentric_weights(x)
    lL, lR = po.lagrange_at(x, bw, -1.0), po.lagrange_at(x, bw, 1.0)
    faces = po.boundary_interp_3d(N, lL, lR, f)
    A, B = np.meshgrid(x, x, indexing="ij")  # (first in-face index, second)
    exp = {0: A * 1 + B ** 3 - A * B,        # south p=-1: (s, q)
           1: A ** 2 + B ** 3 - B,           # east  s=+1: (p, q)
           2: A + B ** 3 - A * B,            # north p=+1
           3: -A ** 2 + B ** 3 + B,          # west  s=-1
           4: A * B ** 2 - 1 + A,            # bottom q=-1: (s, p)
           5: A * B ** 2 + 1 - A}            # top    q=+1
    for k, v in exp.items():
        assert np.abs(faces[k] - v).max() < 1e-12, k
