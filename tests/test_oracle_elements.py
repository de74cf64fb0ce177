"""Pins of the oracle's reference elements and quadrature
(tests/test_p3_accuracy.py:15-79 of the reference, extended to P1/P2 and to
every element type), plus an independent construction of the same bases."""

import itertools
from math import factorial

import numpy as np
import pytest

from oracle import elements as oe
from hybrid import elements as pe

TYPES = ("hexahedron", "wedge", "tetra")
NODES = {("hexahedron", 1): 8, ("hexahedron", 2): 27, ("hexahedron", 3): 64,
         ("wedge", 1): 6, ("wedge", 2): 18, ("wedge", 3): 40,
         ("tetra", 1): 4, ("tetra", 2): 10, ("tetra", 3): 20}


@pytest.mark.parametrize("t,p", list(itertools.product(TYPES, (1, 2, 3))))
def test_kronecker_unity_mass(t, p):
    ref = oe.make_reference_element(t, p)
    n = NODES[(t, p)]
    assert ref.nodes.shape == (n, 3)
    assert np.abs(ref.basis(ref.nodes) - np.eye(n)).max() < 1e-12
    assert np.abs(ref.basis(ref.volume_points).sum(-1) - 1).max() < 1e-12
    assert np.abs(ref.basis_gradients(ref.volume_points).sum(-2)).max() < 1e-12
    M = ref.mass_matrix
    assert np.abs(M - M.T).max() < 1e-14
    assert np.all(np.linalg.eigvalsh(M) > 0)
    assert np.abs(M @ ref.inverse_mass_matrix - np.eye(n)).max() < 1e-11


def _simplex_moment(a, b, c):
    """Integral of x^a y^b z^c over the unit tetrahedron."""
    return factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)


def _tri_moment(a, b):
    return factorial(a) * factorial(b) / factorial(a + b + 2)


def test_p3_tetra_cubic_known_answer():
    """tests/test_p3_accuracy.py:50-79: 1 + 2x - 3y + 4z + 5x^2 - 6xy + 7z^3."""
    ref = oe.make_reference_element("tetra", 3)
    x, y, z = ref.volume_points.T
    f = 1 + 2 * x - 3 * y + 4 * z + 5 * x ** 2 - 6 * x * y + 7 * z ** 3
    exact = 1 / 6 + 2 / 24 - 3 / 24 + 4 / 24 + 5 / 60 - 6 / 120 + 7 / 120
    assert abs(np.sum(ref.volume_weights * f) - exact) < 1e-14


@pytest.mark.parametrize("t,p", list(itertools.product(TYPES, (2, 3))))
def test_volume_rule_exactness(t, p):
    """Exact on the monomials the rule must integrate (degree 2*order-1 per
    collapsed direction) against closed-form moments."""
    X, W = oe.volume_quadrature(t, p)
    for a, b, c in itertools.product(range(4), repeat=3):
        if t == "hexahedron":
            if max(a, b, c) > (5 if p == 3 else 3):
                continue
            exact = 1 / ((a + 1) * (b + 1) * (c + 1))
        elif t == "wedge":
            if a + b > 5 or c > 5:
                continue
            exact = _tri_moment(a, b) / (c + 1)
        else:
            if a + b + c > 5:
                continue
            exact = _simplex_moment(a, b, c)
        got = np.sum(W * X[:, 0] ** a * X[:, 1] ** b * X[:, 2] ** c)
        assert abs(got - exact) < 1e-14, (a, b, c)


@pytest.mark.parametrize("ft,p", list(itertools.product(("quad", "triangle"), (1, 2, 3))))
def test_face_rule(ft, p):
    S, W = oe.face_quadrature(ft, p)
    assert abs(W.sum() - 1) < 1e-14
    assert np.abs(S.sum(1) - 1).max() < 1e-14
    if ft == "triangle" and p >= 2:
        # collapsed Gauss of order n is exact up to total degree 2n - 2
        order = 4 if p <= 2 else 5
        x, y = S[:, 1], S[:, 2]
        for a, b in itertools.product(range(9), repeat=2):
            if a + b <= 2 * order - 2:
                assert abs(0.5 * np.sum(W * x ** a * y ** b) - _tri_moment(a, b)) < 1e-15


@pytest.mark.parametrize("t,p", list(itertools.product(TYPES, (1, 2, 3))))
def test_independent_constructions_agree(t, p):
    """Factor-product bases (oracle) vs Vandermonde bases (device tables)."""
    rng = np.random.default_rng(1)
    x = rng.random((40, 3)) * 0.5
    assert np.abs(oe.basis(t, p, x) - pe.basis(t, p, x)).max() < 1e-12
    assert np.abs(oe.basis_gradients(t, p, x) - pe.basis_gradients(t, p, x)).max() < 1e-12
    assert np.array_equal(oe.reference_nodes(t, p), pe.nodes(t, p))
    Xo, Wo = oe.volume_quadrature(t, p)
    Xp, Wp = pe.volume_quadrature(t, p)
    assert np.array_equal(Xo, Xp) and np.array_equal(Wo, Wp)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    x = rng.random((10, 3)) * 0.4 + 0.05
    h = 1e-6
    for t in TYPES:
        g = oe.basis_gradients(t, 3, x)
        for r in range(3):
            e = np.zeros(3)
            e[r] = h
            fd = (oe.basis(t, 3, x + e) - oe.basis(t, 3, x - e)) / (2 * h)
            assert np.abs(fd - g[..., r]).max() < 1e-7
