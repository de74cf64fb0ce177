"""Reference elements (nodes, bases, gradients, quadrature) restated in numpy.

Follows hybrid/elements.py of the reference.  Every Lagrange basis function is
held as ``coefficient * prod(affine factor)``; each affine factor is
``c0 + cx*x + cy*y + cz*z``.  Values and exact gradients (product rule) come
from the same factor list, so the gradient tables are derived independently of
the reference's hand-expanded derivative formulas
(hybrid/elements.py:277-322, 461-667) while describing the same polynomials.
"""

from dataclasses import dataclass

import numpy as np

ELEMENT_TYPES = ("hexahedron", "wedge", "tetra")

# Line node orderings: P1 (0, 1); P2 (0, 1, 1/2); P3 (0, 1, 1/3, 2/3)
# (hybrid/elements.py:202-207, 325-335).
LINE_NODES = {
    1: (0.0, 1.0),
    2: (0.0, 1.0, 0.5),
    3: (0.0, 1.0, 1.0 / 3.0, 2.0 / 3.0),
}


def _affine(c0=0.0, cx=0.0, cy=0.0, cz=0.0):
    return (float(c0), float(cx), float(cy), float(cz))


def _line_lagrange(axis, degree, index):
    """1-D Lagrange basis on LINE_NODES[degree] along coordinate ``axis``."""
    nodes = LINE_NODES[degree]
    coef = 1.0
    factors = []
    for j, xj in enumerate(nodes):
        if j == index:
            continue
        coef /= nodes[index] - xj
        c = [0.0, 0.0, 0.0]
        c[axis] = 1.0
        factors.append(_affine(-xj, *c))
    return coef, factors


# Barycentric coordinates as affine factors.
def _bary_tet(k):
    return (
        _affine(1.0, -1.0, -1.0, -1.0),
        _affine(0.0, 1.0, 0.0, 0.0),
        _affine(0.0, 0.0, 1.0, 0.0),
        _affine(0.0, 0.0, 0.0, 1.0),
    )[k]


def _bary_tri(k):
    return (
        _affine(1.0, -1.0, -1.0, 0.0),
        _affine(0.0, 1.0, 0.0, 0.0),
        _affine(0.0, 0.0, 1.0, 0.0),
    )[k]


def _scaled(f, a, b):
    """a*f + b for an affine factor f."""
    return (a * f[0] + b, a * f[1], a * f[2], a * f[3])


def _simplex_p1(bary, k):
    return 1.0, [bary(k)]


def _simplex_vertex(bary, k, degree):
    l = bary(k)
    if degree == 2:  # l(2l-1)
        return 1.0, [l, _scaled(l, 2.0, -1.0)]
    # 0.5 l (3l-1)(3l-2)
    return 0.5, [l, _scaled(l, 3.0, -1.0), _scaled(l, 3.0, -2.0)]


def _simplex_edge(bary, a, b, degree, near):
    """Edge node between vertices a and b; ``near`` picks the P3 node closer
    to that vertex."""
    if degree == 2:  # 4 la lb
        return 4.0, [bary(a), bary(b)]
    # 4.5 la lb (3 l_near - 1)
    return 4.5, [bary(a), bary(b), _scaled(bary(near), 3.0, -1.0)]


def _simplex_face(bary, a, b, c):
    return 27.0, [bary(a), bary(b), bary(c)]


def _triangle_functions(degree):
    """Triangle Lagrange functions in node order of hybrid/elements.py:231-233
    (P2) and :384-392 (P3); P1 is the vertex set."""
    t = _bary_tri
    if degree == 1:
        return [_simplex_p1(t, k) for k in range(3)]
    if degree == 2:
        return [
            _simplex_vertex(t, 0, 2), _simplex_vertex(t, 1, 2), _simplex_vertex(t, 2, 2),
            _simplex_edge(t, 0, 1, 2, None), _simplex_edge(t, 1, 2, 2, None),
            _simplex_edge(t, 2, 0, 2, None),
        ]
    return [
        _simplex_vertex(t, 0, 3), _simplex_vertex(t, 1, 3), _simplex_vertex(t, 2, 3),
        _simplex_edge(t, 0, 1, 3, 0), _simplex_edge(t, 0, 1, 3, 1),
        _simplex_edge(t, 1, 2, 3, 1), _simplex_edge(t, 1, 2, 3, 2),
        _simplex_edge(t, 0, 2, 3, 2), _simplex_edge(t, 0, 2, 3, 0),
        _simplex_face(t, 0, 1, 2),
    ]


def _tetra_functions(degree):
    """Tetra Lagrange functions in node order of hybrid/elements.py:37-39
    (P1), :222-229 (P2) and :363-382 (P3)."""
    t = _bary_tet
    if degree == 1:
        return [_simplex_p1(t, k) for k in range(4)]
    if degree == 2:
        out = [_simplex_vertex(t, k, 2) for k in range(4)]
        for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)):
            out.append(_simplex_edge(t, a, b, 2, None))
        return out
    out = [_simplex_vertex(t, k, 3) for k in range(4)]
    for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)):
        out.append(_simplex_edge(t, a, b, 3, a))
        out.append(_simplex_edge(t, a, b, 3, b))
    for a, b, c in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
        out.append(_simplex_face(t, a, b, c))
    return out


def _wedge_functions(degree):
    """Triangle x z-line tensor product; z-level outer, triangle node inner
    (hybrid/elements.py:41-48, 230-237, 383-396, 137-148, 259-264, 439-444)."""
    tri = _triangle_functions(degree)
    out = []
    for iz in range(degree + 1):
        cz, fz = _line_lagrange(2, degree, iz)
        for ct, ft in tri:
            out.append((ct * cz, list(ft) + list(fz)))
    return out


def _hexahedron_functions(degree):
    """Tensor product, x fastest then y then z (hybrid/elements.py:49-56,
    238-243, 397-402)."""
    if degree == 1:
        # P1 hex uses the vertex ordering of hybrid/elements.py:49-56.
        corners = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                   (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
        out = []
        for cx, cy, cz in corners:
            coef = 1.0
            factors = []
            for axis, c in enumerate((cx, cy, cz)):
                cc, ff = _line_lagrange(axis, 1, c)
                coef *= cc
                factors += ff
            out.append((coef, factors))
        return out
    out = []
    n = degree + 1
    for iz in range(n):
        for iy in range(n):
            for ix in range(n):
                coef = 1.0
                factors = []
                for axis, i in enumerate((ix, iy, iz)):
                    cc, ff = _line_lagrange(axis, degree, i)
                    coef *= cc
                    factors += ff
                out.append((coef, factors))
    return out


def _functions(element_type, degree):
    if degree not in (1, 2, 3):
        raise ValueError("Only P1, P2, and P3 are implemented")
    if element_type == "tetra":
        return _tetra_functions(degree)
    if element_type == "wedge":
        return _wedge_functions(degree)
    if element_type == "hexahedron":
        return _hexahedron_functions(degree)
    raise ValueError(f"Unsupported element type {element_type!r}")


def _eval_factor(f, pts):
    return f[0] + f[1] * pts[..., 0] + f[2] * pts[..., 1] + f[3] * pts[..., 2]


def basis(element_type, degree, points):
    """phi_i(points): shape (..., N)."""
    pts = np.asarray(points, dtype=np.float64)
    out = []
    for coef, factors in _functions(element_type, degree):
        value = np.full(pts.shape[:-1], coef)
        for f in factors:
            value = value * _eval_factor(f, pts)
        out.append(value)
    return np.stack(out, axis=-1)


def basis_gradients(element_type, degree, points):
    """d phi_i / d xi_r: shape (..., N, 3)."""
    pts = np.asarray(points, dtype=np.float64)
    out = []
    for coef, factors in _functions(element_type, degree):
        values = [_eval_factor(f, pts) for f in factors]
        grad = np.zeros(pts.shape[:-1] + (3,))
        for k, f in enumerate(factors):
            others = np.full(pts.shape[:-1], coef)
            for j, v in enumerate(values):
                if j != k:
                    others = others * v
            grad = grad + others[..., None] * np.asarray(f[1:4])
        out.append(grad)
    return np.stack(out, axis=-2)


def reference_vertices(element_type):
    """P1 corner coordinates, hybrid/elements.py:35-57."""
    if element_type == "tetra":
        return np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    if element_type == "wedge":
        return np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0],
                         [0, 0, 1], [1, 0, 1], [0, 1, 1]], float)
    if element_type == "hexahedron":
        return np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                         [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    raise ValueError(element_type)


_TRIANGLE_NODES = {
    1: [(0, 0), (1, 0), (0, 1)],
    2: [(0, 0), (1, 0), (0, 1), (0.5, 0), (0.5, 0.5), (0, 0.5)],
    3: [(0, 0), (1, 0), (0, 1), (1 / 3, 0), (2 / 3, 0), (2 / 3, 1 / 3),
        (1 / 3, 2 / 3), (0, 2 / 3), (0, 1 / 3), (1 / 3, 1 / 3)],
}


def reference_nodes(element_type, degree):
    """Solution nodes: hybrid/elements.py:35-57 (P1), 221-244 (P2), 362-403 (P3)."""
    if degree == 1:
        return reference_vertices(element_type)
    if element_type == "tetra":
        if degree == 2:
            return np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                             [.5, 0, 0], [0, .5, 0], [0, 0, .5],
                             [.5, .5, 0], [.5, 0, .5], [0, .5, .5]], float)
        t = 1.0 / 3.0
        s = 2.0 / 3.0
        return np.array([
            [0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
            [t, 0, 0], [s, 0, 0], [0, t, 0], [0, s, 0], [0, 0, t], [0, 0, s],
            [s, t, 0], [t, s, 0], [s, 0, t], [t, 0, s], [0, s, t], [0, t, s],
            [t, t, 0], [t, 0, t], [0, t, t], [t, t, t]], float)
    line = LINE_NODES[degree]
    if element_type == "wedge":
        return np.array([[x, y, z] for z in line for x, y in _TRIANGLE_NODES[degree]],
                        float)
    if element_type == "hexahedron":
        return np.array([[x, y, z] for z in line for y in line for x in line], float)
    raise ValueError(element_type)


def _gauss_01(order):
    raw, raw_w = np.polynomial.legendre.leggauss(order)
    return 0.5 * (raw + 1.0), 0.5 * raw_w


def volume_quadrature(element_type, degree):
    """hybrid/elements.py:60-128: collapsed/tensor Gauss for P2/P3 (order 5
    for P3 or simplex-like types, else 3); fixed low-order rules for P1."""
    if degree in (2, 3):
        order = 5 if (degree == 3 or element_type in ("tetra", "wedge")) else 3
        line, w1 = _gauss_01(order)
        pts, wts = [], []
        if element_type == "tetra":
            for a, wa in zip(line, w1):
                for b, wb in zip(line, w1):
                    for c, wc in zip(line, w1):
                        pts.append((a, (1 - a) * b, (1 - a) * (1 - b) * c))
                        wts.append(wa * wb * wc * (1 - a) ** 2 * (1 - b))
        elif element_type == "wedge":
            for z, wz in zip(line, w1):
                for a, wa in zip(line, w1):
                    for b, wb in zip(line, w1):
                        pts.append((a, (1 - a) * b, z))
                        wts.append(wa * wb * wz * (1 - a))
        elif element_type == "hexahedron":
            for z, wz in zip(line, w1):
                for y, wy in zip(line, w1):
                    for x, wx in zip(line, w1):
                        pts.append((x, y, z))
                        wts.append(wx * wy * wz)
        else:
            raise ValueError(element_type)
        return np.array(pts, float), np.array(wts, float)
    if element_type == "tetra":
        a, b = 0.5854101966249685, 0.1381966011250105
        return (np.array([[b, b, b], [a, b, b], [b, a, b], [b, b, a]]),
                np.full(4, 1.0 / 24.0))
    off = 1.0 / np.sqrt(3.0)
    lp = 0.5 * (1.0 + np.array([-off, off]))
    lw = np.full(2, 0.5)
    if element_type == "wedge":
        tp = np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]])
        pts = [(x, y, z) for z in lp for x, y in tp]
        wts = [wt * wz for wz in lw for wt in np.full(3, 1 / 6)]
        return np.array(pts), np.array(wts)
    if element_type == "hexahedron":
        pts = [(x, y, z) for z in lp for y in lp for x in lp]
        wts = [wx * wy * wz for wz in lw for wy in lw for wx in lw]
        return np.array(pts), np.array(wts)
    raise ValueError(element_type)


def face_quadrature(face_type, degree):
    """hybrid/geometry.py:48-89: (face shape values at points, weights summing
    to one)."""
    if face_type == "triangle" and degree == 1:
        return (np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6],
                          [1 / 6, 1 / 6, 2 / 3]]), np.full(3, 1 / 3))
    if face_type == "triangle":
        order = 4 if degree <= 2 else 5
        line, w1 = _gauss_01(order)
        shapes, wts = [], []
        for a, wa in zip(line, w1):
            for b, wb in zip(line, w1):
                x, y = a, (1 - a) * b
                shapes.append([1 - x - y, x, y])
                wts.append(2 * wa * wb * (1 - a))
        return np.array(shapes), np.array(wts)
    if face_type == "quad":
        order = {1: 2, 2: 3}.get(degree, 4)
        line, w1 = _gauss_01(order)
        shapes, wts = [], []
        for v, wv in zip(line, w1):
            for u, wu in zip(line, w1):
                shapes.append([(1 - u) * (1 - v), u * (1 - v), u * v, (1 - u) * v])
                wts.append(wu * wv)
        return np.array(shapes), np.array(wts)
    raise ValueError(face_type)


@dataclass(frozen=True)
class ReferenceElement:
    """hybrid/elements.py:10-32, 670-700."""
    element_type: str
    polynomial_degree: int
    nodes: np.ndarray
    volume_points: np.ndarray
    volume_weights: np.ndarray
    mass_matrix: np.ndarray
    inverse_mass_matrix: np.ndarray

    def basis(self, points):
        return basis(self.element_type, self.polynomial_degree, points)

    def basis_gradients(self, points):
        return basis_gradients(self.element_type, self.polynomial_degree, points)


def make_reference_element(element_type, degree=1):
    nodes = reference_nodes(element_type, degree)
    pts, wts = volume_quadrature(element_type, degree)
    phi = basis(element_type, degree, pts)
    mass = np.einsum("q,qi,qj->ij", wts, phi, phi)
    return ReferenceElement(element_type, degree, nodes, pts, wts, mass,
                            np.linalg.inv(mass))
