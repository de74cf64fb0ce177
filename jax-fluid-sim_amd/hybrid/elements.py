"""Reference-element tables for the device kernels.

Node sets and quadrature rules are those of the reference's
hybrid/elements.py (nodes :35-57, :221-244, :362-403; quadrature :60-128) and
hybrid/geometry.py:48-89 (faces).  The Lagrange bases are built here as
``monomials(x) @ inv(Vandermonde)`` on each element's polynomial space
(complete P_p for tetrahedra, P_p(triangle) x P_p(z) for prisms, Q_p for
hexahedra), which spans the same space and interpolates the same nodes as the
reference's closed-form bases.
"""

from functools import lru_cache

import numpy as np

TYPES = ("hexahedron", "wedge", "tetra")
TYPE_ID = {t: i for i, t in enumerate(TYPES)}
FACE_TYPES = ("quad", "triangle")

LOCAL_FACES = {
    "tetra": ((0, 2, 1), (0, 1, 3), (1, 2, 3), (2, 0, 3)),
    "wedge": ((0, 2, 1), (3, 4, 5), (0, 1, 4, 3), (1, 2, 5, 4), (2, 0, 3, 5)),
    "hexahedron": ((0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4),
                   (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)),
}

_LINE = {1: (0.0, 1.0), 2: (0.0, 1.0, 0.5), 3: (0.0, 1.0, 1 / 3, 2 / 3)}
_TRI = {
    1: ((0, 0), (1, 0), (0, 1)),
    2: ((0, 0), (1, 0), (0, 1), (0.5, 0), (0.5, 0.5), (0, 0.5)),
    3: ((0, 0), (1, 0), (0, 1), (1 / 3, 0), (2 / 3, 0), (2 / 3, 1 / 3),
        (1 / 3, 2 / 3), (0, 2 / 3), (0, 1 / 3), (1 / 3, 1 / 3)),
}
_CORNERS = {
    "tetra": ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)),
    "wedge": ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)),
    "hexahedron": ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                   (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)),
}


def corners(element_type):
    return np.asarray(_CORNERS[element_type], float)


def nodes(element_type, p):
    if p == 1:
        return corners(element_type)
    if element_type == "hexahedron":
        L = _LINE[p]
        return np.array([(x, y, z) for z in L for y in L for x in L], float)
    if element_type == "wedge":
        return np.array([(x, y, z) for z in _LINE[p] for x, y in _TRI[p]], float)
    a, b = (0.5, None) if p == 2 else (1 / 3, 2 / 3)
    if p == 2:
        return np.array(list(_CORNERS["tetra"]) + [
            (a, 0, 0), (0, a, 0), (0, 0, a), (a, a, 0), (a, 0, a), (0, a, a)], float)
    return np.array(list(_CORNERS["tetra"]) + [
        (a, 0, 0), (b, 0, 0), (0, a, 0), (0, b, 0), (0, 0, a), (0, 0, b),
        (b, a, 0), (a, b, 0), (b, 0, a), (a, 0, b), (0, b, a), (0, a, b),
        (a, a, 0), (a, 0, a), (0, a, a), (a, a, a)], float)


def _powers(x, k):
    x = np.asarray(x, float)
    return np.stack([x ** j for j in range(k + 1)], -1)


def _dpowers(x, k):
    x = np.asarray(x, float)
    return np.stack([j * x ** max(j - 1, 0) for j in range(k + 1)], -1)


@lru_cache(maxsize=None)
def _line_coefficients(p):
    return np.linalg.inv(_powers(np.array(_LINE[p]), p))


def _line(p, x):
    return _powers(x, p) @ _line_coefficients(p)


def _dline(p, x):
    return _dpowers(x, p) @ _line_coefficients(p)


def _tri_exponents(p):
    return [(i, j) for j in range(p + 1) for i in range(p + 1 - j)]


def _tri_monomials(p, x, y, dx=0, dy=0):
    cols = []
    for i, j in _tri_exponents(p):
        cx = (i if dx else 1.0) * np.asarray(x, float) ** max(i - dx, 0) if (i >= dx) else 0 * x
        cy = (j if dy else 1.0) * np.asarray(y, float) ** max(j - dy, 0) if (j >= dy) else 0 * y
        cols.append(cx * cy)
    return np.stack(cols, -1)


@lru_cache(maxsize=None)
def _tri_coefficients(p):
    t = np.array(_TRI[p], float)
    return np.linalg.inv(_tri_monomials(p, t[:, 0], t[:, 1]))


def _tet_exponents(p):
    return [(i, j, k) for k in range(p + 1) for j in range(p + 1 - k)
            for i in range(p + 1 - j - k)]


def _tet_monomials(p, x, d=None):
    x = np.asarray(x, float)
    cols = []
    for e in _tet_exponents(p):
        v = np.ones(x.shape[:-1])
        for axis in range(3):
            k = e[axis]
            if d == axis:
                v = v * (k * x[..., axis] ** max(k - 1, 0) if k > 0 else 0.0)
            else:
                v = v * x[..., axis] ** k
        cols.append(v)
    return np.stack(cols, -1)


@lru_cache(maxsize=None)
def _tet_coefficients(p):
    return np.linalg.inv(_tet_monomials(p, nodes("tetra", p)))


def _hex_order(p):
    if p == 1:  # corner order of the P1 hexahedron
        return [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    return [(i, j, k) for k in range(p + 1) for j in range(p + 1) for i in range(p + 1)]


def basis(element_type, p, x):
    """Lagrange basis values phi_i(x), shape (..., N)."""
    x = np.asarray(x, float)
    if element_type == "hexahedron":
        lx, ly, lz = (_line(p, x[..., a]) for a in range(3))
        return np.stack([lx[..., i] * ly[..., j] * lz[..., k] for i, j, k in _hex_order(p)], -1)
    if element_type == "wedge":
        t = _tri_monomials(p, x[..., 0], x[..., 1]) @ _tri_coefficients(p)
        lz = _line(p, x[..., 2])
        return np.concatenate([t * lz[..., k:k + 1] for k in range(p + 1)], -1)
    return _tet_monomials(p, x) @ _tet_coefficients(p)


def basis_gradients(element_type, p, x):
    """Reference gradients d phi_i / d xi_r, shape (..., N, 3)."""
    x = np.asarray(x, float)
    if element_type == "hexahedron":
        l = [_line(p, x[..., a]) for a in range(3)]
        d = [_dline(p, x[..., a]) for a in range(3)]
        out = []
        for i, j, k in _hex_order(p):
            out.append(np.stack([d[0][..., i] * l[1][..., j] * l[2][..., k],
                                 l[0][..., i] * d[1][..., j] * l[2][..., k],
                                 l[0][..., i] * l[1][..., j] * d[2][..., k]], -1))
        return np.stack(out, -2)
    if element_type == "wedge":
        C = _tri_coefficients(p)
        t = _tri_monomials(p, x[..., 0], x[..., 1]) @ C
        tx = _tri_monomials(p, x[..., 0], x[..., 1], dx=1) @ C
        ty = _tri_monomials(p, x[..., 0], x[..., 1], dy=1) @ C
        lz, dz = _line(p, x[..., 2]), _dline(p, x[..., 2])
        gx = np.concatenate([tx * lz[..., k:k + 1] for k in range(p + 1)], -1)
        gy = np.concatenate([ty * lz[..., k:k + 1] for k in range(p + 1)], -1)
        gz = np.concatenate([t * dz[..., k:k + 1] for k in range(p + 1)], -1)
        return np.stack([gx, gy, gz], -1)
    C = _tet_coefficients(p)
    return np.stack([_tet_monomials(p, x, d) @ C for d in range(3)], -1)


def _gauss01(order):
    x, w = np.polynomial.legendre.leggauss(order)
    return 0.5 * (x + 1.0), 0.5 * w


def volume_quadrature(element_type, p):
    """Points/weights of hybrid/elements.py:60-128."""
    if p >= 2:
        order = 5 if (p == 3 or element_type != "hexahedron") else 3
        g, w = _gauss01(order)
        if element_type == "hexahedron":
            X = np.array([(x, y, z) for z in g for y in g for x in g])
            W = np.array([wx * wy * wz for wz in w for wy in w for wx in w])
        elif element_type == "wedge":
            X = np.array([(a, (1 - a) * b, z) for z in g for a in g for b in g])
            W = np.array([wa * wb * wz * (1 - a) for wz in w for a, wa in zip(g, w) for wb in w])
        else:
            X = np.array([(a, (1 - a) * b, (1 - a) * (1 - b) * c)
                          for a in g for b in g for c in g])
            W = np.array([wa * wb * wc * (1 - a) ** 2 * (1 - b)
                          for a, wa in zip(g, w) for b, wb in zip(g, w) for wc in w])
        return X, W
    if element_type == "tetra":
        a, b = 0.5854101966249685, 0.1381966011250105
        return np.array([[b, b, b], [a, b, b], [b, a, b], [b, b, a]]), np.full(4, 1 / 24)
    s = 1.0 / np.sqrt(3.0)
    l = 0.5 * (1 + np.array([-s, s]))
    if element_type == "wedge":
        t = ((1 / 6, 1 / 6), (2 / 3, 1 / 6), (1 / 6, 2 / 3))
        return (np.array([(x, y, z) for z in l for x, y in t]), np.full(6, 1 / 12))
    return np.array([(x, y, z) for z in l for y in l for x in l]), np.full(8, 1 / 8)


def face_quadrature(face_type, p):
    """(face shape values, normalized weights), hybrid/geometry.py:48-89."""
    if face_type == "triangle":
        if p == 1:
            S = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])
            return S, np.full(3, 1 / 3)
        g, w = _gauss01(4 if p <= 2 else 5)
        S, W = [], []
        for a, wa in zip(g, w):
            for b, wb in zip(g, w):
                x, y = a, (1 - a) * b
                S.append((1 - x - y, x, y))
                W.append(2 * wa * wb * (1 - a))
        return np.array(S), np.array(W)
    g, w = _gauss01({1: 2, 2: 3}.get(p, 4))
    S = [((1 - u) * (1 - v), u * (1 - v), u * v, (1 - u) * v) for v in g for u in g]
    W = [wu * wv for wv in w for wu in w]
    return np.array(S), np.array(W)


@lru_cache(maxsize=None)
def reference_tables(element_type, p):
    """Volume operator tables of one element type (FP64 numpy)."""
    X, W = volume_quadrature(element_type, p)
    B = basis(element_type, p, X)
    D = basis_gradients(element_type, p, X)
    M = np.einsum("q,qi,qj->ij", W, B, B)
    return dict(points=X, weights=W, B=B, D=D, M=M, Minv=np.linalg.inv(M),
                N=B.shape[1], Q=B.shape[0], nodes=nodes(element_type, p))
