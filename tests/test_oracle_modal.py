"""Pins of the oracle's modal shock limiter (hybrid/limiter.py:186-264):
P1 data pass unchanged; a troubled element keeps its mean exactly and
becomes P1 (its own P1 projection); where that P1 state would be negative
at a sample the element falls back to its mean; untroubled elements are
untouched."""

import numpy as np
import pytest

from hybrid.meshgen import box_mesh
from oracle.limiter import ModalShockLimiter, PositivityLimiter
from oracle.mesh import from_arrays
from oracle.operators import EulerOperator
from oracle.solver import CompressibleSolver


def _setup(p=2):
    m = box_mesh(3, split_fraction=0.4, seed=4)
    op = EulerOperator(from_arrays(m), CompressibleSolver(), p)
    pos = PositivityLimiter(op, density_floor=1e-7, pressure_floor=1e-7)
    return op, ModalShockLimiter(op, positivity=pos), pos


def _states(op, f):
    s = op.solver
    return {t: s.primitive_to_conserved(f(x)) for t, x in op.geometry.solution_node_coordinates.items()}


def _linear(x):
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    return np.stack([1 + 0.2 * X, 0.3 * Y, 0 * X, 0.1 + 0 * X, 1 + 0.1 * Z], -1)


@pytest.mark.parametrize("p", [2, 3])
def test_p1_data_unchanged(p):
    op, lim, _ = _setup(p)
    st = _states(op, _linear)
    # conserved linear in rho and p only where u is constant: use the
    # primitive-linear field's conserved P1 projection as the data
    low = {t: np.einsum("ij,ejv->eiv", lim.sensor.p1_projectors[t], U) for t, U in st.items()}
    out = lim(low)
    for t in low:
        assert np.abs(out[t] - low[t]).max() < 1e-13


def test_troubled_elements_become_p1_with_their_mean():
    op, lim, pos = _setup(2)
    st = _states(op, _linear)
    rng = np.random.default_rng(1)
    bumped = {}
    for t, U in st.items():
        U = U.copy()
        sel = rng.random(len(U)) < 0.4
        U[sel, 1, 0] *= 3.0            # one node's density and energy tripled
        U[sel, 1, 4] *= 3.0
        bumped[t] = sel
        st[t] = U
    _, flags = lim.sensor(st)
    out = lim(st)
    nflag = 0
    for bid, b in enumerate(op.mesh.element_blocks):
        t = b.element_type
        f = flags[t]
        assert np.array_equal(f, bumped[t])
        nflag += f.sum()
        assert np.array_equal(out[t][~f], pos(st)[t][~f])
        m0, m1 = pos.mean(bid, st[t])[f], pos.mean(bid, out[t])[f]
        assert np.abs(m1 - m0).max() < 1e-13 * np.abs(m0).max()
        P = lim.sensor.p1_projectors[t]
        assert np.abs(np.einsum("ij,ejv->eiv", P, out[t][f]) - out[t][f]).max() < 1e-13
    assert nflag > 0


def test_severe_elements_fall_back_to_the_mean():
    op, lim, pos = _setup(2)
    st = _states(op, _linear)
    t = "hexahedron"
    U = st[t].copy()
    from hybrid.elements import nodes
    xi = nodes(t, 2)
    # a steep linear part (its P1 projection is negative at the x- face) and
    # a quadratic part that trips the sensor; the mean stays positive
    U[0, :, 0] = 1.0 + 3.0 * (xi[:, 0] - 0.5) + 6.0 * (xi[:, 1] - 0.5) ** 2
    U[0, :, 1:4] = 0.0
    U[0, :, 4] = 10.0
    st[t] = U
    _, flags = lim.sensor(st)
    assert flags[t][0]
    out = lim(st)
    mean = pos.mean(0, U)[0]
    assert np.abs(out[t][0] - mean).max() < 1e-13 * np.abs(mean).max()
