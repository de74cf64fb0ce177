"""Pins of the oracle's time-step control: the element-local dt and its
neighbour smoothing (hybrid/time_integration.py:171-262), the conservation
sums, and the accept / reject loop of run_hybrid_sphere.py:265-346, against
closed forms."""

import numpy as np
import pytest

from hybrid.meshgen import box_mesh
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver
from oracle.time_integration import SSPRK3


def _uniform(op, W):
    s = op.solver
    return {t: np.broadcast_to(s.primitive_to_conserved(np.asarray(W, float)),
                               x.shape[:-1] + (5,)).copy()
            for t, x in op.geometry.solution_node_coordinates.items()}


@pytest.mark.parametrize("p", [1, 2])
def test_local_dt_uniform_closed_form(p):
    """Cubes of side h: L = 2V / A = h / 3 (above V^(1/3) / (4 (p+1))), rate =
    (2p+1)(|u| + c) / L, so every dt is cfl L / ((2p+1)(|u| + c)) and the
    smoothing leaves equal values unchanged; stable_dt clips the same value."""
    h = 0.5
    m = box_mesh(4, lengths=(4 * h,) * 3, periodic=(True,) * 3)
    op = NavierStokesOperator(from_arrays(m), CompressibleSolver(), p)
    rk = SSPRK3(op, cfl=0.3)
    W = [1.2, 0.3, -0.4, 0.1, 0.9]
    st = _uniform(op, W)
    speed = np.linalg.norm(W[1:4]) + np.sqrt(1.4 * W[4] / W[0])
    want = 0.3 * (h / 3) / ((2 * p + 1) * speed)
    d = rk.local_dt(st)["hexahedron"]
    assert d.shape == (64, 1, 1)
    assert np.allclose(d, want, rtol=1e-14, atol=0)
    assert abs(rk.stable_dt(st) - min(want, 1e-3)) <= 1e-15 * want


def test_local_dt_smoothing_and_sliver():
    """One fast element: after two rounds of smoothing its face neighbours
    hold at most 1.3 x its dt and their neighbours at most 1.3^2 x; an
    element far away keeps its own dt; the sliver factor min(1, V/1e-4)
    multiplies the result."""
    for side, sliver in ((5.0, 1.0), (0.2, (0.2 / 5) ** 3 / 1e-4)):
        m = box_mesh(5, lengths=(side,) * 3)
        op = NavierStokesOperator(from_arrays(m), CompressibleSolver(), 1)
        rk = SSPRK3(op, cfl=0.3)
        st = _uniform(op, [1.0, 0.0, 0.0, 0.0, 1.0])
        fast = 62                                     # the centre cell (2, 2, 2)
        U = st["hexahedron"]
        U[fast] = op.solver.primitive_to_conserved(np.array([1.0, 0.0, 0.0, 0.0, 100.0]))
        d = rk.local_dt(st)["hexahedron"][:, 0, 0] / sliver
        base = rk.local_dt(_uniform(op, [1.0, 0, 0, 0, 1.0]))["hexahedron"][0, 0, 0] / sliver
        ijk = np.stack(np.unravel_index(np.arange(125), (5, 5, 5)), 1)[:, ::-1]
        dist = np.abs(ijk - ijk[fast]).sum(1)
        dmin = d[fast]
        assert dmin < base / 5
        assert np.all(d[dist == 1] <= 1.3 * dmin * (1 + 1e-14))
        assert np.all(d[dist == 2] <= 1.3 ** 2 * dmin * (1 + 1e-14))
        assert np.allclose(d[dist >= 4], base, rtol=1e-14)


def test_conserved_totals_closed_form():
    m = box_mesh(3, lengths=(1.0, 2.0, 0.5), split_fraction=0.4, seed=2)
    op = NavierStokesOperator(from_arrays(m), CompressibleSolver(), 2)
    rk = SSPRK3(op)
    W = np.array([1.3, 0.2, -0.1, 0.4, 2.0])
    st = _uniform(op, W)
    U0 = op.solver.primitive_to_conserved(W)
    assert np.allclose(rk.conserved_totals(st), U0 * 1.0, rtol=1e-13)


class _Scripted(SSPRK3):
    """A step is admissible when dt <= 0.3; the CFL dt is 1."""

    def __init__(self):
        self.cfl = 1.0

    def stable_dt(self, states):
        return 1.0

    def checked_step(self, states, dt, boundary_state=None):
        return {"x": states["x"] + dt}, dt <= 0.3, 0.0, 0.0


def test_step_loop_reject_retry_closed_form():
    """run_hybrid_sphere.py:265-346: halve on rejection, grow by 1.25 (up to
    1) on acceptance, clip the last step to the target time."""
    rk = _Scripted()
    out, t, steps, rej = rk.advance({"x": np.zeros(1)}, 1.0)
    # attempts: 1 r, .5 r, .25 a, .3125 r, .15625 a, .1953125 a, .244140625 a,
    # .30517578125 r, .152587890625 a, ... until t = 1
    retry, tt, n, r, seq = 1.0, 0.0, 0, 0, []
    while tt < 1.0:
        dt = min(retry, 1.0 - tt)
        if dt <= 0.3:
            tt += dt
            n += 1
            retry = min(1.0, 1.25 * retry)
            seq.append(dt)
        else:
            r += 1
            retry *= 0.5
    assert seq[:4] == [0.25, 0.15625, 0.1953125, 0.244140625]
    assert (steps, rej) == (n, r) and t == 1.0
    assert abs(out["x"][0] - sum(seq)) < 1e-15
    with pytest.raises(FloatingPointError):
        rk.checked_step = lambda s, dt, b=None: (s, False, 0.0, 0.0)
        rk.advance({"x": np.zeros(1)}, 1.0, maximum_rejections=5)
