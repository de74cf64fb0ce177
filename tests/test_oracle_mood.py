"""Pins of the oracle's troubled-cell sensor, PAD detection, positivity
limiter and subcell finite-volume recompute (hybrid/limiter.py,
hybrid/subcell.py) against closed forms, independent constructions and
conservation (the reference has no golden vectors for them; its own test,
tests/test_subcell_conservation.py, checks conservation to 1e-3 on a free
stream)."""

import numpy as np
import pytest

from conftest import perturbed_primitive

from hybrid.elements import basis as product_basis, volume_quadrature as product_quadrature
from hybrid.meshgen import box_mesh
from oracle.limiter import MOODSubcellLimiter, PositivityLimiter, TroubledCellSensor
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver
from oracle.subcell import SubcellFV, dg_to_subcells, subcells_to_dg


def _op(p=2, split=0.4, n=3, seed=3):
    m = box_mesh(n, periodic=(True,) * 3, split_fraction=split, seed=seed)
    return NavierStokesOperator(from_arrays(m), CompressibleSolver(), p)


def _states(op, prim):
    s = op.solver
    return {t: s.primitive_to_conserved(prim(x)) for t, x in op.geometry.solution_node_coordinates.items()}


def _linear(x):
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    return np.stack([1.0 + 0.2 * X - 0.1 * Y, 0.3 + 0 * X, -0.2 + 0 * X, 0.1 + 0 * X,
                     1.0 + 0.1 * Z + 0.05 * X], -1)


def _total(op, S):
    tot = 0.0
    for bid, b in enumerate(op.mesh.element_blocks):
        ref = op.geometry.reference_elements[b.element_type]
        w = op.geometry.element_metrics[bid].integration_weights
        tot = tot + np.einsum("qi,eiv,eq->v", ref.basis(ref.volume_points), S[b.element_type], w)
    return tot


@pytest.mark.parametrize("p", [2, 3])
def test_sensor_vanishes_on_p1_fields(p):
    op = _op(p)
    sen = TroubledCellSensor(op)
    s, flags = sen(_states(op, _linear))
    for t in s:
        assert np.abs(s[t]).max() < 1e-12 and not flags[t].any()


def test_sensor_single_node_perturbation_closed_form():
    """A P1 density plus delta at one node: the ratio is the squared residual
    of the least-squares P1 fit (computed here with lstsq on an independently
    built P1 Vandermonde) over sum(rho^2); pressure stays P1 (ratio 0)."""
    op = _op(2)
    sen = TroubledCellSensor(op)
    st = _states(op, _linear)
    for t, U in st.items():
        e, k, delta = 1, 7 % U.shape[1], 0.3
        U = U.copy()
        U[e, k, 0] += delta
        # keep the pressure field unchanged: adjust the energy for the new density
        W = op.solver.conserved_to_primitive(st[t])
        W[e, k, 0] += delta
        U[e] = op.solver.primitive_to_conserved(W[e])
        ratio = sen.block_sensor(t, U)[e]
        x, y, z = op.geometry.reference_elements[t].nodes.T
        one = np.ones_like(x)
        # span of the element's P1 space (hybrid/elements.py:131-159):
        # trilinear on hexahedra, linear x bilinear-in-z on wedges, linear on tetrahedra
        V1 = np.column_stack({"hexahedron": [one, x, y, z, x * y, y * z, x * z, x * y * z],
                              "wedge": [one, x, y, z, x * z, y * z],
                              "tetra": [one, x, y, z]}[t])
        rho = W[e, :, 0]
        coef, *_ = np.linalg.lstsq(V1, rho, rcond=None)
        want = np.sum((rho - V1 @ coef) ** 2) / np.sum(rho ** 2)
        assert abs(ratio - want) <= 1e-12 * max(want, 1e-30), (t, ratio, want)
        assert want > 1e-3


def test_pad_detects_sub_floor_values_between_nodes():
    """Nodes above the floor, a volume point below it (independent product
    basis): PAD flags the cell; the same state with a lower floor is clean."""
    op = _op(2, split=0.0)
    st = _states(op, lambda x: _linear(x) * 0 + np.array([1.0, 0, 0, 0, 1.0]))
    U = st["hexahedron"].copy()
    nodes = op.geometry.reference_elements["hexahedron"].nodes
    # a quadratic in x whose minimum (0.01) sits on the first Gauss abscissa
    # of the 3-point rule, while every node value is >= 0.0227
    xr = nodes[:, 0]
    g = 0.5 - np.sqrt(15.0) / 10.0
    rho = 0.01 + (xr - g) ** 2
    W = op.solver.conserved_to_primitive(U)
    W[0, :, 0] = rho
    U[0] = op.solver.primitive_to_conserved(W[0])
    pts, _ = product_quadrature("hexahedron", 2)
    vals = product_basis("hexahedron", 2, pts) @ rho
    assert rho.min() >= 0.015 and vals.min() < 0.015
    for floor, want in ((0.015, True), (0.005, False)):
        pos = PositivityLimiter(op, density_floor=floor, pressure_floor=1e-10)
        lim = MOODSubcellLimiter(op, threshold=1e9, positivity=pos)
        assert lim.troubled({"hexahedron": U})["hexahedron"][0] == want


@pytest.mark.parametrize("p", [1, 2, 3])
def test_subcell_projection_is_a_left_inverse(p):
    op = _op(p)
    fv = SubcellFV(op)
    st = _states(op, lambda x: perturbed_primitive(x, seed=2))
    for t, U in st.items():
        pj = fv.projections[t]
        assert np.abs(pj.projection.sum(-1) - 1.0).max() < 1e-12           # partition of unity
        bid = [b.element_type for b in op.mesh.element_blocks].index(t)
        vol = op.geometry.element_metrics[bid].integration_weights.sum(1)
        assert np.abs(pj.volumes.sum(1) - vol).max() < 1e-13 * vol.max()  # subcells tile the cell
        back = subcells_to_dg(dg_to_subcells(U, pj), pj, bound_limit=False)
        assert np.abs(back - U).max() < 1e-11
        # the bound-limited reconstruction preserves the cell mean exactly
        avg = dg_to_subcells(U, pj) * (1 + 0.1 * np.random.default_rng(0).standard_normal(
            (U.shape[0], 64, U.shape[2])))
        lim = subcells_to_dg(avg, pj)
        m1 = (avg * pj.volumes[..., None]).sum(1)
        m2 = (dg_to_subcells(lim, pj) * pj.volumes[..., None]).sum(1)
        assert np.abs(m1 - m2).max() < 1e-12 * np.abs(m1).max()


@pytest.mark.parametrize("recon", ["constant", "tvd", "weno"])
def test_subcell_recompute_free_stream_and_conservation(recon):
    op = _op(2)
    rng = np.random.default_rng(0)
    uniform = _states(op, lambda x: _linear(x) * 0 + np.array([1.1, 0.3, -0.2, 0.1, 2.0]))
    troubled = {t: rng.random(len(v)) < 0.4 for t, v in uniform.items()}
    # (with the reference's subface indexing a free stream is not preserved:
    # fluxes land on the wrong subcells, see the next test)
    fv = SubcellFV(op, reconstruction=recon, indexing="owner")
    out = fv.recompute(uniform, uniform, 1e-3, troubled)
    for t in out:
        assert np.abs(out[t] - uniform[t]).max() < 1e-12, t
    st = _states(op, lambda x: perturbed_primitive(x, seed=1))
    dt = 1e-3
    inv = op.inviscid_rhs(st)
    cand = {k: st[k] + dt * inv[k] for k in st}
    fv = SubcellFV(op, reconstruction=recon, indexing="owner")
    out = fv.recompute(st, cand, dt, troubled)
    diff = _total(op, out) - _total(op, cand)
    assert np.abs(diff).max() < 1e-12 * np.abs(_total(op, cand)).max()
    # troubled cells really changed, others only by the mean correction
    for t in out:
        d = out[t] - cand[t]
        assert np.abs(d[troubled[t]]).max() > 1e-8
        dn = d[~troubled[t]]
        assert np.abs(dn - dn[:, :1]).max() < 1e-14


def test_reference_indexing_is_not_conservative_on_hexahedra():
    """hybrid/subcell.py:774-776, 805-807, 842-848 index subcells by boundary
    subface id (0..95 on a hexahedron): the recompute then loses conservation,
    which the owner indexing restores."""
    op = _op(2, split=0.0)
    st = _states(op, lambda x: perturbed_primitive(x, seed=1))
    dt = 1e-3
    inv = op.inviscid_rhs(st)
    cand = {k: st[k] + dt * inv[k] for k in st}
    troubled = {"hexahedron": np.arange(len(st["hexahedron"])) % 3 == 0}
    ref = SubcellFV(op, indexing="reference").recompute(st, cand, dt, troubled)
    own = SubcellFV(op, indexing="owner").recompute(st, cand, dt, troubled)
    tot = np.abs(_total(op, cand)).max()
    assert np.abs(_total(op, ref) - _total(op, cand)).max() > 1e-8 * tot
    assert np.abs(_total(op, own) - _total(op, cand)).max() < 1e-12 * tot


def test_positivity_limiter_invariants():
    op = _op(2)
    pos = PositivityLimiter(op, density_floor=1e-6, pressure_floor=1e-6)
    st = _states(op, lambda x: perturbed_primitive(x, seed=4))
    # admissible cells: only round-off change
    out = pos(st)
    for t in st:
        assert np.abs(out[t] - st[t]).max() < 1e-13
    # make some nodes non-admissible (mean stays admissible)
    bad = {}
    for t, U in st.items():
        U = U.copy()
        W = op.solver.conserved_to_primitive(U)
        W[::2, 0, 0] = -0.05
        W[1::3, 1, 4] = -0.02
        bad[t] = op.solver.primitive_to_conserved(W)
    out = pos(bad)
    for bid, b in enumerate(op.mesh.element_blocks):
        t = b.element_type
        m0, m1 = pos.mean(bid, bad[t]), pos.mean(bid, out[t])
        assert np.abs(m0 - m1).max() < 1e-13
        Wm = op.solver.conserved_to_primitive(m0)
        W = op.solver.conserved_to_primitive(pos.sample(t, out[t]))
        rho_floor = np.maximum(1e-6, 0.05 * Wm[:, 0])
        p_floor = np.maximum(1e-6, 0.05 * Wm[:, 4])
        assert np.all(W[..., 0].min(1) >= rho_floor * (1 - 1e-12))
        # bisection to 2^-30: the pressure bound holds at the accepted theta
        assert np.all(W[..., 4].min(1) >= p_floor * (1 - 1e-6))


def _interior_faces(fv, t):
    """Internal faces between two subcells whose faces are all internal (the
    central 2 x 2 x 2 block of a 4 x 4 x 4 hexahedron subdivision): there the
    Green-Gauss gradient over internal faces is exact for linear data."""
    rs = fv.references[t]
    o, nb = rs.internal_owner, rs.internal_neighbor
    nface = np.bincount(np.concatenate([o, nb]), minlength=rs.count)
    inner = nface == 6
    return np.flatnonzero(inner[o] & inner[nb])


def _linear_subcell_averages(fv, t):
    c = fv.projections[t].physical_centers
    X, Y, Z = c[..., 0], c[..., 1], c[..., 2]
    return np.stack([1.0 + 0.6 * X - 0.3 * Y, 0.3 + 0.4 * Z, -0.2 + 0 * X, 0.1 + 0 * X,
                     3.0 + 0.5 * Z + 0.25 * X], -1)


def test_tvd_reconstruction_linear_data_closed_form():
    """Linear subcell data on the inner faces: the gradient is exact, so the
    slope ratio is r = 1 and the limiter phi(r) = (r^2 + 2r) / (r^2 + r + 2)
    of hybrid/subcell.py:600-615 is 3/4: each side's trace is its own value
    moved 3/8 of the jump towards the other."""
    op = _op(2, split=0.0)
    fv = SubcellFV(op, reconstruction="tvd")
    t = "hexahedron"
    f = _interior_faces(fv, t)
    assert len(f) == 12
    U = _linear_subcell_averages(fv, t)
    UL, UR = fv._reconstruct(t, U)
    o, nb = fv.references[t].internal_owner[f], fv.references[t].internal_neighbor[f]
    dU = U[:, nb] - U[:, o]
    assert np.abs(dU).max() > 1e-2
    # the ratio's 1e-12 regulariser moves r off 1 by 1e-12 / dU: 2e-13 here
    np.testing.assert_allclose(UL[:, f], U[:, o] + 0.375 * dU, rtol=0, atol=1e-11)
    np.testing.assert_allclose(UR[:, f], U[:, nb] - 0.375 * dU, rtol=0, atol=1e-11)


def test_tvd_reconstruction_stays_between_value_and_midpoint():
    """For any data, 0 <= phi(r) <= 1 and phi(r) / r <= 1 for r > 0 (and
    phi = 0 at extrema, r <= 0), so each trace lies between its own subcell
    value and the mean of the two (checked on the unclamped density and
    momentum, random data with extrema)."""
    op = _op(2, split=0.4)
    fv = SubcellFV(op, reconstruction="tvd")
    rng = np.random.default_rng(5)
    for t, rs in fv.references.items():
        ne = len(fv.projections[t].volumes)
        U = np.empty((ne, rs.count, 5))
        U[..., 0] = 1.0 + 0.5 * rng.random((ne, rs.count))
        U[..., 1:4] = rng.standard_normal((ne, rs.count, 3))
        U[..., 4] = 20.0 + rng.random((ne, rs.count))
        UL, UR = fv._reconstruct(t, U)
        o, nb = rs.internal_owner, rs.internal_neighbor
        mid = 0.5 * (U[:, o] + U[:, nb])
        tol = 1e-14
        for side, own in ((UL, U[:, o]), (UR, U[:, nb])):
            lo, hi = np.minimum(own, mid), np.maximum(own, mid)
            assert np.all(side[..., :4] >= lo[..., :4] - tol), t
            assert np.all(side[..., :4] <= hi[..., :4] + tol), t
        # some traces are really limited (extrema) and some really moved
        assert np.any(np.abs(UL[..., 0] - U[:, o, 0]) < 1e-15)
        assert np.any(np.abs(UL[..., 0] - U[:, o, 0]) > 1e-3)


def test_weno_reconstruction_linear_data_closed_form():
    """Linear subcell data on the inner faces: both one-sided slopes and the
    jump are equal (s_o = s_n = dU), so the three candidates of
    hybrid/subcell.py:617-645 are the midpoint (central and one-sided) and
    the cell value, with smoothness indicators beta_h = beta_o = dU^2 / 10,
    beta_c = dU^2 and tau = 9 dU^2 / 10.  The weights are then
    0.70 (1 + 81), 0.28 (1 + 81) and 0.02 (1 + 0.81): the trace is the
    cell value moved (1 - w) / 2 of the jump, w = 0.0362 / 80.3962."""
    op = _op(2, split=0.0)
    fv = SubcellFV(op, reconstruction="weno")
    t = "hexahedron"
    f = _interior_faces(fv, t)
    U = _linear_subcell_averages(fv, t)
    UL, UR = fv._reconstruct(t, U)
    o, nb = fv.references[t].internal_owner[f], fv.references[t].internal_neighbor[f]
    dU = U[:, nb] - U[:, o]
    w = 0.02 * 1.81 / (0.70 * 82 + 0.28 * 82 + 0.02 * 1.81)
    big = np.abs(dU) > 1e-3        # the 1e-12 regulariser is negligible there
    assert big.sum() > 100
    want_l = U[:, o] + 0.5 * (1 - w) * dU
    want_r = U[:, nb] - 0.5 * (1 - w) * dU
    assert np.abs((UL[:, f] - want_l)[big]).max() < 1e-12
    assert np.abs((UR[:, f] - want_r)[big]).max() < 1e-12
    # components without a jump keep the cell value
    assert np.abs((UL[:, f] - U[:, o])[np.abs(dU) < 1e-14]).max() < 1e-14


def test_tvd_reconstruction_quadratic_data_closed_form():
    """rho = 1 + (x + 1)^2 on the inner faces: the Green-Gauss gradient at an
    inner subcell centre is exact for a quadratic on the uniform subdivision
    (2 (x_c + 1)), while the jump to the neighbour is not the one-sided slope
    times the distance, so r = dU / (g_o . dc) != 1 and the trace is
    U_o + phi(r) g_o . dc / 2 with the limiter of hybrid/subcell.py:600-615,
    not a multiple of the jump."""
    op = _op(2, split=0.0)
    fv = SubcellFV(op, reconstruction="tvd")
    t = "hexahedron"
    f = _interior_faces(fv, t)
    c = fv.projections[t].physical_centers
    U = np.zeros(c.shape[:2] + (5,))
    U[..., 0] = 1.0 + (c[..., 0] + 1.0) ** 2
    U[..., 4] = 3.0
    UL, UR = fv._reconstruct(t, U)
    o, nb = fv.references[t].internal_owner[f], fv.references[t].internal_neighbor[f]
    dx = c[:, nb, 0] - c[:, o, 0]
    along = np.abs(dx) > 1e-9          # faces normal to x (4 of the 12)
    assert along.sum() == 4 * len(c)

    def phi(r):
        return np.clip(np.where(r > 0, (r * r + 2 * r) / (r * r + r + 2), 0.0), 0.0, 1.0)

    dU = U[:, nb, 0] - U[:, o, 0]
    so = 2 * (c[:, o, 0] + 1) * dx         # exact gradient at the owner centre . dc
    sn = -2 * (c[:, nb, 0] + 1) * dx       # and at the neighbour, towards the owner
    with np.errstate(invalid="ignore", divide="ignore"):   # faces not normal to x: 0 / 0
        want_l = U[:, o, 0] + 0.5 * phi(dU / so) * so
        want_r = U[:, nb, 0] + 0.5 * phi(-dU / sn) * sn
        assert np.abs(dU / so - 1)[along].min() > 1e-2
    assert np.abs(UL[:, f, 0] - want_l)[along].max() < 1e-10
    assert np.abs(UR[:, f, 0] - want_r)[along].max() < 1e-10
    assert np.abs(UL[:, f, 0] - U[:, o, 0])[~along].max() == 0.0


def test_weno_reconstruction_quadratic_data_closed_form():
    """rho = 1 + (x + 1)^2 on the inner faces normal to x: exact one-sided
    slopes s_o = 2 (x_o + 1) dx and s_n = 2 (x_n + 1) dx (Green-Gauss is
    exact there), secant dU = s_o + dx^2, all three different.  The trace is
    the CWENO-Z combination of hybrid/subcell.py:604-638 (central Hermite
    value, one-sided value, cell value; linear weights 0.70 / 0.28 / 0.02;
    beta from the slope differences, tau = |beta_c - beta_h|), written here
    from those lines with the exact slopes."""
    op = _op(2, split=0.0)
    fv = SubcellFV(op, reconstruction="weno")
    t = "hexahedron"
    f = _interior_faces(fv, t)
    c = fv.projections[t].physical_centers
    U = np.zeros(c.shape[:2] + (5,))
    U[..., 0] = 1.0 + (c[..., 0] + 1.0) ** 2
    U[..., 4] = 3.0
    UL, UR = fv._reconstruct(t, U)
    o, nb = fv.references[t].internal_owner[f], fv.references[t].internal_neighbor[f]
    dx = c[:, nb, 0] - c[:, o, 0]
    along = np.abs(dx) > 1e-9
    uo, un = U[:, o, 0], U[:, nb, 0]
    so = 2 * (c[:, o, 0] + 1) * dx
    sn = 2 * (c[:, nb, 0] + 1) * dx
    sec = un - uo
    e = 1e-12
    bh = (sn - so) ** 2 + 0.1 * sec ** 2
    bo = (sec - so) ** 2 + 0.1 * sec ** 2
    bn = (sec - sn) ** 2 + 0.1 * sec ** 2
    bc = sec ** 2
    tau = np.abs(bc - bh)
    central = 0.5 * (uo + un) + 0.125 * (so - sn)
    w = [0.70 * (1 + (tau / (bh + e)) ** 2), 0.28 * (1 + (tau / (bo + e)) ** 2),
         0.02 * (1 + (tau / (bc + e)) ** 2)]
    want_l = (w[0] * central + w[1] * (uo + 0.5 * so) + w[2] * uo) / (w[0] + w[1] + w[2])
    w[1] = 0.28 * (1 + (tau / (bn + e)) ** 2)
    want_r = (w[0] * central + w[1] * (un - 0.5 * sn) + w[2] * un) / (w[0] + w[1] + w[2])
    assert np.abs(sec - so)[along].min() > 1e-4 and np.abs(sn - so)[along].min() > 1e-4
    assert np.abs(UL[:, f, 0] - want_l)[along].max() < 1e-12
    assert np.abs(UR[:, f, 0] - want_r)[along].max() < 1e-12


def test_positivity_limiter_is_tight():
    """The limiter (hybrid/limiter.py:100-147) scales the deviations from the
    mean by the largest theta that keeps every sample admissible: theta_rho
    in closed form puts the lowest sampled density exactly on its floor, and
    the bisection on theta_p leaves the lowest sampled pressure within the
    bisection's resolution of its floor.  So every limited element ends with
    one of the two at its floor (a limiter that shrinks further, to the mean
    say, satisfies the floors too but fails this)."""
    op = _op(2)
    pos = PositivityLimiter(op, density_floor=1e-6, pressure_floor=1e-6)
    st = _states(op, lambda x: perturbed_primitive(x, seed=4))
    bad = {}
    for t, U in st.items():
        W = op.solver.conserved_to_primitive(U.copy())
        W[::2, 0, 0] = -0.05
        W[1::3, 1, 4] = -0.02
        bad[t] = op.solver.primitive_to_conserved(W)
    out = pos(bad)
    limited = 0
    for bid, b in enumerate(op.mesh.element_blocks):
        t = b.element_type
        Wm = op.solver.conserved_to_primitive(pos.mean(bid, bad[t]))
        W = op.solver.conserved_to_primitive(pos.sample(t, out[t]))
        rho_floor = np.maximum(1e-6, 0.05 * Wm[:, 0])
        p_floor = np.maximum(1e-6, 0.05 * Wm[:, 4])
        changed = np.abs(out[t] - bad[t]).max((1, 2)) > 1e-12
        on_rho = np.abs(W[..., 0].min(1) - rho_floor) <= 1e-10 * rho_floor
        on_p = (W[..., 4].min(1) - p_floor) <= 1e-6 * Wm[:, 4]
        assert np.all((on_rho | on_p)[changed]), t
        limited += changed.sum()
    assert limited > 10
