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
