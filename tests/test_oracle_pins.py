"""Pins of the oracle functions that the closed-form operator tests leave
open: the BR2 lift magnitude and the sliver-scaled inverse mass, the boundary
ghosts and the adiabatic-wall energy flux, the HLLC wave and contact speeds,
and the Park mechanism data and its equilibrium constants.

Each check is chosen so that a plausible mistake in the oracle fails it:
* BR2: int_K r_f dV = 1/2 s_K int_f [[v]] (x) n dA with s_K = min(1, V/1e-4)
  (a factor 1 instead of 1/2, a missing M^-1, a missing sliver factor or the
  penalty applied in the volume gradient each break it);
* sliver scaling of the rhs: int_K rhs dV = s_K sum_i R_i;
* slip walls: the HLLC flux against the reflected state carries no mass or
  energy, and its momentum is p* n with the star pressure of the mirror
  problem; a box closed by slip walls conserves mass and energy to round-off;
* no-slip / isothermal / inlet ghosts: face averages in closed form;
* adiabatic walls: the viscous energy flux through the walls is zero, so a box
  closed by walls conserves energy in the viscous residual;
* HLLC: both star pressures agree at the contact speed S_M, the outer speeds
  bound every acoustic speed and equal one of them, and identical states give
  (u - c, u, u + c);
* Park: species and reaction tables equal the fixture read from the
  reference's data file; q = 0 where the chemical potentials balance
  (detailed balance); NASA-7 h and s are consistent integrals of cp and match
  standard-state values.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from conftest import perturbed_primitive

from hybrid.meshgen import box_mesh
from oracle.chemistry import MOLAR_MASS, NASA7, P_STD, R_UNIV, REACTIONS, SPECIES, Mechanism
from oracle.geometry import SLIVER_VOLUME
from oracle.mesh import from_arrays
from oracle.operators import EulerOperator, NavierStokesOperator
from oracle.solver import CompressibleSolver, FarfieldWallBoundary

GOLDEN = Path(__file__).parent / "golden" / "air11_park_tables.json"
ALL_TAGS = (1, 2, 3, 4, 5, 6)


def _states(op, prim):
    s = op.solver
    return {t: s.primitive_to_conserved(prim(x))
            for t, x in op.geometry.solution_node_coordinates.items()}


def _basis_integrals(op, bid):
    """int_K phi_i dV per element, (E, N)."""
    t = op.mesh.element_blocks[bid].element_type
    ref = op.geometry.reference_elements[t]
    return op.geometry.element_metrics[bid].integration_weights @ ref.basis(ref.volume_points)


def _sliver(op, bid):
    v = op.geometry.element_metrics[bid].integration_weights.sum(1)
    return np.minimum(1.0, v / SLIVER_VOLUME)


MESHES = {
    # unit-size elements: s_K = 1
    "large": dict(n=3, lengths=(1.0, 1.2, 0.8), split_fraction=0.5, seed=8),
    # 0.1-box of 4^3: hexahedra 1.6e-5 (s = 0.16), tetrahedra 2.6e-6
    "sliver": dict(n=4, lengths=(0.1, 0.1, 0.1), split_fraction=0.3, seed=3),
}


# -- BR2 lift and the sliver-scaled inverse mass ---------------------------------
@pytest.mark.parametrize("mesh", sorted(MESHES))
@pytest.mark.parametrize("p", [1, 2])
def test_br2_lift_integral_identity(mesh, p):
    m = box_mesh(**MESHES[mesh])
    om = from_arrays(m)
    op = NavierStokesOperator(om, CompressibleSolver(dynamic_viscosity=0.01), p)
    st = _states(op, lambda x: perturbed_primitive(x * (10.0 if mesh == "sliver" else 1.0),
                                                   seed=4, amplitude=0.1))
    bnd = FarfieldWallBoundary(op.solver, [1.1, 0.5, 0.1, 0.0, 1.0], inlet_tags=(1,),
                               wall_tags=(3, 4, 5, 6), no_slip_tags=(5,), slip_tags=(3, 4, 6))
    grads, raw, face_data = op.br2_gradients(st, bnd)
    if mesh == "sliver":
        assert min(_sliver(op, b).max() for b in range(len(om.element_blocks))) < 0.5
    ints = {b: _basis_integrals(op, b) for b in range(len(om.element_blocks))}
    sl = {b: _sliver(op, b) for b in ints}
    expect_total = {t: np.zeros(U.shape[:1] + (4, 3)) for t, U in st.items()}
    worst = 0.0
    for d in face_data:
        w = d["fm"].integration_weights[d["idx"]]
        n = d["fm"].normals[d["idx"]]
        jn = np.einsum("fq,fqm,fqd->fmd", w, d["jump"], n)   # int_f [[v]] (x) n dA
        sides = [("o", d["ob"], d["ocells"])] + ([("n", d["nb"], d["ncells"])]
                                                 if d["nb"] >= 0 else [])
        for side, bid, cells in sides:
            lift = d[side + "lift"]                              # (F, N, 4, 3)
            got = np.einsum("fi,fimd->fmd", ints[bid][cells], lift)
            want = 0.5 * sl[bid][cells][:, None, None] * jn
            scale = np.abs(want).max() + 1e-300
            worst = max(worst, np.abs(got - want).max() / scale)
            np.add.at(expect_total[om.element_blocks[bid].element_type], cells, want)
    assert worst < 1e-12, worst
    # the volume gradient adds the summed lifts without the penalty
    for bid, blk in enumerate(om.element_blocks):
        t = blk.element_type
        wq = op.geometry.element_metrics[bid].integration_weights
        got = np.einsum("eq,eqmd->emd", wq, grads[t] - raw[t])
        want = expect_total[t]
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), t


@pytest.mark.parametrize("mesh", sorted(MESHES))
def test_rhs_applies_sliver_scaled_inverse_mass(mesh):
    m = box_mesh(**MESHES[mesh])
    om = from_arrays(m)
    op = EulerOperator(om, CompressibleSolver(), 2)
    st = _states(op, lambda x: perturbed_primitive(x * (10.0 if mesh == "sliver" else 1.0),
                                                   seed=6))
    R = op.residual(st)
    rhs = op.apply_inverse_mass(R)
    for bid, blk in enumerate(om.element_blocks):
        t = blk.element_type
        got = np.einsum("ei,eiv->ev", _basis_integrals(op, bid), rhs[t])
        want = _sliver(op, bid)[:, None] * R[t].sum(1)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), t


# -- boundary ghosts ------------------------------------------------------------------
def _random_primitive(n, seed):
    rng = np.random.default_rng(seed)
    W = np.stack([0.5 + rng.random(n), rng.standard_normal(n), rng.standard_normal(n),
                  rng.standard_normal(n), 0.5 + rng.random(n)], -1)
    nrm = rng.standard_normal((n, 3))
    return W, nrm / np.linalg.norm(nrm, axis=1, keepdims=True)


def test_slip_wall_flux_is_pure_pressure():
    """HLLC between a state and its slip-wall ghost (reflected normal velocity):
    zero mass and energy flux, momentum flux p* n with
    p* = p + rho u_n (c + |u_n| + u_n) (the left star pressure at S_M = 0,
    S_L = -|u_n| - c)."""
    s = CompressibleSolver()
    W, n = _random_primitive(300, 11)
    U = s.primitive_to_conserved(W)[:, None, :]
    b = FarfieldWallBoundary(s, [1.0, 0, 0, 0, 1.0], wall_tags=(3,))
    tags = np.full(300, 3)
    G = b(U, n[:, None, :], tags)
    F = s.hllc_flux(U, G, n[:, None, :])[:, 0]
    un = np.sum(W[:, 1:4] * n, 1)
    c = np.sqrt(1.4 * W[:, 4] / W[:, 0])
    pstar = W[:, 4] + W[:, 0] * un * (c + np.abs(un) + un)
    scale = np.abs(F).max()
    assert np.abs(F[:, 0]).max() < 1e-13 * scale
    assert np.abs(F[:, 4]).max() < 1e-13 * scale
    assert np.abs(F[:, 1:4] - pstar[:, None] * n).max() < 1e-13 * scale
    # tangential flow: exactly (0, p n, 0)
    Wt = W.copy()
    Wt[:, 1:4] -= un[:, None] * n
    Ut = s.primitive_to_conserved(Wt)[:, None, :]
    Ft = s.hllc_flux(Ut, b(Ut, n[:, None, :], tags), n[:, None, :])[:, 0]
    assert np.abs(Ft[:, 1:4] - W[:, 4:5] * n).max() < 1e-12 * scale


def test_viscous_ghosts_closed_forms():
    s = CompressibleSolver(dynamic_viscosity=0.01, gas_constant=1.3)
    W, n = _random_primitive(200, 12)
    U = s.primitive_to_conserved(W)[:, None, :]
    nn = n[:, None, :]
    far = np.array([1.2, 0.9, 0.1, -0.05, 1.1])
    vin = s.viscous_variables(U)[:, 0]
    vel, Tin = vin[:, :3], vin[:, 3]
    assert np.allclose(Tin, W[:, 4] / (W[:, 0] * 1.3), rtol=1e-14)
    for Tw in (None, 2.5):
        b = FarfieldWallBoundary(s, far, inlet_tags=(1,), wall_tags=(3, 5),
                                 no_slip_tags=(5,), slip_tags=(3,), wall_temperature=Tw)
        # no-slip: the face average of the velocity vanishes; T averages to T_wall
        g = b.viscous_ghost(U, nn, np.full(200, 5))[:, 0]
        assert np.abs(0.5 * (g[:, :3] + vel)).max() == 0.0
        Tavg = 0.5 * (g[:, 3] + Tin)
        assert np.allclose(Tavg, Tin if Tw is None else Tw, rtol=1e-14, atol=0)
        # slip: zero average normal velocity, tangential velocity kept, T kept
        g = b.viscous_ghost(U, nn, np.full(200, 3))[:, 0]
        avg = 0.5 * (g[:, :3] + vel)
        assert np.abs(np.sum(avg * n, 1)).max() < 1e-14 * np.abs(vel).max()
        tang = vel - np.sum(vel * n, 1)[:, None] * n
        assert np.abs(avg - tang).max() < 1e-14 * np.abs(vel).max()
        assert np.array_equal(g[:, 3], Tin)
        # inlet: the farfield's velocity and temperature, in both ghosts
        g = b.viscous_ghost(U, nn, np.full(200, 1))[:, 0]
        assert np.allclose(g, np.broadcast_to([*far[1:4], far[4] / (far[0] * 1.3)], g.shape),
                           rtol=1e-15, atol=0)
        G = b(U, nn, np.full(200, 1))[:, 0]
        assert np.allclose(s.conserved_to_primitive(G), np.broadcast_to(far, G.shape),
                           rtol=1e-15, atol=1e-15)
        # outlet / unknown tags copy the inside state
        assert np.array_equal(b(U, nn, np.full(200, 2)), s.primitive_to_conserved(
            s.conserved_to_primitive(U)))


def _closed_box(viscous, walls):
    m = box_mesh(3, lengths=(1.0, 1.3, 0.9), split_fraction=0.4, seed=5)
    om = from_arrays(m)
    s = CompressibleSolver(dynamic_viscosity=0.02 if viscous else 0.0)
    op = (NavierStokesOperator if viscous else EulerOperator)(om, s, 2)
    b = FarfieldWallBoundary(s, [1.0, 0, 0, 0, 1.0],
                             outlet_tags=() if walls else ALL_TAGS,
                             wall_tags=ALL_TAGS if walls else (),
                             slip_tags=ALL_TAGS if walls else ())
    st = _states(op, lambda x: perturbed_primitive(x, seed=9, amplitude=0.1))
    return op, b, st


def _total(res):
    tot = sum(v.sum((0, 1)) for v in res.values())
    scale = sum(np.abs(v).sum((0, 1)) for v in res.values())
    return tot, scale


def test_slip_walls_conserve_mass_and_energy():
    op, b, st = _closed_box(False, True)
    tot, scale = _total(op.residual(st, b))
    assert abs(tot[0]) < 1e-13 * scale.max() and abs(tot[4]) < 1e-13 * scale.max()
    # control: transmissive boundaries let mass through
    op, b, st = _closed_box(False, False)
    tot, scale = _total(op.residual(st, b))
    assert abs(tot[0]) > 1e-6 * scale.max()


def test_adiabatic_walls_carry_no_viscous_energy_flux():
    op, b, st = _closed_box(True, True)
    tot, scale = _total(op.viscous_residual(st, b))
    assert abs(tot[0]) == 0.0
    assert abs(tot[4]) < 1e-13 * scale.max()
    # control: the same box with outlets has a nonzero viscous energy flux
    op, b, st = _closed_box(True, False)
    tot, scale = _total(op.viscous_residual(st, b))
    assert abs(tot[4]) > 1e-6 * scale.max()


# -- HLLC wave and contact speeds ---------------------------------------------------------
def test_hllc_contact_speed_balances_star_pressures():
    s = CompressibleSolver()
    WL, n = _random_primitive(500, 21)
    WR, _ = _random_primitive(500, 22)
    UL, UR = s.primitive_to_conserved(WL), s.primitive_to_conserved(WR)
    sl, sm, sr = s.hllc_wave_speeds(UL, UR, n)
    ul, ur = np.sum(WL[:, 1:4] * n, 1), np.sum(WR[:, 1:4] * n, 1)
    pl = WL[:, 4] + WL[:, 0] * (sl - ul) * (sm - ul)
    pr = WR[:, 4] + WR[:, 0] * (sr - ur) * (sm - ur)
    assert np.abs(pl - pr).max() < 1e-12 * np.abs(pl).max()
    # the star states move with S_M and keep the tangential velocity
    SL = s.hllc_star_state(UL, WL, sl, sm, n)
    vs = SL[:, 1:4] / SL[:, :1]
    assert np.abs(np.sum(vs * n, 1) - sm).max() < 1e-12 * np.abs(sm).max()
    tl = WL[:, 1:4] - ul[:, None] * n
    assert np.abs(vs - np.sum(vs * n, 1)[:, None] * n - tl).max() < 1e-12


def test_hllc_outer_speeds_bound_both_sides():
    s = CompressibleSolver()
    WL, n = _random_primitive(500, 23)
    WR, _ = _random_primitive(500, 24)
    WR[:, 4] *= 4.0     # sound speeds differ widely between the sides
    sl, sm, sr = s.hllc_wave_speeds(s.primitive_to_conserved(WL), s.primitive_to_conserved(WR), n)
    cands = []
    for W in (WL, WR):
        u = np.sum(W[:, 1:4] * n, 1)
        c = np.sqrt(1.4 * W[:, 4] / W[:, 0])
        cands.append((u - c, u + c))
    lo = np.stack([cands[0][0], cands[1][0]])
    hi = np.stack([cands[0][1], cands[1][1]])
    assert np.all(sl <= lo.min(0) + 1e-14 * np.abs(lo).max(0))
    assert np.all(sr >= hi.max(0) - 1e-14 * np.abs(hi).max(0))
    assert np.all(np.isclose(lo, sl, rtol=1e-15, atol=1e-15).any(0))
    assert np.all(np.isclose(hi, sr, rtol=1e-15, atol=1e-15).any(0))
    # the left estimate takes the right state's speed in some cells
    assert np.any(~np.isclose(lo[0], sl, rtol=1e-15, atol=1e-15))


def test_hllc_identical_states_speeds():
    s = CompressibleSolver()
    W, n = _random_primitive(100, 25)
    U = s.primitive_to_conserved(W)
    sl, sm, sr = s.hllc_wave_speeds(U, U, n)
    u = np.sum(W[:, 1:4] * n, 1)
    c = np.sqrt(1.4 * W[:, 4] / W[:, 0])
    assert np.allclose(sl, u - c, rtol=1e-14, atol=1e-14)
    assert np.allclose(sm, u, rtol=1e-12, atol=1e-12)
    assert np.allclose(sr, u + c, rtol=1e-14, atol=1e-14)


# -- Park mechanism data ----------------------------------------------------------------------
def _golden():
    return json.loads(GOLDEN.read_text())


def _fixture_arrays():
    g = _golden()
    ns, nr = len(g["species"]), len(g["reactions"])
    mode = {"T": 0, "Ta": 1, "Tv": 2}
    nu1, nu2, alpha = np.zeros((nr, ns)), np.zeros((nr, ns)), np.zeros((nr, ns))
    rx = np.zeros((nr, 6))
    for r, x in enumerate(g["reactions"]):
        for s, c in x["reactants"]:
            nu1[r, s] += c
        for s, c in x["products"]:
            nu2[r, s] += c
        if x["third_body"]:
            alpha[r] = [0.0 if sp["name"] == "e-" else 1.0 for sp in g["species"]]
            for s, e in x["efficiencies"]:
                alpha[r, s] = e
        rx[r] = [x["A"], x["b"], x["theta"], float(x["third_body"]), mode[x["fwd_temp"]],
                 mode[x["rev_temp"]]]
    return g, nu1, nu2, alpha, rx


def test_oracle_mechanism_matches_reference_data():
    g, nu1, nu2, alpha, rx = _fixture_arrays()
    assert all(x["reversible"] for x in g["reactions"])   # the oracle treats all as reversible
    assert tuple(sp["name"] for sp in g["species"]) == SPECIES
    assert np.array_equal(MOLAR_MASS, [sp["molar_mass"] for sp in g["species"]])
    for (lo, mid, hi, cl, ch), sp in zip(NASA7, g["species"]):
        f = sp["nasa7"]
        assert (lo, mid, hi) == (f["t_low"], f["t_mid"], f["t_high"])
        assert np.array_equal(np.asarray(cl, float), f["low"])
        assert np.array_equal(np.asarray(ch, float), f["high"])
    m = Mechanism()
    assert np.array_equal(m.nu1, nu1) and np.array_equal(m.nu2, nu2)
    assert np.array_equal(m.alpha, alpha)
    assert np.array_equal(np.stack([m.A, m.b, m.theta, m.third, m.fwd_mode, m.rev_mode], 1), rx)


def test_product_mechanism_table_matches_reference_data():
    from hybrid.chemistry.air11_park import mechanism_table
    g, nu1, nu2, alpha, rx = _fixture_arrays()
    ns, nr = nu1.shape[1], nu1.shape[0]
    tab = mechanism_table()
    nasa = np.concatenate([[f["t_low"], f["t_mid"], f["t_high"], *f["low"], *f["high"]]
                           for f in (sp["nasa7"] for sp in g["species"])])
    want = np.concatenate([[sp["molar_mass"] for sp in g["species"]], nasa, rx.ravel(),
                           nu1.ravel(), nu2.ravel(), alpha.ravel()])
    # the reference data, then 3 derived values per reaction, 18 per species
    # (pinned in test_oracle_chemistry) and the stoichiometry fingerprint
    assert len(tab) == len(want) + 3 * nr + 18 * ns + 1 and np.array_equal(tab[:len(want)], want)
    assert ns == 11 and nr == 22


def _chemical_potentials(mech, C, T):
    """mu_s / (R T) of ideal-gas species at concentration C (mol/m^3)."""
    return mech.g_over_RT(T) + np.log(C * R_UNIV * T / P_STD)


@pytest.mark.parametrize("T", [2500.0, 7000.0, 11000.0])
def test_detailed_balance(T):
    """At concentrations where sum nu'' mu = sum nu' mu (chemical equilibrium of
    reaction r), the reaction's net rate of progress vanishes: K_c must be the
    equilibrium constant of the mechanism's own thermodynamics."""
    rng = np.random.default_rng(int(T))
    full = Mechanism()
    for r in range(full.nr):
        one = Mechanism(reactions=(REACTIONS[r],))
        C = np.exp(rng.uniform(np.log(1e-3), np.log(10.0), full.ns))
        # solve for the first product's concentration (it appears with nu'' >= 1)
        k = int(np.flatnonzero(one.nu2[0] > one.nu1[0])[0])
        dnu = one.dnu[0]
        mu = _chemical_potentials(full, C, T)
        imbalance = np.dot(dnu, mu)
        C[k] *= np.exp(-imbalance / dnu[k])
        assert abs(np.dot(dnu, _chemical_potentials(full, C, T))) < 1e-10
        omega, _ = one.net_production_rates(C * MOLAR_MASS, T)
        kf = one.forward_rate_constants(T, T)[0]
        fwd = kf * np.prod(C ** one.nu1[0]) * (one.alpha[0] @ C if one.third[0] else 1.0)
        assert np.abs(omega).max() <= 1e-9 * fwd * np.abs(dnu).max(), (r, omega, fwd)
        # and out of balance the sign follows the affinity
        C2 = C.copy()
        C2[k] *= 2.0
        omega2, _ = one.net_production_rates(C2 * MOLAR_MASS, T)
        assert omega2[k] < 0.0


def test_nasa7_integral_consistency():
    """d(h/R)/dT = cp/R and d(s/R)/dT = cp/(R T) inside each fit range."""
    m = Mechanism()
    for T in (500.0, 800.0, 1500.0, 3000.0, 5500.0):
        h = 1e-3 * T
        cp = m.cp_over_R(T)
        dh = ((m.h_over_RT(T + h) * (T + h)) - (m.h_over_RT(T - h) * (T - h))) / (2 * h)
        ds = (m.s_over_R(T + h) - m.s_over_R(T - h)) / (2 * h)
        assert np.allclose(dh, cp, rtol=1e-5), T     # central difference: O((h/T)^2)
        assert np.allclose(ds * T, cp, rtol=1e-5), T
    # clamped above t_high: cp frozen at its t_high value
    assert np.array_equal(m.cp_over_R(9000.0), m.cp_over_R(6000.0))


def test_nasa7_standard_state_values():
    """Standard enthalpies of formation (kJ/mol) and entropies (J/mol/K) at
    298.15 K (NIST-JANAF / Gurvich tables) within 1 %."""
    m = Mechanism()
    T = 298.15
    h = m.h_over_RT(T) * R_UNIV * T / 1e3
    s = m.s_over_R(T) * R_UNIV
    idx = {n: i for i, n in enumerate(SPECIES)}
    for name, hf in (("N", 472.68), ("O", 249.18), ("NO", 91.27), ("N+", 1882.1),
                     ("O+", 1568.8)):
        assert abs(h[idx[name]] - hf) < 0.01 * hf, (name, h[idx[name]])
    for name in ("N2", "O2"):
        assert abs(h[idx[name]]) < 0.05, (name, h[idx[name]])
    for name, s0 in (("N2", 191.61), ("O2", 205.15), ("N", 153.30), ("O", 161.06),
                     ("NO", 210.76)):
        assert abs(s[idx[name]] - s0) < 0.01 * s0, (name, s[idx[name]])
