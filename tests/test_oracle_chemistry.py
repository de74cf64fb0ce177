"""Pins of the chemistry oracle: the reference's own known-answer and
invariant tests (tests/test_air11_chemistry.py, tests/test_reacting_coupler.py)
restated, plus a finite-difference check of the analytic Newton Jacobian and
the small-dt linearisation of the implicit step."""

import numpy as np
import pytest

from oracle.chemistry import MOLAR_MASS, R_UNIV, Mechanism, ReactingFlowCoupler, integrator_step

MECH = Mechanism()


def test_thermodynamics():
    """tests/test_air11_chemistry.py:19-37."""
    assert MECH.ns == 11 and MECH.nr == 22
    for T in (300.0, 1000.0, 4000.0, 8000.0):
        assert np.all(np.isfinite(MECH.g_over_RT(T)))
        rho = np.zeros(11)
        rho[9], rho[10] = 0.8, 0.2
        pr = MECH.mixture_properties(rho, T)
        assert abs(pr["density"] - 1.0) < 1e-6
        assert pr["cp"] > pr["cv"] > 0
        assert 1.0 < pr["gamma_frozen"] < 1.7
        assert pr["speed_of_sound"] > 200.0


def test_park_rate_known_answer():
    """tests/test_air11_chemistry.py:40-62: omega_N = 2 kf C_N2 (C_N2 + C_O2)."""
    T = 8000.0
    cN2, cO2 = 0.8 / (2 * 0.0140067), 0.4 / (2 * 0.015999)
    kf = 7.0e15 * T ** -1.6 * np.exp(-113200.0 / T)
    expect = 2.0 * kf * cN2 * (cN2 + cO2)
    rho = np.zeros(11)
    rho[9], rho[10] = 0.8, 0.4
    omega, _ = MECH.net_production_rates(rho, T)
    assert abs(omega[6] - expect) / expect < 1e-3


def test_mass_conservation():
    """tests/test_air11_chemistry.py:65-86."""
    states = [np.array([0, 0, 0, 0, 0, 0, 0, 0, 0, 0.79, 0.21]),
              np.array([1e-6, 1e-5, 1e-5, 1e-4, 1e-5, 1e-5, 0.2, 0.1, 0.05, 0.4, 0.1]),
              np.array([1e-4, 0.05, 0.05, 0.05, 0.02, 0.02, 0.3, 0.2, 0.01, 0.1, 0.05])]
    for s in states:
        for T in (500.0, 3000.0, 8000.0, 15000.0):
            _, w = MECH.net_production_rates(s, T)
            mx = np.abs(w).max()
            if mx > 1e-10:
                assert abs(w.sum()) / mx < 1e-14
            else:
                assert abs(w.sum()) < 1e-12


def test_point_implicit_integrator():
    """tests/test_air11_chemistry.py:109-129."""
    rho0 = np.zeros(11)
    rho0[9], rho0[10] = 0.8, 0.4
    r, _, _ = integrator_step(MECH, rho0, 8000.0, 1e-6, iters=16, constant_temperature=True)
    assert np.all(r >= 0)
    assert abs(r.sum() - rho0.sum()) < 1e-12
    assert r[6] > 0 and r[7] > 0
    assert r[9] < rho0[9] and r[10] < rho0[10]


def test_generic_mechanism_extensibility():
    """tests/test_air11_chemistry.py:132-172: a toy H2/O2 mechanism."""
    nasa = (
        (200.0, 1000.0, 6000.0, (3.298, 8.249e-4, -8.143e-7, -9.475e-11, 4.134e-13, -1012.5, -3.294),
         (2.932, 8.266e-4, -1.464e-7, 1.541e-11, -6.888e-16, -813.0, -1.024)),
        (200.0, 1000.0, 6000.0, (3.782, -2.996e-3, 9.847e-6, -9.681e-9, 3.243e-12, -1063.9, 3.657),
         (3.660, 6.563e-4, -1.411e-7, 2.057e-11, -1.299e-15, -1215.9, 3.415)),
        (200.0, 1000.0, 6000.0, (4.198, -2.036e-3, 6.520e-6, -5.487e-9, 1.771e-12, -30293.7, -0.849),
         (2.677, 2.973e-3, -7.737e-7, 9.443e-11, -4.269e-15, -29885.8, 6.882)))
    rx = ((((0, 2), (1, 1)), ((2, 2),), 1.0e11, 0.0, 15000.0, False, (), "T", "T"),)
    m = Mechanism(nasa, rx, (0.002016, 0.031998, 0.018015), ("H2", "O2", "H2O"))
    _, w = m.net_production_rates(np.array([0.2, 0.8, 0.01]), 2000.0)
    assert w[2] > 0 and w[0] < 0 and w[1] < 0 and abs(w.sum()) < 1e-12


@pytest.mark.parametrize("T", [3000.0, 8000.0, 12000.0])
def test_jacobian_matches_finite_differences(T):
    rng = np.random.default_rng(int(T))
    rho = rng.dirichlet(np.full(11, 2.0)) * 0.05
    _, w, J = MECH.net_production_rates(rho, T, jacobian=True)
    for j in range(11):
        h = 1e-4 * rho[j]  # central difference; larger steps lose to round-off
        e = np.zeros(11)
        e[j] = h
        fd = (MECH.net_production_rates(rho + e, T)[1] - MECH.net_production_rates(rho - e, T)[1]) / (2 * h)
        scale = np.abs(fd).max() + 1e-300
        assert np.abs(fd - J[:, j]).max() <= 1e-6 * scale, j


def test_implicit_step_solves_backward_euler():
    """After the fixed Newton iterations the result satisfies
    rho - rho0 - dt w(rho) = 0 (the backward-Euler equation), and for a tiny
    step it is the linearisation rho0 + (I - dt J)^-1 dt w."""
    rho = np.array([1e-8, 1e-6, 1e-6, 1e-5, 1e-6, 1e-6, 0.01, 0.02, 0.005, 0.6, 0.15])
    T = 6000.0
    for dt in (1e-12, 1e-9):
        r, _, _ = integrator_step(MECH, rho, T, dt, iters=6, constant_temperature=True)
        _, w = MECH.net_production_rates(r, T)
        assert np.abs(r - rho - dt * w).max() <= 1e-9 * np.abs(dt * w).max()
    _, w0, J = MECH.net_production_rates(rho, T, jacobian=True)
    r1, _, _ = integrator_step(MECH, rho, T, 1e-12, iters=1, constant_temperature=True)
    lin = rho + np.linalg.solve(np.eye(11) - 1e-12 * J, 1e-12 * w0)
    assert np.abs(r1 - lin).max() <= 1e-9 * np.abs(1e-12 * w0).max()


def test_temperature_inversion_roundtrip():
    """tests/test_reacting_coupler.py:20-33."""
    rho = np.zeros(11)
    rho[9], rho[10] = 0.767, 0.233
    for T in (500.0, 1500.0, 4000.0, 8000.0):
        e = MECH.mixture_properties(rho, T)["internal_energy"]
        assert abs(MECH.recover_temperature_from_energy(rho, e, 1000.0) - T) < 1.0


def test_coupler_block_invariants():
    """tests/test_reacting_coupler.py:35-91 (block (4, 10, 16), 6000 K,
    u = 2000 m/s, dt = 1e-7): species sum = rho, rho/momentum/energy
    unchanged, positivity, O produced."""
    c = ReactingFlowCoupler(MECH, max_newton_iters=16)
    rho, u, T = 1.2, 2000.0, 6000.0
    Y = np.zeros(11)
    Y[9], Y[10] = 0.767, 0.233
    e = MECH.mixture_properties(rho * Y, T)["internal_energy"]
    node = np.zeros(16)
    node[0], node[1], node[4] = rho, rho * u, rho * (e + 0.5 * u * u)
    node[5:] = rho * Y
    U = np.broadcast_to(node, (4, 10, 16)).copy()
    out = c.react_block(U[:1, :2], 1e-7)
    np.testing.assert_allclose(out[..., 5:].sum(-1), rho, rtol=1e-7)
    np.testing.assert_allclose(out[..., :5], U[:1, :2, :5], rtol=1e-7)
    assert np.all(out[..., 5:] >= 0)
    assert out[0, 0, 5 + 7] > 1e-10


def test_device_table_transcription_matches_oracle():
    """The device mechanism table (product) and the oracle's mechanism were
    transcribed separately from hybrid/chemistry/air11_park.py; they must hold
    identical numbers."""
    from hybrid.chemistry.air11_park import mechanism_table
    tab = mechanism_table()
    ns, nr = 11, 22
    assert np.array_equal(tab[:ns], MOLAR_MASS)
    nasa = tab[ns:ns + 17 * ns].reshape(ns, 17)
    for s, (lo, mid, hi, cl, ch) in enumerate(MECH.nasa):
        assert np.array_equal(nasa[s], np.array([lo, mid, hi, *cl, *ch], float)), s
    o = ns + 17 * ns
    rx = tab[o:o + 6 * nr].reshape(nr, 6)
    assert np.array_equal(rx[:, 0], MECH.A) and np.array_equal(rx[:, 1], MECH.b)
    assert np.array_equal(rx[:, 2], MECH.theta)
    assert np.array_equal(rx[:, 3].astype(bool), MECH.third)
    assert np.array_equal(rx[:, 4], MECH.fwd_mode) and np.array_equal(rx[:, 5], MECH.rev_mode)
    o += 6 * nr
    for arr in (MECH.nu1, MECH.nu2, MECH.alpha):
        assert np.array_equal(tab[o:o + nr * ns].reshape(nr, ns), arr)
        o += nr * ns
    # derived per reaction: ln A, sum(nu'' - nu'), sum M (nu'' - nu')
    der = tab[o:o + 3 * nr].reshape(nr, 3)
    d = MECH.nu2 - MECH.nu1
    assert np.allclose(np.exp(der[:, 0]), MECH.A, rtol=1e-14, atol=0)
    assert np.array_equal(der[:, 1], d.sum(1))
    assert np.allclose(der[:, 2], d @ MOLAR_MASS, rtol=0, atol=1e-15)
    o += 3 * nr
    # derived per species: 1/M, R/M, 1/t_low, 1/t_high, divided NASA-7
    # coefficients per fit range; the Horner forms they feed equal the power
    # sums of the fits
    sp = tab[o:o + 18 * ns].reshape(ns, 18)
    assert o + 18 * ns + 1 == len(tab)   # then the stoichiometry fingerprint
    assert np.allclose(sp[:, 0] * MOLAR_MASS, 1.0, rtol=1e-15, atol=0)
    assert np.allclose(sp[:, 1] * MOLAR_MASS, R_UNIV, rtol=1e-15, atol=0)
    for s, (lo, mid, hi, cl, ch) in enumerate(MECH.nasa):
        assert sp[s, 2] * lo == pytest.approx(1.0, rel=1e-15)
        assert sp[s, 3] * hi == pytest.approx(1.0, rel=1e-15)
        for a, d7, t in ((cl, sp[s, 4:11], 0.5 * (lo + mid)), (ch, sp[s, 11:18], 0.5 * (mid + hi))):
            h_power = (a[0] + a[1] * t / 2 + a[2] * t ** 2 / 3 + a[3] * t ** 3 / 4
                       + a[4] * t ** 4 / 5 + a[5] / t)
            h_horner = a[0] + t * (d7[0] + t * (d7[1] + t * (d7[2] + t * d7[3]))) + a[5] / t
            s_power = (a[0] * np.log(t) + a[1] * t + a[2] * t ** 2 / 2 + a[3] * t ** 3 / 3
                       + a[4] * t ** 4 / 4 + a[6])
            s_horner = a[0] * np.log(t) + t * (a[1] + t * (d7[4] + t * (d7[5] + t * d7[6]))) + a[6]
            assert h_horner == pytest.approx(h_power, rel=1e-12, abs=1e-12), s
            assert s_horner == pytest.approx(s_power, rel=1e-12, abs=1e-12), s


def test_be_check_pins():
    """The integrator's check of a prediction: its own converged isothermal
    result passes with a negligible error estimate; the unchanged input
    (identity prediction) has R = -dt w(rho_0) exactly; a prediction whose
    change of a reacting species is off by more than the relative bar is
    flagged."""
    from oracle.chemistry import be_check
    rho0 = np.zeros(11)
    rho0[9], rho0[10], rho0[7] = 0.7, 0.25, 0.05
    T, dt = 5000.0, 1e-7
    good, _, _ = integrator_step(MECH, rho0, T, dt, iters=16, constant_temperature=True)
    flag, err = be_check(MECH, rho0, T, dt, good, constant_temperature=True)
    assert not flag and err < 1e-12
    flag, err = be_check(MECH, rho0, T, dt, np.maximum(rho0, 1e-20), constant_temperature=True)
    _, w, dw = MECH.net_production_rates(np.maximum(rho0, 1e-20), T, T, jacobian=True)
    est = np.abs(dt * w) / np.maximum(np.abs(1.0 - dt * np.diag(dw)), 1.0)
    assert abs(err - est.max() / np.maximum(rho0, 1e-20).sum()) <= 1e-15 * err
    # the bar is relative to each species' predicted change: 20 % off the
    # change of O fails, 5 % passes
    r0 = np.maximum(rho0, 1e-20)
    for frac, flagged in ((1.2, True), (1.05, False)):
        pred = good.copy()
        pred[7] = r0[7] + frac * (good[7] - r0[7])
        assert be_check(MECH, rho0, T, dt, pred, constant_temperature=True)[0] == flagged
