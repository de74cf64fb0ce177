"""GPU Park-11 chemistry against the CPU FP64 oracle.

Bar, per cell: max_s |rho_gpu - rho_oracle| <= max(1e-10, 4 * floor) * sum(rho),
where floor is the oracle's own change (max over 8 trials, relative to
sum(rho)) when its forward and equilibrium rate constants carry 2e-14
relative random noise (about 100 ulp): the device forms k_f prod C^nu' and
k_r prod C^nu'' each as one exponential of a summed logarithm whose
argument reaches ~100 in magnitude, and the rounding of that argument moves
the result by about |argument| ulp (the reference's own exp(-theta/T) has
an error of the same kind).  Three trials underestimated the spread on
sensitive cells by up to 4x (measured on the coupler batch), hence eight.  The point-implicit scheme with temperature recovery inside
the Newton loop (hybrid/chemistry/integrator.py:88-95) is not a contraction
far from equilibrium: after a few iterations T jumps between iterates and the
result of 12 iterations moves by 1-60 % under 1e-16 input noise (measured
on this batch), so those cells are compared against their floor, and each
test also asserts that most cells sit at the absolute 1e-10 bar (short
Newton counts, constant T) so the comparison pins the arithmetic.  The
invariants (species clip, total-mass rescale, flow variables untouched) are
checked on every cell at round-off.  Cells whose floor exceeds 1e-8 are not
compared at all (a few noisy trials do not bound a chaotic spread); every
test prints and asserts the fraction of cells that are compared.

On the random far-from-equilibrium batch, 12 coupled iterations leave no cell
below that floor, so that case checks the invariants only.  The coupled
12-iteration scheme that the reacting flow runs is pinned on the states where
it is well conditioned: air (N2/O2) at 1500-3000 K, the reacting flow's
inflow regime, and cells first brought to their isothermal equilibrium at
3000-5500 K.  There every cell must match at the absolute 1e-10 bar."""

import numpy as np
import pytest
import torch

from hybrid.chemistry import Air11ParkMechanism, PointImplicitIntegrator, ReactingFlowCoupler
from hybrid.chemistry.batch import make_batch
from oracle.chemistry import Mechanism, ReactingFlowCoupler as OracleCoupler, integrator_step

pytestmark = pytest.mark.gpu
TOL = 1e-10
UNPINNED = 1e-8


NOISE = 2e-14


class NoisyMechanism(Mechanism):
    """Oracle mechanism with NOISE relative noise on k_f and K_c."""

    def __init__(self, seed):
        super().__init__()
        self.rng = np.random.default_rng(seed)

    def forward_rate_constants(self, T, Tv):
        k = super().forward_rate_constants(T, Tv)
        return k * (1.0 + NOISE * self.rng.standard_normal(k.shape))

    def equilibrium_constants(self, T, Tv):
        k = super().equilibrium_constants(T, Tv)
        return k * (1.0 + NOISE * self.rng.standard_normal(k.shape))


def _bars(ref_fn, refs, scale, trials=8):
    """Per-cell bars max(TOL, 4 floor) * scale from re-running the oracle
    (infinite where the floor exceeds UNPINNED)."""
    floors = np.zeros(len(refs))
    with np.errstate(all="ignore"):
        for j in range(trials):
            mech = NoisyMechanism(j)
            for c, ref in enumerate(refs):
                # A perturbed run that overflows, or whose Newton matrix turns
                # singular, leaves that cell unpinned.
                try:
                    d = np.abs(ref_fn(mech, c) - ref).max() / scale[c]
                except np.linalg.LinAlgError:
                    d = np.inf
                floors[c] = max(floors[c], d if np.isfinite(d) else np.inf)
    # A floor above 1e-8 marks a cell where the scheme amplifies round-off
    # chaotically: a few noisy trials do not bound the spread there, so the
    # cell is not compared (its invariants still are).
    floors[floors > UNPINNED] = np.inf
    return np.maximum(TOL, 4.0 * floors) * scale


def air_batch(n, seed, T_range=(1500.0, 3000.0)):
    """Air (N2 0.767, O2 0.233 by mass) at T uniform in T_range and p
    log-uniform in [1e3, 1e5] Pa."""
    rng = np.random.default_rng(seed)
    om = Mechanism()
    T = rng.uniform(*T_range, n)
    p = np.exp(rng.uniform(np.log(1e3), np.log(1e5), n))
    Y = np.zeros((n, 11))
    Y[:, 9], Y[:, 10] = 0.767, 0.233
    R = 8.31446261815324 * np.sum(Y / om.molar_masses, 1)
    return (p / (R * T))[:, None] * Y, T


def equilibrated_batch(n, seed, T_range=(3000.0, 5500.0)):
    """make_batch cells brought close to their isothermal equilibrium by the
    oracle's backward Euler with a long step (dt 1e-3 s, 60 iterations)."""
    rho, T = make_batch(n, seed=seed, T_range=T_range)
    om = Mechanism()
    return np.array([integrator_step(om, rho[c], T[c], 1e-3, 60, True)[0]
                     for c in range(n)]), T


def _report(name, bars, scale):
    finite = float(np.mean(np.isfinite(bars)))
    pinned = float(np.mean(bars <= TOL * scale * (1 + 1e-12)))
    print(f"{name}: compared {finite:.3f}, at the 1e-10 bar {pinned:.3f}", flush=True)
    return finite, pinned


@pytest.mark.parametrize("batch,const_T,iters,pinned", [
    ("random", True, 12, 0.95), ("random", False, 1, 0.85), ("random", False, 2, 0.6),
    ("random", False, 12, 0.0), ("air", False, 12, 1.0), ("equilibrated", False, 12, 1.0)])
def test_integrator_matches_oracle(batch, const_T, iters, pinned):
    if batch == "random":
        rho, T = make_batch(96, seed=3)
    elif batch == "air":
        rho, T = air_batch(48, seed=3)
    else:
        rho, T = equilibrated_batch(24, seed=5)
    mech = Air11ParkMechanism()
    integ = PointImplicitIntegrator(mech, max_newton_iters=iters)
    got, Tg, _ = integ.step(rho, T, 1e-7, constant_temperature=const_T)
    got = got.cpu().numpy()
    Tg = Tg.cpu().numpy()
    om = Mechanism()
    refs = [integrator_step(om, rho[c], T[c], 1e-7, iters, const_T)[0] for c in range(len(T))]
    scale = rho.sum(1)
    bars = _bars(lambda m, c: integrator_step(m, rho[c], T[c], 1e-7, iters, const_T)[0],
                 refs, scale)
    assert all(np.all(np.isfinite(r)) for r in refs)
    err = np.array([np.abs(got[c] - refs[c]).max() for c in range(len(T))])
    bad = np.nonzero(err > bars)[0]
    assert bad.size == 0, [(int(c), float(err[c]), float(bars[c])) for c in bad[:5]]
    finite, at_bar = _report(f"integrator {batch} const_T={const_T} iters={iters}", bars, scale)
    assert at_bar >= pinned
    if batch != "random":
        assert finite == 1.0 and at_bar == 1.0
    # Invariants: clip, total-mass rescale, temperature bounds.
    assert np.all(got >= 0) and np.all(np.isfinite(got))
    np.testing.assert_allclose(got.sum(1), np.maximum(rho, 1e-20).sum(1), rtol=1e-13)
    if const_T:
        assert np.array_equal(Tg, T)
    else:
        assert np.all((Tg >= 100.0) & (Tg <= 50000.0))


def _random_block(dimensionless, seed=7, E=6, N=10, air=False):
    """Charge-neutral air batch (T 3000-12000 K, p 1e3-1e5 Pa; or N2/O2 air at
    1500-3000 K) as flow nodes with random velocities up to 1500 m/s;
    non-dimensionalised with the coupler's reference scales when asked."""
    rho_s, T = air_batch(E * N, seed) if air else make_batch(E * N, seed=seed)
    rng = np.random.default_rng(seed)
    om = Mechanism()
    c = OracleCoupler(om, dimensionless=True)
    U = np.zeros((E * N, 16))
    for i in range(E * N):
        rho = rho_s[i].sum()
        u = rng.uniform(-1500, 1500, 3)
        e_th = om.mixture_properties(rho_s[i], T[i])["internal_energy"]
        U[i, 0] = rho
        U[i, 1:4] = rho * u
        U[i, 4] = rho * (e_th + 0.5 * np.sum(u * u))
        U[i, 5:] = rho_s[i]
    if dimensionless:
        U[:, 0] /= c.rho_ref
        U[:, 5:] /= c.rho_ref
        U[:, 1:4] /= c.rho_ref * c.u_ref
        U[:, 4] /= c.rho_ref * c.e_ref
    return U.reshape(E, N, 16)


def _check_coupler(U, dt, iters, dimensionless, pinned):
    dev = ReactingFlowCoupler(max_newton_iters=iters, dimensionless=dimensionless)
    got = dev.react_block(torch.as_tensor(U, device="cuda"), dt).cpu().numpy()
    ref = OracleCoupler(Mechanism(), max_newton_iters=iters,
                        dimensionless=dimensionless).react_block(U, dt)
    assert np.all(np.isfinite(ref))
    flat_U = U.reshape(-1, U.shape[-1])
    refs = list(ref.reshape(-1, U.shape[-1]))
    scale = flat_U[:, 0]
    bars = _bars(lambda m, c: OracleCoupler(m, max_newton_iters=iters,
                                            dimensionless=dimensionless).react_cell(flat_U[c], dt),
                 refs, scale)
    err = np.abs(got - ref).reshape(len(refs), -1).max(1)
    bad = np.nonzero(err > bars)[0]
    assert bad.size == 0, [(int(c), float(err[c]), float(bars[c])) for c in bad[:5]]
    _, at_bar = _report(f"coupler iters={iters} dimensionless={dimensionless}", bars, scale)
    assert at_bar >= pinned
    # hybrid/chemistry/coupling.py:170-177: flow variables untouched, species
    # non-negative and summing to rho.
    assert np.array_equal(got[..., :5], U[..., :5])
    assert np.all(got[..., 5:] >= 0)
    np.testing.assert_allclose(got[..., 5:].sum(-1), U[..., 0], rtol=1e-13)
    return got


@pytest.mark.parametrize("dimensionless", [False, True])
@pytest.mark.parametrize("iters,pinned,air", [(1, 0.85, False), (2, 0.55, False),
                                              (12, 0.0, False), (12, 1.0, True)])
def test_coupler_matches_oracle(dimensionless, iters, pinned, air):
    U = _random_block(dimensionless, air=air)
    dt = 1e-7 if not dimensionless else 1e-7 / OracleCoupler(dimensionless=True).t_ref
    _check_coupler(U, dt, iters, dimensionless, pinned)


def test_coupler_reference_state():
    """tests/test_reacting_coupler.py:35-91 of the reference: N2/O2 at 6000 K,
    u = 2000 m/s, dt = 1e-7 s, 16 Newton iterations, 4 x 10 nodes; plus
    cooler nodes (3000 and 4500 K, well-conditioned) that pin the result."""
    om = Mechanism()
    Y = np.zeros(11)
    Y[9], Y[10] = 0.767, 0.233
    rho, u = 1.2, 2000.0
    U = np.zeros((4, 10, 16))
    temps = np.array([6000.0, 3000.0, 4500.0, 6000.0])
    for e, T in enumerate(temps):
        e_th = om.mixture_properties(rho * Y, T)["internal_energy"]
        U[e, :, 0] = rho
        U[e, :, 1] = rho * u
        U[e, :, 4] = rho * (e_th + 0.5 * u * u)
        U[e, :, 5:] = rho * Y
    got = _check_coupler(U, 1e-7, 16, False, 0.45)
    assert got[0, 0, 5 + 7] > 1e-10           # O2 dissociates at 6000 K


@pytest.mark.parametrize("const_T", [True, False])
def test_integrator_check_matches_oracle(const_T):
    """fsg_chem_check (the backward-Euler residual of a prediction, scaled by
    the Jacobian diagonal) against the oracle's be_check: the error
    estimates to 1e-6 relative (the residual is FP64; the diagonal, which
    only sets the scale max(|1 - dt dw/drho|, 1), is accumulated in FP32:
    measured 2e-9) and identical flags (cells within 1e-6 of the bar
    excluded; the test asserts most cells are compared)."""
    import ctypes
    from hybrid import _native as nat
    from oracle.chemistry import be_check
    # air at 3000-5000 K over 1e-6 s: the 12-iteration scheme converges (its
    # own result has a residual at round-off) and the predictions, the
    # integrator's changes with 5 % noise, are partly flagged
    rho, T = air_batch(64, seed=3, T_range=(3000.0, 5000.0))
    om = Mechanism()
    dt = 1e-6
    good = np.array([integrator_step(om, rho[c], T[c], dt, 12, const_T)[0] for c in range(64)])
    rng = np.random.default_rng(2)
    r0 = np.maximum(rho, 1e-20)
    pred = r0 + (good - r0) * (1 + 0.05 * rng.standard_normal(good.shape))
    assert not any(be_check(om, rho[c], T[c], dt, good[c], const_T)[0] for c in range(64))
    mech = Air11ParkMechanism()
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    flag = torch.zeros(64, dtype=torch.int32, device="cuda")
    err = torch.empty(64, dtype=torch.float64, device="cuda")
    Rd, Td, Pd = dev(rho), dev(T), dev(pred)
    nat.check_chem(nat.lib().fsg_chem_check(nat.ptr(mech.table), nat.ptr(Rd), nat.ptr(Td), None,
                                            dt, int(const_T), nat.ptr(Pd), 0.1, 1e-8,
                                            nat.ptr(flag), nat.ptr(err), 64,
                                            nat.stream_handle(Rd.device)))
    got_f, got_e = flag.cpu().numpy().astype(bool), err.cpu().numpy()
    want = [be_check(om, rho[c], T[c], dt, pred[c], const_T) for c in range(64)]
    want_f = np.array([w[0] for w in want])
    want_e = np.array([w[1] for w in want])
    np.testing.assert_allclose(got_e, want_e, rtol=1e-6)
    # cells far enough from the bar to have an unambiguous flag
    margin = np.array([
        be_check(om, rho[c], T[c], dt, pred[c], const_T, 0.1 * (1 - 1e-6), 1e-8 * (1 - 1e-6))[0]
        == be_check(om, rho[c], T[c], dt, pred[c], const_T, 0.1 * (1 + 1e-6), 1e-8 * (1 + 1e-6))[0]
        for c in range(64)])
    print(f"check: flagged {want_f.mean():.2f}, compared {margin.mean():.2f}", flush=True)
    assert margin.mean() > 0.9 and 0 < want_f.mean() < 1
    assert np.array_equal(got_f[margin], want_f[margin])


def _other_fingerprint(mech):
    """The same mechanism with another stoichiometry fingerprint: the
    kernels then take the table-driven instantiations."""
    other = Air11ParkMechanism()
    other.table = mech.table.clone()
    other.table[-1] = 0.0
    return other


@pytest.mark.parametrize("const_T,iters", [(True, 4), (False, 3)])
def test_table_driven_integrator_equals_compiled_stoichiometry(const_T, iters):
    """The integrator with the Park-11 stoichiometry compiled in and the
    table-driven one compute the same arithmetic: bit-identical results."""
    mech = Air11ParkMechanism()
    other = _other_fingerprint(mech)
    rho, T = make_batch(4096, seed=5)
    a = PointImplicitIntegrator(mech, max_newton_iters=iters).step(rho, T, 1e-7, constant_temperature=const_T)
    b = PointImplicitIntegrator(other, max_newton_iters=iters).step(rho, T, 1e-7, constant_temperature=const_T)
    assert torch.isfinite(a[0]).all()
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    U = torch.as_tensor(_random_block(True, seed=11), device="cuda")
    ua = ReactingFlowCoupler(mech, max_newton_iters=iters, dimensionless=True).react_block(U, 1e-6)
    ub = ReactingFlowCoupler(other, max_newton_iters=iters, dimensionless=True).react_block(U, 1e-6)
    assert torch.equal(ua, ub)


def test_check_refuses_a_table_it_was_not_compiled_for():
    """fsg_chem_check judges with the compiled-in stoichiometry only: a table
    whose fingerprint differs has every cell flagged for the integrator."""
    import ctypes
    from hybrid import _native as nat
    mech = Air11ParkMechanism()
    other = _other_fingerprint(mech)
    rho, T = air_batch(256, seed=4, T_range=(1500.0, 3000.0))
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    Rd, Td = dev(rho), dev(T)
    for table, want in ((mech.table, False), (other.table, True)):
        flag = torch.zeros(256, dtype=torch.int32, device="cuda")
        err = torch.empty(256, dtype=torch.float64, device="cuda")
        # the prediction "no change" over a 1e-20 s step passes the check
        nat.check_chem(nat.lib().fsg_chem_check(nat.ptr(table), nat.ptr(Rd), nat.ptr(Td), None,
                                                1e-20, 1, nat.ptr(Rd), 0.1, 1e-8, nat.ptr(flag),
                                                nat.ptr(err), 256, nat.stream_handle(Rd.device)))
        f, e = flag.cpu().numpy().astype(bool), err.cpu().numpy()
        if want:
            assert f.all() and np.isinf(e).all()
        else:
            assert not f.any() and np.isfinite(e).all()
