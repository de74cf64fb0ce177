"""MOOD on the GPU against the CPU FP64 oracle (hybrid/limiter.py,
hybrid/subcell.py).

* Troubled sets: identical to the oracle's (elements whose sensor lies
  within 1e-9 of the threshold are excluded; the test asserts there are
  none on these states); sensor values to 1e-10 relative.
* Positivity, subcell recompute and the MOOD stage: relative L2 per
  variable <= max(1e-12, 4 x the oracle's own change under 1-ulp input
  noise) in FP64, the residual tests' bar.  The pressure bisection of the
  positivity limiter takes discrete decisions at p == floor; the states used
  keep every decision away from ties (the floor check makes that visible).
* Conservation of the owner-indexed recompute on the GPU to 1e-12.
"""

import numpy as np
import pytest
import torch

from conftest import PLANE_NAMES, perturbed_primitive, rel_l2

from hybrid import (FarfieldWallBoundary, HybridCompressibleSolver, HybridMOODSubcellLimiter,
                    HybridNavierStokesOperator, HybridPositivityLimiter, HybridSSPRK3, box_mesh)
from oracle.limiter import MOODSubcellLimiter, PositivityLimiter
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver, FarfieldWallBoundary as OracleBoundary
from oracle.subcell import SubcellFV
from oracle.time_integration import SSPRK3

pytestmark = pytest.mark.gpu

TOL = 1e-12


def shock_primitive(x, seed=0):
    """A smooth perturbed field with a density/pressure jump across x = 0.52
    (inside a layer of elements) and a thin low-pressure pocket."""
    W = perturbed_primitive(x, seed=seed, amplitude=0.01)
    left = x[..., 0] < 0.52
    W[..., 0] = np.where(left, W[..., 0], 0.3 * W[..., 0])
    W[..., 4] = np.where(left, W[..., 4], 0.25 * W[..., 4])
    pocket = np.linalg.norm(x - np.array([0.3, 0.6, 0.4]), axis=-1) < 0.12
    W[..., 4] = np.where(pocket, 0.2 * W[..., 4], W[..., 4])
    return W


def _unit(name):
    """Coordinate factor that maps the case's box onto the unit box the
    shock fields are defined on."""
    return 10.0 if name == "sliver" else 1.0


def _dt(name):
    return 2e-4 if name == "sliver" else 2e-3


def _case(name, p, recon="weno", indexing="reference", dtype=torch.float64):
    boundary = name != "periodic"
    if name == "periodic":
        m = box_mesh(4, periodic=(True,) * 3)
    elif name == "split":
        m = box_mesh(4, split_fraction=0.4, seed=7, plane_names=PLANE_NAMES)
    elif name == "split_periodic":
        m = box_mesh(4, periodic=(True,) * 3, split_fraction=0.5, seed=3)
    elif name == "sliver":
        # the split box scaled to side 0.1: M^-1 carries min(1, V/1e-4) < 1
        m = box_mesh(4, lengths=(0.1,) * 3, split_fraction=0.4, seed=7, plane_names=PLANE_NAMES)
    else:
        raise KeyError(name)
    mu = 0.005
    solver = HybridCompressibleSolver(m, polynomial_degree=p, dynamic_viscosity=mu, prandtl=0.71,
                                      dtype=dtype)
    op = HybridNavierStokesOperator(solver)
    osolver = CompressibleSolver(gamma=1.4, dynamic_viscosity=mu, prandtl=0.71)
    oop = NavierStokesOperator(from_arrays(m), osolver, p)
    states = {t: osolver.primitive_to_conserved(shock_primitive(x * _unit(name)))
              for t, x in oop.geometry.solution_node_coordinates.items()}
    b = ob = None
    if boundary:
        far = np.array([1.2, 0.9, 0.1, -0.05, 1.1])
        b = FarfieldWallBoundary(solver, far)
        ob = OracleBoundary(osolver, far, inlet_tags=sorted(b.inlet_tags),
                            outlet_tags=sorted(b.outlet_tags), wall_tags=sorted(b.wall_tags),
                            no_slip_tags=sorted(b.no_slip_tags), slip_tags=sorted(b.slip_tags))
    lim = HybridMOODSubcellLimiter(op, reconstruction=recon, indexing=indexing)
    olim = MOODSubcellLimiter(oop, reconstruction=recon)
    olim.subcell = SubcellFV(oop, reconstruction=recon, indexing=indexing)
    return solver, op, b, lim, oop, ob, olim, states


def _dev(solver, d):
    return solver.to_device(d) if hasattr(solver, "to_device") else \
        {k: torch.as_tensor(v, device="cuda", dtype=solver.dtype) for k, v in d.items()}


def _np(d):
    return {k: v.double().cpu().numpy() for k, v in d.items()}


def _bar(fn, states, tol=TOL):
    rng = np.random.default_rng(5)
    base = fn(states)
    worst = 0.0
    for _ in range(2):
        noisy = {k: v * (1 + 2.2e-16 * rng.standard_normal(v.shape)) for k, v in states.items()}
        worst = np.maximum(worst, rel_l2(fn(noisy), base))
    return np.maximum(tol, 4.0 * worst)


CASES = [("periodic", 2), ("split", 2), ("split", 3), ("split_periodic", 2), ("sliver", 2)]


@pytest.mark.parametrize("name,p", CASES)
def test_troubled_sets_match(name, p):
    solver, op, b, lim, oop, ob, olim, states = _case(name, p)
    sensor, _ = olim.sensor(states)
    for t, s in sensor.items():
        assert np.all(np.abs(s - olim.sensor.threshold) > 1e-9), t
    want = olim.troubled(states)
    got = lim.troubled(_dev(solver, states))
    n = 0
    for t in want:
        assert np.array_equal(got[t].cpu().numpy(), want[t]), t
        n += int(want[t].sum())
    assert 0 < n < sum(len(v) for v in want.values())
    gs, _ = lim.sensor(_dev(solver, states))
    for t in sensor:
        np.testing.assert_allclose(gs[t].cpu().numpy(), sensor[t], rtol=1e-10, atol=1e-14)
    assert lim.troubled_count() == n


@pytest.mark.parametrize("name,p", [("split", 2), ("split", 3)])
def test_positivity_matches(name, p):
    solver, op, b, lim, oop, ob, olim, states = _case(name, p)
    bad = {}
    for t, U in states.items():
        W = oop.solver.conserved_to_primitive(U)
        W[::3, 0, 0] = -0.05
        W[1::4, 2, 4] = -0.02
        bad[t] = oop.solver.primitive_to_conserved(W)
    opos = PositivityLimiter(oop)
    want = opos(bad)
    got = _np(HybridPositivityLimiter(op)(_dev(solver, bad)))
    assert np.all(rel_l2(got, want) <= _bar(opos, bad))


@pytest.mark.parametrize("recon", ["constant", "tvd", "weno"])
@pytest.mark.parametrize("indexing", ["reference", "owner"])
@pytest.mark.parametrize("name,p", [("split", 2), ("split_periodic", 2), ("split", 3),
                                    ("sliver", 2)])
def test_recompute_matches(name, p, recon, indexing):
    solver, op, b, lim, oop, ob, olim, states = _case(name, p, recon, indexing)
    dt = _dt(name)
    inv = oop.inviscid_rhs(states, ob)
    visc = oop.viscous_rhs(states, ob)
    cand = {k: states[k] + dt * (inv[k] + visc[k]) for k in states}
    mask = olim.troubled(cand)

    def oracle(st):
        return olim.recompute(st, cand, dt, ob, visc, troubled_mask=mask)

    want = oracle(states)
    got = lim.recompute(_dev(solver, states), _dev(solver, cand), dt, b, _dev(solver, visc),
                        troubled_mask={k: torch.as_tensor(v, device="cuda") for k, v in mask.items()})
    err = rel_l2(_np(got), want)
    assert np.all(err <= _bar(oracle, states)), err
    # detection inside recompute gives the same answer as the given mask
    got2 = lim.recompute(_dev(solver, states), _dev(solver, cand), dt, b, _dev(solver, visc))
    assert np.all(rel_l2(_np(got2), _np(got)) == 0.0)


def _total(oop, S):
    tot = 0.0
    for bid, blk in enumerate(oop.mesh.element_blocks):
        ref = oop.geometry.reference_elements[blk.element_type]
        w = oop.geometry.element_metrics[bid].integration_weights
        tot = tot + np.einsum("qi,eiv,eq->v", ref.basis(ref.volume_points), S[blk.element_type], w)
    return tot


@pytest.mark.parametrize("recon", ["constant", "weno"])
def test_owner_indexing_conserves(recon):
    solver, op, b, lim, oop, ob, olim, states = _case("split_periodic", 2, recon, "owner")
    dt = 2e-3
    inv = oop.inviscid_rhs(states)
    cand = {k: states[k] + dt * inv[k] for k in states}
    got = _np(lim.recompute(_dev(solver, states), _dev(solver, cand), dt))
    assert lim.troubled_count() > 0
    tot = np.abs(_total(oop, cand)).max()
    assert np.abs(_total(oop, got) - _total(oop, cand)).max() <= 1e-12 * tot


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("name,p", [("split", 2), ("periodic", 2), ("split", 3), ("sliver", 2)])
def test_mood_stage_matches(name, p, fused):
    """Both stage forms against the oracle: one residual pass with the
    viscous rhs re-evaluated on the troubled elements (fused), and the
    reference's separate inviscid and viscous passes."""
    print(f"case {name} p={p} fused={fused}", flush=True)
    solver, op, b, lim, oop, ob, olim, states = _case(name, p)
    curr = {t: oop.solver.primitive_to_conserved(shock_primitive(x * _unit(name), seed=9))
            for t, x in oop.geometry.solution_node_coordinates.items()}
    dt = _dt(name)
    rk = SSPRK3(oop, limiter=olim)

    def oracle(c):
        return rk.rk_stage(states, c, dt, 0.75, 0.25, ob)[0]

    want, wvalid, wr, wp = rk.rk_stage(states, curr, dt, 0.75, 0.25, ob)
    drk = HybridSSPRK3(op, limiter=lim)
    drk.fused_mood = fused
    got, valid, rmin, pmin = drk.rk_stage(_dev(solver, states), _dev(solver, curr), dt, 0.75, 0.25, b)
    assert lim.troubled_count() > 0
    assert bool(valid) == bool(wvalid)
    assert abs(float(rmin) - wr) <= 1e-12 * abs(wr) and abs(float(pmin) - wp) <= 1e-10 * abs(wp)
    assert np.all(rel_l2(_np(got), want) <= _bar(oracle, curr))


def test_checked_step_with_mood():
    solver, op, b, lim, oop, ob, olim, states = _case("split", 2)
    dt = 1e-3
    want, wvalid, _, _ = SSPRK3(oop, limiter=olim).checked_step(states, dt, ob)
    got, valid, _, _ = HybridSSPRK3(op, limiter=lim).checked_step(_dev(solver, states), dt, b)
    assert bool(valid) == bool(wvalid)
    assert np.all(rel_l2(_np(got), want) <= 1e-11)


FP32_TOL = 1e-4


@pytest.mark.parametrize("name,p", [("split", 2), ("split_periodic", 2), ("sliver", 2)])
def test_fp32_path(name, p):
    """FP32 MOOD kernels: identical troubled sets (every sensor value is at
    least 4e-4 away from the threshold on these states, far above FP32
    rounding of the ratio) and the stage within 1e-4 relative L2."""
    solver, op, b, lim, oop, ob, olim, states = _case(name, p, dtype=torch.float32)
    sensor, _ = olim.sensor(states)
    for t, s in sensor.items():
        assert np.all(np.abs(s - olim.sensor.threshold) > 1e-4), t
    want = olim.troubled(states)
    got = lim.troubled(_dev(solver, states))
    for t in want:
        assert np.array_equal(got[t].cpu().numpy(), want[t]), t
    curr = {t: oop.solver.primitive_to_conserved(shock_primitive(x * _unit(name), seed=9))
            for t, x in oop.geometry.solution_node_coordinates.items()}
    dt = _dt(name)
    w, wv, _, _ = SSPRK3(oop, limiter=olim).rk_stage(states, curr, dt, 0.75, 0.25, ob)
    g, gv, _, _ = HybridSSPRK3(op, limiter=lim).rk_stage(_dev(solver, states), _dev(solver, curr),
                                                         dt, 0.75, 0.25, b)
    assert bool(gv) == bool(wv)
    assert np.all(rel_l2(_np(g), w) <= FP32_TOL)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("name,p", [("split", 2), ("split", 3)])
def test_masked_viscous_rhs_is_bitwise(name, p, dtype):
    """The viscous rhs restricted to marked elements equals the full viscous
    rhs on those elements, on every element path (the masked pass
    re-evaluates every face of a marked element): bit for bit in FP64; in
    FP32 within 1e-6 of the block's largest value (the restricted pass is a
    separate kernel instantiation, a loop over the marked elements, and the
    compiler may contract its FP32 multiply-adds differently)."""
    print(f"case {name} p={p} {dtype} masked viscous rhs", flush=True)
    solver, op, b, lim, oop, ob, olim, states = _case(name, p, dtype=dtype)
    U = _dev(solver, states)
    g = torch.Generator().manual_seed(3)
    masks = {t: (torch.rand(U[t].shape[0], generator=g) < 0.3).to(torch.int32).cuda()
             for t in op.block_types}
    dm = op.device_mesh
    try:
        for mfma, sf in ((None, None), (True, False), (False, False), (False, True)):
            dm.set_element_path(mfma=mfma)
            dm.set_hex_sum_factorisation(sf)
            full = op.viscous_rhs(U, b)
            op.rhs(U, b)     # leaves combined face records behind
            part = op.viscous_rhs_masked(U, masks, b)
            for t in op.block_types:
                m = masks[t].bool()
                assert int(m.sum()) > 0
                if dtype == torch.float64:
                    assert torch.equal(part[t][m], full[t][m]), (t, mfma, sf)
                else:
                    d = float((part[t][m] - full[t][m]).abs().max())
                    ref = float(full[t].abs().max())
                    print(f"{t} mfma={mfma} sf={sf}: max |diff| / max |rhs| {d / ref:.2e}",
                          flush=True)
                    assert d <= 1e-6 * ref, (t, mfma, sf, d, ref)
    finally:
        dm.set_element_path(None)
        dm.set_hex_sum_factorisation(None)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fused_stage_matches_split_stage(dtype):
    """One residual pass per MOOD stage against the reference's two passes:
    the same troubled elements, and states equal to the rounding of
    inv + visc (1e-13 relative in FP64, 1e-5 in FP32)."""
    solver, op, b, lim, oop, ob, olim, states = _case("split", 2, dtype=dtype)
    curr = {t: oop.solver.primitive_to_conserved(shock_primitive(x, seed=9))
            for t, x in oop.geometry.solution_node_coordinates.items()}
    S, C = _dev(solver, states), _dev(solver, curr)
    drk = HybridSSPRK3(op, limiter=lim)
    out, flags = {}, {}
    for fused in (True, False):
        drk.fused_mood = fused
        out[fused] = _np(drk.rk_stage(S, C, 2e-3, 0.75, 0.25, b)[0])
        flags[fused] = {t: v.clone() for t, v in
                        ((t, lim._ctx().tables.flags[k]) for k, t in enumerate(op.block_types))}
    for t in op.block_types:
        assert torch.equal(flags[True][t], flags[False][t]), t
    assert sum(int(f.sum()) for f in flags[True].values()) > 0
    tol = 1e-13 if dtype == torch.float64 else 1e-5
    assert np.all(rel_l2(out[True], out[False]) <= tol), rel_l2(out[True], out[False])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fused_stage_from_total_matches_euler_candidate(dtype):
    """fsg_mood_stage_fused from the total rhs (detection forms curr + dt rhs
    itself) against the form the stepper uses (the residual's epilogue
    writes that Euler candidate, no total passed): the same troubled
    elements, states equal to rounding, the same diagnostics."""
    from hybrid import _native as nat
    solver, op, b, lim, oop, ob, olim, states = _case("split", 2, dtype=dtype)
    curr = {t: oop.solver.primitive_to_conserved(shock_primitive(x, seed=9))
            for t, x in oop.geometry.solution_node_coordinates.items()}
    S, C = _dev(solver, states), _dev(solver, curr)
    drk = HybridSSPRK3(op, limiter=lim)
    drk.fused_mood = True
    got, valid, rmin, pmin = drk.rk_stage(S, C, 2e-3, 0.75, 0.25, b)
    flags = {t: lim._ctx().tables.flags[k].clone() for k, t in enumerate(op.block_types)}
    stream = nat.stream_handle(solver.device)
    drk._diag.reset(stream)
    total = op.rhs(C, b)
    visc = {t: torch.empty_like(C[t]) for t in op.block_types}
    work = {t: torch.empty_like(C[t]) for t in op.block_types}
    nxt = {t: torch.empty_like(C[t]) for t in op.block_types}
    lim.stage_fused(S, C, total, visc, work, nxt, 2e-3, 0.75, 0.25, b, nat.ptr(drk._diag.buf))
    valid2, rmin2, pmin2 = drk._diag.values(drk.density_floor, drk.pressure_floor)
    for k, t in enumerate(op.block_types):
        assert torch.equal(lim._ctx().tables.flags[k], flags[t]), t
    assert sum(int(f.sum()) for f in flags.values()) > 0
    tol = 1e-13 if dtype == torch.float64 else 1e-5
    assert np.all(rel_l2(_np(nxt), _np(got)) <= tol), rel_l2(_np(nxt), _np(got))
    assert bool(valid) == bool(valid2)
    assert abs(float(rmin) - float(rmin2)) <= tol * abs(float(rmin))
    assert abs(float(pmin) - float(pmin2)) <= 1e3 * tol * abs(float(pmin))


@pytest.mark.parametrize("name,p", [("split", 2), ("split", 3), ("sliver", 2)])
def test_modal_shock_limiter_matches(name, p):
    """HybridModalShockLimiter (hybrid/limiter.py:186-264) against the oracle,
    alone and as the limiter of an SSP-RK3 step (the stage candidate limited,
    admissibility of the limited state)."""
    from hybrid import HybridModalShockLimiter
    from oracle.limiter import ModalShockLimiter
    print(f"case modal {name} p={p}", flush=True)
    solver, op, b, lim, oop, ob, olim, states = _case(name, p)
    opos = PositivityLimiter(oop, density_floor=1e-7, pressure_floor=1e-7)
    omod = ModalShockLimiter(oop, positivity=opos)
    dmod = HybridModalShockLimiter(op, positivity=HybridPositivityLimiter(
        op, density_floor=1e-7, pressure_floor=1e-7))
    sensor, flags = omod.sensor(states)
    for t, sv in sensor.items():
        assert np.all(np.abs(sv - 0.05) > 1e-9), t
    assert sum(int(f.sum()) for f in flags.values()) > 0
    want = omod(states)
    got = _np(dmod(_dev(solver, states)))
    assert np.all(rel_l2(got, want) <= _bar(omod, states)), rel_l2(got, want)
    dt = _dt(name) / 2
    ork = SSPRK3(oop, limiter=omod)
    want, wv, wr, wp = ork.checked_step(states, dt, ob)
    got, gv, gr, gp = HybridSSPRK3(op, limiter=dmod).checked_step(_dev(solver, states), dt, b)
    assert bool(gv) == bool(wv)
    assert abs(float(gr) - wr) <= 1e-12 * abs(wr)
    assert np.all(rel_l2(_np(got), want) <= 1e-11), rel_l2(_np(got), want)
