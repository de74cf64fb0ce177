"""Reacting flow on the GPU: the coupler's split path (prepare -> substep ->
finish) used for FP32 states and for the surrogate, and one operator-split
step of a reacting hybrid-mesh flow (SSP-RK3 with MOOD, then the Park-11
coupler, run_hybrid_sphere.py:280-312 of the reference) against the CPU FP64
oracle.

* The split path in FP64 is bit-identical to the one-kernel coupler; on an
  FP32 state it equals the FP64 coupler on the same (widened) values rounded
  once to FP32, bit for bit.
* With a surrogate that flags every cell, the coupler returns the plain
  integrator's result bit for bit (the device-side fallback list covers every
  cell).
* One reacting step on a 2^3 box with 40 % of hexahedra split, p=2, 11
  species, Mach 3 air at 1500 K with a 1 % density perturbation: relative L2
  per conserved variable <= 1e-11 against the oracle (the flow step's bar with
  MOOD active; the chemistry at 1500 K is smooth and well conditioned).
"""

import numpy as np
import pytest
import torch

from conftest import PLANE_NAMES, rel_l2
from test_gpu_chemistry import _random_block

from hybrid import (FarfieldWallBoundary, HybridCompressibleSolver, HybridMOODSubcellLimiter,
                    HybridNavierStokesOperator, HybridSSPRK3, box_mesh)
from hybrid.chemistry import Air11ParkMechanism, PointImplicitIntegrator, ReactingFlowCoupler
from hybrid.chemistry.air11_park import SPECIES_NAMES
from hybrid.chemistry.batch import make_batch
from hybrid.chemistry.surrogate import ChemistrySurrogate
from oracle.chemistry import Mechanism, ReactingFlowCoupler as OracleCoupler
from oracle.limiter import MOODSubcellLimiter
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver, FarfieldWallBoundary as OracleBoundary
from oracle.time_integration import SSPRK3

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dimensionless", [False, True])
def test_split_path_matches_fused(dimensionless):
    U = torch.as_tensor(_random_block(dimensionless, seed=3), device="cuda")
    dt = 1e-7 if not dimensionless else 1e-7 / OracleCoupler(dimensionless=True).t_ref
    c = ReactingFlowCoupler(max_newton_iters=12, dimensionless=dimensionless)
    fused = c.react_block(U, dt)
    c.fused = False
    split = c.react_block(U, dt)
    assert torch.equal(fused, split)
    # per-element dt
    dte = dt * (1 + 0.5 * torch.rand(U.shape[0], device="cuda", dtype=torch.float64))
    c.fused = True
    fused = c.react_block(U, dte)
    c.fused = False
    assert torch.equal(fused, c.react_block(U, dte))


@pytest.mark.parametrize("dimensionless", [False, True])
def test_fp32_state(dimensionless):
    U = torch.as_tensor(_random_block(dimensionless, seed=4), device="cuda")
    U32 = U.float()
    dt = 1e-7 if not dimensionless else 1e-7 / OracleCoupler(dimensionless=True).t_ref
    c = ReactingFlowCoupler(max_newton_iters=12, dimensionless=dimensionless)
    got = c.react_block(U32, dt)
    assert got.dtype == torch.float32
    want = c.react_block(U32.double(), dt).float()
    assert torch.equal(got, want)


def test_surrogate_fallback_covers_every_flagged_cell():
    mech = Air11ParkMechanism()
    integ = PointImplicitIntegrator(mech, max_newton_iters=12)
    sur = ChemistrySurrogate(integ, constant_temperature=False, seed=1)
    rho, T = make_batch(2048, seed=5)
    X, O, _ = sur.dataset(rho, T, 1e-7)
    sur.set_normalisation(X, O)
    # an empty hull: every cell is out of it
    sur.stats["hi"] = sur.stats["lo"] - 1e6
    U = torch.as_tensor(_random_block(True, seed=6), device="cuda")
    dt = 1e-7 / OracleCoupler(dimensionless=True).t_ref
    c = ReactingFlowCoupler(mech, max_newton_iters=12, dimensionless=True, surrogate=sur)
    got = c.react_block(U, dt)
    assert bool(c.last_fallback.all())
    plain = ReactingFlowCoupler(mech, max_newton_iters=12, dimensionless=True)
    assert torch.equal(got, plain.react_block(U, dt))


def _air_stream(x, seed=5):
    """Mach 3 air at 1500 K in the coupler's units (rho 1, p 5), 1 % smooth
    density perturbation at constant pressure."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 4, size=(4, 3))
    ph = rng.uniform(0, 2 * np.pi, 4)
    rho = 1.0 + 0.01 * sum(np.sin(2 * np.pi * (x @ kk) + p) for kk, p in zip(k, ph)) / 4
    W = np.zeros(x.shape[:-1] + (16,))
    W[..., 0] = rho
    W[..., 1] = 3.0 * np.sqrt(1.4 * 5.0)
    W[..., 4] = 5.0
    W[..., 5 + 9], W[..., 5 + 10] = 0.767, 0.233
    return W


def test_reacting_step_matches_oracle():
    m = box_mesh(2, split_fraction=0.4, seed=7, plane_names=PLANE_NAMES)
    mu, p = 1e-3, 2
    solver = HybridCompressibleSolver(m, species_names=SPECIES_NAMES, polynomial_degree=p,
                                      dynamic_viscosity=mu, prandtl=0.71)
    op = HybridNavierStokesOperator(solver)
    osolver = CompressibleSolver(gamma=1.4, dynamic_viscosity=mu, prandtl=0.71, num_species=11)
    oop = NavierStokesOperator(from_arrays(m), osolver, p)
    states = {t: osolver.primitive_to_conserved(_air_stream(x))
              for t, x in oop.geometry.solution_node_coordinates.items()}
    far = _air_stream(np.zeros((1, 3)))[0]
    far[0] = 1.0
    b = FarfieldWallBoundary(solver, far)
    ob = OracleBoundary(osolver, far, inlet_tags=sorted(b.inlet_tags),
                        outlet_tags=sorted(b.outlet_tags), wall_tags=sorted(b.wall_tags),
                        no_slip_tags=sorted(b.no_slip_tags), slip_tags=sorted(b.slip_tags))
    rk = HybridSSPRK3(op, cfl=0.3, limiter=HybridMOODSubcellLimiter(op))
    ork = SSPRK3(oop, cfl=0.3, limiter=MOODSubcellLimiter(oop))
    dt = float(ork.stable_dt(states))
    coupler = ReactingFlowCoupler(max_newton_iters=12, dimensionless=True)
    ocoupler = OracleCoupler(Mechanism(), max_newton_iters=12, dimensionless=True)

    cand, valid, _, _ = rk.checked_step(solver.to_device(states), dt, b)
    got = coupler.advance(cand, dt)
    ocand, ovalid, _, _ = ork.checked_step(states, dt, ob)
    want = ocoupler.advance(ocand, dt)
    assert bool(valid) and bool(ovalid)
    got = {k: v.cpu().numpy() for k, v in got.items()}
    err = rel_l2(got, want)
    assert np.all(err <= 1e-11), err
    # the chemistry moved the species (O2 -> O at 1500 K) by more than the bar
    moved = rel_l2({k: v[..., 5:] for k, v in want.items()},
                   {k: v[..., 5:] for k, v in ocand.items()})
    assert np.max(moved) > 0


def test_fast_path_error_after_100_steps():
    """The fast path (FP32 residual + surrogate chemistry with its integrator
    fallback) against the plain FP64 GPU path, which matches the oracle to
    1e-11 per step (above), after 100 operator-split steps of the Mach 3 air
    stream on a 4^3 box with 40 % of hexahedra split, p=2.  The surrogate is
    trained here on integrator trajectories started from the flow's own cells
    (temperature spread by +-10 %, density by +-5 %).

    Bars: density and energy relative L2 <= 1e-4; the RMS error of every
    mass fraction Y_s = rho_s / rho <= 1e-6 (the integrator check's absolute
    bar is 1e-8 of the cell density per step, FP32 transport adds ~1e-7).
    The check bites here (the coarse mesh's dt lets the chemistry move): a
    share of the cells fall back to the integrator.  With its bars tightened
    to rtol 1e-2, atol 1e-15 more cells fall back and the O and NO errors do
    not grow."""
    import ctypes
    from hybrid import _native as nat
    m = box_mesh(4, split_fraction=0.4, seed=7, plane_names=PLANE_NAMES)
    far = _air_stream(np.zeros((1, 3)))[0]
    far[0] = 1.0
    mu = far[1] / 1.0e4
    mech = Air11ParkMechanism()
    out, fallbacks, dt = {}, {}, None
    for key, dtype, check in (("plain", torch.float64, None), ("fast", torch.float32, (0.1, 1e-8)),
                              ("strict", torch.float32, (1e-2, 1e-15))):
        print(f"case {key}", flush=True)
        solver = HybridCompressibleSolver(m, species_names=SPECIES_NAMES, polynomial_degree=2,
                                          dynamic_viscosity=mu, prandtl=0.71, dtype=dtype)
        op = HybridNavierStokesOperator(solver)
        rk = HybridSSPRK3(op, cfl=0.3, limiter=HybridMOODSubcellLimiter(op))
        b = FarfieldWallBoundary(solver, far)
        U = solver.state_blocks_from_primitive_function(_air_stream)
        if dt is None:
            dt = float(rk.stable_dt(U))
        coupler = ReactingFlowCoupler(mech, max_newton_iters=12, dimensionless=True)
        if check is not None:
            g = torch.Generator(device="cuda").manual_seed(3)
            Ud = torch.cat([v.double().reshape(-1, 16) for v in U.values()])
            n = 8192
            pick = torch.randint(0, Ud.shape[0], (n,), device="cuda", generator=g)
            rho = torch.empty(n, 11, dtype=torch.float64, device="cuda")
            T0 = torch.empty(n, dtype=torch.float64, device="cuda")
            prm = coupler._params()
            nat.check_chem(nat.lib().fsg_chem_prepare(
                nat.ptr(mech.table), nat.ptr(Ud[pick].contiguous()), nat.F64, n,
                ctypes.byref(prm), nat.ptr(rho), nat.ptr(T0), nat.stream_handle(Ud.device)))
            spread = lambda a: 1 + a * (2 * torch.rand(n, device="cuda", generator=g,
                                                       dtype=torch.float64) - 1)
            T0 = T0 * spread(0.10)
            rho = rho * spread(0.05)[:, None]
            sur = ChemistrySurrogate(coupler.integrator, constant_temperature=False, seed=3,
                                     check_rtol=check[0], check_atol=check[1])
            X, O, _ = sur.dataset(rho, T0, dt * coupler.t_ref, trajectory_steps=24)
            sur.set_normalisation(X, O)
            sur.train(X, O, steps=1500, batch=4096)
            coupler.surrogate = sur
        fallback = torch.zeros((), dtype=torch.int64, device="cuda")
        cells = 0
        for _ in range(100):
            U, valid, _, _ = rk.checked_step(U, dt, b)
            assert bool(valid)
            nxt = {}
            for t, v in U.items():
                nxt[t] = coupler.react_block(v, dt)
                if check is not None:
                    fallback += coupler.last_fallback.sum()
                    cells += v.shape[0] * v.shape[1]
            U = nxt
        out[key] = torch.cat([v.double().reshape(-1, 16) for v in U.values()])
        fallbacks[key] = int(fallback) / max(cells, 1)
    ref = out["plain"]
    rho_norm = torch.linalg.vector_norm(ref[:, 0])
    species = {}
    for key in ("fast", "strict"):
        a = out[key]
        err = (torch.linalg.vector_norm(a - ref, dim=0) / torch.linalg.vector_norm(ref, dim=0)).cpu()
        # RMS error of the mass fractions Y_s = rho_s / rho (the density
        # error of the FP32 flow path divides out)
        ymf = (a[:, 5:] / a[:, :1] - ref[:, 5:] / ref[:, :1]).pow(2).mean(0).sqrt().cpu()
        species[key] = (ymf, err[5:])
        print(f"{key}: fallback {fallbacks[key]:.4f}, rel L2 rho {float(err[0]):.3e} "
              f"rho_E {float(err[4]):.3e}; species mass-fraction L2 {ymf.numpy()}; "
              f"relative L2 O {float(err[5 + 7]):.3e} NO {float(err[5 + 8]):.3e}", flush=True)
        assert float(err[0]) <= 1e-4 and float(err[4]) <= 1e-4, err
        assert float(ymf.max()) <= 1e-6, ymf
    assert fallbacks["strict"] > fallbacks["fast"] > 0
    for s in (7, 8):                     # O, NO
        assert species["strict"][0][s] <= species["fast"][0][s] * (1 + 1e-6)
    # the chemistry moved the species over the run
    assert float(torch.linalg.vector_norm(ref[:, 5 + 10] - 0.233 * ref[:, 0])) > 0
