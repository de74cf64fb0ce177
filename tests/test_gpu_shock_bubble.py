"""The Mach 2.5 shock-bubble workload (the unit-box MOOD benchmark) down-scaled
to a few hundred elements, one MOOD SSP-RK3 step on the GPU against the CPU
FP64 oracle.

The box is 6^3 hexahedra with a seeded 20 % split into wedges and
tetrahedra, like the benchmark mesh, scaled to a side of 0.1 so that every
element is below the sliver volume 1e-4 and M^-1 is scaled by
min(1, V/1e-4) < 1 everywhere (0.046 on hexahedra), as on the full-size
64^3 mesh (0.038).  The flow is the benchmark's own initial condition
(``bench.shock_bubble_primitive``) in the scaled coordinates: post-shock
farfield inflow at x-, transmissive outflow at x+, slip walls elsewhere,
Re 1e4 on the box side and the post-shock speed, MOOD with the sensor 0.05,
WENO subcells and positivity floors 1e-7; dt from the oracle's CFL 0.3
estimate.

Bars: per stage, identical troubled sets (no sensor value within 1e-9 of
the threshold; in FP32, compared at stages where every sensor value is at
least 1e-4 from it); after the step, relative L2 per variable <= 1e-11 in FP64
(three MOOD stages; 1e-12 per stage) and 1e-4 in FP32; identical validity.
"""

import numpy as np
import pytest
import torch

from conftest import rel_l2

import bench
from hybrid import (FarfieldWallBoundary, HybridCompressibleSolver, HybridMOODSubcellLimiter,
                    HybridNavierStokesOperator, HybridPositivityLimiter, HybridSSPRK3, box_mesh)
from oracle.limiter import MOODSubcellLimiter, PositivityLimiter
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver, FarfieldWallBoundary as OracleBoundary
from oracle.time_integration import RK3_WEIGHTS, SSPRK3

pytestmark = pytest.mark.gpu

SIDE = 0.1
NAMES = {"xmin": "inlet", "xmax": "outlet", "ymin": "walls", "ymax": "walls",
         "zmin": "walls", "zmax": "walls"}


def _setup(dtype):
    mesh = box_mesh(6, lengths=(SIDE,) * 3, split_fraction=0.2, seed=1, plane_names=NAMES)
    _, far = bench.shock_bubble_primitive(np.zeros((1, 3)))
    mu = far[1] * SIDE / 1.0e4
    gamma, pr = bench.GAMMA, bench.PR
    solver = HybridCompressibleSolver(mesh, polynomial_degree=2, gamma=gamma,
                                      dynamic_viscosity=mu, prandtl=pr, dtype=dtype)
    op = HybridNavierStokesOperator(solver)
    pos = HybridPositivityLimiter(op, density_floor=1e-7, pressure_floor=1e-7)
    lim = HybridMOODSubcellLimiter(op, threshold=0.05, positivity=pos, reconstruction="weno")
    bnd = FarfieldWallBoundary(solver, far)
    osolver = CompressibleSolver(gamma=gamma, dynamic_viscosity=mu, prandtl=pr)
    oop = NavierStokesOperator(from_arrays(mesh), osolver, 2)
    opos = PositivityLimiter(oop, density_floor=1e-7, pressure_floor=1e-7)
    olim = MOODSubcellLimiter(oop, threshold=0.05, positivity=opos, reconstruction="weno")
    ob = OracleBoundary(osolver, far, inlet_tags=sorted(bnd.inlet_tags),
                        outlet_tags=sorted(bnd.outlet_tags), wall_tags=sorted(bnd.wall_tags),
                        no_slip_tags=sorted(bnd.no_slip_tags), slip_tags=sorted(bnd.slip_tags))
    states = {t: osolver.primitive_to_conserved(bench.shock_bubble_primitive(x / SIDE)[0])
              for t, x in oop.geometry.solution_node_coordinates.items()}
    return solver, op, lim, bnd, oop, olim, ob, states


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_shock_bubble_mood_step(dtype):
    fp64 = dtype == torch.float64
    print(f"case shock-bubble 6^3 side {SIDE} {dtype}", flush=True)
    solver, op, lim, bnd, oop, olim, ob, states = _setup(dtype)
    scale = [float((t["scale"] * t["detw"]).max()) for t in op.device_mesh.block_tables]
    assert max(scale) < 0.1, scale          # sliver-scaled M^-1 on every block
    ork = SSPRK3(oop, cfl=0.3, limiter=olim, density_floor=1e-7, pressure_floor=1e-7)
    dt = ork.stable_dt(states)
    # record the oracle's troubled set of each stage (detection on the stage candidate)
    seen, margins = [], []
    detect = olim.troubled

    def recording(candidate):
        sensor, _ = olim.sensor(candidate)
        margin = min(np.abs(s - olim.sensor.threshold).min() for s in sensor.values())
        assert margin > 1e-9
        out = detect(candidate)
        seen.append(out)
        margins.append(margin)
        return out

    olim.troubled = recording
    drk = HybridSSPRK3(op, cfl=0.3, limiter=lim, density_floor=1e-7, pressure_floor=1e-7)
    S0 = solver.to_device(states)
    o, d = states, S0
    total = 0
    for k, (a, b) in enumerate(RK3_WEIGHTS):
        o, ov, orr, opp = ork.rk_stage(states, o, dt, a, b, ob)
        d, dv, drr, dpp = drk.rk_stage(S0, d, dt, a, b, bnd)
        flags = lim._ctx().tables.flags
        # FP32 sensor values carry ~1e-6 relative rounding: compare the sets
        # where every sensor value is at least 1e-4 from the threshold
        for bi, t in enumerate(op.block_types) if fp64 or margins[k] > 1e-4 else ():
            got = flags[bi].cpu().numpy() != 0
            want = seen[k][t]
            assert np.array_equal(got, want), (k, t, int(got.sum()), int(want.sum()))
            total += int(want.sum())
        assert bool(dv) == bool(ov)
    print(f"dt {dt:.3e}, troubled per stage {[int(sum(v.sum() for v in s.values())) for s in seen]}",
          flush=True)
    assert total > 0
    err = rel_l2({k: v.double().cpu().numpy() for k, v in d.items()}, o)
    print(f"rel L2 after one step {err}", flush=True)
    assert np.all(err <= (1e-11 if fp64 else 1e-4)), err
