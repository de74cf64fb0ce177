"""Time-step control on the GPU against the CPU FP64 oracle: the CFL dt, the
element-local dt with its neighbour smoothing, the FP64 conservation sums,
and the device-resident step loop (accept / reject, retry factor, target
time) against the oracle's restatement of run_hybrid_sphere.py:265-346.

Bars: dt values 1e-13 relative (element maxima and means of the same FP64
values, summed in another order); totals 1e-13 relative to the largest
total; the step loop: the same number of accepted steps and rejections, the
same time to 1e-14, states to 1e-11 relative L2 (a few stages of 1e-12)."""

import numpy as np
import pytest
import torch

from conftest import PLANE_NAMES, perturbed_primitive, rel_l2

import bench
from hybrid import (FarfieldWallBoundary, HybridCompressibleSolver, HybridMOODSubcellLimiter,
                    HybridNavierStokesOperator, HybridPositivityLimiter, HybridSSPRK3, box_mesh)
from oracle.limiter import MOODSubcellLimiter, PositivityLimiter
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver, FarfieldWallBoundary as OracleBoundary
from oracle.time_integration import SSPRK3

pytestmark = pytest.mark.gpu

NAMES = {"xmin": "inlet", "xmax": "outlet", "ymin": "walls", "ymax": "walls",
         "zmin": "walls", "zmax": "walls"}


def _case(name, dtype=torch.float64, mood=False, cfl=0.3, pfloor=1e-7, wale=False):
    if name == "split":
        m = box_mesh(4, split_fraction=0.35, seed=7, plane_names=PLANE_NAMES)
        mu = 0.01
        prim = lambda x: perturbed_primitive(x, seed=2)
        far = np.array([1.2, 0.9, 0.1, -0.05, 1.1])
    elif name == "sliver":
        m = box_mesh(4, lengths=(0.1,) * 3, split_fraction=0.2, seed=5, plane_names=PLANE_NAMES)
        mu = 0.001
        prim = lambda x: perturbed_primitive(10.0 * x, seed=6)
        far = np.array([1.2, 0.9, 0.1, -0.05, 1.1])
    elif name == "shock":
        m = box_mesh(3, split_fraction=0.3, seed=2, plane_names=NAMES)
        _, far = bench.shock_bubble_primitive(np.zeros((1, 3)))
        mu = far[1] / 1e3
        prim = lambda x: bench.shock_bubble_primitive(x)[0]
    else:
        raise KeyError(name)
    ds = HybridCompressibleSolver(m, polynomial_degree=2, dynamic_viscosity=mu, prandtl=0.71,
                                  use_wale=wale, dtype=dtype)
    dop = HybridNavierStokesOperator(ds)
    os_ = CompressibleSolver(dynamic_viscosity=mu, prandtl=0.71, use_wale=wale)
    oop = NavierStokesOperator(from_arrays(m), os_, 2)
    db = FarfieldWallBoundary(ds, far)
    ob = OracleBoundary(os_, far, inlet_tags=sorted(db.inlet_tags),
                        outlet_tags=sorted(db.outlet_tags), wall_tags=sorted(db.wall_tags),
                        no_slip_tags=sorted(db.no_slip_tags), slip_tags=sorted(db.slip_tags))
    dlim = olim = None
    if mood:
        dlim = HybridMOODSubcellLimiter(dop, positivity=HybridPositivityLimiter(
            dop, density_floor=1e-7, pressure_floor=1e-7))
        olim = MOODSubcellLimiter(oop, positivity=PositivityLimiter(
            oop, density_floor=1e-7, pressure_floor=1e-7))
    drk = HybridSSPRK3(dop, cfl=cfl, limiter=dlim, pressure_floor=pfloor)
    ork = SSPRK3(oop, cfl=cfl, limiter=olim, pressure_floor=pfloor)
    st = {t: os_.primitive_to_conserved(prim(x))
          for t, x in oop.geometry.solution_node_coordinates.items()}
    return ds, drk, db, ork, ob, st


def _np(d):
    return {k: v.double().cpu().numpy() for k, v in d.items()}


@pytest.mark.parametrize("name", ["split", "sliver"])
def test_stable_and_local_dt(name):
    print(f"case {name} dt", flush=True)
    ds, drk, db, ork, ob, st = _case(name)
    U = ds.to_device(st)
    got = float(drk.stable_dt(U))
    want = ork.stable_dt(st)
    assert abs(got - want) <= 1e-13 * want, (got, want)
    gl = _np(drk.local_dt(U))
    wl = ork.local_dt(st)
    for t in wl:
        assert gl[t].shape == wl[t].shape
        assert np.abs(gl[t] - wl[t]).max() <= 1e-13 * wl[t].max(), t


@pytest.mark.parametrize("name", ["split", "sliver"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_dt_with_wale(name, dtype):
    """The CFL and element-local dt with the WALE viscosity of the resolved
    gradient (hybrid/time_integration.py:119-148) in the viscous rate; the
    WALE term moves the oracle's dt, so the comparison sees it.  FP64 states:
    1e-13; FP32 states: the FP64 rate of the FP32 state against the oracle on
    the same rounded state, 1e-7 (the basis tables and metrics rounded to
    FP32, through the power laws of the model; measured 3e-9)."""
    print(f"case {name} dt WALE {dtype}", flush=True)
    ds, drk, db, ork, ob, st = _case(name, dtype=dtype, wale=True)
    _, _, _, ork0, _, _ = _case(name)
    U = ds.to_device(st)
    st = _np(U)
    want = ork.stable_dt(st)
    assert abs(ork0.stable_dt(st) - want) > 1e-3 * want
    tol = 1e-13 if dtype == torch.float64 else 1e-7
    got = float(drk.stable_dt(U))
    print(f"dt {got!r} oracle {want!r} rel {abs(got - want) / want:.2e}", flush=True)
    assert abs(got - want) <= tol * want, (got, want)
    gl = _np(drk.local_dt(U))
    wl = ork.local_dt(st)
    for t in wl:
        assert np.abs(gl[t] - wl[t]).max() <= tol * wl[t].max(), t


def test_conserved_totals_and_conservation():
    """The totals match the oracle's; on a periodic mesh SSP-RK3 steps keep
    them to round-off in FP64 (1e-13) and to FP32 round-off in FP32."""
    m = box_mesh(3, lengths=(2 * np.pi,) * 3, periodic=(True,) * 3, split_fraction=0.4, seed=6)
    for dtype, tol in ((torch.float64, 1e-13), (torch.float32, 2e-6)):
        ds = HybridCompressibleSolver(m, polynomial_degree=2, dynamic_viscosity=0.01,
                                      dtype=dtype)
        rk = HybridSSPRK3(HybridNavierStokesOperator(ds), cfl=0.3)
        os_ = CompressibleSolver(dynamic_viscosity=0.01)
        oop = NavierStokesOperator(from_arrays(m), os_, 2)
        st = {t: os_.primitive_to_conserved(perturbed_primitive(x, seed=5))
              for t, x in oop.geometry.solution_node_coordinates.items()}
        U = ds.to_device(st)
        want = SSPRK3(oop).conserved_totals({k: v.double().cpu().numpy() for k, v in U.items()})
        tot0 = rk.conserved_totals(U).cpu().numpy()
        scale = np.abs(want).max()
        assert np.abs(tot0 - want).max() <= 1e-13 * scale
        dt = float(rk.stable_dt(U))
        for _ in range(4):
            U, valid, _, _ = rk.checked_step(U, dt)
        tot = rk.conserved_totals(U).cpu().numpy()
        print(f"{dtype}: totals change {np.abs(tot - tot0) / scale}", flush=True)
        assert bool(valid)
        assert np.abs(tot - tot0).max() <= tol * scale


def test_step_loop_with_rejections_matches_oracle():
    """A pressure floor of 0.95 (the pre-shock pressure is 1) makes attempts
    at the CFL dt fail as the expansion behind the shock develops; the loop
    halves dt on each rejection and grows it by 1.25 after each accepted
    step.  Every attempt's p_min is at least 1e-9 from the floor, so the
    decisions do not hinge on round-off."""
    ds, drk, db, ork, ob, st = _case("shock", cfl=1.0, pfloor=0.95)
    attempts = []
    step = ork.checked_step

    def recording(states, dt, b=None):
        out = step(states, dt, b)
        attempts.append(out[3])
        return out

    ork.checked_step = recording
    want, wt, wn, wr = ork.advance(st, 3e-3, ob, maximum_steps=6)
    margin = min(abs(p - 0.95) for p in attempts)
    print(f"oracle: {wn} steps, {wr} rejections, t {wt:.6e}, p_min margin {margin:.2e}",
          flush=True)
    assert wr > 0 and margin > 1e-9
    got, gt, gn, gr = drk.advance(ds.to_device(st), 3e-3, db, maximum_steps=6, sync_every=4)
    assert (gn, gr) == (wn, wr)
    assert abs(gt - wt) <= 1e-14 * wt
    assert np.all(rel_l2(_np(got), want) <= 1e-11)


def test_step_loop_mood_lts_matches_oracle():
    ds, drk, db, ork, ob, st = _case("sliver", mood=True, cfl=0.3)
    want, wt, wn, wr = ork.advance(st, 1.0, ob, lts=True, maximum_steps=2)
    got, gt, gn, gr = drk.advance(ds.to_device(st), 1.0, db, lts=True, maximum_steps=2)
    assert (gn, gr) == (wn, wr) == (2, 0)
    assert abs(gt - wt) <= 1e-13 * wt
    assert np.all(rel_l2(_np(got), want) <= 1e-11)
