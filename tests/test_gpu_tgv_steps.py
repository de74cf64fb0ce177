"""Multi-step Taylor-Green runs on the GPU.

* The 8^3, p=2 configuration in full: 10 SSP-RK3 steps at CFL 0.3 (dt from
  the oracle's estimate, the same dt for both), FP64 GPU against the CPU
  FP64 oracle: relative L2 <= 1e-11 per variable (30 stages of the 1e-12
  per-stage bar; the Mach 0.1 pressure cancellation amplifies round-off), and
  the FP32 path within 1e-5 for rho and rho E.
* 100 steps of the p=3 path on 8^3 elements: FP32 against FP64 on the GPU,
  rho and rho E within the 1e-4 error budget of the FP32 configuration.
"""

import numpy as np
import pytest
import torch

from conftest import rel_l2, tgv_primitive

from hybrid import HybridCompressibleSolver, HybridNavierStokesOperator, HybridSSPRK3, box_mesh
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver
from oracle.time_integration import SSPRK3

pytestmark = pytest.mark.gpu


def _np(d):
    return {k: v.double().cpu().numpy() for k, v in d.items()}


def _device(m, p, dtype):
    s = HybridCompressibleSolver(m, polynomial_degree=p, dynamic_viscosity=1 / 1600,
                                 prandtl=0.71, dtype=dtype)
    return s, HybridSSPRK3(HybridNavierStokesOperator(s), cfl=0.3)


def test_tgv_8_p2_ten_steps_against_oracle():
    m = box_mesh(8, lengths=(2 * np.pi,) * 3, periodic=(True,) * 3)
    os_ = CompressibleSolver(dynamic_viscosity=1 / 1600, prandtl=0.71)
    oop = NavierStokesOperator(from_arrays(m), os_, 2)
    ork = SSPRK3(oop, cfl=0.3)
    st = {t: os_.primitive_to_conserved(tgv_primitive(x))
          for t, x in oop.geometry.solution_node_coordinates.items()}
    dt = ork.stable_dt(st)
    print(f"case TGV 8^3 p=2, dt {dt:.4e}", flush=True)
    ref = st
    for _ in range(10):
        ref, ok, _, _ = ork.checked_step(ref, dt)
        assert ok
    for dtype, tol in ((torch.float64, 1e-11), (torch.float32, 1e-5)):
        s, rk = _device(m, 2, dtype)
        U = s.to_device(st)
        for _ in range(10):
            U, valid, _, _ = rk.checked_step(U, dt)
        err = rel_l2(_np(U), ref)
        print(f"{dtype}: rel L2 after 10 steps {err}", flush=True)
        assert bool(valid)
        if dtype == torch.float64:
            assert np.all(err <= tol), err
        else:
            assert err[0] <= tol and err[4] <= tol, err


def test_tgv_p3_fp32_hundred_steps_within_budget():
    m = box_mesh(8, lengths=(2 * np.pi,) * 3, periodic=(True,) * 3)
    out, dt = {}, None
    for dtype in (torch.float64, torch.float32):
        s, rk = _device(m, 3, dtype)
        U = s.state_blocks_from_primitive_function(tgv_primitive)
        if dt is None:
            dt = float(rk.stable_dt(U))
        for _ in range(100):
            U, valid, _, _ = rk.checked_step(U, dt)
        assert bool(valid)
        out[dtype] = _np(U)
    err = rel_l2(out[torch.float32], out[torch.float64])
    print(f"FP32 vs FP64 after 100 steps: {err}", flush=True)
    assert err[0] <= 1e-4 and err[4] <= 1e-4, err
