"""GPU Park-11 chemistry against the CPU FP64 oracle.

Bar: per cell, max_s |rho_gpu - rho_oracle| <= 1e-10 * sum(rho) (FP64; the
two sides differ only in summation order and the LU pivot arithmetic,
amplified by at most the 12-16 Newton iterations)."""

import numpy as np
import pytest
import torch

from hybrid.chemistry import Air11ParkMechanism, PointImplicitIntegrator, ReactingFlowCoupler
from hybrid.chemistry.batch import make_batch
from oracle.chemistry import Mechanism, ReactingFlowCoupler as OracleCoupler, integrator_step

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.mark.parametrize("const_T", [True, False])
def test_integrator_matches_oracle(const_T):
    rho, T = make_batch(96, seed=3)
    mech = Air11ParkMechanism()
    integ = PointImplicitIntegrator(mech, max_newton_iters=12)
    got, Tg, _ = integ.step(rho, T, 1e-7, constant_temperature=const_T)
    got = got.cpu().numpy()
    om = Mechanism()
    for c in range(len(T)):
        ref, Tr, _ = integrator_step(om, rho[c], T[c], 1e-7, 12, const_T)
        assert np.abs(got[c] - ref).max() <= TOL * ref.sum(), c
        assert abs(float(Tg[c]) - Tr) <= 1e-9 * Tr


@pytest.mark.parametrize("dimensionless", [False, True])
def test_coupler_matches_oracle(dimensionless):
    rng = np.random.default_rng(7)
    E, N = 6, 10
    om = Mechanism()
    U = np.zeros((E, N, 16))
    rho = 1.2 if not dimensionless else 1.0
    for e in range(E):
        for i in range(N):
            Y = rng.dirichlet(np.full(11, 2.0))
            T = rng.uniform(4000, 9000) if not dimensionless else rng.uniform(8, 25)
            u = rng.uniform(-1500, 1500, 3) if not dimensionless else rng.uniform(-2, 2, 3)
            if dimensionless:
                c = OracleCoupler(om, dimensionless=True)
                e_th = om.mixture_properties(rho * c.rho_ref * Y, T * c.T_ref)["internal_energy"] / c.e_ref
            else:
                e_th = om.mixture_properties(rho * Y, T)["internal_energy"]
            U[e, i, 0] = rho
            U[e, i, 1:4] = rho * u
            U[e, i, 4] = rho * (e_th + 0.5 * np.sum(u * u))
            U[e, i, 5:] = rho * Y
    dt = 1e-7 if not dimensionless else 0.05
    dev = ReactingFlowCoupler(dimensionless=dimensionless)
    got = dev.react_block(torch.as_tensor(U, device="cuda"), dt).cpu().numpy()
    ref = OracleCoupler(om, max_newton_iters=12, dimensionless=dimensionless).react_block(U, dt)
    assert np.array_equal(got[..., :5], U[..., :5])
    assert np.abs(got - ref).max() <= TOL * rho
    assert np.all(got[..., 5:] >= 0)
