"""Pins of the oracle's WALE eddy viscosity (hybrid/solver.py:317-355) and
its use in the viscous flux and the operator (hybrid/navier_stokes.py:
220-290): closed forms for pure shear (nu_t = 0, the model's defining
property), solid-body rotation ((C_w D)^2 (2/3)^(1/4) w) and irrotational
compression (switched off by the Ducros sensor); the flux adds rho nu_t to
mu and rho nu_t c_p / Pr_t to kappa; a linear velocity field on a mesh keeps
the affine-exactness of the BR2 operator with the eddy viscosity of its
constant gradient."""

import numpy as np
import pytest

from hybrid.meshgen import box_mesh
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver

S = CompressibleSolver(use_wale=True, wale_constant=0.5, dynamic_viscosity=0.01)


def _grad(G):
    g = np.zeros((1, 4, 3))
    g[0, :3] = G
    return g


def _nu(G, width=0.3, rho=1.7):
    U = np.array([[rho, 0, 0, 0, 2.5]])
    return S.wale_dynamic_viscosity(U, _grad(G), width)[0] / rho


def test_pure_shear_has_no_eddy_viscosity():
    assert _nu(np.array([[0.0, 3.0, 0.0], [0, 0, 0], [0, 0, 0]])) == 0.0


def test_solid_body_rotation_closed_form():
    w = 2.5
    G = np.array([[0.0, -w, 0.0], [w, 0.0, 0.0], [0.0, 0.0, 0.0]])
    want = (0.5 * 0.3) ** 2 * (2.0 / 3.0) ** 0.25 * w * (4 * w * w / (4 * w * w + 1e-12))
    assert abs(_nu(G) - want) < 1e-14 * want


def test_irrotational_compression_is_switched_off():
    G = np.diag([-2.0, -1.0, 0.5])
    assert _nu(G) < 1e-10


def test_flux_adds_the_eddy_viscosity():
    rng = np.random.default_rng(3)
    G = rng.standard_normal((3, 3))
    U = np.array([[1.3, 0.2, -0.1, 0.4, 3.0]])
    g = _grad(G)
    g[0, 3] = [0.3, -0.2, 0.1]
    mt = S.wale_dynamic_viscosity(U, g, 0.2)[0]
    assert mt > 0
    plain = CompressibleSolver(dynamic_viscosity=0.01 + mt)
    F = S.viscous_physical_flux(U, g, 0.2)
    Fp = plain.viscous_physical_flux(U, g)
    # momentum rows: (mu + mu_t) tau; energy: u.tau + (kappa + mu_t cp / Prt) grad T
    assert np.allclose(F[..., 1:4], Fp[..., 1:4], rtol=1e-14, atol=1e-16)
    dk = (S.thermal_conductivity + mt * S.specific_heat_cp / S.turbulent_prandtl
          - plain.thermal_conductivity)
    assert np.allclose(F[..., 4], Fp[..., 4] + dk * g[0, 3], rtol=1e-13, atol=1e-15)
    # without a filter width (or with WALE off) the molecular flux is unchanged
    assert np.array_equal(S.viscous_physical_flux(U, g),
                          CompressibleSolver(dynamic_viscosity=0.01).viscous_physical_flux(U, g))


@pytest.mark.parametrize("p", [2, 3])
def test_affine_field_with_wale(p):
    """Affine u and T, constant rho, equal hexahedra: every lift vanishes and
    the raw gradient is exact, so the viscous rhs is (0, 0, 0, 0, grad u : tau) with tau built
    from mu + rho nu_t of the constant gradient, at every node."""
    s = CompressibleSolver(dynamic_viscosity=0.02, prandtl=0.7, use_wale=True)
    G = np.array([[0.1, -0.6, 0.3], [0.5, 0.2, -0.1], [-0.15, 0.1, 0.05]])
    gT = np.array([0.2, -0.1, 0.3])

    def prim(x):
        u = 0.3 + np.einsum("id,...d->...i", G, x)
        T = 2.0 + x @ gT
        rho = np.full(x.shape[:-1], 1.5)
        return np.concatenate([rho[..., None], u, (rho * T)[..., None]], -1)

    # equal elements: one filter width, so rho nu_t is the same on both sides
    # of every face (a jump in it would leave a face flux difference)
    m = box_mesh(2, lengths=(1.0, 1.2, 0.8))
    op = NavierStokesOperator(from_arrays(m), s, p)
    st = {t: s.primitive_to_conserved(prim(x)) for t, x in op.geometry.solution_node_coordinates.items()}
    r = op.viscous_rhs(st)
    for t, v in r.items():
        # 8 hexahedra of 0.5 x 0.6 x 0.4: width V^(1/3) / (p + 1)
        w = np.full(len(v), 0.12 ** (1.0 / 3.0) / (p + 1))
        assert np.allclose(op.filter_widths[t], w, rtol=1e-12)
        mt = s.wale_dynamic_viscosity(np.array([[1.5, 0, 0, 0, 1.0]]), _grad(G), w)   # (E,)
        div = np.trace(G)
        tau = (s.dynamic_viscosity + mt)[:, None, None] * (G + G.T - 2 / 3 * div * np.eye(3))
        want = np.sum(G * tau, axis=(1, 2))
        assert np.abs(v[..., :4]).max() < 1e-9
        assert np.abs(v[..., 4] - want[:, None]).max() < 1e-9 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("p", [1, 2])
def test_eddy_viscosity_comes_from_the_resolved_gradient(p):
    """A piecewise-constant state has a zero polynomial gradient inside every
    element, so the WALE viscosity (taken from the resolved gradient, not the
    lifted one) vanishes everywhere and the viscous rhs equals the molecular
    one, although the BR2 lifts of the jumps are far from zero."""
    m = box_mesh(2, split_fraction=0.5, seed=4)
    rng = np.random.default_rng(11)
    mol = CompressibleSolver(dynamic_viscosity=0.02)
    wal = CompressibleSolver(dynamic_viscosity=0.02, use_wale=True)
    op_m = NavierStokesOperator(from_arrays(m), mol, p)
    op_w = NavierStokesOperator(from_arrays(m), wal, p)
    st = {}
    for t, x in op_m.geometry.solution_node_coordinates.items():
        W = np.empty(x.shape[:-1] + (5,))
        E = x.shape[0]
        W[..., 0] = rng.uniform(0.8, 1.2, E)[:, None]
        W[..., 1:4] = rng.uniform(-1.0, 1.0, (E, 1, 3))
        W[..., 4] = rng.uniform(0.8, 1.2, E)[:, None]
        st[t] = mol.primitive_to_conserved(W)
    rm, rw = op_m.viscous_rhs(st), op_w.viscous_rhs(st)
    for t in rm:
        assert np.abs(rm[t]).max() > 1e-2
        assert np.allclose(rw[t], rm[t], rtol=1e-13, atol=1e-13)
