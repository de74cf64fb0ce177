"""Pins of the oracle's fluxes, geometry, Euler/BR2 operators and SSP-RK3
against closed forms and invariants (the reference has no golden vectors for
these functions)."""

import numpy as np
import pytest

from conftest import PLANE_NAMES, perturbed_primitive, tgv_primitive

from hybrid.meshgen import box_mesh, smoke_mesh
from oracle.mesh import from_arrays
from oracle.operators import EulerOperator, NavierStokesOperator
from oracle.solver import CompressibleSolver
from oracle.time_integration import SSPRK3


def _states(oop, prim):
    s = oop.solver
    return {t: s.primitive_to_conserved(prim(x))
            for t, x in oop.geometry.solution_node_coordinates.items()}


# -- pointwise fluxes -------------------------------------------------------------
def _random_states(n, seed):
    rng = np.random.default_rng(seed)
    s = CompressibleSolver()
    W = np.stack([0.5 + rng.random(n), rng.standard_normal(n), rng.standard_normal(n),
                  rng.standard_normal(n), 0.5 + rng.random(n)], -1)
    nrm = rng.standard_normal((n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return s, s.primitive_to_conserved(W), nrm


def test_hllc_consistency_and_mirror_symmetry():
    s, UL, n = _random_states(200, 0)
    _, UR, _ = _random_states(200, 1)
    assert np.array_equal(s.hllc_flux(UL, UL, n), s.normal_physical_flux(UL, n))
    F1 = s.hllc_flux(UL, UR, n)
    F2 = s.hllc_flux(UR, UL, -n)
    assert np.abs(F1 + F2).max() < 1e-12 * np.abs(F1).max()


def test_hllc_resolves_isolated_contact_exactly():
    """A stationary contact (equal p, zero normal velocity, different rho and
    tangential velocity) has the exact flux (0, p n, 0)."""
    s = CompressibleSolver()
    n = np.array([[0.0, 0.0, 1.0]])
    WL = np.array([[1.0, 0.3, -0.2, 0.0, 2.0]])
    WR = np.array([[0.25, -0.7, 0.4, 0.0, 2.0]])
    F = s.hllc_flux(s.primitive_to_conserved(WL), s.primitive_to_conserved(WR), n)
    assert np.allclose(F, [[0.0, 0.0, 0.0, 2.0, 0.0]], atol=1e-14)


def test_rusanov_closed_form_and_fallback():
    s = CompressibleSolver()
    n = np.array([[1.0, 0.0, 0.0]])
    UL = s.primitive_to_conserved(np.array([[1.0, 0.5, 0.0, 0.0, 1.0]]))
    UR = s.primitive_to_conserved(np.array([[0.5, -0.5, 0.0, 0.0, 0.4]]))
    cl, cr = np.sqrt(1.4 / 1.0), np.sqrt(1.4 * 0.4 / 0.5)
    smax = max(0.5 + cl, 0.5 + cr)
    expect = 0.5 * (s.normal_physical_flux(UL, n) + s.normal_physical_flux(UR, n)) \
        - 0.5 * smax * (UR - UL)
    assert np.allclose(s.rusanov_flux(UL, UR, n), expect, rtol=1e-15, atol=1e-15)
    # negative pressure on one side -> HLLC falls back to Rusanov
    bad = UR.copy()
    bad[0, 4] = 0.01
    assert np.array_equal(s.hllc_flux(UL, bad, n), s.rusanov_flux(UL, bad, n))


# -- geometry ------------------------------------------------------------------------
@pytest.mark.parametrize("mesh", ["split", "smoke", "periodic"])
def test_geometry_invariants(mesh):
    if mesh == "split":
        m = box_mesh(3, lengths=(1.0, 2.0, 0.5), split_fraction=0.5, seed=3,
                     plane_names=PLANE_NAMES)
        volume = 1.0
    elif mesh == "smoke":
        m = smoke_mesh()
        volume = 1.0 + 0.5 + 1.0 / 6.0
    else:
        m = box_mesh(3, lengths=(2.0, 2.0, 2.0), periodic=(True, True, False),
                     split_fraction=0.4, seed=9, plane_names=PLANE_NAMES)
        volume = 8.0
    om = from_arrays(m)
    op = EulerOperator(om, CompressibleSolver(), 2)
    geo = op.geometry
    total = sum(geo.element_metrics[b].integration_weights.sum() for b in geo.element_metrics)
    assert abs(total - volume) < 1e-12
    for ftype, fb in om.face_blocks.items():
        fm = geo.face_metrics[ftype]
        assert np.abs(np.linalg.norm(fm.normals, axis=-1) - 1).max() < 1e-14
        assert np.abs(fm.integration_weights.sum(1) - fb.areas).max() < 1e-13
        inner = np.flatnonzero(fb.internal)
        # neighbour reference points map to the owner's physical points
        for bid in np.unique(fb.neighbor_block[inner]):
            sel = inner[fb.neighbor_block[inner] == bid]
            block = om.element_blocks[bid]
            g = geo.geometry_reference_elements[block.element_type]
            X = om.points[block.connectivity[fb.neighbor_local[sel]]]
            x = np.einsum("fqi,fid->fqd", g.basis(fm.neighbor_reference_points[sel]), X)
            d = x - fm.physical_points[sel]
            if om.period is not None:
                for k in range(3):
                    if om.period[k] > 0:
                        d[..., k] -= om.period[k] * np.round(d[..., k] / om.period[k])
            assert np.abs(d).max() < 1e-13


# -- Euler operator ------------------------------------------------------------------
def test_free_stream_residual_is_exactly_zero():
    m = box_mesh(3, periodic=(True,) * 3, split_fraction=0.4, seed=4)
    op = EulerOperator(from_arrays(m), CompressibleSolver(), 3)
    W = np.array([1.3, 0.4, -0.2, 0.7, 2.1])
    st = {t: np.broadcast_to(op.solver.primitive_to_conserved(W),
                             x.shape[:-1] + (5,)).copy()
          for t, x in op.geometry.solution_node_coordinates.items()}
    r = op.residual(st)
    assert all(np.abs(v).max() == 0.0 for v in r.values())


@pytest.mark.parametrize("p", [1, 2, 3])
def test_periodic_conservation(p):
    """On a periodic mesh sum_i R_i (sum of the weak residual over all basis
    functions) vanishes to round-off for Euler and BR2 terms."""
    m = box_mesh(3, lengths=(2 * np.pi,) * 3, periodic=(True,) * 3, split_fraction=0.4, seed=6)
    op = NavierStokesOperator(from_arrays(m), CompressibleSolver(dynamic_viscosity=0.01), p)
    st = _states(op, lambda x: perturbed_primitive(x, seed=5))
    for res in (op.residual(st), op.viscous_residual(st)):
        tot = sum(v.sum((0, 1)) for v in res.values())
        scale = sum(np.abs(v).sum((0, 1)) for v in res.values())
        assert np.all(np.abs(tot) <= 1e-13 * scale.max())


def test_euler_rhs_converges_to_exact_divergence():
    """rhs -> -div F(U) for a smooth periodic field as the mesh is refined."""
    s = CompressibleSolver()

    def prim(x):
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        return np.stack([1 + 0.2 * np.sin(X), 0.5 + 0.1 * np.cos(Y), 0.1 * np.sin(Z),
                         0.2 + 0 * X, 1 + 0.1 * np.cos(X + Y)], -1)

    def exact_dudt(x, h=1e-5):
        out = np.zeros(x.shape[:-1] + (5,))
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            Fp = s.physical_flux_tensor(s.primitive_to_conserved(prim(x + e)))[..., d, :]
            Fm = s.physical_flux_tensor(s.primitive_to_conserved(prim(x - e)))[..., d, :]
            out -= (Fp - Fm) / (2 * h)
        return out

    errs = []
    for n in (3, 6):
        m = box_mesh(n, lengths=(2 * np.pi,) * 3, periodic=(True,) * 3)
        op = EulerOperator(from_arrays(m), s, 2)
        st = _states(op, prim)
        x = op.geometry.solution_node_coordinates["hexahedron"]
        errs.append(np.abs(op.rhs(st)["hexahedron"] - exact_dudt(x)).max())
    assert errs[1] < errs[0] / 3.5, errs


# -- BR2 ---------------------------------------------------------------------------------
@pytest.mark.parametrize("p", [2, 3])
def test_br2_affine_field_exact(p):
    """rho constant, u affine and T affine: traces match, every lift is zero,
    the gradient is exact and viscous_rhs = (0, 0, 0, 0, grad u : tau) at every
    node, boundary elements included (boundary_state=None copies the trace)."""
    s = CompressibleSolver(dynamic_viscosity=0.3, prandtl=0.7)
    G = np.array([[0.1, -0.2, 0.3], [0.05, 0.2, -0.1], [-0.15, 0.1, 0.05]])
    gT = np.array([0.2, -0.1, 0.3])

    def prim(x):
        u = 0.3 + np.einsum("id,...d->...i", G, x)
        T = 2.0 + x @ gT
        rho = np.full(x.shape[:-1], 1.5)
        return np.concatenate([rho[..., None], u, (rho * T)[..., None]], -1)

    m = box_mesh(2, lengths=(1.0, 1.2, 0.8), split_fraction=0.5, seed=8,
                 plane_names=PLANE_NAMES)
    op = NavierStokesOperator(from_arrays(m), s, p)
    st = _states(op, prim)
    r = op.viscous_rhs(st)
    div = np.trace(G)
    tau = s.dynamic_viscosity * (G + G.T - 2 / 3 * div * np.eye(3))
    expect = np.array([0, 0, 0, 0, np.sum(G * tau)])
    for v in r.values():
        assert np.abs(v - expect).max() < 1e-9  # round-off through M^-1 at p=3


# -- SSP-RK3 ---------------------------------------------------------------------------
class _Linear:
    """dU/dt = lam U on one dummy block."""

    def __init__(self, lam):
        self.lam = lam
        self.solver = CompressibleSolver()
        self.degree = 1

    def rhs(self, states, boundary_state=None):
        return {k: self.lam * v for k, v in states.items()}


def test_ssprk3_stability_polynomial():
    lam, dt = -0.7, 0.3
    op = _Linear(lam)
    rk = SSPRK3.__new__(SSPRK3)
    rk.operator, rk.solver, rk.limiter = op, op.solver, None
    rk.density_floor = rk.pressure_floor = -np.inf
    U0 = {"hexahedron": np.full((1, 1, 5), 2.0)}
    out, ok, _, _ = rk.checked_step(U0, dt)
    z = lam * dt
    assert np.allclose(out["hexahedron"], 2.0 * (1 + z + z ** 2 / 2 + z ** 3 / 6),
                       rtol=1e-15)
