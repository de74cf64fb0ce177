"""Parity of the HIP residual / rhs / RK-stage path with the CPU FP64 oracle.

FP64 bar: relative L2 error <= 1e-12 per conserved variable for the Euler
residual, the total rhs and the stage (the plain-path tolerance), raised only
where the oracle itself is less well conditioned: the bar is
max(1e-12, 4 x the oracle's own change under 1-ulp input noise), measured in
the test (p=3 Taylor-Green at Mach 0.1 has a floor of ~2.5e-12).  The
viscous-only residual is checked to 1e-10: on the Taylor-Green field its
rho*w component is zero analytically (only discretisation error remains) and
its energy component differentiates T ~ 71 +- 0.06, so two FP64 summation
orders differ there at the 1e-12..1e-11 level.

FP32 bar: relative L2 <= 5e-5 per variable on the Mach 0.5-1 fields; the
Mach 0.1 Taylor-Green field differentiates p ~ 71 +- 0.06 and is held to
5e-3 (single-precision cancellation, eps32 * 71 / 0.06 ~ 1e-4 per operation).
"""

import numpy as np
import pytest
import torch

from conftest import PLANE_NAMES, perturbed_primitive, rel_l2, tgv_primitive

from hybrid import FarfieldWallBoundary, HybridCompressibleSolver, HybridNavierStokesOperator, \
    HybridSSPRK3, box_mesh, smoke_mesh
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver, FarfieldWallBoundary as OracleBoundary
from oracle.time_integration import SSPRK3

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _floor(fn, states):
    """Conditioning floor of an oracle operator: its own relative change per
    variable under 1-ulp relative noise on the input (max of two draws)."""
    rng = np.random.default_rng(123)
    base = fn(states)
    worst = 0.0
    for _ in range(2):
        noisy = {k: v * (1 + 2.2e-16 * rng.standard_normal(v.shape)) for k, v in states.items()}
        worst = np.maximum(worst, rel_l2(fn(noisy), base))
    return worst


def _bar(fn, states, tol):
    """Parity bar: tol, or 4x the oracle's own conditioning floor where the
    problem is worse conditioned than tol (p=3 at Mach 0.1: ~2.5e-12)."""
    return np.maximum(tol, 4.0 * _floor(fn, states))
VISCOUS_TOL = 1e-10
FP32_TOL = {"tgv": 5e-3, "split": 5e-5, "species": 5e-5, "sliver": 5e-5}


def _case(name, p, dtype=torch.float64, wale=False, mu_override=None):
    mu = 0.0
    nspecies = 0
    boundary = False
    if name == "tgv":
        m = box_mesh(4, lengths=(2 * np.pi,) * 3, periodic=(True,) * 3)
        mu = 1.0 / 1600.0
        prim = tgv_primitive
    elif name == "smoke":
        m = smoke_mesh()
        mu = 0.02
        boundary = True
        prim = lambda x: perturbed_primitive(x, seed=1)
    elif name == "split":
        m = box_mesh(4, split_fraction=0.35, seed=7, plane_names=PLANE_NAMES)
        mu = 0.01
        boundary = True
        prim = lambda x: perturbed_primitive(x, seed=2)
    elif name == "split_periodic":
        m = box_mesh((4, 3, 4), lengths=(1.0, 2.0, 1.5), periodic=(True, True, True),
                     split_fraction=0.4, seed=11)
        mu = 0.005
        prim = lambda x: perturbed_primitive(x, seed=3)
    elif name == "sliver":
        # 0.1-box: every element below the sliver volume 1e-4, so M^-1 (and
        # the BR2 lifts) carry the factor min(1, V/1e-4) < 1
        m = box_mesh(4, lengths=(0.1,) * 3, split_fraction=0.2, seed=5, plane_names=PLANE_NAMES)
        mu = 0.001
        boundary = True
        prim = lambda x: perturbed_primitive(10.0 * x, seed=6)
    elif name == "species":
        m = box_mesh(3, split_fraction=0.4, seed=5, plane_names=PLANE_NAMES)
        nspecies = 11
        mu = 0.01
        boundary = True
        prim = lambda x: perturbed_primitive(x, seed=4, nspecies=11)
    else:
        raise KeyError(name)
    if mu_override is not None:
        mu = mu_override
    names = tuple(f"s{k}" for k in range(nspecies))
    dev_solver = HybridCompressibleSolver(m, species_names=names, polynomial_degree=p,
                                          dynamic_viscosity=mu, prandtl=0.71, use_wale=wale,
                                          dtype=dtype)
    dev_op = HybridNavierStokesOperator(dev_solver)
    om = from_arrays(m)
    osolver = CompressibleSolver(gamma=1.4, dynamic_viscosity=mu, prandtl=0.71,
                                 num_species=nspecies, use_wale=wale)
    oop = NavierStokesOperator(om, osolver, p)
    states = {t: osolver.primitive_to_conserved(prim(x))
              for t, x in oop.geometry.solution_node_coordinates.items()}
    dev_b = ob = None
    if boundary:
        far = np.concatenate([[1.2, 0.9, 0.1, -0.05, 1.1],
                              np.full(nspecies, 1.0 / max(nspecies, 1))])
        dev_b = FarfieldWallBoundary(dev_solver, far)
        ob = OracleBoundary(osolver, far, inlet_tags=sorted(dev_b.inlet_tags),
                            outlet_tags=sorted(dev_b.outlet_tags),
                            wall_tags=sorted(dev_b.wall_tags),
                            no_slip_tags=sorted(dev_b.no_slip_tags),
                            slip_tags=sorted(dev_b.slip_tags))
    return dev_solver, dev_op, dev_b, oop, ob, states


def _np(d):
    return {k: v.double().cpu().numpy() for k, v in d.items()}


CASES = [("smoke", 1), ("smoke", 2), ("smoke", 3), ("tgv", 1), ("tgv", 2), ("tgv", 3),
         ("split", 2), ("split", 3), ("split_periodic", 2), ("species", 2), ("sliver", 1),
         ("sliver", 2), ("sliver", 3)]


@pytest.mark.parametrize("name,p", CASES)
def test_residual_and_rhs_fp64(name, p):
    print(f"case {name} p={p}", flush=True)
    ds, dop, db, oop, ob, states = _case(name, p)
    U = ds.to_device(states)
    got = _np(dop.residual(U, db))
    ref = oop.residual(states, ob)
    bar = _bar(lambda s: oop.residual(s, ob), states, FP64_TOL)
    assert np.all(rel_l2(got, ref) <= bar), (rel_l2(got, ref), bar)
    got = _np(dop.viscous_residual(U, db))
    ref = oop.viscous_residual(states, ob)
    bar = _bar(lambda s: oop.viscous_residual(s, ob), states, VISCOUS_TOL)
    assert np.all(rel_l2(got, ref) <= bar), (rel_l2(got, ref), bar)
    got = _np(dop.rhs(U, db))
    ref = oop.rhs(states, ob)
    bar = _bar(lambda s: oop.rhs(s, ob), states, FP64_TOL)
    assert np.all(rel_l2(got, ref) <= bar), (rel_l2(got, ref), bar)


@pytest.mark.parametrize("name,p", [("tgv", 2), ("split", 2), ("smoke", 3), ("sliver", 2)])
def test_rk_stage_and_step_fp64(name, p):
    print(f"case {name} p={p}", flush=True)
    ds, dop, db, oop, ob, states = _case(name, p)
    U = ds.to_device(states)
    dt = 1e-4 if name == "sliver" else 1e-3
    oint = SSPRK3(oop)
    dint = HybridSSPRK3(dop)
    base = {k: v * 1.001 for k, v in states.items()}
    ref, ok, rmin, pmin = oint.rk_stage(base, states, dt, 0.75, 0.25, ob)
    got, gok, grmin, gpmin = dint.rk_stage(ds.to_device(base), U, dt, 0.75, 0.25, db)
    assert np.all(rel_l2(_np(got), ref) <= FP64_TOL)
    assert bool(gok) == ok
    assert abs(float(grmin) - rmin) <= 1e-13 * abs(rmin)
    assert abs(float(gpmin) - pmin) <= 1e-12 * abs(pmin)
    ref, ok, rmin, pmin = oint.checked_step(states, dt, ob)
    got, gok, grmin, gpmin = dint.checked_step(U, dt, db)
    assert np.all(rel_l2(_np(got), ref) <= FP64_TOL)
    assert bool(gok) == ok
    assert abs(float(grmin) - rmin) <= 1e-12 * abs(rmin)


@pytest.mark.parametrize("name,p", [("tgv", 3), ("split", 2), ("species", 2), ("sliver", 2)])
def test_rhs_fp32(name, p):
    print(f"case {name} p={p} fp32", flush=True)
    ds, dop, db, oop, ob, states = _case(name, p, dtype=torch.float32)
    got = _np(dop.rhs(ds.to_device(states), db))
    ref = oop.rhs(states, ob)
    err = rel_l2(got, ref)
    assert np.all(err <= FP32_TOL[name]), err


def test_free_stream_is_exact():
    """A uniform state has an exactly zero Euler residual on a periodic mesh
    (anchored traces + the identical-state HLLC shortcut)."""
    m = box_mesh(3, periodic=(True,) * 3, split_fraction=0.3, seed=2)
    ds = HybridCompressibleSolver(m, polynomial_degree=2)
    op = HybridNavierStokesOperator(ds)
    U = ds.uniform_state_blocks(1.3, (0.4, -0.2, 0.7), 2.1)
    r = op.residual(U)
    assert all(float(v.abs().max()) == 0.0 for v in r.values())


def test_in_place_stage_matches_out_of_place():
    ds, dop, db, oop, ob, states = _case("split", 2)
    U = ds.to_device(states)
    dint = HybridSSPRK3(dop)
    a, _, _, _ = dint.rk_stage(U, U, 1e-3, 0.0, 1.0, db)
    V = {k: v.clone() for k, v in U.items()}
    import ctypes
    from hybrid import _native as nat
    dint._stage(V, V, V, 1e-3, 0.0, 1.0, db, True)
    for k in a:
        assert torch.equal(a[k], V[k])


@pytest.mark.parametrize("name,p", [("split", 2), ("species", 2), ("sliver", 2)])
def test_euler_candidate_stage_without_base(name, p):
    """fsg_rk_stage with no base state and no diagnostics (the Euler
    candidate curr + dt rhs(curr) a MOOD stage starts from) equals the same
    stage with the base weight 0, bit for bit, and the oracle's candidate to
    the FP64 bar; the element paths (sum-factorised hexahedra, dense wedges
    and tetrahedra) all write it."""
    from hybrid import _native as nat
    ds, dop, db, oop, ob, states = _case(name, p)
    U = ds.to_device(states)
    dt = 1e-4 if name == "sliver" else 1e-3
    dint = HybridSSPRK3(dop)
    cand = {k: torch.empty_like(v) for k, v in U.items()}
    dint._stage_kernel(None, U, cand, dt, 0.0, 1.0, db, None)
    with_base = {k: torch.empty_like(v) for k, v in U.items()}
    dint._diag.reset(nat.stream_handle(ds.device))
    dint._stage_kernel(U, U, with_base, dt, 0.0, 1.0, db, dint._diag)
    for k in U:
        assert torch.equal(cand[k], with_base[k]), k
    rhs = oop.rhs(states, ob)
    ref = {k: states[k] + dt * rhs[k] for k in states}
    assert np.all(rel_l2(_np(cand), ref) <= FP64_TOL)


@pytest.mark.parametrize("name,p", [("smoke", 2), ("split", 3), ("tgv", 3), ("species", 2),
                                    ("sliver", 2), ("sliver", 3)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_both_element_paths(name, p, dtype):
    """The element paths compute the same operator: batched matrix-core
    products, one workgroup per element with the dense reference operators,
    and (hexahedra at p >= 2) sum factorisation over the 1-D factors.  Each
    matches the oracle, in FP64 at the FP64 bar and in FP32 at the FP32 bar."""
    print(f"case {name} p={p} {dtype} all element paths", flush=True)
    ds, dop, db, oop, ob, states = _case(name, p, dtype=dtype)
    U = ds.to_device(states)
    ref = oop.rhs(states, ob)
    if dtype == torch.float64:
        bar = _bar(lambda s: oop.rhs(s, ob), states, FP64_TOL)
    else:
        bar = FP32_TOL.get(name, 5e-5) if name != "smoke" else 5e-5
    out = {}
    dm = dop.device_mesh
    try:
        for path, mfma, sf in (("mfma", True, False), ("dense", False, False),
                               ("sum-factorised", False, True)):
            dm.set_element_path(mfma=mfma)
            dm.set_hex_sum_factorisation(sf)
            out[path] = _np(dop.rhs(U, db))
    finally:
        dm.set_element_path(None)
        dm.set_hex_sum_factorisation(None)
    for path, got in out.items():
        assert np.all(rel_l2(got, ref) <= bar), (path, rel_l2(got, ref), bar)


def test_slip_overrides_no_slip_on_shared_tags():
    """A tag listed as both no-slip and slip takes the slip ghost
    (hybrid/boundary.py:80-95: the slip reflection of the inside velocity
    replaces the no-slip ghost); with a wall temperature set."""
    ds, dop, db, oop, ob, states = _case("split", 2)
    far = np.array([1.2, 0.9, 0.1, -0.05, 1.1])
    names = dict(no_slip_names=("sphere_wall", "walls"), slip_names=("walls", "symmetry"),
                 wall_temperature=1.3)
    db = FarfieldWallBoundary(ds, far, **names)
    assert db.no_slip_tags & db.slip_tags
    ob = OracleBoundary(oop.solver, far, inlet_tags=sorted(db.inlet_tags),
                        outlet_tags=sorted(db.outlet_tags), wall_tags=sorted(db.wall_tags),
                        no_slip_tags=sorted(db.no_slip_tags), slip_tags=sorted(db.slip_tags),
                        wall_temperature=1.3)
    U = ds.to_device(states)
    got = _np(dop.viscous_residual(U, db))
    ref = oop.viscous_residual(states, ob)
    bar = _bar(lambda s: oop.viscous_residual(s, ob), states, VISCOUS_TOL)
    assert np.all(rel_l2(got, ref) <= bar), (rel_l2(got, ref), bar)


def test_xla_custom_call_target_is_rhs():
    """fsg_xla_rhs, called the way XLA calls an untyped custom-call target
    (stream, buffer pointer array, opaque string), writes the operator's rhs
    bit for bit."""
    import ctypes
    from hybrid import _native as nat
    ds, dop, db, oop, ob, states = _case("split", 2)
    U = ds.to_device(states)
    want = dop.rhs(U, db)
    outs = {t: torch.full_like(U[t], float("nan")) for t in dop.block_types}
    bufs = nat.ptr_array([U[t] for t in dop.block_types] + [outs[t] for t in dop.block_types])
    text = dop.xla_opaque(db)
    nat.lib().fsg_xla_rhs(nat.stream_handle(ds.device), bufs, text, len(text))
    torch.cuda.synchronize()
    for t in dop.block_types:
        assert torch.equal(outs[t], want[t]), t


WALE_CASES = [("split", 1), ("split", 2), ("split", 3), ("split_periodic", 2), ("tgv", 2),
              ("tgv", 3), ("species", 2), ("sliver", 2)]


@pytest.mark.parametrize("name,p", WALE_CASES)
def test_wale_viscous_term_fp64(name, p):
    """WALE eddy viscosity alone (mu = 0, hybrid/solver.py:317-406): the
    viscous residual and the total rhs match the oracle to 1e-12 (measured
    1.6e-15 .. 1e-14), the Mach 0.1 Taylor-Green field to the viscous bar
    (measured 6e-12 at p=3, the cancellation described above).  The eddy
    viscosity is the whole viscous term here, so the comparison sees it."""
    print(f"case {name} p={p} WALE", flush=True)
    ds, dop, db, oop, ob, states = _case(name, p, wale=True, mu_override=0.0)
    U = ds.to_device(states)
    ref = oop.viscous_residual(states, ob)
    assert np.all(np.array([np.abs(v).max() for v in ref.values()]) > 0)
    tol = VISCOUS_TOL if name == "tgv" else 1e-12
    got = _np(dop.viscous_residual(U, db))
    bar = _bar(lambda s: oop.viscous_residual(s, ob), states, tol)
    err = rel_l2(got, ref)
    print(f"viscous residual rel L2 {err.max():.2e} (bar {np.max(bar):.2e})", flush=True)
    assert np.all(err <= bar), (err, bar)
    got = _np(dop.rhs(U, db))
    ref = oop.rhs(states, ob)
    bar = _bar(lambda s: oop.rhs(s, ob), states, tol)
    assert np.all(rel_l2(got, ref) <= bar), (rel_l2(got, ref), bar)


def test_wale_adds_to_molecular_viscosity():
    """mu > 0 with WALE: the device rhs matches the oracle, and differs from
    the molecular-only rhs by much more than the bar (the eddy viscosity is
    not negligible on this field)."""
    ds, dop, db, oop, ob, states = _case("split", 2, wale=True)
    U = ds.to_device(states)
    ref = oop.viscous_residual(states, ob)
    got = _np(dop.viscous_residual(U, db))
    bar = _bar(lambda s: oop.viscous_residual(s, ob), states, VISCOUS_TOL)
    assert np.all(rel_l2(got, ref) <= bar), (rel_l2(got, ref), bar)
    _, dop0, db0, oop0, ob0, _ = _case("split", 2)
    plain = oop0.viscous_residual(states, ob0)
    assert rel_l2(plain, ref).max() > 1e-3


@pytest.mark.parametrize("name,p", [("split", 2), ("tgv", 3), ("species", 2), ("sliver", 2)])
def test_wale_rhs_fp32(name, p):
    print(f"case {name} p={p} WALE fp32", flush=True)
    ds, dop, db, oop, ob, states = _case(name, p, dtype=torch.float32, wale=True,
                                         mu_override=0.0)
    got = _np(dop.rhs(ds.to_device(states), db))
    ref = oop.rhs(states, ob)
    err = rel_l2(got, ref)
    print(f"rel L2 {err}", flush=True)
    assert np.all(err <= FP32_TOL[name]), err


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_wale_runs_the_vector_paths(dtype):
    """With WALE the element kernel is the vector one and the face gradients
    the dense ones whatever path is requested: asking for the matrix-core
    element kernel or sum factorisation changes no bit of the rhs."""
    ds, dop, db, oop, ob, states = _case("tgv", 3, dtype=dtype, wale=True)
    U = ds.to_device(states)
    dm = dop.device_mesh
    base = dop.rhs(U, db)
    try:
        for mfma, sf in ((True, False), (False, True), (True, True)):
            dm.set_element_path(mfma=mfma)
            dm.set_hex_sum_factorisation(sf)
            got = dop.rhs(U, db)
            for k in base:
                assert torch.equal(got[k], base[k]), (mfma, sf, k)
    finally:
        dm.set_element_path(None)
        dm.set_hex_sum_factorisation(None)


def test_wale_masked_viscous_rhs_is_bitwise():
    """viscous_rhs_masked with WALE equals viscous_rhs on every marked
    element, bit for bit."""
    ds, dop, db, oop, ob, states = _case("split", 2, wale=True)
    U = ds.to_device(states)
    full = dop.viscous_rhs(U, db)
    rng = np.random.default_rng(4)
    masks = {t: torch.as_tensor((rng.random(u.shape[0]) < 0.3).astype(np.int32),
                                device=u.device) for t, u in U.items()}
    part = dop.viscous_rhs_masked(U, masks, db)
    for t in full:
        sel = masks[t].bool()
        assert torch.equal(part[t][sel], full[t][sel]), t
