"""Benchmark of the hybrid DG residual + SSP-RK3 stage on one MI355X.

Workload (the metric's configuration that fits one GPU): Taylor-Green vortex
on a periodic [0, 2 pi]^3 box of 32^3 hexahedra, DG p=3 (64 nodes/element,
2,097,152 DOF per variable), 5 conserved variables, gamma 1.4, Pr 0.71,
Mach 0.1, Re 1600, HLLC + BR2, SSP-RK3 at CFL 0.3, no limiter, no chemistry.
Synthetic data: the analytic TGV initial condition at the solution nodes.

A timed step is one full SSP-RK3 step (three fused residual + RK-epilogue
stages) on state already resident in HBM.  The plain FP64 path and the FP32
path run the same number of steps from the same initial state with the same
dt; ``value`` is the DOF-updates/s per RK stage of the FP32 path when its
relative L2 error against the FP64 path stays within 1e-4 for density and
energy, else of the FP64 path.  ``roofline`` reports the face kernel (the
longest launch, bound by HBM latency and bandwidth) timed with HIP events on
its own stream, ``element_kernel`` the sum-factorised element kernel; ``cpu_baseline``
times the CPU FP64 oracle on a bounded sample (8^3 elements, 2 stages).
"""

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "jax-fluid-sim_amd"))

PEAK = {"f64": 78.6, "f32": 157.3}          # TFLOP/s, MI355X dense
HBM_PEAK = 8000.0                            # GB/s
GAMMA, MACH, RE, PR = 1.4, 0.1, 1600.0, 0.71


def tgv_primitive(x):
    """TGV at Mach 0.1 (rho = 1, U = 1, p0 = 1/(gamma M^2))."""
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    p = 1.0 / (GAMMA * MACH ** 2) + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2) / 16
    return np.stack([np.ones_like(X), np.sin(X) * np.cos(Y) * np.cos(Z),
                     -np.cos(X) * np.sin(Y) * np.cos(Z), np.zeros_like(X), p], -1)


def element_flops(N, Q, V):
    """Algorithmic flops of one element in the element kernel (dense operator
    formulation): interpolation 2QNV, raw viscous gradient 2QN*12 + 2Q*36,
    BR2 lift M^-1 2N^2*12 and interpolation 2QN*12, weak divergence
    2N*3Q*V, inverse mass 2N^2 V."""
    return (2 * Q * N * V + 2 * Q * N * 12 + 2 * Q * 36 + 2 * N * N * 12 + 2 * Q * N * 12
            + 2 * N * 3 * Q * V + 2 * N * N * V)


def hex_sf_flops(P, V):
    """Algorithmic flops of one hexahedron in the sum-factorised element
    kernel (viscous terms on): the 1-D passes at 2 flops per multiply-add --
    lift M^-1 (3 passes x 12 columns), interpolation x / y / z passes of the
    state, lift and viscous-variable columns, the three transposed divergence
    passes, M^-1 on the residual -- plus 370 flops per volume point for the
    pointwise Euler and viscous fluxes and the pull-back."""
    N1, Q1 = P + 1, (5 if P == 3 else 3)
    N, Q, CI = N1 ** 3, Q1 ** 3, V + 16
    return (3 * N * 12 * N1 * 2 + N1 * N1 * (CI + 4) * Q1 * N1 * 2
            + N1 * Q1 * (CI + 8) * Q1 * N1 * 2 + Q1 * Q1 * (V + 24) * Q1 * N1 * 2
            + Q1 * Q1 * 3 * V * N1 * Q1 * 2 + N1 * Q1 * V * N1 * Q1 * 6
            + N1 * N1 * V * N1 * Q1 * 4 + 3 * N * V * N1 * 2 + 370 * Q)


def face_bytes(N, NF, V, itemsize):
    """Algorithmic HBM bytes of one interior face in the face kernel: both
    elements' nodal states read, both sides' records (V surface values and
    12 lift values per support node) written."""
    return (2 * N * V + 2 * NF * (V + 12)) * itemsize


def measured_traffic(kernel, key):
    """Calibrated HBM bytes per launch from the committed PMC summary (32^3, p=3)."""
    import re
    path = ROOT / "profiles" / "hbm_traffic_tgv32_p3.md"
    if not path.exists():
        return None
    m = re.search(rf"traffic_bytes_per_launch {kernel} {key} = (\d+)", path.read_text())
    return int(m.group(1)) if m else None


def run_path(topo, dtype, steps, warmup, dt, element_path="auto"):
    import torch
    from hybrid import HybridCompressibleSolver, HybridNavierStokesOperator, HybridSSPRK3
    from hybrid import _native as nat
    solver = HybridCompressibleSolver(topo, polynomial_degree=3, gamma=GAMMA,
                                      dynamic_viscosity=1.0 / RE, prandtl=PR, dtype=dtype)
    op = HybridNavierStokesOperator(solver)
    op.device_mesh.set_element_path({"auto": None, "mfma": True, "vector": False}[element_path])
    rk = HybridSSPRK3(op, cfl=0.3)
    U0 = solver.state_blocks_from_primitive_function(tgv_primitive)
    if dt is None:
        dt = float(rk.stable_dt(U0))
    U = U0
    for _ in range(warmup):
        U, valid, _, _ = rk.checked_step(U, dt)
    torch.cuda.synchronize()
    ctx = op.device_mesh.ctx
    nat.check(nat.lib().fsg_timing(ctx, 1))
    U = {k: v.clone() for k, v in U0.items()}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        U, valid, rmin, pmin = rk.checked_step(U, dt)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    import ctypes
    gms = (ctypes.c_double * len(op.device_mesh.groups))()
    bms = (ctypes.c_double * len(op.device_mesh.block_tables))()
    gn = (ctypes.c_int64 * len(op.device_mesh.groups))()
    bn = (ctypes.c_int64 * len(op.device_mesh.block_tables))()
    nat.check(nat.lib().fsg_timing_read(ctx, ctypes.addressof(gms), ctypes.addressof(bms),
                                        ctypes.addressof(gn), ctypes.addressof(bn)))
    nat.check(nat.lib().fsg_timing(ctx, 0))
    return dict(state=U, elapsed=elapsed, dt=dt, valid=bool(valid), rho_min=float(rmin),
                p_min=float(pmin), elem_ms=bms[0] / max(bn[0], 1), elem_launches=bn[0],
                face_ms=gms[0] / max(gn[0], 1), face_launches=gn[0])


def shock_bubble_primitive(x):
    """Mach 2.5 planar shock at x = 0.1 moving into rho = 1, p = 1, u = 0
    (Rankine-Hugoniot post-shock state behind it) and a Gaussian density
    bubble rho = 1 + 4 exp(-|x - (0.4, 0.5, 0.5)|^2 / 0.15^2) in pressure
    equilibrium."""
    M, g = 2.5, GAMMA
    r2 = (g + 1) * M * M / ((g - 1) * M * M + 2)
    p2 = 1 + 2 * g / (g + 1) * (M * M - 1)
    u2 = M * np.sqrt(g) * (1 - 1 / r2)
    post = x[..., 0] < 0.1
    d2 = ((x - np.array([0.4, 0.5, 0.5])) ** 2).sum(-1)
    rho = np.where(post, r2, 1.0 + 4.0 * np.exp(-d2 / 0.15 ** 2))
    zero = np.zeros_like(rho)
    return np.stack([rho, np.where(post, u2, 0.0), zero, zero, np.where(post, p2, 1.0)], -1), \
        np.array([r2, u2, 0.0, 0.0, p2])


AIR_Y = np.zeros(11)
AIR_Y[9], AIR_Y[10] = 0.767, 0.233        # N2, O2 (species order e-, N+, ..., N2, O2)


def reacting_inflow_primitive(x, mach=3.0, seed=5, amplitude=0.01):
    """Uniform Mach 3 stream of air at 1500 K in the coupler's dimensionless
    units (rho = 1, p = 5, so T = (p / rho) T_ref with T_ref = 300 K) along
    +x, with a seeded smooth density perturbation of 1 % amplitude (eight
    random Fourier modes, wavenumbers 1-3, normalised to unit peak) at
    constant pressure; N2/O2 mass fractions 0.767/0.233.  Returns the
    (..., 16) primitive field and the inflow farfield."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 4, size=(8, 3))
    phase = rng.uniform(0, 2 * np.pi, 8)
    amp = rng.uniform(0.5, 1.0, 8)
    field = sum(a * np.sin(2 * np.pi * (x @ kk) + ph) for a, kk, ph in zip(amp, k, phase))
    rho = 1.0 + amplitude * field / amp.sum()
    p = 5.0
    u = mach * np.sqrt(GAMMA * p / 1.0)
    W = np.zeros(x.shape[:-1] + (16,))
    W[..., 0] = rho
    W[..., 1] = u
    W[..., 4] = p
    W[..., 5:] = AIR_Y
    far = np.concatenate([[1.0, u, 0.0, 0.0, p], AIR_Y])
    return W, far


def mood_rate(n=64, split=0.2, steps=5, seed=1):
    """Shock-bubble on the unit box of n^3 hexahedra with a seeded fraction
    split into wedges / tetrahedra, p=2, HLLC + BR2 (Re 1e4 on the post-shock
    speed), MOOD (sensor 0.05, WENO, 64 subcells, reference subface
    indexing), positivity floors 1e-7 / relative 0.05; inflow = post-shock
    farfield, outflow transmissive, slip walls.  Same dt for FP64 and FP32."""
    import torch
    from hybrid import (FarfieldWallBoundary, HybridCompressibleSolver, HybridMOODSubcellLimiter,
                        HybridNavierStokesOperator, HybridPositivityLimiter, HybridSSPRK3,
                        box_mesh)
    names = {"xmin": "inlet", "xmax": "outlet", "ymin": "walls", "ymax": "walls",
             "zmin": "walls", "zmax": "walls"}
    mesh = box_mesh(n, split_fraction=split, seed=seed, plane_names=names)
    out = {"workload": f"shock-bubble, unit box {n}^3 hexahedra, {int(split * 100)}% split "
                       "(seeded, wedges/tetrahedra), p=2, MOOD + positivity, Mach 2.5"}
    res, dt = {}, None
    for key, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        t0 = time.perf_counter()
        _, far = shock_bubble_primitive(np.zeros((1, 3)))
        solver = HybridCompressibleSolver(mesh, polynomial_degree=2, gamma=GAMMA,
                                          dynamic_viscosity=far[1] / 1.0e4, prandtl=PR, dtype=dtype)
        op = HybridNavierStokesOperator(solver)
        pos = HybridPositivityLimiter(op, density_floor=1e-7, pressure_floor=1e-7)
        lim = HybridMOODSubcellLimiter(op, threshold=0.05, positivity=pos, reconstruction="weno")
        rk = HybridSSPRK3(op, cfl=0.3, limiter=lim)
        bnd = FarfieldWallBoundary(solver, far)
        U0 = solver.state_blocks_from_primitive_function(lambda x: shock_bubble_primitive(x)[0])
        if dt is None:
            dt = float(rk.stable_dt(U0))
        U, valid, _, _ = rk.checked_step(U0, dt, bnd)        # warm-up (builds the MOOD tables)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        U = U0
        (U, valid, rmin, pmin), ms = _timed(lambda: _steps(rk, U0, dt, bnd, steps))
        E = sum(b.count for b in solver.mesh.blocks)
        dof = sum(v.shape[0] * v.shape[1] for v in U.values())
        res[key] = dict(state=U, ms_per_step=ms / steps,
                        dof_updates_per_s_per_stage=dof * 3 * steps / (ms * 1e-3),
                        troubled_last_stage=lim.troubled_count(), elements=E, valid=bool(valid),
                        setup_s=setup)
        out["elements"] = {b.element_type: b.count for b in solver.mesh.blocks}
        out["dof_per_var"] = dof
        print(f"[bench] MOOD {key}: {ms / steps:.3f} ms/step, troubled "
              f"{res[key]['troubled_last_stage']} / {E}, valid {bool(valid)}", file=sys.stderr,
              flush=True)
        del rk, lim, pos, op, solver
        torch.cuda.empty_cache()
    for key in ("f64", "f32"):
        r = res[key]
        out[key] = {k: r[k] for k in ("ms_per_step", "dof_updates_per_s_per_stage",
                                      "troubled_last_stage", "valid", "setup_s")}
    a, b = res["f32"]["state"], res["f64"]["state"]
    A = torch.cat([a[k].double().reshape(-1, 5) for k in b])
    B = torch.cat([b[k].reshape(-1, 5) for k in b])
    err = (torch.linalg.vector_norm(A - B, dim=0) / torch.linalg.vector_norm(B, dim=0)).cpu()
    out["fp32_rel_l2_vs_fp64"] = dict(zip(["rho", "rho_u", "rho_v", "rho_w", "rho_E"],
                                          map(float, err)))
    out["steps"], out["dt"] = steps, dt
    return out


def reacting_rate(n=86, split=0.2, steps=5, seed=1, train_cells=1 << 15, train_traj=24,
                  train_steps=3000):
    """Reacting hybrid-mesh flow: the shock-bubble mesh family scaled up (unit
    box of n^3 hexahedra, a seeded 20 % split into wedges / tetrahedra; n = 86
    gives about 1.0M elements), p=2, 5 + 11 variables, Mach 3 air at 1500 K
    with a 1 % density perturbation (Re 1e4 on the inflow speed), MOOD +
    positivity active, and the Park-11 coupler after every accepted SSP-RK3
    step (operator splitting as in run_hybrid_sphere.py:280-312).
    Plain path: FP64 residual, FP64 point-implicit chemistry (12 Newton
    iterations, temperature re-solved inside).  Fast path: FP32 residual and
    the BF16-MFMA surrogate of the same substep with its integrator fallback,
    the surrogate trained here on integrator trajectories started from the
    flow's own cells.  Same dt for both; a timed step is the RK step and the
    chemistry."""
    import torch
    from hybrid import (FarfieldWallBoundary, HybridCompressibleSolver, HybridMOODSubcellLimiter,
                        HybridNavierStokesOperator, HybridPositivityLimiter, HybridSSPRK3,
                        box_mesh)
    from hybrid import _native as nat
    from hybrid.chemistry import Air11ParkMechanism, ReactingFlowCoupler
    from hybrid.chemistry.air11_park import SPECIES_NAMES
    from hybrid.chemistry.surrogate import ChemistrySurrogate
    from hybrid.topology import Topology
    import ctypes
    names = {"xmin": "inlet", "xmax": "outlet", "ymin": "walls", "ymax": "walls",
             "zmin": "walls", "zmax": "walls"}
    t0 = time.perf_counter()
    mesh = Topology(box_mesh(n, split_fraction=split, seed=seed, plane_names=names))
    _, far = reacting_inflow_primitive(np.zeros((1, 3)))
    mech = Air11ParkMechanism()
    out = {"workload": f"reacting air, unit box {n}^3 hexahedra, {int(split * 100)}% split "
                       "(seeded, wedges/tetrahedra), p=2, 16 variables, Mach 3 at 1500 K, "
                       "MOOD + positivity, Park-11 chemistry after every step",
           "mesh_s": time.perf_counter() - t0}
    res, dt = {}, None
    for key, dtype, use_sur in (("f64_plain", torch.float64, False),
                                ("f32_surrogate", torch.float32, True)):
        t0 = time.perf_counter()
        solver = HybridCompressibleSolver(mesh, species_names=SPECIES_NAMES, polynomial_degree=2,
                                          gamma=GAMMA, dynamic_viscosity=far[1] / 1.0e4,
                                          prandtl=PR, dtype=dtype)
        op = HybridNavierStokesOperator(solver)
        pos = HybridPositivityLimiter(op, density_floor=1e-7, pressure_floor=1e-7)
        lim = HybridMOODSubcellLimiter(op, threshold=0.05, positivity=pos, reconstruction="weno")
        rk = HybridSSPRK3(op, cfl=0.3, limiter=lim)
        bnd = FarfieldWallBoundary(solver, far)
        U0 = solver.state_blocks_from_primitive_function(
            lambda x: reacting_inflow_primitive(x)[0])
        if dt is None:
            dt = float(rk.stable_dt(U0))
        coupler = ReactingFlowCoupler(mech, max_newton_iters=12, dimensionless=True, gamma=GAMMA)
        entry = {}
        if use_sur:
            # training data: integrator trajectories from a sample of the
            # flow's cells, temperature and density spread by +-10 % / +-5 %
            ts = time.perf_counter()
            Ud = torch.cat([v.double().reshape(-1, 16) for v in U0.values()])
            g = torch.Generator(device="cuda").manual_seed(seed)
            pick = torch.randint(0, Ud.shape[0], (train_cells,), device="cuda", generator=g)
            rho_s = torch.empty(train_cells, 11, dtype=torch.float64, device="cuda")
            T0 = torch.empty(train_cells, dtype=torch.float64, device="cuda")
            p = coupler._params()
            nat.check_chem(nat.lib().fsg_chem_prepare(
                nat.ptr(mech.table), nat.ptr(Ud[pick].contiguous()), nat.F64, train_cells,
                ctypes.byref(p), nat.ptr(rho_s), nat.ptr(T0), nat.stream_handle(Ud.device)))
            u = lambda a: 1 + a * (2 * torch.rand(train_cells, device="cuda", generator=g,
                                                  dtype=torch.float64) - 1)
            T0 = T0 * u(0.10)
            rho_s = rho_s * u(0.05)[:, None]
            sur = ChemistrySurrogate(coupler.integrator, constant_temperature=False, seed=seed)
            X, O, _ = sur.dataset(rho_s, T0, dt * coupler.t_ref, trajectory_steps=train_traj)
            sur.set_normalisation(X, O)
            hist = sur.train(X, O, steps=train_steps, batch=4096)
            torch.cuda.synchronize()
            coupler.surrogate = sur
            entry["surrogate_train"] = dict(samples=int(X.shape[0]), steps=train_steps,
                                            loss_first_last=[hist[0], hist[-1]],
                                            seconds=time.perf_counter() - ts)
            del X, O
        U, valid, _, _ = rk.checked_step(U0, dt, bnd)        # warm-up (builds the MOOD tables)
        U = coupler.advance(U, dt)
        torch.cuda.synchronize()
        entry["setup_s"] = time.perf_counter() - t0
        chem_ev = []
        U = U0
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        rejected = fallback = 0
        for _ in range(steps):
            cand, valid, rmin, pmin = rk.checked_step(U, dt, bnd)
            if not bool(valid):
                rejected += 1
                continue
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            U = coupler.advance(cand, dt)
            b.record()
            chem_ev.append((a, b))
            if use_sur:
                fallback += int(coupler.last_fallback.sum())
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1)
        chem_ms = sum(a.elapsed_time(b) for a, b in chem_ev)
        dof = sum(v.shape[0] * v.shape[1] for v in U.values())
        E = sum(b.count for b in solver.mesh.blocks)
        entry.update(ms_per_step=ms / steps, chemistry_ms_per_step=chem_ms / steps,
                     dof_updates_per_s_per_stage=dof * 3 * steps / (ms * 1e-3),
                     rejected_steps=rejected, troubled_last_stage=lim.troubled_count())
        if use_sur:
            entry["fallback_fraction"] = fallback / max(dof * (steps - rejected), 1)
        res[key] = (entry, U)
        out["elements"] = {b.element_type: b.count for b in solver.mesh.blocks}
        out["elements_total"], out["dof_per_var"] = E, dof
        print(f"[bench] reacting {key}: {ms / steps:.2f} ms/step (chemistry "
              f"{chem_ms / steps:.2f}), {E} elements, troubled {entry['troubled_last_stage']}",
              file=sys.stderr, flush=True)
        del rk, lim, pos, op, solver, coupler
        torch.cuda.empty_cache()
    for key in res:
        out[key] = res[key][0]
    a, b = res["f32_surrogate"][1], res["f64_plain"][1]
    A = torch.cat([a[k].double().reshape(-1, 16) for k in b])
    B = torch.cat([b[k].reshape(-1, 16) for k in b])
    diff = torch.linalg.vector_norm(A - B, dim=0)
    norm = torch.linalg.vector_norm(B, dim=0)
    labels = ["rho", "rho_u", "rho_v", "rho_w", "rho_E"] + [f"rho_{s}" for s in SPECIES_NAMES]
    # relative L2 per variable for the variables with O(1) norms; v, w and the
    # species are measured against the density norm (their own norms are
    # round-off or tiny in this stream)
    out["fast_rel_l2_vs_plain_fp64"] = {labels[i]: float(diff[i] / norm[i]) for i in (0, 1, 4)}
    out["fast_l2_over_rho_l2"] = {k: float(v) for k, v in zip(labels, (diff / norm[0]).cpu())}
    out["species_rel_l2"] = {f"rho_{s}": float(diff[5 + i] / norm[5 + i])
                             for i, s in enumerate(SPECIES_NAMES) if s in ("N2", "O2", "O", "NO")}
    out["speedup_fast_vs_plain"] = (out["f64_plain"]["ms_per_step"]
                                    / out["f32_surrogate"]["ms_per_step"])
    out["steps"], out["dt"] = steps, dt
    return out


def _steps(rk, U, dt, bnd, steps):
    for _ in range(steps):
        U, valid, rmin, pmin = rk.checked_step(U, dt, bnd)
    return U, valid, rmin, pmin


def _timed(fn):
    import torch
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    out = fn()
    ev1.record()
    torch.cuda.synchronize()
    return out, ev0.elapsed_time(ev1)


def chemistry_rate(cells, seed=11, train_cells=1 << 18, train_steps=3000):
    """0-D air-11 batch: 2^20 cells, T in [3000, 12000] K, p log-uniform in
    [1e3, 1e5] Pa, Dirichlet(2) mass fractions made charge-neutral, dt 1e-7 s.
    Plain FP64 point-implicit step (12 Newton iterations, T recovered at
    constant internal energy) and the BF16-MFMA surrogate (trained here on
    2^18 integrator results from another seed) with its integrator fallback."""
    import torch
    from hybrid.chemistry import Air11ParkMechanism, PointImplicitIntegrator
    from hybrid.chemistry.batch import make_batch
    from hybrid.chemistry.surrogate import ChemistrySurrogate
    rho, T = make_batch(cells, seed=seed)
    mech = Air11ParkMechanism()
    integ = PointImplicitIntegrator(mech, max_newton_iters=12)
    r = torch.as_tensor(rho, device="cuda")
    t = torch.as_tensor(T, device="cuda")
    integ.step(r[:4096], t[:4096], 1e-7, constant_temperature=False)
    ref, ms = _timed(lambda: integ.step(r, t, 1e-7, constant_temperature=False)[0])
    ok = bool(torch.isfinite(ref).all()) and bool((ref >= 0).all())
    out = {"cells": cells, "plain_fp64_ms": ms,
           "plain_fp64_cell_updates_per_s": cells / (ms * 1e-3), "finite_nonnegative": ok}
    print(f"[bench] chemistry plain FP64: {ms:.3f} ms", file=sys.stderr, flush=True)

    # Surrogate of the isothermal substep: data from the integrator itself
    # (initial states and one step along their trajectories), FP32 training
    # kernels, BF16 MFMA inference.
    (ref_ct, ms_ct) = _timed(lambda: integ.step(r, t, 1e-7, constant_temperature=True)[0])
    out["plain_fp64_isothermal_ms"] = ms_ct
    rho_tr, T_tr = make_batch(train_cells, seed=seed + 1)
    sur = ChemistrySurrogate(integ, constant_temperature=True, seed=0)
    t0 = time.perf_counter()
    X, O, _ = sur.dataset(rho_tr, T_tr, 1e-7, trajectory_steps=2)
    sur.set_normalisation(X, O)
    hist = sur.train(X, O, steps=train_steps, batch=4096)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    print(f"[bench] surrogate trained in {train_s:.1f} s, loss {hist[0]:.4f} -> {hist[-1]:.4f}",
          file=sys.stderr, flush=True)
    sur.infer(r[:4096], t[:4096], 1e-7)
    sur.step(r[:4096], t[:4096], 1e-7)
    (pred, fb), ms_inf = _timed(lambda: sur.infer(r, t, 1e-7))
    (full, _, _), ms_step = _timed(lambda: sur.step(r, t, 1e-7))
    Yref = ref_ct / ref_ct.sum(1, keepdim=True)
    l1 = lambda x: float((x / x.sum(1, keepdim=True) - Yref).abs().sum(1).mean())
    # Sensitivity of the coupled-temperature integrator to ~1 ulp of input noise.
    pert = r * (1 + 2e-16 * torch.randn(r.shape, generator=torch.Generator(device="cuda").manual_seed(1),
                                        device="cuda", dtype=torch.float64))
    ref2 = integ.step(pert, t, 1e-7, constant_temperature=False)[0]
    Yc = ref / ref.sum(1, keepdim=True)
    l1_c = lambda x: float((x / x.sum(1, keepdim=True) - Yc).abs().sum(1).mean())
    out["surrogate"] = {
        "train_cells": train_cells, "train_steps": train_steps, "train_s": train_s,
        "train_loss_first_last": [hist[0], hist[-1]],
        "infer_ms": ms_inf, "infer_cell_updates_per_s": cells / (ms_inf * 1e-3),
        "step_with_fallback_ms": ms_step,
        "step_cell_updates_per_s": cells / (ms_step * 1e-3),
        "speedup_vs_plain_fp64_isothermal": ms_ct / ms_step,
        "fallback_fraction": float(fb.float().mean()),
        "mean_L1_massfraction_error_vs_integrator": l1(full),
        "mean_L1_identity_predictor": l1(r),
        "target": "isothermal point-implicit substep (12 Newton iterations), dt 1e-7 s",
        "coupled_T_integrator_1ulp_sensitivity_mean_L1": l1_c(ref2),
    }
    print(f"[bench] surrogate: infer {ms_inf:.3f} ms, step+fallback {ms_step:.3f} ms, "
          f"fallback {out['surrogate']['fallback_fraction']:.4f}, L1 {out['surrogate']['mean_L1_massfraction_error_vs_integrator']:.4f}",
          file=sys.stderr, flush=True)
    return out


def cpu_baseline():
    """CPU FP64 oracle, one thread, on 8^3 elements for 2 RK stages."""
    code = r"""
import sys, time, json
sys.path.insert(0, %r); sys.path.insert(0, %r)
import numpy as np
from bench import tgv_primitive, GAMMA, RE, PR
from hybrid.meshgen import box_mesh
from oracle.mesh import from_arrays
from oracle.solver import CompressibleSolver
from oracle.operators import NavierStokesOperator
from oracle.time_integration import SSPRK3
m = box_mesh(8, lengths=(2 * np.pi,) * 3, periodic=(True,) * 3)
s = CompressibleSolver(gamma=GAMMA, dynamic_viscosity=1.0 / RE, prandtl=PR)
op = NavierStokesOperator(from_arrays(m), s, 3)
st = {k: s.primitive_to_conserved(tgv_primitive(x)) for k, x in op.geometry.solution_node_coordinates.items()}
rk = SSPRK3(op, cfl=0.3)
dt = rk.stable_dt(st)
t0 = time.perf_counter()
s1 = rk.rk_stage(st, st, dt, 0.0, 1.0)[0]
s2 = rk.rk_stage(st, s1, dt, 0.75, 0.25)[0]
t = (time.perf_counter() - t0) / 2
print(json.dumps({"dof": 512 * 64, "stage_s": t}))
""" % (str(ROOT), str(ROOT / "jax-fluid-sim_amd"))
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       cwd=str(ROOT), timeout=600)
    if r.returncode != 0:
        return {"value": None, "error": r.stderr[-400:]}
    d = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": d["dof"] / d["stage_s"], "unit": "DOF-updates/s per RK stage",
            "cores": 1, "kind": "port",
            "sample": "CPU FP64 oracle (numpy), TGV 8^3 hexahedra p=3 (32,768 DOF/var), "
                      "2 SSP-RK3 stages, 1 thread; per-DOF rate"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=32, help="elements per box edge")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-chem", action="store_true")
    ap.add_argument("--skip-mood", action="store_true")
    ap.add_argument("--only", choices=["f64", "f32", "mood", "chem", "reacting"], default=None,
                    help="run one path of the residual benchmark, or only the MOOD or "
                         "chemistry section (for profiling; prints that section's JSON)")
    ap.add_argument("--mood-steps", type=int, default=5)
    ap.add_argument("--skip-reacting", action="store_true")
    ap.add_argument("--reacting-steps", type=int, default=5)
    ap.add_argument("--reacting-n", type=int, default=86)
    ap.add_argument("--element-path", choices=["auto", "mfma", "vector"], default="auto",
                    help="element operators on the matrix cores or the vector units "
                         "(auto: the library default, vector units)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is a design, not a build)", file=sys.stderr)
        sys.exit(2)
    if args.only == "mood":
        print(json.dumps(mood_rate(steps=args.mood_steps)))
        return
    if args.only == "reacting":
        print(json.dumps(reacting_rate(n=args.reacting_n, steps=args.reacting_steps)))
        return
    if args.only == "chem":
        print(json.dumps(chemistry_rate(1 << 20)))
        return

    import torch
    from hybrid import Topology, box_mesh
    topo = Topology(box_mesh(args.n, lengths=(2 * np.pi,) * 3, periodic=(True,) * 3))
    E = topo.blocks[0].count
    N, Q, V = 64, 125, 5
    dof = E * N
    print(f"[bench] TGV {args.n}^3 p=3: {E} elements, {dof} DOF/var", file=sys.stderr, flush=True)

    res = {}
    dt = None
    for name, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        if args.only and args.only != name:
            continue
        r = run_path(topo, dtype, args.steps, args.warmup, dt, args.element_path)
        dt = r["dt"]
        r["rate"] = dof * 3 * args.steps / r["elapsed"]
        res[name] = r
        print(f"[bench] {name}: {r['elapsed'] * 1e3 / args.steps:.3f} ms/step, "
              f"{r['rate']:.4e} DOF/s/stage, element {r['elem_ms']:.3f} ms, "
              f"face {r['face_ms']:.3f} ms, valid={r['valid']}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()

    err = None
    if "f64" in res and "f32" in res:
        a = res["f32"]["state"]["hexahedron"].double()
        b = res["f64"]["state"]["hexahedron"]
        err = (torch.linalg.vector_norm((a - b).reshape(-1, V), dim=0)
               / torch.linalg.vector_norm(b.reshape(-1, V), dim=0)).cpu().numpy()
    within = err is not None and err[0] <= 1e-4 and err[4] <= 1e-4
    key = "f32" if within else ("f64" if "f64" in res else "f32")
    r = res[key]
    nfaces = 3 * E
    itemsize = 4 if key == "f32" else 8
    fbytes = face_bytes(N, 16, V, itemsize) * nfaces
    face_gbs = fbytes / (r["face_ms"] * 1e-3) / 1e9
    eflops = hex_sf_flops(3, V) * E
    elem_tf = eflops / (r["elem_ms"] * 1e-3) / 1e12
    traffic = measured_traffic("face", key) if args.n == 32 else None
    out = {
        "metric": "DOF-updates/s per RK stage, plain FP64 vs FP32+surrogate chem, at fixed L2 error",
        "value": r["rate"], "unit": "DOF-updates/s per RK stage", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["elapsed"] * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": key, "data": "synthetic",
        "config": {"workload": f"Taylor-Green vortex, periodic {args.n}^3 hexahedra, DG p=3, "
                               "HLLC + BR2, SSP-RK3, CFL 0.3, Mach 0.1, Re 1600, Pr 0.71",
                   "elements": E, "dof_per_var": dof, "dt": dt,
                   "dt_clipped": bool(dt >= 1e-3 or dt <= 1e-7)},
        "fp64_dof_updates_per_s": res.get("f64", {}).get("rate"),
        "fp32_dof_updates_per_s": res.get("f32", {}).get("rate"),
        "fp32_rel_l2_vs_fp64": None if err is None else
            dict(zip(["rho", "rho_u", "rho_v", "rho_w", "rho_E"], map(float, err))),
        "roofline": {"kernel": f"face_kernel<quad,hex,hex,p3,V5,{key}>", "bound": "hbm",
                     "achieved": face_gbs, "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": face_gbs / HBM_PEAK, "traffic": traffic,
                     "bytes_per_launch": fbytes, "faces": nfaces, "avg_launch_ms": r["face_ms"]},
        "element_kernel": {"kernel": f"hex_sf_kernel<p3,V5,{key}> (sum factorisation)",
                           "avg_launch_ms": r["elem_ms"], "flops_per_launch": eflops,
                           "achieved_tflops": elem_tf, "peak_tflops": PEAK[key],
                           "frac": elem_tf / PEAK[key],
                           "traffic": measured_traffic("element", key) if args.n == 32 else None,
                           "dense_operator_flops_per_launch": element_flops(N, Q, V) * E},
    }
    if not args.skip_chem:
        out["chemistry"] = chemistry_rate(1 << 20)
    if not args.skip_mood:
        out["mood"] = mood_rate(steps=args.mood_steps)
    if not args.skip_reacting:
        out["reacting"] = reacting_rate(n=args.reacting_n, steps=args.reacting_steps)
    if not args.skip_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
