"""Benchmark of the hybrid DG hot path on one MI355X.

Headline workload (the metric's reacting configuration, which fits one
GPU): the reacting hybrid-mesh flow -- unit box of 86^3 hexahedra with a
seeded 20 % split into wedges and tetrahedra (1,016,939 elements,
19.8M nodes), DG p=2, 16 variables (5 flow + 11 species), Mach 3 air at
1500 K with a seeded 1 % density perturbation, HLLC + BR2 at Re 1e4, MOOD +
positivity, SSP-RK3 at CFL 0.3, and the Park-11 coupler after every accepted
step.  Synthetic data: the analytic inflow state at the solution nodes.

A timed step is one full SSP-RK3 step (three MOOD stages) plus the chemistry
substep, on state already resident in HBM.  The plain path (FP64 residual,
FP64 point-implicit chemistry) and the fast path (FP32 residual, BF16-MFMA
chemistry surrogate with the FP64 integrator as fallback) run the same steps
with the same dt; ``value`` is the fast path's DOF-updates/s per RK stage,
with the plain path's rate, the speed-up and the fast path's relative L2
error against the plain FP64 path after ``--error-steps`` steps beside it.
``roofline`` is the fast path's longest-running kernel over the timed
region (HIP events around each launch on its stream; full residual passes
only), with its algorithmic work per launch.  ``cpu_baseline`` times the
CPU FP64 oracle on a bounded sample of the same workload.  Extra keys: the
Taylor-Green vortex (32^3, p=3), the 0-D chemistry batch (2^20 cells) and
the shock-bubble MOOD workload (64^3, 20 % split).
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


def measured_traffic_reacting(kernel):
    """Calibrated HBM bytes per launch of one kernel of the reacting workload
    from the committed PMC summary (FETCH_SIZE / WRITE_SIZE runs), or None."""
    import re
    path = ROOT / "profiles" / "hbm_traffic_reacting_air_86.md"
    if not path.exists():
        return None
    m = re.search(rf"traffic_bytes_per_launch {re.escape(kernel)} = (\d+)", path.read_text())
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


def _timing_start(op, mode=2):
    from hybrid import _native as nat
    nat.check(nat.lib().fsg_timing(op.device_mesh.ctx, mode))


def _timing_read(op):
    """Summed ms and launch counts per face group and element block since
    _timing_start (fsg_timing / fsg_timing_read: HIP events on the launch
    stream around each kernel launch)."""
    import ctypes
    from hybrid import _native as nat
    dm = op.device_mesh
    gms = (ctypes.c_double * max(len(dm.groups), 1))()
    bms = (ctypes.c_double * len(dm.block_tables))()
    gn = (ctypes.c_int64 * max(len(dm.groups), 1))()
    bn = (ctypes.c_int64 * len(dm.block_tables))()
    nat.check(nat.lib().fsg_timing_read(dm.ctx, ctypes.addressof(gms), ctypes.addressof(bms),
                                        ctypes.addressof(gn), ctypes.addressof(bn)))
    nat.check(nat.lib().fsg_timing(dm.ctx, 0))
    return list(gms)[:len(dm.groups)], list(gn)[:len(dm.groups)], list(bms), list(bn)


TYPE_SHORT = {"hexahedron": "hex", "wedge": "wedge", "tetra": "tet"}


def kernel_table(op, p, V, itemsize, timing):
    """Per timed kernel (face group or element block): name, launches, average
    ms, and the algorithmic work per launch -- HBM bytes for face kernels
    (per face: read 2 N_f V B of traces + 4 Q_f B of geometry, write
    2 N_f V B; one side on boundary faces) and flops for element kernels
    (the sum-factorised count for hexahedra at p >= 2, the dense-operator
    count otherwise)."""
    from hybrid.elements import face_quadrature, reference_tables
    gms, gn, bms, bn = timing
    dm = op.device_mesh
    prec = "f64" if itemsize == 8 else "f32"
    rows = []
    for g, ms, n in zip(dm.groups, gms, gn):
        if not n:
            continue
        to = TYPE_SHORT[dm.topo.blocks[g["owner_block"]].element_type]
        tn = ("boundary" if g["neighbor_block"] < 0
              else TYPE_SHORT[dm.topo.blocks[g["neighbor_block"]].element_type])
        ft = ("quad", "tri")[g["face_type"]]
        qf = len(face_quadrature(("quad", "triangle")[g["face_type"]], p)[1])
        per_face = 2 * (g["FO"] + g["FN"]) * V * itemsize + 4 * qf * itemsize
        rows.append(dict(kernel=f"face_kernel<{ft},{to},{tn}{',sub' if g['sub'] else ''},"
                                f"p{p},V{V},{prec}>", kind="face", launches=int(n),
                         total_ms=ms, avg_ms=ms / n, count=int(g["count"]),
                         bytes_per_launch=int(per_face * g["count"])))
    for b, t, ms, n in zip(dm.topo.blocks, dm.block_tables, bms, bn):
        if not n:
            continue
        R = reference_tables(b.element_type, p)
        sf = b.element_type == "hexahedron" and p >= 2
        fl = hex_sf_flops(p, V) if sf else element_flops(R["N"], R["Q"], V)
        name = (f"hex_sf_kernel<p{p},V{V},{prec}>" if sf else
                f"element_kernel<{TYPE_SHORT[b.element_type]},p{p},V{V},{prec}>")
        rows.append(dict(kernel=name, kind="element", launches=int(n), total_ms=ms,
                         avg_ms=ms / n, count=int(b.count), flops_per_launch=int(fl * b.count)))
    return sorted(rows, key=lambda r: -r["total_ms"])


def roofline_of(row, prec, traffic=None):
    """The roofline object of one kernel_table row (face: HBM bytes; element:
    FP32/FP64 flops against the dense matrix-core peak)."""
    if row["kind"] == "face":
        gbs = row["bytes_per_launch"] / (row["avg_ms"] * 1e-3) / 1e9
        return {"kernel": row["kernel"], "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK,
                "unit": "GB/s", "frac": gbs / HBM_PEAK, "traffic": traffic,
                "bytes_per_launch": row["bytes_per_launch"], "faces": row["count"],
                "avg_launch_ms": row["avg_ms"], "launches": row["launches"]}
    tf = row["flops_per_launch"] / (row["avg_ms"] * 1e-3) / 1e12
    return {"kernel": row["kernel"], "bound": "mfma", "achieved": tf, "peak": PEAK[prec],
            "unit": "TFLOP/s", "frac": tf / PEAK[prec], "traffic": traffic,
            "flops_per_launch": row["flops_per_launch"], "elements": row["count"],
            "avg_launch_ms": row["avg_ms"], "launches": row["launches"]}


def reacting_rate(n=86, split=0.2, steps=20, warmup=2, error_steps=100, seed=1,
                  train_cells=1 << 15, train_traj=24, train_steps=3000,
                  paths=("f64_plain", "f32_surrogate")):
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
    flow's own cells.  Same dt for both.  A timed step is the RK step and the
    chemistry; `steps` steps are timed after `warmup` untimed ones, both
    paths then continue untimed to `error_steps` steps for the L2 error of
    the fast path against the plain FP64 path.  ``paths`` may name one path
    only (profiling passes): then no comparison is made."""
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
    error_steps = max(error_steps, steps)
    for key, dtype, use_sur in (("f64_plain", torch.float64, False),
                                ("f32_surrogate", torch.float32, True)):
        if key not in paths:
            continue
        print(f"[bench] reacting {key}: setting up", file=sys.stderr, flush=True)
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

        fallback = torch.zeros((), dtype=torch.int64, device="cuda")
        chem_ev = []
        rejected = [0]

        def step(U, record=False):
            cand, valid, _, _ = rk.checked_step(U, dt, bnd)
            if not bool(valid):                  # the reference's one host sync per step
                rejected[0] += 1
                return U
            a = torch.cuda.Event(enable_timing=True) if record else None
            b = torch.cuda.Event(enable_timing=True) if record else None
            if record:
                a.record()
            nxt = {}
            for t, v in cand.items():
                nxt[t] = coupler.react_block(v, dt)
                if use_sur and coupler.last_fallback is not None:
                    fallback.add_(coupler.last_fallback.sum())
            if record:
                b.record()
                chem_ev.append((a, b))
            return nxt

        U = U0
        for _ in range(warmup):                  # builds the MOOD tables, warms the caches
            U = step(U)
        torch.cuda.synchronize()
        entry["setup_s"] = time.perf_counter() - t0
        fallback.zero_()
        rejected[0] = 0
        _timing_start(op, 2)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        U = U0
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        ev0.record()
        for _ in range(steps):
            U = step(U, record=True)
        ev1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        ms = ev0.elapsed_time(ev1)
        timing = _timing_read(op)
        chem_ms = sum(a.elapsed_time(b) for a, b in chem_ev)
        dof = sum(v.shape[0] * v.shape[1] for v in U.values())
        E = sum(b.count for b in solver.mesh.blocks)
        itemsize = 8 if dtype == torch.float64 else 4
        ktab = kernel_table(op, 2, 16, itemsize, timing)
        entry.update(ms_per_step=ms / steps, wall_ms_per_step=wall * 1e3 / steps,
                     chemistry_ms_per_step=chem_ms / steps,
                     dof_updates_per_s_per_stage=dof * 3 * steps / (ms * 1e-3),
                     rejected_steps=rejected[0], troubled_last_stage=lim.troubled_count(),
                     kernels=ktab[:6])
        if use_sur:
            entry["fallback_fraction"] = int(fallback) / max(dof * (steps - rejected[0]), 1)
        at_timed = {k: v.clone() for k, v in U.items()}
        for _ in range(error_steps - steps):
            U = step(U)
        torch.cuda.synchronize()
        res[key] = (entry, at_timed, U)
        out["elements"] = {b.element_type: b.count for b in solver.mesh.blocks}
        out["elements_total"], out["dof_per_var"] = E, dof
        print(f"[bench] reacting {key}: {ms / steps:.2f} ms/step (chemistry "
              f"{chem_ms / steps:.2f}), {E} elements, troubled {entry['troubled_last_stage']}, "
              f"top kernel {ktab[0]['kernel']} {ktab[0]['avg_ms']:.3f} ms", file=sys.stderr,
              flush=True)
        del rk, lim, pos, op, solver, coupler
        torch.cuda.empty_cache()
    for key in res:
        out[key] = res[key][0]
    out["steps"], out["warmup"], out["error_steps"], out["dt"] = steps, warmup, error_steps, dt
    if len(res) < 2:
        return out
    labels = ["rho", "rho_u", "rho_v", "rho_w", "rho_E"] + [f"rho_{s}" for s in SPECIES_NAMES]

    def errors(a, b):
        A = torch.cat([a[k].double().reshape(-1, 16) for k in b])
        B = torch.cat([b[k].reshape(-1, 16) for k in b])
        diff = torch.linalg.vector_norm(A - B, dim=0)
        norm = torch.linalg.vector_norm(B, dim=0)
        # relative L2 per variable for the flow variables with O(1) norms and
        # the species; v and w are measured against the rho u norm (their
        # own norms are the 1 % perturbation's)
        d = {labels[i]: float(diff[i] / norm[i]) for i in (0, 1, 4)}
        d["rho_v"], d["rho_w"] = float(diff[2] / norm[1]), float(diff[3] / norm[1])
        sp = {f"rho_{s}": float(diff[5 + i] / norm[5 + i])
              for i, s in enumerate(SPECIES_NAMES) if s in ("N2", "O2", "O", "NO", "N")}
        # RMS error of the mass fractions Y_s = rho_s / rho (trace species
        # whose values sit far below 1e-12 make the relative L2 above large)
        ym = (A[:, 5:] / A[:, :1] - B[:, 5:] / B[:, :1]).pow(2).mean(0).sqrt().cpu()
        sp["mass_fraction_rms"] = {s: float(ym[i]) for i, s in enumerate(SPECIES_NAMES)
                                   if s in ("N2", "O2", "O", "NO", "N")}
        return d, sp

    out["fast_rel_l2_vs_plain_fp64"], out["species_rel_l2"] = errors(
        res["f32_surrogate"][1], res["f64_plain"][1])
    out["fast_rel_l2_vs_plain_fp64_at_error_steps"], out["species_rel_l2_at_error_steps"] = \
        errors(res["f32_surrogate"][2], res["f64_plain"][2])
    out["speedup_fast_vs_plain"] = (out["f64_plain"]["ms_per_step"]
                                    / out["f32_surrogate"]["ms_per_step"])
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
    (full, _, fb_step), ms_step = _timed(lambda: sur.step(r, t, 1e-7))
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
        # cells the step hands to the integrator: the inference's own checks
        # (hull, projection residual, finiteness) or the integrator-judged
        # error check of the prediction
        "fallback_fraction": float(fb_step.float().mean()),
        "inference_flag_fraction": float(fb.float().mean()),
        "mean_L1_massfraction_error_vs_integrator": l1(full),
        "mean_L1_prediction_error_vs_integrator": l1(pred),
        "mean_L1_identity_predictor": l1(r),
        "target": "isothermal point-implicit substep (12 Newton iterations), dt 1e-7 s",
        "coupled_T_integrator_1ulp_sensitivity_mean_L1": l1_c(ref2),
    }
    print(f"[bench] surrogate: infer {ms_inf:.3f} ms, step+fallback {ms_step:.3f} ms, "
          f"fallback {out['surrogate']['fallback_fraction']:.4f}, L1 {out['surrogate']['mean_L1_massfraction_error_vs_integrator']:.4f}",
          file=sys.stderr, flush=True)
    return out


def cpu_baseline():
    """CPU FP64 oracle, one thread, on the Taylor-Green vortex (8^3 elements,
    p=3) for 2 RK stages."""
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


REACTING_CPU_SAMPLE = r"""
import sys, time, json
sys.path.insert(0, %r); sys.path.insert(0, %r)
import numpy as np
from bench import reacting_inflow_primitive, GAMMA, PR
from hybrid.meshgen import box_mesh
from oracle.mesh import from_arrays
from oracle.solver import CompressibleSolver, FarfieldWallBoundary
from oracle.operators import NavierStokesOperator
from oracle.limiter import MOODSubcellLimiter, PositivityLimiter
from oracle.time_integration import SSPRK3
from oracle.chemistry import ReactingFlowCoupler
names = {"xmin": "inlet", "xmax": "outlet", "ymin": "walls", "ymax": "walls",
         "zmin": "walls", "zmax": "walls"}
m = box_mesh(3, split_fraction=0.2, seed=1, plane_names=names)
_, far = reacting_inflow_primitive(np.zeros((1, 3)))
s = CompressibleSolver(gamma=GAMMA, dynamic_viscosity=far[1] / 1e4, prandtl=PR, num_species=11)
op = NavierStokesOperator(from_arrays(m), s, 2)
lim = MOODSubcellLimiter(op, positivity=PositivityLimiter(op, density_floor=1e-7,
                                                          pressure_floor=1e-7))
pn = m.physical_names
ob = FarfieldWallBoundary(s, far, inlet_tags=[pn["inlet"]], outlet_tags=[pn["outlet"]],
                          wall_tags=[pn["walls"]], slip_tags=[pn["walls"]])
st = {t: s.primitive_to_conserved(reacting_inflow_primitive(x)[0])
      for t, x in op.geometry.solution_node_coordinates.items()}
rk = SSPRK3(op, cfl=0.3, limiter=lim)
dt = rk.stable_dt(st)
cp = ReactingFlowCoupler(max_newton_iters=12, dimensionless=True, gamma=GAMMA)
t0 = time.perf_counter()
u, ok, _, _ = rk.checked_step(st, dt, ob)
t1 = time.perf_counter()
flat = np.concatenate([v.reshape(-1, v.shape[-1]) for v in u.values()])
nodes = %d
cp.react_block(flat[None, :nodes], dt)
t2 = time.perf_counter()
print(json.dumps({"dof": int(flat.shape[0]), "flow_s": t1 - t0, "chem_s": t2 - t1,
                  "chem_nodes": nodes}))
"""


def cpu_baseline_reacting(nodes=160):
    """The CPU FP64 oracle on a bounded sample of the reacting workload: one
    SSP-RK3 step with MOOD + positivity on a 3^3 box with the same 20 % split
    (22 hexahedra + 30 tetrahedra, 894 nodes, 16 variables), and the Park-11
    coupler on `nodes` of its nodes, its time extrapolated per node to all
    of them.  Run once alone on one thread, then as one single-thread copy
    per available core at once (all-core throughput: the sum of the copies'
    rates, each measured while all run)."""
    code = REACTING_CPU_SAMPLE % (str(ROOT), str(ROOT / "jax-fluid-sim_amd"), nodes)
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")

    def launch(k):
        return [subprocess.Popen([sys.executable, "-c", code], stdout=subprocess.PIPE,
                                 stderr=subprocess.PIPE, text=True, env=env, cwd=str(ROOT))
                for _ in range(k)]

    def rates(procs):
        out = []
        for pr in procs:
            so, se = pr.communicate(timeout=900)
            if pr.returncode != 0:
                raise RuntimeError(se[-400:])
            d = json.loads(so.strip().splitlines()[-1])
            t = d["flow_s"] + d["chem_s"] * d["dof"] / d["chem_nodes"]
            out.append((d["dof"] * 3 / t, d))
        return out

    cores = len(os.sched_getaffinity(0))
    try:
        one = rates(launch(1))[0]
        allc = rates(launch(cores))
    except Exception as e:                              # noqa: BLE001
        return {"value": None, "error": str(e)}
    d = one[1]
    print(f"[bench] CPU oracle reacting sample: {one[0]:.1f} DOF/s/stage on 1 thread, "
          f"{sum(r for r, _ in allc):.1f} on {cores} cores", file=sys.stderr, flush=True)
    return {"value": sum(r for r, _ in allc), "unit": "DOF-updates/s per RK stage",
            "cores": cores, "kind": "port", "one_thread_value": one[0],
            "sample": f"CPU FP64 oracle (numpy) on the reacting workload down-scaled: 3^3 box, "
                      f"20 % split (22 hex + 30 tet, {d['dof']} nodes x 16 variables), one "
                      f"SSP-RK3 step with MOOD + positivity ({d['flow_s']:.2f} s on one thread) "
                      f"and the Park-11 coupler on {nodes} nodes ({d['chem_s']:.2f} s), "
                      "extrapolated per node; value = sum of one single-thread copy per core "
                      "running at once; one_thread_value = one copy alone"}


def tgv_rate(n, steps, warmup, element_path="auto"):
    """Taylor-Green vortex, periodic 32^3 hexahedra, p=3 (the plain FP64 and
    the FP32 residual paths with the same dt), with the face and element
    kernels' own launch times."""
    import torch
    from hybrid import Topology, box_mesh
    topo = Topology(box_mesh(n, lengths=(2 * np.pi,) * 3, periodic=(True,) * 3))
    E = topo.blocks[0].count
    N, Q, V = 64, 125, 5
    dof = E * N
    print(f"[bench] TGV {n}^3 p=3: {E} elements, {dof} DOF/var", file=sys.stderr, flush=True)
    res, dt = {}, None
    for name, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        r = run_path(topo, dtype, steps, warmup, dt, element_path)
        dt = r["dt"]
        r["rate"] = dof * 3 * steps / r["elapsed"]
        res[name] = r
        print(f"[bench] TGV {name}: {r['elapsed'] * 1e3 / steps:.3f} ms/step, "
              f"{r['rate']:.4e} DOF/s/stage, element {r['elem_ms']:.3f} ms, "
              f"face {r['face_ms']:.3f} ms, valid={r['valid']}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    a = res["f32"]["state"]["hexahedron"].double()
    b = res["f64"]["state"]["hexahedron"]
    err = (torch.linalg.vector_norm((a - b).reshape(-1, V), dim=0)
           / torch.linalg.vector_norm(b.reshape(-1, V), dim=0)).cpu().numpy()
    out = {"workload": f"Taylor-Green vortex, periodic {n}^3 hexahedra, DG p=3, HLLC + BR2, "
                       "SSP-RK3, CFL 0.3, Mach 0.1, Re 1600, Pr 0.71",
           "elements": E, "dof_per_var": dof, "dt": dt, "steps": steps,
           "dt_clipped": bool(dt >= 1e-3 or dt <= 1e-7)}
    nfaces = 3 * E
    for key in ("f64", "f32"):
        r = res[key]
        itemsize = 4 if key == "f32" else 8
        fb = (2 * 16 * V * 2 + 4 * 16) * itemsize * nfaces          # per face: traces + geometry
        eflops = hex_sf_flops(3, V) * E
        out[key] = {
            "ms_per_step": r["elapsed"] * 1e3 / steps, "dof_updates_per_s_per_stage": r["rate"],
            "face_kernel": {"kernel": f"face_kernel<quad,hex,hex,p3,V5,{key}>",
                            "avg_launch_ms": r["face_ms"], "bytes_per_launch": fb,
                            "achieved_GBs": fb / (r["face_ms"] * 1e-3) / 1e9,
                            "frac_hbm": fb / (r["face_ms"] * 1e-3) / 1e9 / HBM_PEAK,
                            "traffic": measured_traffic("face", key) if n == 32 else None},
            "element_kernel": {"kernel": f"hex_sf_kernel<p3,V5,{key}> (sum factorisation)",
                               "avg_launch_ms": r["elem_ms"], "flops_per_launch": eflops,
                               "achieved_tflops": eflops / (r["elem_ms"] * 1e-3) / 1e12,
                               "frac": eflops / (r["elem_ms"] * 1e-3) / 1e12 / PEAK[key],
                               "traffic": measured_traffic("element", key) if n == 32 else None,
                               "dense_operator_flops_per_launch": element_flops(N, Q, V) * E}}
    out["fp32_rel_l2_vs_fp64"] = dict(zip(["rho", "rho_u", "rho_v", "rho_w", "rho_E"],
                                          map(float, err)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100,
                    help="timed steps of the headline workload (the reacting configuration's "
                         "100 steps; about 8 s fast path and 38 s plain path)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--error-steps", type=int, default=100,
                    help="steps after which the fast path's L2 error is measured (untimed "
                         "beyond --steps)")
    ap.add_argument("--reacting-n", type=int, default=86)
    ap.add_argument("--tgv-n", type=int, default=32)
    ap.add_argument("--tgv-steps", type=int, default=20)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-chem", action="store_true")
    ap.add_argument("--skip-mood", action="store_true")
    ap.add_argument("--skip-tgv", action="store_true")
    ap.add_argument("--only", choices=["tgv", "mood", "chem", "reacting"], default=None,
                    help="run one section only and print its JSON (for profiling)")
    ap.add_argument("--mood-steps", type=int, default=5)
    ap.add_argument("--train-steps", type=int, default=3000,
                    help="Adam steps of the reacting workload's surrogate training")
    ap.add_argument("--reacting-paths", choices=["both", "fast", "plain"], default="both",
                    help="with --only reacting: run one path only (profiling passes)")
    ap.add_argument("--element-path", choices=["auto", "mfma", "vector"], default="auto",
                    help="TGV element operators on the matrix cores or the vector units "
                         "(auto: the library default, vector units)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is a design, not a build)", file=sys.stderr)
        sys.exit(2)
    if args.only == "mood":
        print(json.dumps(mood_rate(steps=args.mood_steps)))
        return
    if args.only == "tgv":
        print(json.dumps(tgv_rate(args.tgv_n, args.tgv_steps, args.warmup, args.element_path)))
        return
    if args.only == "chem":
        print(json.dumps(chemistry_rate(1 << 20)))
        return
    paths = {"both": ("f64_plain", "f32_surrogate"), "fast": ("f32_surrogate",),
             "plain": ("f64_plain",)}[args.reacting_paths]
    if args.only != "reacting":
        paths = ("f64_plain", "f32_surrogate")
    r = reacting_rate(n=args.reacting_n, steps=args.steps, warmup=args.warmup,
                      error_steps=args.error_steps, train_steps=args.train_steps, paths=paths)
    if args.only == "reacting":
        print(json.dumps(r))
        return
    fast, plain = r["f32_surrogate"], r["f64_plain"]
    top = fast["kernels"][0]
    traffic = measured_traffic_reacting(top["kernel"])
    err = r["fast_rel_l2_vs_plain_fp64_at_error_steps"]
    out = {
        "metric": "DOF-updates/s per RK stage, plain FP64 vs FP32+surrogate chem, at fixed L2 error",
        "value": fast["dof_updates_per_s_per_stage"], "unit": "DOF-updates/s per RK stage",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": fast["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": r["workload"], "elements": r["elements"],
                   "elements_total": r["elements_total"], "dof_per_var": r["dof_per_var"],
                   "dt": r["dt"], "chemistry": "BF16-MFMA surrogate of the 12-iteration "
                   "point-implicit substep (FP64 projection, FP64 integrator fallback)"},
        "plain_fp64_dof_updates_per_s": plain["dof_updates_per_s_per_stage"],
        "plain_fp64_ms_per_step": plain["ms_per_step"],
        "speedup_fast_vs_plain": r["speedup_fast_vs_plain"],
        "rel_l2_vs_plain_fp64": {"after_steps": r["error_steps"], **err,
                                 "species": r["species_rel_l2_at_error_steps"]},
        "rel_l2_vs_plain_fp64_after_timed_steps": {**r["fast_rel_l2_vs_plain_fp64"],
                                                   "species": r["species_rel_l2"]},
        "within_1e-4": bool(err["rho"] <= 1e-4 and err["rho_E"] <= 1e-4),
        "roofline": roofline_of(top, "f32", traffic),
        "reacting": r,
    }
    if not args.skip_tgv:
        out["tgv"] = tgv_rate(args.tgv_n, args.tgv_steps, args.warmup, args.element_path)
    if not args.skip_chem:
        out["chemistry"] = chemistry_rate(1 << 20)
    if not args.skip_mood:
        out["mood"] = mood_rate(steps=args.mood_steps)
    if not args.skip_cpu:
        out["cpu_baseline"] = cpu_baseline_reacting()
        if "tgv" in out:
            out["tgv"]["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
