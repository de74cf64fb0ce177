"""Shared test set-up: the ``gpu`` marker, import paths, and builders that
produce the same mesh/state for the oracle (CPU, FP64) and the device path."""

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "jax-fluid-sim_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def pytest_collection_modifyitems(config, items):
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


# Box plane names used by the boundary tests: x- inflow, x+ outflow,
# y/z planes no-slip ("sphere_wall") or slip ("walls") walls.
PLANE_NAMES = {"xmin": "inlet", "xmax": "outlet", "ymin": "walls", "ymax": "walls",
               "zmin": "sphere_wall", "zmax": "symmetry"}


def tgv_primitive(x, gamma=1.4, mach=0.1):
    """Taylor-Green vortex initial condition (the reference's legacy driver,
    tetra_dg.py:2289-2362): rho = 1, p = 1/(gamma M^2) + perturbation."""
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    rho = np.ones_like(X)
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(X)
    p = 1.0 / (gamma * mach ** 2) + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2) / 16
    return np.stack([rho, u, v, w, p], -1)


def perturbed_primitive(x, seed=0, nspecies=0, amplitude=0.05):
    """Smooth non-uniform primitive state plus seeded noise (all positive)."""
    rng = np.random.default_rng(seed)
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    rho = 1.0 + 0.2 * np.sin(2 * X + 1) * np.cos(Y) + amplitude * rng.standard_normal(X.shape)
    u = 0.8 + 0.3 * np.cos(Z) + amplitude * rng.standard_normal(X.shape)
    v = 0.2 * np.sin(X + Y) + amplitude * rng.standard_normal(X.shape)
    w = -0.1 + 0.2 * np.cos(X) + amplitude * rng.standard_normal(X.shape)
    p = 1.0 + 0.3 * np.cos(X - Z) + amplitude * rng.standard_normal(X.shape)
    parts = [rho, u, v, w, p]
    if nspecies:
        Y = rng.dirichlet(np.full(nspecies, 2.0), size=X.shape)
        parts += [Y[..., s] for s in range(nspecies)]
    return np.stack(parts, -1)


def rel_l2(a, b):
    """Relative L2 error per variable over all blocks; variables whose
    reference norm is zero are measured against the largest variable norm."""
    A = np.concatenate([np.asarray(a[k]).reshape(-1, np.asarray(a[k]).shape[-1]) for k in b])
    B = np.concatenate([np.asarray(b[k]).reshape(-1, np.asarray(b[k]).shape[-1]) for k in b])
    nb = np.linalg.norm(B, axis=0)
    scale = np.where(nb > 0, nb, max(nb.max(), 1e-300))
    return np.linalg.norm(A - B, axis=0) / scale
