"""Synthetic 0-D air batches for the chemistry throughput configuration.

Cells have T uniform in [T_lo, T_hi] K, p log-uniform in [p_lo, p_hi] Pa and
mass fractions drawn from a symmetric Dirichlet(alpha), then made
charge-neutral by setting the electron mass fraction to
M_e * sum_ions(Y_i / M_i) and renormalising; rho = p / (R_mix T).
"""

import numpy as np

from .air11_park import CHARGES, MOLAR_MASSES

R_UNIV = 8.31446261815324


def make_batch(n, seed=0, T_range=(3000.0, 12000.0), p_range=(1.0e3, 1.0e5), alpha=2.0):
    rng = np.random.default_rng(seed)
    M = np.asarray(MOLAR_MASSES)
    q = np.asarray(CHARGES)
    T = rng.uniform(*T_range, size=n)
    p = np.exp(rng.uniform(np.log(p_range[0]), np.log(p_range[1]), size=n))
    Y = rng.dirichlet(np.full(11, alpha), size=n)
    ions = q > 0
    Y[:, 0] = M[0] * np.sum(Y[:, ions] / M[ions], axis=1)
    Y /= Y.sum(1, keepdims=True)
    R_mix = R_UNIV * np.sum(Y / M, axis=1)
    rho = p / (R_mix * T)
    return rho[:, None] * Y, T
