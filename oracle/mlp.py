"""CPU FP64 reference of the chemistry surrogate (no counterpart in the
reference solver; parity of the GPU kernels with this file is pinned by
finite differences of the loss and by closed forms of the projection).

Network: x (14) -> [W1 x + b1, SiLU] -> 128 -> [SiLU] -> 128 -> [SiLU] -> 128
-> W4 h + b4 (11).  Loss: mean squared error over a batch and the 11
outputs.  Adam: the standard update (Kingma & Ba 2015) with bias correction.

Projection: given a positive prediction Y and the element/charge/mass
constraints A Y = c (rows: N atoms per kg, O atoms per kg, charge per kg,
total mass fraction), the minimum-KL point Y' = Y exp(A^T lam) / normaliser
solves the constraints; lam comes from Newton on A (Y exp(A^T lam)) = c.
"""

import numpy as np

from .chemistry import CHARGE, MOLAR_MASS

# Element content (N atoms, O atoms) of e-, N+, O+, NO+, N2+, O2+, N, O, NO, N2, O2.
N_ATOMS = np.array([0, 1, 0, 1, 2, 0, 1, 0, 1, 2, 0], float)
O_ATOMS = np.array([0, 0, 1, 1, 0, 2, 0, 1, 1, 0, 2], float)


def constraint_matrix():
    """Rows per unit mass: mol N, mol O, mol charge, mass fraction sum."""
    return np.stack([N_ATOMS / MOLAR_MASS, O_ATOMS / MOLAR_MASS, CHARGE / MOLAR_MASS,
                     np.ones(11)])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x):
    """Returns outputs and the per-layer (input, pre-activation) cache."""
    h = x
    cache = []
    n = len(params) // 2
    for k in range(n):
        W, b = params[2 * k], params[2 * k + 1]
        z = h @ W.T + b
        cache.append((h, z))
        h = silu(z) if k < n - 1 else z
    return h, cache


def loss_and_grads(params, x, target):
    out, cache = forward(params, x)
    diff = out - target
    loss = np.mean(diff ** 2)
    g = 2.0 * diff / diff.size
    grads = [None] * len(params)
    n = len(params) // 2
    for k in reversed(range(n)):
        h, z = cache[k]
        if k < n - 1:
            g = g * dsilu(z)
        grads[2 * k] = g.T @ h
        grads[2 * k + 1] = g.sum(0)
        g = g @ params[2 * k]
    return loss, grads


def adam_step(params, grads, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    out = []
    for k, (p, g) in enumerate(zip(params, grads)):
        m[k] = b1 * m[k] + (1 - b1) * g
        v[k] = b2 * v[k] + (1 - b2) * g * g
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        out.append(p - lr * mh / (np.sqrt(vh) + eps))
    return out


def project(Y, c, iters=8):
    """Minimum-KL projection of positive Y onto A Y' = c (Newton on lam)."""
    A = constraint_matrix()
    lam = np.zeros(4)
    for _ in range(iters):
        Yp = Y * np.exp(A.T @ lam)
        r = A @ Yp - c
        J = (A * Yp) @ A.T
        lam = lam - np.linalg.solve(J, r)
    return Y * np.exp(A.T @ lam)
