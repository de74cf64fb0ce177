"""CPU FP64 reference of the chemistry surrogate (no counterpart in the
reference solver; parity of the GPU kernels with this file is pinned by
finite differences of the loss and by closed forms of the projection).

Network: x (14) -> [W1 x + b1, SiLU] -> 128 -> [SiLU] -> 128 -> [SiLU] -> 128
-> W4 h + b4 (11).  Loss: mean squared error over a batch and the 11
outputs.  Adam: the standard update (Kingma & Ba 2015) with bias correction.

Projection: given a positive prediction Y and the element/charge
constraints A Y = c (rows: N atoms per kg, O atoms per kg, charge per kg),
the minimum-KL point Y' = Y exp(A^T lam) solves them; lam comes from a damped
Newton iteration on A (Y exp(A^T lam)) = c.  The total mass is restored
afterwards by normalising Y' (the integrator's own final rescale): with the
species table giving ions the neutral molar mass, sum(Y) equals
M_N (N row) + M_O (O row) + Y_e, so a mass row would freeze Y_e.
"""

import numpy as np

from .chemistry import CHARGE, MOLAR_MASS

# Element content (N atoms, O atoms) of e-, N+, O+, NO+, N2+, O2+, N, O, NO, N2, O2.
N_ATOMS = np.array([0, 1, 0, 1, 2, 0, 1, 0, 1, 2, 0], float)
O_ATOMS = np.array([0, 0, 1, 1, 0, 2, 0, 1, 1, 0, 2], float)


def constraint_matrix():
    """Rows per unit mass: mol N, mol O, mol charge."""
    return np.stack([N_ATOMS / MOLAR_MASS, O_ATOMS / MOLAR_MASS, CHARGE / MOLAR_MASS])


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


def project(Y, c, iters=32):
    """Minimum-KL projection of positive Y onto A Y' = c: Newton on the dual
    variable lam, each step scaled so that no species' log changes by more
    than 2 (the electron column of A is ~1e6 times the others), stopping
    once the step is below 1e-14."""
    A = constraint_matrix()
    lam = np.zeros(3)
    for _ in range(iters):
        Yp = Y * np.exp(A.T @ lam)
        r = A @ Yp - c
        J = (A * Yp) @ A.T
        d = np.linalg.solve(J, r)
        step = np.max(np.abs(A.T @ d))
        lam = lam - (2.0 / step if step > 2.0 else 1.0) * d
        if not step > 1e-14:
            break
    return Y * np.exp(A.T @ lam)
