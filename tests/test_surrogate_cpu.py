"""Surrogate host logic on the CPU: the MLP oracle against finite differences
and closed forms, and the packed BF16 weight layout through an emulation of
the 16x16x32 matrix-core operand/accumulator lane layout."""

import numpy as np
import pytest
import torch

from oracle import mlp as omlp
from hybrid.chemistry.surrogate import (HIDDEN, N_IN, N_OUT, _pack_layer, pack_weights,
                                        perm_k)


def _params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    shapes = [(HIDDEN, N_IN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (N_OUT, HIDDEN)]
    out = []
    for o, i in shapes:
        out += [rng.uniform(-1, 1, (o, i)) * scale / np.sqrt(i), rng.uniform(-0.1, 0.1, o)]
    return out


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    params = _params(1)
    x = rng.standard_normal((5, N_IN))
    t = rng.standard_normal((5, N_OUT))
    loss, grads = omlp.loss_and_grads(params, x, t)
    for k in (0, 1, 4, 7):
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in params[k].shape)
            h = 1e-6
            pp = [p.copy() for p in params]
            pp[k][idx] += h
            pm = [p.copy() for p in params]
            pm[k][idx] -= h
            fd = (omlp.loss_and_grads(pp, x, t)[0] - omlp.loss_and_grads(pm, x, t)[0]) / (2 * h)
            assert abs(fd - grads[k][idx]) <= 1e-7 * max(1.0, abs(fd)), (k, idx)


def test_oracle_adam_first_step_is_signed_lr():
    """With bias correction the first Adam step is -lr * sign(g) (|g| >> eps)."""
    p = [np.array([1.0, -2.0, 3.0])]
    g = [np.array([0.5, -4.0, 1e-3])]
    out = omlp.adam_step(p, g, [np.zeros(3)], [np.zeros(3)], 1, lr=0.01)
    np.testing.assert_allclose(out[0], p[0] - 0.01 * np.sign(g[0]), rtol=0, atol=1e-7)


def test_oracle_projection_conserves_and_is_identity_on_feasible_points():
    A = omlp.constraint_matrix()
    rng = np.random.default_rng(2)
    Y0 = rng.dirichlet(np.full(11, 2.0))
    c = A @ Y0
    # identity on a point that already satisfies the constraints
    np.testing.assert_allclose(omlp.project(Y0, c), Y0, rtol=1e-12)
    # a perturbed point lands back on the constraint set, with the
    # exponential-family form Y' = Y exp(A^T lam)
    Yp = Y0 * np.exp(0.2 * rng.standard_normal(11))
    Yq = omlp.project(Yp, c)
    np.testing.assert_allclose(A @ Yq, c, rtol=1e-12, atol=1e-14)
    lam, *_ = np.linalg.lstsq(A.T, np.log(Yq / Yp), rcond=None)
    np.testing.assert_allclose(A.T @ lam, np.log(Yq / Yp), atol=1e-6)   # A is badly scaled (e-)


def _mfma(A_lanes, B_lanes):
    """16x16x32 MFMA on lane-held operands (64 lanes x 8 slots) -> C (16, 16):
    lane l holds A row l % 16 and B column l % 16; lane group l // 16 and slot
    j name the same K index in both operands."""
    C = np.zeros((16, 16))
    for g in range(4):
        a = A_lanes[16 * g:16 * g + 16]          # (16 rows, 8)
        b = B_lanes[16 * g:16 * g + 16]          # (16 cols, 8)
        C += a @ b.T
    return C


def _c_to_lanes(C):
    """Accumulator layout: lane l, element j holds C[(l // 16) * 4 + j, l % 16]."""
    return np.array([[C[(l // 16) * 4 + j, l % 16] for j in range(4)] for l in range(64)])


def _emulate(params, X):
    """Network on 16 cells through the packed layout and the C->B handoff."""
    packed, bias = pack_weights([torch.as_tensor(p) for p in params])
    blk = 64 * 8
    offs = np.cumsum([0, 8 * blk, 32 * blk, 32 * blk])
    Wl = [packed[offs[i]:offs[i] + n * blk].reshape(n, 64, 8)
          for i, n in enumerate((8, 32, 32, 4))]
    xin = np.zeros((64, 8))
    for l in range(64):
        for j in range(8):
            k = perm_k(l, j)
            xin[l, j] = X[l % 16, k] if k < N_IN else 0.0
    silu = lambda z: z / (1 + np.exp(-z))

    def handoff(tiles, b):
        out = []
        for s in range(4):
            lo, hi = _c_to_lanes(tiles[2 * s]), _c_to_lanes(tiles[2 * s + 1])
            r = np.zeros((64, 8))
            for l in range(64):
                base = (l // 16) * 4
                for j in range(4):
                    r[l, j] = silu(lo[l, j] + b[32 * s + base + j])
                    r[l, 4 + j] = silu(hi[l, j] + b[32 * s + 16 + base + j])
            out.append(r)
        return out

    h = [_mfma(Wl[0][o], xin) for o in range(8)]
    B = handoff(h, bias[:HIDDEN])
    for layer in (1, 2):
        h = [sum(_mfma(Wl[layer][o * 4 + s], B[s]) for s in range(4)) for o in range(8)]
        B = handoff(h, bias[HIDDEN * layer:HIDDEN * (layer + 1)])
    out = sum(_mfma(Wl[3][s], B[s]) for s in range(4))
    return out[:N_OUT].T + bias[3 * HIDDEN:3 * HIDDEN + N_OUT]


def test_packed_layout_reproduces_the_network():
    params = _params(3)
    params = [p.astype(np.float32).astype(np.float64) for p in params]
    X = np.random.default_rng(4).standard_normal((16, N_IN))
    ref, _ = omlp.forward(params, X)
    got = _emulate(params, X)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5)


def test_pack_layer_zero_pads():
    W = np.ones((11, 128), np.float32)
    p = _pack_layer(W, 1, 4).reshape(4, 64, 8)
    # rows 11..15 of the single output tile are padding
    assert np.all(p[:, 11:16, :] == 0) and np.all(p[:, 27:32, :] == 0)
    assert np.all(p[:, 0:11, :] == 1)
    W1 = np.ones((128, 14), np.float32)
    p1 = _pack_layer(W1, 8, 1).reshape(8, 64, 8)
    ks = np.array([[perm_k(l, j) for j in range(8)] for l in range(64)])
    assert np.all(p1[:, ks >= 14] == 0) and np.all(p1[:, ks < 14] == 1)
