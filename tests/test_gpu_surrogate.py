"""Chemistry surrogate on the GPU: the FP32 training kernels against the FP64
MLP oracle, BF16 MFMA inference against an exact emulation of its rounding
points, exact element/charge/mass conservation of the projection, the hull
fallback, and the ablation against the trivial (no-change) predictor."""

import numpy as np
import pytest
import torch

from hybrid.chemistry import Air11ParkMechanism, PointImplicitIntegrator
from hybrid.chemistry.air11_park import CHARGES, MOLAR_MASSES
from hybrid.chemistry.batch import make_batch
from hybrid.chemistry.surrogate import (HIDDEN, N_IN, N_OUT, ChemistrySurrogate,
                                        log_ratio_targets, raw_features)
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def integrator():
    return PointImplicitIntegrator(Air11ParkMechanism(), max_newton_iters=12)


def _as_np(params):
    return [p.double().cpu().numpy() for p in params]


def test_training_kernels_match_oracle(integrator):
    sur = ChemistrySurrogate(integrator, seed=5)
    rng = np.random.default_rng(0)
    M = 300                                     # not a multiple of the 64-row tile
    Xn = torch.as_tensor(rng.standard_normal((M, N_IN)), dtype=torch.float32, device="cuda")
    On = torch.as_tensor(rng.standard_normal((M, N_OUT)), dtype=torch.float32, device="cuda")
    loss, grads = sur.loss_and_grads(Xn, On)
    ref_loss, ref_grads = omlp.loss_and_grads(_as_np(sur.params), Xn.double().cpu().numpy(),
                                              On.double().cpu().numpy())
    assert abs(float(loss) - ref_loss) <= 1e-5 * ref_loss
    for k, (g, r) in enumerate(zip(grads, ref_grads)):
        err = np.linalg.norm(g.double().cpu().numpy() - r) / np.linalg.norm(r)
        assert err <= 1e-5, (k, err)


def test_adam_kernel_matches_oracle(integrator):
    sur = ChemistrySurrogate(integrator, seed=6)
    rng = np.random.default_rng(1)
    Xn = torch.as_tensor(rng.standard_normal((128, N_IN)), dtype=torch.float32, device="cuda")
    On = torch.as_tensor(rng.standard_normal((128, N_OUT)), dtype=torch.float32, device="cuda")
    p = _as_np(sur.params)
    m = [np.zeros_like(x) for x in p]
    v = [np.zeros_like(x) for x in p]
    for t in range(1, 4):
        _, grads = sur.loss_and_grads(Xn, On)
        gr = [g.double().cpu().numpy() for g in grads]
        sur.adam(grads, lr=1e-3)
        p = omlp.adam_step(p, gr, m, v, t, lr=1e-3)
        for a, b in zip(sur.params, p):
            np.testing.assert_allclose(a.double().cpu().numpy(), b, rtol=0, atol=2e-6)


def _bf16(x):
    return torch.as_tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()


def _emulated_infer(sur, rho, T, dt):
    """Reference of the inference kernel: BF16 weights and BF16 activations at
    the same points, wide accumulation, then the FP64 softmax and projection."""
    s = {k: v.double().cpu().numpy() for k, v in sur.stats.items()}
    X = raw_features(rho, T, dt).cpu().numpy()
    x = _bf16(((X.astype(np.float32) - s["mean"].astype(np.float32))
               / s["std"].astype(np.float32)))
    W = [_bf16(p.cpu().numpy()) for p in sur.params[0::2]]
    b = [p.double().cpu().numpy() for p in sur.params[1::2]]
    h = x
    for k in range(3):
        z = h @ W[k].T + b[k]
        h = _bf16(z / (1 + np.exp(-z)))
    o = (h @ W[3].T + b[3]) * s["out_std"] + s["out_mean"]
    r = rho.cpu().numpy()
    rt = r.sum(1, keepdims=True)
    Y0 = r / rt
    lp = np.log(Y0 + sur.eps_y) + o
    Yp = np.exp(lp - lp.max(1, keepdims=True))
    Yp /= Yp.sum(1, keepdims=True)
    A = omlp.constraint_matrix()
    out = np.stack([omlp.project(Yp[c], A @ Y0[c]) for c in range(len(r))])
    return out / out.sum(1, keepdims=True) * rt


def _fitted(integrator, n=4096, seed=2, train_steps=0):
    rho, T = make_batch(n, seed=seed)
    sur = ChemistrySurrogate(integrator, seed=seed)
    X, O, nxt = sur.dataset(rho, T, 1e-7)
    sur.set_normalisation(X, O)
    if train_steps:
        sur.train(X, O, steps=train_steps, batch=2048)
    return sur, torch.as_tensor(rho, device="cuda"), torch.as_tensor(T, device="cuda"), nxt


def test_mfma_inference_matches_bf16_emulation(integrator):
    sur, rho, T, _ = _fitted(integrator, n=1000, train_steps=50)
    got, fb = sur.infer(rho, T, 1e-7)
    ref = _emulated_infer(sur, rho, T, 1e-7)
    got, fb = got.cpu().numpy(), fb.cpu().numpy()
    assert fb.mean() < 0.05
    ok = ~fb
    rt = rho.sum(1).cpu().numpy()[:, None]
    err = np.abs(got - ref)[ok] / (np.abs(ref[ok]) + 1e-12 * rt[ok])
    assert np.quantile(err, 0.99) <= 1e-3 and err.max() <= 2e-2, (np.quantile(err, 0.99), err.max())


def test_projection_conserves_elements_charge_and_mass(integrator):
    """Charge and total mass exact; N and O content moved by exactly the
    relative change the electron mass fraction forces, |Ye0 - Ye|/(1 - Ye0)
    (the integrator's own invariants, see the surrogate module)."""
    sur, rho, T, _ = _fitted(integrator, n=2048, train_steps=0)   # untrained: far from feasible
    got, fb = sur.infer(rho, T, 1e-7)
    A = omlp.constraint_matrix()
    r = rho.cpu().numpy()
    g = got.cpu().numpy()
    ok = ~fb.cpu().numpy()
    assert ok.mean() > 0.9
    Y0 = r / r.sum(1, keepdims=True)
    Y1 = g / g.sum(1, keepdims=True)
    scale = np.abs(A) @ Y0.T                                    # (3, n) per-row magnitudes
    res = np.abs(A @ (Y1 - Y0).T) / scale
    drift = np.abs(Y0[:, 0] - Y1[:, 0]) / (1 - Y0[:, 0])
    assert res[2, ok].max() <= 1e-9                             # charge
    assert np.abs(res[:2, ok] - drift[ok]).max() <= 1e-9        # N, O
    assert np.any(Y1[ok, 0] != Y0[ok, 0])                       # electrons do move
    np.testing.assert_allclose(g.sum(1), r.sum(1), rtol=1e-13)
    assert np.all(g[ok] > 0)
    # flagged cells are returned unchanged
    assert np.array_equal(g[~ok], r[~ok])


def test_out_of_hull_cells_fall_back_to_the_integrator(integrator):
    """Cells outside the training hull fall back: with the integrator-judged
    error check off, exactly the hull-flagged cells take the integrator's
    result and the others the network's; with it on (the default), the
    fallback set still holds every hull-flagged cell and every fallback cell
    carries the integrator's result."""
    sur, rho, T, _ = _fitted(integrator, n=512)
    T2 = T.clone()
    T2[::7] = 60000.0                           # far outside the 3000-12000 K training hull
    out, fb = sur.infer(rho, T2, 1e-7)
    fb = fb.cpu().numpy()
    assert fb[::7].all()
    assert np.array_equal(out.cpu().numpy()[::7], rho.cpu().numpy()[::7])
    check = sur.check_rtol
    sur.check_rtol = None
    try:
        full, _, fb2 = sur.step(rho, T2, 1e-7)
    finally:
        sur.check_rtol = check
    ref = integrator.step(rho[::7], T2[::7], 1e-7,
                          constant_temperature=sur.constant_temperature)[0]
    assert torch.equal(full[::7], ref)
    assert torch.equal(fb2.cpu(), torch.as_tensor(fb))
    keep = ~fb2
    assert torch.equal(full[keep], out[keep])
    full, _, fb3 = sur.step(rho, T2, 1e-7)
    fb3 = fb3.cpu()
    assert bool(fb3[torch.as_tensor(fb)].all())
    sel = fb3.nonzero().flatten().to(rho.device)
    ref = integrator.step(rho[sel], T2[sel], 1e-7,
                          constant_temperature=sur.constant_temperature)[0]
    assert torch.equal(full[sel], ref)


def test_neutral_cells_pass_the_projection_check(integrator):
    """Cells with no charged species (neutral air) have a zero charge row; the
    projection's residual is measured against the predicted traces, so such
    cells are not sent to the integrator (hull check switched off here)."""
    sur, rho, T, _ = _fitted(integrator, n=2048, train_steps=200)
    sur.stats["lo"] = torch.full_like(sur.stats["lo"], -1e30)
    sur.stats["hi"] = torch.full_like(sur.stats["hi"], 1e30)
    sur._packed = None
    n = 256
    Y = torch.zeros(n, 11, dtype=torch.float64, device="cuda")
    Y[:, 9], Y[:, 10] = 0.767, 0.233
    rt = torch.logspace(-3, -1, n, dtype=torch.float64, device="cuda")
    Tn = torch.linspace(3000.0, 9000.0, n, dtype=torch.float64, device="cuda")
    out, fb = sur.infer(rt[:, None] * Y, Tn, 1e-7)
    assert not bool(fb.any())
    # charge stays zero to round-off of the traces, mass is kept
    q = (out[:, :6] * torch.as_tensor(CHARGES[:6], device="cuda", dtype=torch.float64)
         / torch.as_tensor(MOLAR_MASSES[:6], device="cuda", dtype=torch.float64)).sum(1)
    trace = (out[:, :6] / torch.as_tensor(MOLAR_MASSES[:6], device="cuda",
                                          dtype=torch.float64)).sum(1)
    assert bool((q.abs() <= 1e-9 * trace + 1e-300).all())
    torch.testing.assert_close(out.sum(1), rt, rtol=1e-13, atol=0)


def test_surrogate_ablation(integrator):
    """Training helps: the trained surrogate is closer to the plain integrator
    than the no-change predictor and than the untrained network, on held-out
    cells; the projection holds the element content exactly."""
    rho, T = make_batch(1 << 15, seed=11)
    sur = ChemistrySurrogate(integrator, seed=0)
    X, O, _ = sur.dataset(rho, T, 1e-7, trajectory_steps=2)
    sur.set_normalisation(X, O)
    rho_t, T_t = make_batch(8192, seed=12)
    rho_t = torch.as_tensor(rho_t, device="cuda")
    T_t = torch.as_tensor(T_t, device="cuda")
    ref = integrator.step(rho_t, T_t, 1e-7, constant_temperature=sur.constant_temperature)[0]
    Yref = ref / ref.sum(1, keepdim=True)

    def l1(pred):
        return float((pred / pred.sum(1, keepdim=True) - Yref).abs().sum(1).mean())

    e_identity = l1(rho_t)
    e_untrained = l1(sur.infer(rho_t, T_t, 1e-7)[0])
    hist = sur.train(X, O, steps=1500, batch=4096)
    pred, fb = sur.infer(rho_t, T_t, 1e-7)
    e_trained = l1(pred)
    print(f"[surrogate ablation] L1(Y) identity {e_identity:.4f} untrained {e_untrained:.4f} "
          f"trained {e_trained:.4f} fallback {float(fb.float().mean()):.4f} loss {hist[0]:.3f}->{hist[-1]:.3f}",
          flush=True)
    assert hist[-1] < 0.5 * hist[0]
    assert e_trained < 0.7 * e_identity and e_trained < 0.7 * e_untrained
