"""Learned surrogate of the stiff air-11 chemistry substep (no counterpart in
the reference solver).

The network maps the 14 features of a cell, [log T, log rho, Y_1..Y_11,
log dt] (standardised), through 14 -> 128 -> 128 -> 128 -> 11 (SiLU) to
log-ratios o_s; the prediction is Y' = softmax(log(Y + eps) + o), then the
minimum-KL projection onto the cell's nitrogen, oxygen and charge content,
A Y' = A Y (rows N_s/M_s, O_s/M_s, q_s/M_s), then the total-mass rescale of
the integrator (hybrid/chemistry/integrator.py:102).  The output conserves
charge and total mass to FP64 round-off and N and O to the change of the
electron mass fraction, exactly the invariants of the integrator itself (the
species table gives ions the neutral molar mass, so sum(Y) = M_N (N row) +
M_O (O row) + Y_e).  It is trained on states and results of the plain
point-implicit integrator (hybrid/chemistry/integrator.py:31-104, run on the
GPU by PointImplicitIntegrator) with the hand-written FP32 GEMM
forward/backward and Adam kernels; inference runs on BF16 MFMA with FP32
accumulation.

The default target is the isothermal substep (constant_temperature=True):
with the temperature re-solved inside every Newton iteration
(constant_temperature=False, the coupler's mode) the integrator's result on
far-from-equilibrium air moves by O(10 %) under 1e-16 input noise, a map no
regression can follow; that mode can still be selected.

A cell whose features leave the training hull (plus a margin), whose
projection misses the tolerance, or whose prediction fails the integrator's
own check -- the backward-Euler residual of the prediction, one rate
evaluation scaled by the Jacobian diagonal into an error estimate, above
check_rtol x the predicted change + check_atol x the cell density -- is
advanced by the plain integrator instead.
"""

import ctypes
import math

import numpy as np
import torch

from .. import _native as nat
from .air11_park import CHARGES, MOLAR_MASSES

N_IN, HIDDEN, N_OUT = 14, 128, 11
# Element content (N atoms, O atoms) of e-, N+, O+, NO+, N2+, O2+, N, O, NO, N2, O2.
N_ATOMS = (0, 1, 0, 1, 2, 0, 1, 0, 1, 2, 0)
O_ATOMS = (0, 0, 1, 1, 0, 2, 0, 1, 1, 0, 2)


class InferConsts(ctypes.Structure):
    """Mirror of the constant block the inference kernel takes by value."""
    _fields_ = [("mean", ctypes.c_float * N_IN), ("inv_std", ctypes.c_float * N_IN),
                ("out_mean", ctypes.c_float * 16), ("out_std", ctypes.c_float * 16),
                ("hull_lo", ctypes.c_float * N_IN), ("hull_hi", ctypes.c_float * N_IN),
                ("eps_y", ctypes.c_float), ("hull_margin", ctypes.c_float),
                ("proj_tol", ctypes.c_double),
                ("molar_mass", ctypes.c_double * N_OUT), ("charge", ctypes.c_double * N_OUT),
                ("n_atoms", ctypes.c_double * N_OUT), ("o_atoms", ctypes.c_double * N_OUT)]


def perm_k(lane, j):
    """K index held by (lane, slot j) of a 16x16x32 MFMA operand: slots 0-3
    take the C-layout rows of the first 16-row tile, slots 4-7 of the second."""
    return (lane // 16) * 4 + j if j < 4 else 16 + (lane // 16) * 4 + (j - 4)


def _pack_layer(W, out_tiles, k_steps):
    """W (out, in) float -> (out_tiles * k_steps, 64, 8) in A-operand order,
    zero-padded: block (o, s), lane l, slot j holds W[16 o + l % 16, 32 s + perm_k(l, j)]."""
    lanes = np.arange(64)
    rows = (np.arange(out_tiles)[:, None, None, None] * 16 + (lanes % 16)[None, None, :, None])
    ks = np.array([[perm_k(l, j) for j in range(8)] for l in range(64)])
    cols = np.arange(k_steps)[None, :, None, None] * 32 + ks[None, None, :, :]
    rows, cols = np.broadcast_arrays(rows, cols)
    Wp = np.zeros((W.shape[0] + 32, W.shape[1] + 32), np.float32)
    Wp[:W.shape[0], :W.shape[1]] = W
    return Wp[rows, cols].reshape(-1)


def pack_weights(params):
    """Packed bf16 weights and the 400-float bias array of the inference kernel."""
    W1, b1, W2, b2, W3, b3, W4, b4 = [p.detach().float().cpu().numpy() for p in params]
    packed = np.concatenate([_pack_layer(W1, 8, 1), _pack_layer(W2, 8, 4),
                             _pack_layer(W3, 8, 4), _pack_layer(W4, 1, 4)])
    bias = np.zeros(3 * HIDDEN + 16, np.float32)
    bias[:HIDDEN], bias[HIDDEN:2 * HIDDEN], bias[2 * HIDDEN:3 * HIDDEN] = b1, b2, b3
    bias[3 * HIDDEN:3 * HIDDEN + N_OUT] = b4
    return packed, bias


def raw_features(rho, T, dt):
    """(n, 14) float64: [log T, log rho, Y_1..Y_11, log dt]."""
    rt = rho.sum(-1, keepdim=True)
    dt = torch.as_tensor(dt, dtype=torch.float64, device=rho.device).expand(rho.shape[:-1])
    return torch.cat([torch.log(T)[:, None], torch.log(rt), rho / rt, torch.log(dt)[:, None]], 1)


def log_ratio_targets(rho, rho_next, eps):
    Y0 = rho / rho.sum(-1, keepdim=True)
    Y1 = rho_next / rho_next.sum(-1, keepdim=True)
    return torch.log(Y1 + eps) - torch.log(Y0.clamp_min(0) + eps)


class ChemistrySurrogate:
    """Learned replacement for ``PointImplicitIntegrator.step`` over a batch.

    ``step(rho, T, dt)`` has the integrator's signature and returns
    (rho_new, T, fallback_mask); rho_new holds the projected prediction, or
    the plain integrator's result where the cell was flagged."""

    def __init__(self, integrator, constant_temperature=True, seed=0, eps_y=1e-12,
                 hull_margin=0.25, proj_tol=1e-9, device=None, check_rtol=0.1,
                 check_atol=1e-8):
        self.integrator = integrator
        self.constant_temperature = bool(constant_temperature)
        # the integrator's error check of each prediction (fsg_chem_check):
        # fall back where the estimated error of a species exceeds
        # check_rtol x its predicted change + check_atol x the cell density
        # (check_rtol None: no check)
        self.check_rtol = None if check_rtol is None else float(check_rtol)
        self.check_atol = float(check_atol)
        self.last_error = None
        self.device = integrator.mechanism.device if device is None else torch.device(device)
        self.eps_y = float(eps_y)
        self.hull_margin = float(hull_margin)
        self.proj_tol = float(proj_tol)
        g = torch.Generator().manual_seed(seed)
        self.params = []
        for fan_in, fan_out in ((N_IN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, N_OUT)):
            bound = 1.0 / math.sqrt(fan_in)
            W = (torch.rand(fan_out, fan_in, generator=g) * 2 - 1) * bound
            b = (torch.rand(fan_out, generator=g) * 2 - 1) * bound
            self.params += [W.to(self.device).contiguous(), b.to(self.device).contiguous()]
        self.adam_m = [torch.zeros_like(p) for p in self.params]
        self.adam_v = [torch.zeros_like(p) for p in self.params]
        self.adam_t = 0
        self.stats = None
        self._packed = None

    # -- data ------------------------------------------------------------------
    def reference_step(self, rho, T, dt):
        out, _, _ = self.integrator.step(rho, T, dt, constant_temperature=self.constant_temperature)
        return out

    def set_normalisation(self, X, O):
        """Feature/target mean and std and the raw feature hull.  A feature's
        std is floored at 1e-6 + 1e-3 |mean|: a (nearly) constant feature such
        as a fixed log dt would otherwise be standardised into FP32 rounding
        noise."""
        mean, om = X.mean(0), O.mean(0)
        std = torch.maximum(X.std(0), 1e-6 + 1e-3 * mean.abs())
        osd = torch.maximum(O.std(0), 1e-6 + 1e-3 * om.abs())
        self.stats = dict(mean=mean.float(), std=std.float(), out_mean=om.float(),
                          out_std=osd.float(), lo=X.min(0).values.float(),
                          hi=X.max(0).values.float())
        self._packed = None

    def dataset(self, rho, T, dt, trajectory_steps=1):
        """Features and log-ratio targets from the plain integrator, over the
        first ``trajectory_steps`` states of each cell's trajectory (the
        initial state, then the integrator's own results)."""
        rho = torch.as_tensor(rho, dtype=torch.float64, device=self.device)
        T = torch.as_tensor(T, dtype=torch.float64, device=self.device)
        Xs, Os = [], []
        cur = rho
        for _ in range(trajectory_steps):
            nxt = self.reference_step(cur, T, dt)
            Xs.append(raw_features(cur, T, dt))
            Os.append(log_ratio_targets(cur, nxt, self.eps_y))
            cur = nxt
        return torch.cat(Xs), torch.cat(Os), nxt

    def _standardise(self, X, O=None):
        s = self.stats
        Xn = ((X.float() - s["mean"]) / s["std"]).contiguous()
        if O is None:
            return Xn
        return Xn, ((O.float() - s["out_mean"]) / s["out_std"]).contiguous()

    # -- training (hand-written FP32 kernels) ----------------------------------
    def _forward(self, Xn):
        L = nat.lib()
        st = nat.stream_handle(self.device)
        M = Xn.shape[0]
        acts = [Xn]
        zs = []
        h = Xn
        for k in range(4):
            W, b = self.params[2 * k], self.params[2 * k + 1]
            N, K = W.shape
            Z = torch.empty(M, N, device=self.device)
            H = torch.empty(M, N, device=self.device)
            nat.check_mlp(L.fsg_mlp_linear_forward(nat.ptr(h), nat.ptr(W), nat.ptr(b), nat.ptr(Z),
                                                   nat.ptr(H), M, N, K, int(k < 3), st))
            zs.append(Z)
            acts.append(H)
            h = H
        return h, acts, zs

    def forward(self, X):
        """Standardised network output for raw features X (FP32 path)."""
        return self._forward(self._standardise(X))[0]

    def loss_and_grads(self, Xn, On):
        """Mean squared error and parameter gradients (FP32 kernels)."""
        L = nat.lib()
        st = nat.stream_handle(self.device)
        out, acts, zs = self._forward(Xn)
        M = Xn.shape[0]
        g = torch.empty_like(out)
        nat.check_mlp(L.fsg_mlp_mse_grad(nat.ptr(out), nat.ptr(On), nat.ptr(g), out.numel(),
                                         2.0 / out.numel(), st))
        grads = [None] * 8
        for k in reversed(range(4)):
            W = self.params[2 * k]
            N, K = W.shape
            dW = torch.empty_like(W)
            db = torch.empty_like(self.params[2 * k + 1])
            dX = torch.empty(M, K, device=self.device) if k > 0 else None
            work = torch.empty(L.fsg_mlp_backward_work_size(M, N, K), device=self.device)
            nat.check_mlp(L.fsg_mlp_linear_backward(nat.ptr(g), nat.ptr(zs[k]), nat.ptr(acts[k]),
                                                    nat.ptr(W), nat.ptr(dW), nat.ptr(db),
                                                    nat.ptr(dX), nat.ptr(work), M, N, K,
                                                    int(k < 3), st))
            grads[2 * k], grads[2 * k + 1] = dW, db
            g = dX
        loss = torch.mean((out - On) ** 2)
        return loss, grads

    def adam(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        L = nat.lib()
        st = nat.stream_handle(self.device)
        self.adam_t += 1
        for p, gr, m, v in zip(self.params, grads, self.adam_m, self.adam_v):
            nat.check_mlp(L.fsg_mlp_adam(nat.ptr(p), nat.ptr(gr), nat.ptr(m), nat.ptr(v),
                                         p.numel(), lr, b1, b2, eps, self.adam_t, st))
        self._packed = None

    def train(self, X, O, steps=2000, batch=4096, lr=2e-3, lr_final=1e-4, seed=0):
        """Adam on minibatches of (X, O) with a cosine learning-rate decay;
        returns the loss history (one value per 100 steps)."""
        if self.stats is None:
            self.set_normalisation(X, O)
        Xn, On = self._standardise(X, O)
        g = torch.Generator(device=self.device).manual_seed(seed)
        hist = []
        n = Xn.shape[0]
        for it in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), device=self.device, generator=g)
            a = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * it / max(steps, 1)))
            loss, grads = self.loss_and_grads(Xn[idx].contiguous(), On[idx].contiguous())
            self.adam(grads, lr=a)
            if it % 100 == 0 or it == steps - 1:
                hist.append(float(loss))
        return hist

    # -- inference (BF16 MFMA) ----------------------------------------------------
    def _pack(self):
        if self._packed is not None:
            return self._packed
        if self.stats is None:
            raise RuntimeError("surrogate has no normalisation: call set_normalisation or train first")
        L = nat.lib()
        packed, bias = pack_weights(self.params)
        if packed.size != L.fsg_mlp_packed_weight_count():
            raise RuntimeError("packed weight layout does not match the library")
        c = InferConsts()
        if ctypes.sizeof(c) != L.fsg_mlp_consts_size():
            raise RuntimeError("surrogate constant block does not match the library")
        s = {k: v.cpu().numpy() for k, v in self.stats.items()}
        for i in range(N_IN):
            c.mean[i] = s["mean"][i]
            c.inv_std[i] = 1.0 / s["std"][i]
            c.hull_lo[i] = s["lo"][i]
            c.hull_hi[i] = s["hi"][i]
        for i in range(16):
            c.out_mean[i] = s["out_mean"][i] if i < N_OUT else 0.0
            c.out_std[i] = s["out_std"][i] if i < N_OUT else 1.0
        c.eps_y, c.hull_margin, c.proj_tol = self.eps_y, self.hull_margin, self.proj_tol
        for i in range(N_OUT):
            c.molar_mass[i] = MOLAR_MASSES[i]
            c.charge[i] = CHARGES[i]
            c.n_atoms[i] = N_ATOMS[i]
            c.o_atoms[i] = O_ATOMS[i]
        w = torch.as_tensor(packed).to(torch.bfloat16).to(self.device).contiguous()
        b = torch.as_tensor(bias).to(self.device).contiguous()
        self._packed = (w, b, c)
        return self._packed

    def infer(self, rho, T, dt, blocks=2048):
        """Projected predictions and fallback flags (no integrator call)."""
        out, flag = self._infer(rho, T, dt, blocks)
        return out, flag.bool()

    def _infer(self, rho, T, dt, blocks=2048):
        w, b, c = self._pack()
        rho = torch.as_tensor(rho, dtype=torch.float64, device=self.device).contiguous()
        T = torch.as_tensor(T, dtype=torch.float64, device=self.device).contiguous()
        n = rho.shape[0]
        if rho.shape != (n, N_OUT) or T.shape != (n,):
            raise ValueError("expected rho (n, 11) and T (n,)")
        dtt, dts = None, 0.0
        if torch.is_tensor(dt) and dt.numel() > 1:
            dtt = dt.to(self.device, torch.float64).reshape(n).contiguous()
        else:
            dts = float(dt)
        out = torch.empty_like(rho)
        flag = torch.empty(n, dtype=torch.int32, device=self.device)
        nat.check_mlp(nat.lib().fsg_mlp_infer(nat.ptr(w), nat.ptr(b), ctypes.byref(c), nat.ptr(rho),
                                              nat.ptr(T), nat.ptr(dtt), dts, nat.ptr(out),
                                              nat.ptr(flag), n, blocks,
                                              nat.stream_handle(self.device)))
        return out, flag

    def step(self, rho, T, dt):
        """Surrogate step with the plain integrator on flagged cells, all on
        the device (flagged cells are compacted by the library, no host
        synchronisation).  Returns (rho_new, T, fallback_mask)."""
        rho = torch.as_tensor(rho, dtype=torch.float64, device=self.device).contiguous()
        T = torch.as_tensor(T, dtype=torch.float64, device=self.device).contiguous()
        out, flag = self._infer(rho, T, dt)
        n = rho.shape[0]
        dtt, dts = None, 0.0
        if torch.is_tensor(dt) and dt.numel() > 1:
            dtt = dt.to(self.device, torch.float64).reshape(n).contiguous()
        else:
            dts = float(dt)
        integ = self.integrator
        if self.check_rtol is not None:
            self.last_error = torch.empty(n, dtype=torch.float64, device=self.device)
            nat.check_chem(nat.lib().fsg_chem_check(
                nat.ptr(integ.mechanism.table), nat.ptr(rho), nat.ptr(T), nat.ptr(dtt), dts,
                int(self.constant_temperature), nat.ptr(out), self.check_rtol, self.check_atol,
                nat.ptr(flag), nat.ptr(self.last_error), n, nat.stream_handle(self.device)))
        work = torch.empty(n + 1, dtype=torch.int32, device=self.device)
        nat.check_chem(nat.lib().fsg_chem_integrate_flagged(
            nat.ptr(integ.mechanism.table), nat.ptr(rho), nat.ptr(T), nat.ptr(dtt), dts,
            integ.max_newton_iters, int(self.constant_temperature), nat.ptr(flag),
            nat.ptr(work[1:]), nat.ptr(work[:1]), nat.ptr(out), None, n,
            nat.stream_handle(self.device)))
        return out, T, flag.bool()
