"""Park 11-species air chemistry, restated from hybrid/chemistry/{nasa,
mechanism,air11_park,integrator,coupling}.py (numpy FP64, one cell at a time
or vectorised over cells).

The reference's Newton Jacobian is jax.jacobian of R(rho) = rho - rho0 -
dt w(rho, T) at fixed T (hybrid/chemistry/integrator.py:78); here it is the
analytic derivative of the same expression, including JAX's convention that
d max(x, y) splits the gradient in half when x == y (which matters at
rho_s = 0 after the non-negativity clip).  Checked against central finite
differences in tests/test_oracle_chemistry.py.
"""

import numpy as np

R_UNIV = 8.31446261815324   # hybrid/chemistry/nasa.py:13
P_STD = 101325.0            # hybrid/chemistry/nasa.py:15

MW_E = 5.48579909065e-7     # hybrid/chemistry/air11_park.py:12-17
MW_N = 0.0140067
MW_O = 0.015999
MW_NO = MW_N + MW_O
MW_N2 = 2.0 * MW_N
MW_O2 = 2.0 * MW_O

SPECIES = ("e-", "N+", "O+", "NO+", "N2+", "O2+", "N", "O", "NO", "N2", "O2")
E, NP, OP, NOP, N2P, O2P, N, O, NO, N2, O2 = range(11)
MOLAR_MASS = np.array([MW_E, MW_N, MW_O, MW_NO, MW_N2, MW_O2, MW_N, MW_O, MW_NO, MW_N2, MW_O2])
CHARGE = np.array([-1.0, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0])

# NASA-7 (t_low, t_mid, t_high, low[7], high[7]), hybrid/chemistry/air11_park.py:20-143.
NASA7 = (
    (298.15, 1000.0, 6000.0, (2.5, 0, 0, 0, 0, -745.375, -11.7208122),
     (2.5, 0, 0, 0, 0, -745.375, -11.7208122)),
    (298.15, 1000.0, 6000.0,
     (2.79781357, -1.41287178e-03, 2.68686998e-06, -2.31615591e-09, 7.49585578e-13, 225548.731, 3.59928845),
     (2.51210704, 1.75367440e-06, -1.50557538e-08, 7.06834189e-12, -6.31861166e-16, 225597.004, 4.92236293)),
    (298.15, 1000.0, 6000.0, (2.5, 0, 0, 0, 0, 187935.284, 4.39337676),
     (2.48542028, 2.56978695e-05, -1.28833378e-08, 1.65525487e-12, 1.09933344e-16, 187940.874, 4.47425446)),
    (298.15, 1000.0, 6000.0,
     (3.69301231, -1.34229158e-03, 2.67343395e-06, -1.02609308e-09, -6.95610492e-14, 118103.055, 3.09126691),
     (2.94587702, 1.40325260e-03, -4.95503196e-07, 7.95948973e-11, -4.72076668e-15, 118244.340, 6.70644634)),
    (298.15, 1000.0, 6000.0,
     (3.77540711, -2.06459157e-03, 4.75752301e-06, -3.15664228e-09, 6.70509973e-13, 180481.115, 2.69322178),
     (3.58661363, 2.53071949e-04, 1.84778214e-07, -4.55257223e-11, 3.26818029e-15, 180390.994, 3.09584143)),
    (298.15, 1000.0, 6000.0,
     (4.61017167, -6.35951952e-03, 1.42425624e-05, -1.20997923e-08, 3.70956878e-12, 139742.229, -0.201326941),
     (3.31675922, 1.11522244e-03, -3.83492556e-07, 5.72784687e-11, -2.77648381e-15, 139876.823, 5.44726469)),
    (200.0, 1000.0, 6000.0, (2.5, 0, 0, 0, 0, 56104.638, 4.1939088),
     (2.4159429, 1.74890650e-04, -1.19023690e-07, 3.02262440e-11, -2.03609830e-15, 56133.775, 4.6496095)),
    (200.0, 1000.0, 6000.0,
     (3.16826710, -3.27931884e-03, 6.64306396e-06, -6.12806624e-09, 2.11265971e-12, 29122.2592, 2.05193346),
     (2.54363697, -2.73162486e-05, -4.19029520e-09, 4.95481845e-12, -4.79553694e-16, 29226.012, 4.92229457)),
    (200.0, 1000.0, 6000.0,
     (4.21859896, -4.63988124e-03, 1.10443049e-05, -9.34055507e-09, 2.80554874e-12, 9845.09964, 2.28061001),
     (3.26071234, 1.19101135e-03, -4.29122646e-07, 6.94481463e-11, -4.03295681e-15, 9921.43132, 6.36900518)),
    (200.0, 1000.0, 6000.0,
     (3.53100528, -1.23660988e-04, -5.02999433e-07, 2.43530612e-09, -1.40881235e-12, -1046.97628, 2.96747038),
     (2.95257637, 1.39690040e-03, -4.92631603e-07, 7.86010195e-11, -4.60755204e-15, -923.948688, 5.87188762)),
    (200.0, 1000.0, 6000.0,
     (3.78245636, -2.99673416e-03, 9.84730201e-06, -9.68129509e-09, 3.24372837e-12, -1063.94356, 3.65767573),
     (3.66096065, 6.56365811e-04, -1.41149627e-07, 2.05797935e-11, -1.29913436e-15, -1215.97718, 3.41536279)),
)

# Reactions (reactants, products, A, b, theta, third_body, efficiencies,
# fwd_temp, rev_temp), hybrid/chemistry/air11_park.py:149-311.
REACTIONS = (
    (((N2, 1),), ((N, 2),), 7.0e15, -1.6, 113200.0, True, ((O, 4.285714285714286), (N, 4.285714285714286)), "Ta", "T"),
    (((N2, 1), (E, 1)), ((N, 2), (E, 1)), 3.0e18, -1.6, 113200.0, False, (), "Tv", "Tv"),
    (((O2, 1),), ((O, 2),), 2.0e15, -1.5, 59360.0, True, ((O, 5.0), (N, 5.0)), "Ta", "T"),
    (((NO, 1),), ((N, 1), (O, 1)), 5.0e9, 0.0, 75500.0, True,
     ((NO, 22.0), (N, 22.0), (O, 22.0), (NP, 22.0), (OP, 22.0)), "Ta", "T"),
    (((N2, 1), (O, 1)), ((NO, 1), (N, 1)), 5.7e6, 0.42, 42938.0, False, (), "T", "T"),
    (((NO, 1), (O, 1)), ((O2, 1), (N, 1)), 8.4e6, 0.0, 19400.0, False, (), "T", "T"),
    (((N, 1), (O, 1)), ((NOP, 1), (E, 1)), 5.3e6, 0.0, 31900.0, False, (), "T", "Tv"),
    (((O, 2),), ((O2P, 1), (E, 1)), 7.1e-4, 2.7, 80600.0, False, (), "T", "Tv"),
    (((N, 2),), ((N2P, 1), (E, 1)), 4.4e1, 1.5, 67500.0, False, (), "T", "Tv"),
    (((NOP, 1), (O, 1)), ((NP, 1), (O2, 1)), 1.0e6, 0.5, 77200.0, False, (), "T", "T"),
    (((NP, 1), (N2, 1)), ((N2P, 1), (N, 1)), 1.0e6, 0.5, 12200.0, False, (), "T", "T"),
    (((O2P, 1), (N, 1)), ((NP, 1), (O2, 1)), 8.7e7, 0.14, 28600.0, False, (), "T", "T"),
    (((OP, 1), (NO, 1)), ((NP, 1), (O2, 1)), 1.4e-1, 1.9, 26600.0, False, (), "T", "T"),
    (((O2P, 1), (N2, 1)), ((N2P, 1), (O2, 1)), 9.9e6, 0.0, 40700.0, False, (), "T", "T"),
    (((O2P, 1), (O, 1)), ((OP, 1), (O2, 1)), 4.0e6, 0.09, 18000.0, False, (), "T", "T"),
    (((NOP, 1), (N, 1)), ((OP, 1), (N2, 1)), 3.4e7, -1.08, 12800.0, False, (), "T", "T"),
    (((NOP, 1), (O2, 1)), ((O2P, 1), (NO, 1)), 2.4e7, 0.41, 32600.0, False, (), "T", "T"),
    (((NOP, 1), (O, 1)), ((O2P, 1), (N, 1)), 7.2e6, 0.29, 48600.0, False, (), "T", "T"),
    (((OP, 1), (N2, 1)), ((N2P, 1), (O, 1)), 9.1e5, 0.36, 22800.0, False, (), "T", "T"),
    (((NOP, 1), (N, 1)), ((N2P, 1), (O, 1)), 7.2e7, 0.0, 35500.0, False, (), "T", "T"),
    (((O, 1), (E, 1)), ((OP, 1), (E, 2)), 3.9e27, -3.78, 158500.0, False, (), "Tv", "T"),
    (((N, 1), (E, 1)), ((NP, 1), (E, 2)), 2.5e28, -3.82, 168200.0, False, (), "Tv", "T"),
)


class Mechanism:
    """hybrid/chemistry/mechanism.py:69-287 for a given species/reaction set."""

    def __init__(self, nasa=NASA7, reactions=REACTIONS, molar_mass=MOLAR_MASS, names=SPECIES):
        self.ns = len(molar_mass)
        self.nr = len(reactions)
        self.molar_masses = np.asarray(molar_mass, float)
        self.names = tuple(names)
        self.nasa = nasa
        nu1 = np.zeros((self.nr, self.ns))
        nu2 = np.zeros((self.nr, self.ns))
        alpha = np.zeros((self.nr, self.ns))
        A, b, th, tb, fm, rm = [], [], [], [], [], []
        modes = {"T": 0, "Ta": 1, "Tv": 2}
        for r, (rea, pro, a_, b_, t_, third, eff, ft, rt) in enumerate(reactions):
            for s, c in rea:
                nu1[r, s] += c
            for s, c in pro:
                nu2[r, s] += c
            if third:
                for s in range(self.ns):
                    if self.names[s] != "e-":
                        alpha[r, s] = 1.0
                for s, e_ in eff:
                    alpha[r, s] = e_
            A.append(a_)
            b.append(b_)
            th.append(t_)
            tb.append(third)
            fm.append(modes[ft])
            rm.append(modes[rt])
        self.nu1, self.nu2, self.dnu = nu1, nu2, nu2 - nu1
        self.dn = self.dnu.sum(1)
        self.alpha = alpha
        self.A, self.b, self.theta = np.array(A), np.array(b), np.array(th)
        self.third = np.array(tb)
        self.fwd_mode, self.rev_mode = np.array(fm), np.array(rm)

    # -- NASA-7 (hybrid/chemistry/nasa.py:47-119), T clamped inside ----------
    def _coeffs(self, T):
        T = np.asarray(T, float)
        out = np.empty(T.shape + (self.ns, 7))
        Tc = np.empty(T.shape + (self.ns,))
        for s, (lo, mid, hi, cl, ch) in enumerate(self.nasa):
            t = np.clip(T, lo, hi)
            Tc[..., s] = t
            out[..., s, :] = np.where((t < mid)[..., None], np.asarray(cl), np.asarray(ch))
        return Tc, out

    def cp_over_R(self, T):
        t, a = self._coeffs(T)
        return a[..., 0] + a[..., 1] * t + a[..., 2] * t ** 2 + a[..., 3] * t ** 3 + a[..., 4] * t ** 4

    def h_over_RT(self, T):
        t, a = self._coeffs(T)
        return (a[..., 0] + a[..., 1] * t / 2 + a[..., 2] * t ** 2 / 3 + a[..., 3] * t ** 3 / 4
                + a[..., 4] * t ** 4 / 5 + a[..., 5] / t)

    def s_over_R(self, T):
        t, a = self._coeffs(T)
        return (a[..., 0] * np.log(t) + a[..., 1] * t + a[..., 2] * t ** 2 / 2
                + a[..., 3] * t ** 3 / 3 + a[..., 4] * t ** 4 / 4 + a[..., 6])

    def g_over_RT(self, T):
        return self.h_over_RT(T) - self.s_over_R(T)

    # -- rates --------------------------------------------------------------------
    @staticmethod
    def _select(mode, T, Tv):
        """hybrid/chemistry/mechanism.py:244-247."""
        Ta = np.sqrt(np.maximum(T * Tv, 1.0))
        return np.where(mode == 1, Ta, np.where(mode == 2, Tv, T))

    def forward_rate_constants(self, T, Tv):
        Te = np.maximum(self._select(self.fwd_mode, T, Tv), 10.0)
        return self.A * Te ** self.b * np.exp(-self.theta / Te)

    def equilibrium_constants(self, T, Tv):
        """hybrid/chemistry/mechanism.py:260-276."""
        tr = np.maximum(self._select(self.rev_mode, T, Tv), 10.0)
        out = np.empty(self.nr)
        for r in range(self.nr):
            g = self.g_over_RT(tr[r])
            dG = np.dot(self.dnu[r], g)
            out[r] = np.exp(-np.clip(dG, -700.0, 700.0)) * (P_STD / (R_UNIV * tr[r])) ** self.dn[r]
        return out

    def net_production_rates(self, rho, T, Tv=None, jacobian=False):
        """hybrid/chemistry/mechanism.py:159-220; optionally d w_dot / d rho."""
        if Tv is None:
            Tv = T
        M = self.molar_masses
        x = rho / M
        C = np.maximum(x, 0.0)
        kf = self.forward_rate_constants(T, Tv)
        Kc = self.equilibrium_constants(T, Tv)
        kr = kf / np.maximum(Kc, 1e-300)
        Cs = np.maximum(C, 1e-100)
        lc = np.log(Cs)
        lf = self.nu1 @ lc
        lr = self.nu2 @ lc
        fwd = np.exp(np.clip(lf, -700.0, 700.0))
        rev = np.exp(np.clip(lr, -700.0, 700.0))
        Meff = self.alpha @ C
        col = np.where(self.third, Meff, 1.0)
        q = kf * fwd * col - kr * rev * col
        omega = self.dnu.T @ q
        w = omega * M
        srho = np.sum(rho)
        tot = np.maximum(srho, 1e-14)
        Y = np.maximum(rho, 0.0) / tot
        S = np.sum(w)
        wdot = w - Y * S
        if not jacobian:
            return omega, wdot
        # d/d rho through the max/clip ties of JAX (half gradient at equality).
        dC = np.where(x > 0, 1.0, np.where(x == 0, 0.5, 0.0)) / M
        dlc = np.where(C > 1e-100, 1.0, np.where(C == 1e-100, 0.5, 0.0)) / Cs   # d lc / d C
        cf = np.where(np.abs(lf) < 700.0, 1.0, np.where(np.abs(lf) == 700.0, 0.5, 0.0))
        cr = np.where(np.abs(lr) < 700.0, 1.0, np.where(np.abs(lr) == 700.0, 0.5, 0.0))
        dfwd = (fwd * cf)[:, None] * self.nu1 * dlc[None, :]      # (R, S) d fwd / d C
        drev = (rev * cr)[:, None] * self.nu2 * dlc[None, :]
        dcol = np.where(self.third[:, None], self.alpha, 0.0)
        dq = (kf[:, None] * (dfwd * col[:, None] + fwd[:, None] * dcol)
              - kr[:, None] * (drev * col[:, None] + rev[:, None] * dcol))
        dq = dq * dC[None, :]                                        # d q / d rho
        dw = M[:, None] * (self.dnu.T @ dq)                          # (S, S)
        dtot = 1.0 if srho > 1e-14 else (0.5 if srho == 1e-14 else 0.0)
        dpos = np.where(rho > 0, 1.0, np.where(rho == 0, 0.5, 0.0))
        dY = np.diag(dpos / tot) - np.outer(np.maximum(rho, 0.0) / tot ** 2, np.full(self.ns, dtot))
        dwdot = dw - dY * S - np.outer(Y, dw.sum(0))
        return omega, wdot, dwdot

    def mixture_properties(self, rho, T):
        """hybrid/chemistry/mechanism.py:222-260."""
        M = self.molar_masses
        tot = np.maximum(np.sum(rho), 1e-14)
        Y = np.maximum(rho, 0.0) / tot
        mw = 1.0 / np.sum(Y / M)
        Rm = R_UNIV / mw
        cp = np.sum(Y * self.cp_over_R(T) * R_UNIV / M)
        cv = np.sum(Y * (self.cp_over_R(T) - 1.0) * R_UNIV / M)
        h = np.sum(Y * self.h_over_RT(T) * R_UNIV * T / M)
        e = np.sum(Y * (self.h_over_RT(T) - 1.0) * R_UNIV * T / M)
        gam = cp / max(cv, 1e-10)
        return dict(density=tot, mean_molecular_weight=mw, gas_constant=Rm, pressure=tot * Rm * T,
                    cp=cp, cv=cv, enthalpy=h, internal_energy=e, gamma_frozen=gam,
                    speed_of_sound=np.sqrt(gam * Rm * T))

    def recover_temperature_from_energy(self, rho, e_target, T_init=1000.0, num_iters=10):
        """hybrid/chemistry/mechanism.py:262-287."""
        M = self.molar_masses
        tot = np.maximum(np.sum(rho), 1e-14)
        Y = np.maximum(rho, 0.0) / tot
        T = float(T_init)
        for _ in range(num_iters):
            Ts = min(max(T, 50.0), 50000.0)
            e_s = (self.h_over_RT(Ts) - 1.0) * (R_UNIV * Ts / M)
            cv_s = (self.cp_over_R(Ts) - 1.0) * (R_UNIV / M)
            res = np.sum(Y * e_s) - e_target
            dT = -res / max(np.sum(Y * cv_s), 1e-5)
            T = min(max(T + min(max(dT, -0.5 * T), 0.5 * T), 100.0), 50000.0)
        return T


def integrator_step(mech, rho, T, dt, iters=16, constant_temperature=True):
    """hybrid/chemistry/integrator.py:31-104: backward Euler on the species
    with a fixed number of Newton iterations, non-negativity clip, optional
    temperature recovery at constant internal energy, and the total-mass
    rescale."""
    rho0 = np.maximum(np.asarray(rho, float), 1e-20)
    total = np.sum(rho0)
    e0 = mech.mixture_properties(rho0, T)["internal_energy"]
    r, Tk, Tvk = rho0.copy(), float(T), float(T)
    for _ in range(iters):
        _, w, dw = mech.net_production_rates(r, Tk, Tvk, jacobian=True)
        res = r - rho0 - dt * w
        J = np.eye(mech.ns) - dt * dw
        d = np.linalg.solve(J, -res)
        r = np.maximum(r + d, 0.0)
        if not constant_temperature:
            Tk = mech.recover_temperature_from_energy(r, e0, T_init=Tk)
            Tvk = Tk
    r = r * (total / np.sum(r))
    return r, Tk, Tvk


class ReactingFlowCoupler:
    """hybrid/chemistry/coupling.py:19-209 (dimensional and dimensionless)."""

    def __init__(self, mechanism=None, max_newton_iters=12, dimensionless=False,
                 reference_temperature=300.0, reference_pressure=101325.0, reference_length=1.0,
                 gamma=1.4):
        self.mech = Mechanism() if mechanism is None else mechanism
        self.iters = max_newton_iters
        self.dimensionless = dimensionless
        self.gamma = gamma
        self.T_ref = reference_temperature
        self.p_ref = reference_pressure
        mean_mw = 1.0 / (0.767 / (2 * 0.0140067) + 0.233 / (2 * 0.015999))
        self.R_ref = 8.31446261815324 / mean_mw
        self.rho_ref = self.p_ref / (self.R_ref * self.T_ref)
        self.u_ref = float(np.sqrt(self.R_ref * self.T_ref))
        self.e_ref = self.u_ref ** 2
        self.t_ref = reference_length / self.u_ref

    def react_cell(self, u, dt):
        """hybrid/chemistry/coupling.py:116-177."""
        ns = self.mech.ns
        rt = max(u[0], 1e-14)
        m = u[1:4]
        rs = np.maximum(u[5:5 + ns], 0.0)
        rs = rs * (rt / max(rs.sum(), 1e-20))
        ek = 0.5 * np.sum(m ** 2) / rt ** 2
        if self.dimensionless:
            eth_s = max(u[4] / rt - ek, 1e-4)
            eth = eth_s * self.e_ref
            rp = rs * self.rho_ref
            Tg = max((self.gamma - 1.0) * eth_s * self.T_ref, 200.0)
            T0 = self.mech.recover_temperature_from_energy(rp, eth, T_init=Tg, num_iters=8)
            rn, _, _ = integrator_step(self.mech, rp, T0, dt * self.t_ref, self.iters, False)
            rn = rn / self.rho_ref
        else:
            eth = max(u[4] / rt - ek, 1e3)
            T0 = self.mech.recover_temperature_from_energy(rs, eth, T_init=1000.0, num_iters=8)
            rn, _, _ = integrator_step(self.mech, rs, T0, dt, self.iters, False)
        rn = np.maximum(rn, 0.0)
        rn = rn * (rt / max(rn.sum(), 1e-20))
        out = np.array(u, float)
        out[5:5 + ns] = rn
        return out

    def react_block(self, U, dt):
        """hybrid/chemistry/coupling.py:179-197."""
        U = np.asarray(U, float)
        out = np.empty_like(U)
        dts = np.broadcast_to(np.asarray(dt, float).reshape(-1) if np.ndim(dt) else dt,
                              U.shape[:1])
        for e in range(U.shape[0]):
            for i in range(U.shape[1]):
                out[e, i] = self.react_cell(U[e, i], float(dts[e]) if np.ndim(dts) else dts)
        return out

    def advance(self, states, dt):
        """hybrid/chemistry/coupling.py:199-209."""
        return {k: self.react_block(v, dt) for k, v in states.items()}


def be_check(mech, rho_in, T_in, dt, rho_pred, constant_temperature=False, rtol=0.1,
             atol=1e-8):
    """The integrator's judgement of a predicted substep: the backward-Euler
    residual R = rho_p - rho_0 - dt w(rho_p, T_p) of integrator_step's scheme
    (hybrid/chemistry/integrator.py:31-104), T_p recovered from rho_p at the
    initial internal energy (10 iterations from T_in) unless the temperature
    is held, scaled by the Jacobian diagonal: est_s = |R_s| / max(|1 - dt
    dw_s/drho_s|, 1).  Returns (flag, max_s est_s / sum(rho_0)); flag where est_s >
    rtol |rho_p,s - rho_0,s| + atol sum(rho_0) for some s."""
    r0 = np.maximum(np.asarray(rho_in, float), 1e-20)
    rp = np.asarray(rho_pred, float)
    total = r0.sum()
    T = float(T_in)
    if not constant_temperature:
        e0 = mech.mixture_properties(r0, T)["internal_energy"]
        T = mech.recover_temperature_from_energy(rp, e0, T_init=T, num_iters=10)
    _, w, dw = mech.net_production_rates(rp, T, T, jacobian=True)
    R = rp - r0 - dt * w
    est = np.abs(R) / np.maximum(np.abs(1.0 - dt * np.diag(dw)), 1.0)
    tol = rtol * np.abs(rp - r0) + atol * total
    return bool(np.any(~(est <= tol))), float(est.max() / total)
