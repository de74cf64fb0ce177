"""Thermochemistry on the GPU, mirroring hybrid/chemistry of the reference.

``Air11ParkMechanism`` uploads the Park-11 table once; ``PointImplicitIntegrator``
(hybrid/chemistry/integrator.py:9-104) and ``ReactingFlowCoupler``
(hybrid/chemistry/coupling.py:19-209) run the FP64 kernels of
libfluid_gfx950.so on batches of cells.
"""

import ctypes

import numpy as np
import torch

from .. import _native as nat
from .air11_park import CHARGES, MOLAR_MASSES, SPECIES_NAMES, mechanism_table


class Air11ParkMechanism:
    """hybrid/chemistry/air11_park.py:314-318."""

    def __init__(self, device="cuda"):
        self.species_names = SPECIES_NAMES
        self.num_species = 11
        self.num_reactions = 22
        self.molar_masses = np.asarray(MOLAR_MASSES)
        self.charges = np.asarray(CHARGES)
        table = mechanism_table()
        if len(table) != nat.lib().fsg_chem_table_size():
            raise RuntimeError("mechanism table layout does not match the library")
        self.table = torch.as_tensor(table, dtype=torch.float64, device=device)
        self.device = self.table.device


class PointImplicitIntegrator:
    """hybrid/chemistry/integrator.py:9-104 over a batch of cells."""

    def __init__(self, mechanism, relaxation=None, max_newton_iters=16, tolerance=1.0e-8):
        if relaxation is not None:
            raise NotImplementedError("vibrational relaxation is not used by step()")
        self.mechanism = mechanism
        self.max_newton_iters = int(max_newton_iters)
        self.tolerance = tolerance
        self.num_species = mechanism.num_species

    def step(self, rho_species, T, dt, Tv=None, constant_temperature=True):
        """rho_species (..., 11), T (...), dt scalar or (...); returns
        (rho, T, Tv) like the reference (Tv == T)."""
        if Tv is not None:
            raise NotImplementedError("a separate initial Tv is not supported")
        dev = self.mechanism.device
        rho = torch.as_tensor(rho_species, dtype=torch.float64, device=dev)
        shape = rho.shape[:-1]
        rho = rho.reshape(-1, 11).contiguous()
        n = rho.shape[0]
        Tt = torch.as_tensor(T, dtype=torch.float64, device=dev).expand(shape).reshape(-1).contiguous()
        dtt = None
        dts = 0.0
        if torch.is_tensor(dt) and dt.numel() > 1:
            dtt = dt.to(dev, torch.float64).expand(shape).reshape(-1).contiguous()
        else:
            dts = float(dt)
        out = torch.empty_like(rho)
        Tout = torch.empty_like(Tt)
        nat.check_chem(nat.lib().fsg_chem_integrate(
            nat.ptr(self.mechanism.table), nat.ptr(rho), nat.ptr(Tt), nat.ptr(dtt), dts,
            self.max_newton_iters, int(bool(constant_temperature)), nat.ptr(out), nat.ptr(Tout),
            n, nat.stream_handle(dev)))
        out = out.reshape(shape + (11,))
        Tout = Tout.reshape(shape)
        return out, Tout, Tout


class ReactingFlowCoupler:
    """hybrid/chemistry/coupling.py:19-209.

    With ``surrogate`` (a trained ``ChemistrySurrogate``) the integrator
    substep is replaced by the surrogate with its integrator fallback; the
    species clip/rescale and temperature recovery around it are the
    coupler's own (coupling.py:116-177).  The flow state may be FP64 or FP32;
    chemistry always runs in FP64."""

    def __init__(self, mechanism=None, relaxation=None, max_newton_iters=12, tolerance=1.0e-7,
                 dimensionless=False, reference_temperature=300.0,
                 reference_pressure=101325.0, reference_length=1.0, gamma=1.4, surrogate=None):
        self.mechanism = Air11ParkMechanism() if mechanism is None else mechanism
        self.integrator = PointImplicitIntegrator(self.mechanism, relaxation, max_newton_iters,
                                                  tolerance)
        self.surrogate = surrogate
        self.last_fallback = None
        # FP64 states without a surrogate run the whole cell update in one
        # kernel; False sends them through prepare -> integrator -> finish
        # (same arithmetic).
        self.fused = True
        self.num_species = 11
        self.species_names = SPECIES_NAMES
        self.dimensionless = bool(dimensionless)
        self.gamma = float(gamma)
        # Non-dimensional scales, hybrid/chemistry/coupling.py:55-67.
        self.T_ref = float(reference_temperature)
        self.p_ref = float(reference_pressure)
        self.L_ref = float(reference_length)
        mean_mw = 1.0 / (0.767 / (2.0 * 0.0140067) + 0.233 / (2.0 * 0.015999))
        self.R_ref = 8.31446261815324 / mean_mw
        self.rho_ref = self.p_ref / (self.R_ref * self.T_ref)
        self.u_ref = float(np.sqrt(self.R_ref * self.T_ref))
        self.e_ref = float(self.u_ref ** 2)
        self.t_ref = float(self.L_ref / self.u_ref)

    def _params(self):
        p = nat.ChemParams()
        p.newton_iters = self.integrator.max_newton_iters
        p.dimensionless = int(self.dimensionless)
        p.gamma, p.T_ref, p.rho_ref, p.e_ref, p.t_ref = (
            self.gamma, self.T_ref, self.rho_ref, self.e_ref, self.t_ref)
        return p

    def react_block(self, U_block, dt):
        """(E, N, 16) -> (E, N, 16): species advanced, rho/momentum/energy kept."""
        if U_block.shape[-1] < 5 + self.num_species:
            raise ValueError(f"State has {U_block.shape[-1]} variables; expected at least "
                             f"{5 + self.num_species} for reacting mixture.")
        if U_block.shape[-1] != 5 + self.num_species:
            raise ValueError("the device coupler takes exactly 5 + 11 variables")
        if U_block.dtype not in (torch.float64, torch.float32):
            raise ValueError("flow state must be float64 or float32")
        out = U_block.contiguous().clone()
        E, N = out.shape[0], out.shape[1]
        dte = None
        dts = 0.0
        if torch.is_tensor(dt) and dt.numel() > 1:
            dte = dt.to(out.device, torch.float64).reshape(-1).contiguous()
        else:
            dts = float(dt)
        p = self._params()
        st = nat.stream_handle(out.device)
        if self.surrogate is None and out.dtype == torch.float64 and self.fused:
            nat.check_chem(nat.lib().fsg_chem_react(
                nat.ptr(self.mechanism.table), nat.ptr(out), E * N, N, nat.ptr(dte), dts,
                ctypes.byref(p), st))
            return out
        # prepare -> substep -> finish (coupling.py:116-154, the substep, :165-177)
        n = E * N
        dtype = nat.F64 if out.dtype == torch.float64 else nat.F32
        rho = torch.empty(n, 11, dtype=torch.float64, device=out.device)
        T0 = torch.empty(n, dtype=torch.float64, device=out.device)
        nat.check_chem(nat.lib().fsg_chem_prepare(nat.ptr(self.mechanism.table), nat.ptr(out),
                                                  dtype, n, ctypes.byref(p), nat.ptr(rho),
                                                  nat.ptr(T0), st))
        scale = self.t_ref if self.dimensionless else 1.0
        if dte is not None:
            h = (dte * scale).repeat_interleave(N)
        else:
            h = dts * scale
        if self.surrogate is None:
            new, _, _ = self.integrator.step(rho, T0, h, constant_temperature=False)
        else:
            new, _, fb = self.surrogate.step(rho, T0, h)
            self.last_fallback = fb
        new = new.contiguous()
        nat.check_chem(nat.lib().fsg_chem_finish(nat.ptr(out), dtype, n, ctypes.byref(p),
                                                 nat.ptr(new), st))
        return out

    def advance(self, states, dt):
        return {k: self.react_block(v, dt) for k, v in states.items()}
