"""DG operators backed by the gfx950 kernels.

HybridEulerOperator mirrors hybrid/spatial.py:10-249 and
HybridNavierStokesOperator mirrors hybrid/navier_stokes.py:9-352: the same
method names, argument meaning and return layout (dict of (E, N, V) arrays),
with every evaluation a call into libfluid_gfx950.so.
"""

import ctypes

import torch

from . import _native as nat
from .boundary import no_boundary
from .device import DeviceMesh


def polynomial_br2_penalty(degree):
    """hybrid/navier_stokes.py:350-352."""
    return (degree + 1) * (degree + 2) / 2


class HybridEulerOperator:
    def __init__(self, solver, br2_stabilization=None):
        self.solver = solver
        self.mesh = solver.mesh
        self.polynomial_degree = solver.polynomial_degree
        self.br2_stabilization = float(polynomial_br2_penalty(solver.polynomial_degree)
                                       if br2_stabilization is None else br2_stabilization)
        p = nat.Params()
        p.gamma = solver.gamma
        p.gas_constant = solver.gas_constant
        p.dynamic_viscosity = solver.dynamic_viscosity
        p.thermal_conductivity = solver.thermal_conductivity
        p.br2_penalty = self.br2_stabilization
        p.inviscid_flux = 0 if solver.inviscid_flux == "hllc" else 1
        p.nvars = solver.layout.number_of_variables
        p.degree = solver.polynomial_degree
        p.dtype = nat.F64 if solver.dtype == torch.float64 else nat.F32
        p.use_wale = int(getattr(solver, "use_wale", False))
        p.wale_constant = getattr(solver, "wale_constant", 0.5)
        p.turbulent_prandtl = getattr(solver, "turbulent_prandtl", 0.9)
        self.device_mesh = DeviceMesh(solver.mesh, solver.polynomial_degree, p.nvars, p,
                                      solver.dtype, solver.device)
        self.block_types = [b.element_type for b in solver.mesh.blocks]
        self._boundary_key = None

    # -- plumbing ------------------------------------------------------------
    def _boundary(self, boundary_state):
        if boundary_state is None:
            return no_boundary()
        if self._boundary_key is not boundary_state:
            self.device_mesh.set_boundary_flags(boundary_state.flags_of_tag)
            self._boundary_key = boundary_state
        return boundary_state.native()

    def _inputs(self, states):
        ts = []
        for t in self.block_types:
            U = states[t]
            if (U.dtype != self.solver.dtype or not U.is_contiguous()
                    or U.device != self.solver.device):
                raise ValueError(f"{t} state must be a contiguous {self.solver.dtype} device tensor")
            ts.append(U)
        return ts

    def _call(self, states, mode, terms, boundary_state):
        U = self._inputs(states)
        out = [torch.empty_like(u) for u in U]
        b = self._boundary(boundary_state)
        nat.check(nat.lib().fsg_residual(self.device_mesh.ctx, nat.ptr_array(U),
                                         nat.ptr_array(out), mode, terms, ctypes.byref(b),
                                         nat.stream_handle(self.solver.device)))
        return dict(zip(self.block_types, out))

    def rhs_terms(self):
        return nat.TERM_EULER

    def xla_opaque(self, boundary_state=None, mode=nat.MODE_RHS, terms=None):
        """The opaque string of the fsg_xla_rhs custom call for this operator
        (the boundary struct it names is kept alive by the operator)."""
        b = self._boundary(boundary_state)
        self._xla_boundary = b
        t = self.rhs_terms() if terms is None else int(terms)
        return (f"ctx={self.device_mesh.ctx.value:x};bnd={ctypes.addressof(b):x};mode={int(mode)};"
                f"terms={t};nblocks={len(self.block_types)}").encode()

    # -- reference API -----------------------------------------------------------
    def residual(self, states, boundary_state=None):
        """Weak-form Euler residual before M^-1 (hybrid/spatial.py:137-234)."""
        return self._call(states, nat.MODE_RESIDUAL, nat.TERM_EULER, boundary_state)

    def rhs(self, states, boundary_state=None):
        """M^-1 residual (hybrid/spatial.py:236-249)."""
        return self._call(states, nat.MODE_RHS, nat.TERM_EULER, boundary_state)


class HybridNavierStokesOperator(HybridEulerOperator):
    """hybrid/navier_stokes.py:9-347: BR2 viscous term, with the WALE eddy
    viscosity of the resolved gradient when the solver has use_wale."""

    def viscous_residual(self, states, boundary_state=None):
        return self._call(states, nat.MODE_RESIDUAL, nat.TERM_VISCOUS, boundary_state)

    def viscous_rhs(self, states, boundary_state=None):
        return self._call(states, nat.MODE_RHS, nat.TERM_VISCOUS, boundary_state)

    def inviscid_rhs(self, states, boundary_state=None):
        return HybridEulerOperator.rhs(self, states, boundary_state)

    def rhs_terms(self):
        """Terms of rhs(): the viscous part is skipped when mu == 0 and WALE
        is off (hybrid/navier_stokes.py:340-341)."""
        if self.solver.dynamic_viscosity == 0.0 and not self.solver.use_wale:
            return nat.TERM_EULER
        return nat.TERM_EULER | nat.TERM_VISCOUS

    def rhs(self, states, boundary_state=None):
        return self._call(states, nat.MODE_RHS, self.rhs_terms(), boundary_state)

    def viscous_rhs_masked(self, states, masks, boundary_state=None):
        """viscous_rhs evaluated on the elements whose int32 mask entry is
        non-zero (bitwise equal to viscous_rhs there; other entries are
        unspecified)."""
        U = self._inputs(states)
        out = [torch.empty_like(u) for u in U]
        ms = []
        for u, t in zip(U, self.block_types):
            m = masks[t]
            if m.dtype != torch.int32 or m.numel() != u.shape[0] or not m.is_contiguous():
                raise ValueError(f"{t} mask must be a contiguous int32 tensor of {u.shape[0]}")
            ms.append(m)
        b = self._boundary(boundary_state)
        nat.check(nat.lib().fsg_residual_masked(
            self.device_mesh.ctx, nat.ptr_array(U), nat.ptr_array(out), nat.MODE_RHS,
            nat.TERM_VISCOUS, nat.ptr_array(ms), ctypes.byref(b),
            nat.stream_handle(self.solver.device)))
        return dict(zip(self.block_types, out))
