"""HybridCompressibleSolver facade (hybrid/solver.py:12-470 of the reference).

Holds the physics constants and the mesh; state blocks are dicts
{element type: torch tensor (E, N, 5 + species)} on the GPU, the layout of
hybrid/state.py:7-37.  The pointwise flux functions live in the HIP kernels;
the helpers here only build states.
"""

import os

import numpy as np
import torch

from .elements import reference_tables
from .gmsh import read_gmsh
from .meshgen import MeshArrays
from .topology import Topology


class CompressibleStateLayout:
    """hybrid/state.py:7-37."""

    def __init__(self, species_names=()):
        self.species_names = tuple(species_names)
        self.density = 0
        self.total_energy = 4

    @property
    def momentum(self):
        return slice(1, 4)

    @property
    def species(self):
        return slice(5, 5 + len(self.species_names))

    @property
    def number_of_variables(self):
        return 5 + len(self.species_names)

    def index(self, name):
        try:
            return 5 + self.species_names.index(name)
        except ValueError as error:
            raise KeyError(f"Unknown species {name!r}") from error


class HybridCompressibleSolver:
    def __init__(self, mesh, species_names=(), gamma=1.4, gas_constant=1.0,
                 polynomial_degree=2, inviscid_flux="hllc", dynamic_viscosity=0.0,
                 prandtl=0.72, use_wale=False, wale_constant=0.5, turbulent_prandtl=0.9,
                 dtype=torch.float64, device="cuda"):
        if isinstance(mesh, (str, os.PathLike)):
            # the reference's mesh_path argument (hybrid/solver.py:20-35)
            mesh = read_gmsh(mesh)
        if isinstance(mesh, MeshArrays):
            mesh = Topology(mesh)
        if not isinstance(mesh, Topology):
            raise TypeError("mesh must be a Gmsh file path, MeshArrays or Topology")
        if inviscid_flux not in ("hllc", "rusanov"):
            raise ValueError("inviscid_flux must be 'hllc' or 'rusanov'")
        self.mesh = mesh
        self.layout = CompressibleStateLayout(species_names)
        self.gamma = float(gamma)
        self.gas_constant = float(gas_constant)
        self.polynomial_degree = int(polynomial_degree)
        self.inviscid_flux = inviscid_flux
        self.dynamic_viscosity = float(dynamic_viscosity)
        self.prandtl = float(prandtl)
        # hybrid/solver.py:30-47: WALE eddy viscosity in the viscous term
        self.use_wale = bool(use_wale)
        self.wale_constant = float(wale_constant)
        self.turbulent_prandtl = float(turbulent_prandtl)
        # hybrid/solver.py:48-53
        self.specific_heat_cp = self.gamma * self.gas_constant / (self.gamma - 1.0)
        self.thermal_conductivity = (self.dynamic_viscosity * self.specific_heat_cp
                                     / self.prandtl)
        self.dtype = dtype
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev

    # -- state helpers (hybrid/solver.py:63-98, 421-470) -------------------
    def primitive_to_conserved(self, W):
        rho = W[..., 0:1]
        vel = W[..., 1:4]
        ke = (vel * vel).sum(-1)[..., None]
        E = W[..., 4:5] / (self.gamma - 1.0) + 0.5 * rho * ke
        parts = [rho, rho * vel, E]
        if self.layout.species_names:
            parts.append(rho * W[..., 5:])
        return torch.cat(parts, -1) if torch.is_tensor(W) else np.concatenate(parts, -1)

    def conserved_to_primitive(self, U):
        rho = U[..., 0:1]
        safe = rho.clamp_min(1e-14) if torch.is_tensor(U) else np.maximum(rho, 1e-14)
        vel = U[..., 1:4] / safe
        ke = 0.5 * (U[..., 1:4] * U[..., 1:4]).sum(-1)[..., None] / safe
        p = (self.gamma - 1.0) * (U[..., 4:5] - ke)
        parts = [rho, vel, p]
        if self.layout.species_names:
            parts.append(U[..., 5:] / safe)
        return torch.cat(parts, -1) if torch.is_tensor(U) else np.concatenate(parts, -1)

    def solution_node_coordinates(self):
        """Physical node coordinates per block (hybrid/geometry.py:115-137)."""
        out = {}
        for b in self.mesh.blocks:
            xi = reference_tables(b.element_type, self.polynomial_degree)["nodes"]
            out[b.element_type] = b.coords[:, :1, :] + np.einsum("edr,qr->eqd", b.jac, xi)
        return out

    def to_device(self, states):
        return {k: torch.as_tensor(np.ascontiguousarray(v), dtype=self.dtype,
                                   device=self.device) for k, v in states.items()}

    def uniform_state_blocks(self, density, velocity, pressure, mass_fractions=None):
        W = [density, *velocity, pressure]
        if self.layout.species_names:
            W.extend(mass_fractions)
        U = self.primitive_to_conserved(np.asarray(W, float))
        out = {}
        for b in self.mesh.blocks:
            N = reference_tables(b.element_type, self.polynomial_degree)["N"]
            out[b.element_type] = np.broadcast_to(U, (b.count, N, U.size)).copy()
        return self.to_device(out)

    def state_blocks_from_primitive_function(self, function):
        return self.to_device({k: self.primitive_to_conserved(np.asarray(function(x)))
                               for k, x in self.solution_node_coordinates().items()})
