"""FarfieldWallBoundary facade (hybrid/boundary.py:6-102 of the reference).

The ghost states themselves are computed in the face kernel; this object maps
physical tag names to the per-face flags the kernel reads and carries the
farfield primitive state and the optional wall temperature.
"""

import ctypes

import numpy as np

from . import _native as nat


class FarfieldWallBoundary:
    def __init__(self, solver, farfield_primitive, inlet_names=("inlet",),
                 outlet_names=("outlet",),
                 wall_names=("walls", "sphere_wall", "cylinder_wall", "airfoil_wall", "symmetry"),
                 no_slip_names=("sphere_wall", "cylinder_wall", "airfoil_wall"),
                 slip_names=("walls", "symmetry"), wall_temperature=None):
        self.solver = solver
        self.farfield_primitive = np.asarray(farfield_primitive, float)
        if self.farfield_primitive.shape != (solver.layout.number_of_variables,):
            raise ValueError("farfield_primitive has the wrong state size")
        self.wall_temperature = None if wall_temperature is None else float(wall_temperature)
        names = solver.mesh.mesh.physical_names

        def tags(ns):
            return {names[n] for n in ns if n in names}

        self.inlet_tags = tags(inlet_names)
        self.outlet_tags = tags(outlet_names)
        self.wall_tags = tags(wall_names)
        self.no_slip_tags = tags(no_slip_names)
        self.slip_tags = tags(slip_names)

    def flags_of_tag(self, tag):
        f = 0
        if tag in self.wall_tags:
            f |= nat.BND_WALL
        if tag in self.inlet_tags:
            f |= nat.BND_INLET
        if tag in self.no_slip_tags:
            f |= nat.BND_NOSLIP
        if tag in self.slip_tags:
            f |= nat.BND_SLIP
        return f

    def native(self):
        b = nat.Boundary()
        b.has_boundary = 1
        b.has_wall_temperature = int(self.wall_temperature is not None)
        b.wall_temperature = 0.0 if self.wall_temperature is None else self.wall_temperature
        for i, v in enumerate(self.farfield_primitive):
            b.farfield[i] = float(v)
        return b


def no_boundary():
    """boundary_state=None: every boundary face copies the owner trace."""
    return nat.Boundary()
