"""MI355X-native hybrid DG hot path.

Host mirror of the reference's ``hybrid`` package for the residual / RK-stage
path: the classes keep the reference names and signatures and call the gfx950
kernels of libfluid_gfx950.so.
"""

from .boundary import FarfieldWallBoundary
from .gmsh import read_gmsh, write_gmsh
from .meshgen import box_mesh, smoke_mesh
from .mood import (HybridModalShockLimiter, HybridMOODSubcellLimiter, HybridPositivityLimiter,
                   HybridTroubledCellSensor)
from .solver import CompressibleStateLayout, HybridCompressibleSolver
from .spatial import HybridEulerOperator, HybridNavierStokesOperator
from .time_integration import HybridSSPRK3
from .topology import Topology

__all__ = ["FarfieldWallBoundary", "box_mesh", "smoke_mesh", "read_gmsh", "write_gmsh",
           "CompressibleStateLayout", "HybridCompressibleSolver", "HybridEulerOperator", "HybridNavierStokesOperator",
           "HybridSSPRK3", "Topology", "HybridMOODSubcellLimiter", "HybridPositivityLimiter",
           "HybridTroubledCellSensor", "HybridModalShockLimiter"]
