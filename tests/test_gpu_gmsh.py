"""A solver built from a Gmsh file path (the reference's mesh_path argument,
hybrid/solver.py:20-35) runs the same operator as one built from the mesh in
memory: a split box with inlet, outlet and wall tags is written as MSH 4.1
binary and MSH 2.2 ASCII, and the rhs from the path equals the in-memory rhs
bit for bit and the oracle's on the read mesh at the FP64 bar (1e-12)."""

import numpy as np
import pytest
import torch

from conftest import PLANE_NAMES, perturbed_primitive, rel_l2

from hybrid import (FarfieldWallBoundary, HybridCompressibleSolver, HybridNavierStokesOperator,
                    box_mesh, read_gmsh, write_gmsh)
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver, FarfieldWallBoundary as OracleBoundary

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("version,binary", [("4.1", True), ("2.2", False)])
def test_rhs_from_a_gmsh_path(tmp_path, version, binary):
    print(f"gmsh {version} binary={binary}", flush=True)
    m = box_mesh(3, split_fraction=0.4, seed=3, plane_names=PLANE_NAMES)
    path = tmp_path / "box.msh"
    write_gmsh(path, m, version=version, binary=binary)
    far = np.array([1.2, 0.9, 0.1, -0.05, 1.1])
    out = []
    for mesh in (str(path), m):
        ds = HybridCompressibleSolver(mesh, polynomial_degree=2, dynamic_viscosity=0.01)
        op = HybridNavierStokesOperator(ds)
        db = FarfieldWallBoundary(ds, far)
        xs = ds.solution_node_coordinates()
        st = {t: ds.primitive_to_conserved(perturbed_primitive(x, seed=2)) for t, x in xs.items()}
        out.append((op.rhs(ds.to_device(st), db), st, db))
    for t in out[0][0]:
        assert torch.equal(out[0][0][t], out[1][0][t]), t
    got, st, db = out[0]
    os_ = CompressibleSolver(dynamic_viscosity=0.01)
    oop = NavierStokesOperator(from_arrays(read_gmsh(path)), os_, 2)
    ob = OracleBoundary(os_, far, inlet_tags=sorted(db.inlet_tags),
                        outlet_tags=sorted(db.outlet_tags), wall_tags=sorted(db.wall_tags),
                        no_slip_tags=sorted(db.no_slip_tags), slip_tags=sorted(db.slip_tags))
    ref = oop.rhs(st, ob)
    err = rel_l2({k: v.cpu().numpy() for k, v in got.items()}, ref)
    assert np.all(err <= 1e-12), err
