"""The product's subcell tables (hybrid/subcells.py) against the oracle's
restatement of hybrid/subcell.py: identical subcell / subface numbering
(the reference indexes subcell arrays with subface numbers, so the numbers
are behaviour), projection and reconstruction operators, and subface
pairings across every face of a mixed mesh with sub-faces; and the MOOD
device tables build with one face entry per element face."""

import numpy as np
import pytest

from conftest import PLANE_NAMES

from hybrid import HybridCompressibleSolver, HybridNavierStokesOperator, box_mesh
from hybrid.mood import MAX_FACES, MoodTables
from hybrid.subcells import pair_subfaces, subcell_tables
from hybrid.topology import Topology
from oracle.mesh import from_arrays
from oracle.operators import NavierStokesOperator
from oracle.solver import CompressibleSolver
from oracle.subcell import ReferenceSubcells, SubcellFV


@pytest.mark.parametrize("t", ["hexahedron", "wedge", "tetra"])
def test_numbering_matches(t):
    a, b = subcell_tables(t, 2), ReferenceSubcells(t)
    assert np.allclose(a.vertices, b.vertices) and np.array_equal(a.cells, b.cells)
    assert np.array_equal(a.internal_owner, b.internal_owner)
    assert np.array_equal(a.internal_neighbor, b.internal_neighbor)
    assert np.array_equal(a.boundary_owner, b.boundary_owner)
    assert np.array_equal(a.boundary_parent_face, b.boundary_parent_face)


@pytest.mark.parametrize("p", [2, 3])
def test_projection_and_pairings_match(p):
    m = box_mesh(3, periodic=(True,) * 3, split_fraction=0.5, seed=3)
    op = NavierStokesOperator(from_arrays(m), CompressibleSolver(), p)
    fv = SubcellFV(op, indexing="owner")
    for t, pj in fv.projections.items():
        a = subcell_tables(t, p)
        assert np.abs(a.P - pj.projection[0]).max() < 1e-13
        assert np.abs(a.R - pj.reconstruction[0]).max() < 1e-11
        assert np.abs(a.volume / a.volume.sum() - pj.volumes[0] / pj.volumes[0].sum()).max() < 1e-15
    topo = Topology(m)
    f = topo.faces
    types = [b.element_type for b in topo.blocks]
    n = 0
    for g in fv.face_groups:
        fb = fv.mesh.face_blocks[g["ftype"]]
        for r, fi in enumerate(g["idx"]):
            ob = g["ob"]
            k = np.flatnonzero((f.owner_block == ob) & (f.owner_local == fb.owner_local[fi])
                               & (f.owner_face == fb.owner_face[fi])
                               & (f.face_type == (0 if g["ftype"] == "quad" else 1))
                               & (f.neighbor_block == g["nb"])
                               & ((f.neighbor_local == fb.neighbor_local[fi]) if g["nb"] >= 0
                                  else True))
            assert len(k) == 1
            k = k[0]
            tn = types[f.neighbor_block[k]] if f.neighbor_block[k] >= 0 else None
            pr = pair_subfaces(types[ob], int(f.owner_face[k]), f.owner_corners[k], tn,
                               int(f.neighbor_face[k]), f.neighbor_corners[k], p, bool(f.sub[k]))
            assert np.array_equal(pr["osub"], g["osub"][r])
            assert np.allclose(pr["opt"], g["opt"][r], atol=1e-12)
            if tn is not None:
                assert np.array_equal(pr["nsub"], g["nsub"][r])
                assert np.allclose(pr["npt"], g["npt"][r], atol=1e-12)
            n += 1
    assert n == sum(len(g["idx"]) for g in fv.face_groups)


def test_mood_tables_build_with_all_face_entries():
    import ctypes
    m = box_mesh(3, split_fraction=0.5, seed=3, plane_names=PLANE_NAMES)
    s = HybridCompressibleSolver(m, polynomial_degree=2, dynamic_viscosity=0.01, device="cpu")
    mt = MoodTables(HybridNavierStokesOperator(s), "reference")
    f = s.mesh.faces
    for b, blk in enumerate(s.mesh.blocks):
        raw = (ctypes.c_int32 * (blk.count * MAX_FACES * 2)).from_address(mt.blocks[b].elem_faces)
        ef = np.ctypeslib.as_array(raw).reshape(blk.count, MAX_FACES, 2)
        have = (ef[:, :, 0] >= 0).sum(1)
        want = (np.bincount(f.owner_local[f.owner_block == b], minlength=blk.count)
                + np.bincount(f.neighbor_local[f.neighbor_block == b], minlength=blk.count))
        assert np.array_equal(have, want)
        assert mt.types[["hexahedron", "wedge", "tetra"].index(blk.element_type)].N > 0
