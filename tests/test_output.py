"""Restart files and VTK output (run_hybrid_sphere.py:36-45, 123-131;
hybrid/output.py:274-441, 743-757): the restart archive keeps the
reference's keys and round-trips bitwise; the linear sub-cells tile each
reference element (their volumes sum to it, all positively oriented); the
tessellated file carries every node's primitive state; the midplane cut of
a box covers the box's cross-section and interpolates a linear field
exactly."""

import numpy as np
import pytest
import torch

from hybrid import HybridCompressibleSolver, HybridNavierStokesOperator, box_mesh
from hybrid.elements import nodes
from hybrid.output import (VTK_HEXAHEDRON, VTK_TETRA, VTK_WEDGE, load_restart,
                           midplane_polygons, read_vtu, save_restart, sub_cells,
                           write_midplane_slice_vtu, write_pvd, write_tessellated_vtu)

REF_VOLUME = {"hexahedron": 1.0, "wedge": 0.5, "tetra": 1.0 / 6.0}


def _volume(vt, q):
    if vt == VTK_TETRA:
        return np.dot(np.cross(q[1] - q[0], q[2] - q[0]), q[3] - q[0]) / 6.0
    if vt == VTK_WEDGE:                       # three tetrahedra of the prism
        t = [(0, 1, 2, 3), (1, 2, 3, 4), (2, 3, 4, 5)]
        return -sum(_volume(VTK_TETRA, q[list(k)]) for k in t)
    t = [(0, 1, 3, 4), (1, 2, 3, 6), (1, 3, 4, 6), (1, 4, 5, 6), (3, 4, 6, 7)]
    return sum(_volume(VTK_TETRA, q[list(k)]) for k in t)


@pytest.mark.parametrize("t", ["hexahedron", "wedge", "tetra"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_sub_cells_tile_the_reference_element(t, p):
    vt, cells = sub_cells(t, p)
    assert len(cells) == p ** 3
    x = nodes(t, p)
    vols = np.array([_volume(vt, x[c]) for c in cells])
    assert np.all(vols > 0), vols
    assert abs(vols.sum() - REF_VOLUME[t]) < 1e-13
    assert set(np.unique(cells)) == set(range(len(x)))


def _solver(species=()):
    m = box_mesh(3, lengths=(1.0, 1.5, 2.0), origin=(0.0, 0.0, -1.0), split_fraction=0.4,
                 seed=2)
    return HybridCompressibleSolver(m, species_names=species, polynomial_degree=2,
                                    dynamic_viscosity=0.01, device="cpu")


def _linear_state(s):
    def prim(x):
        W = np.zeros(x.shape[:-1] + (5 + len(s.layout.species_names),))
        W[..., 0] = 1.0 + 0.1 * x[..., 0]
        W[..., 1] = 0.2 * x[..., 1]
        W[..., 4] = 1.0 + 0.05 * x[..., 2]
        if s.layout.species_names:
            W[..., 5:] = 1.0 / len(s.layout.species_names)
        return W
    return s.state_blocks_from_primitive_function(prim)


def test_restart_round_trip(tmp_path):
    s = _solver()
    U = _linear_state(s)
    path = save_restart(tmp_path / "restart.npz", U, 0.125, 7)
    with np.load(path, allow_pickle=False) as r:
        assert sorted(r.files) == sorted(["time", "step"] + [f"U_{t}" for t in U])
        assert r["U_hexahedron"].dtype == np.float64
    V, t, step = load_restart(path, s)
    assert (t, step) == (0.125, 7)
    for k in U:
        assert torch.equal(U[k], V[k])
    assert not (tmp_path / "restart.tmp.npz").exists()


def test_tessellated_vtu(tmp_path):
    from hybrid.chemistry.air11_park import SPECIES_NAMES
    s = _solver(species=SPECIES_NAMES)
    U = _linear_state(s)
    path = write_tessellated_vtu(tmp_path / "f.vtu", HybridNavierStokesOperator(s), U)
    pts, conn, offs, types, data = read_vtu(path)
    E = {b.element_type: b.count for b in s.mesh.blocks}
    N = {"hexahedron": 27, "wedge": 18, "tetra": 10}
    assert len(pts) == sum(E[t] * N[t] for t in E)
    assert len(types) == 8 * sum(E.values())
    assert set(np.unique(types)) == {VTK_HEXAHEDRON, VTK_WEDGE, VTK_TETRA}
    assert offs[-1] == len(conn)
    assert np.allclose(data["density"], 1.0 + 0.1 * pts[:, 0], rtol=1e-14)
    assert np.allclose(data["velocity"][:, 0], 0.2 * pts[:, 1], rtol=1e-13, atol=1e-15)
    assert np.all(data["velocity"][:, 1:] == 0.0)
    assert np.allclose(data["Y_O_plus"], 1.0 / 11.0, rtol=1e-14)
    assert np.allclose(data["Y_e_minus"], 1.0 / 11.0, rtol=1e-14)
    # every sub-cell positively oriented in physical space: the mesh volume
    vol = 0.0
    start = 0
    for o, vt in zip(offs, types):
        vol += _volume(vt, pts[conn[start:o]])
        start = o
    assert abs(vol - 1.0 * 1.5 * 2.0) < 1e-12


def test_midplane_slice(tmp_path):
    s = _solver()
    op = HybridNavierStokesOperator(s)
    U = _linear_state(s)
    for z in (0.0, -0.3):
        v, polys, W = midplane_polygons(op, U, plane_z=z)
        area = 0.0
        for q in polys:
            xy = v[q, :2]
            area += 0.5 * abs(np.dot(xy[:, 0], np.roll(xy[:, 1], -1))
                              - np.dot(xy[:, 1], np.roll(xy[:, 0], -1)))
        assert abs(area - 1.5) < 1e-10, (z, area)
        assert np.allclose(v[:, 2], z, atol=1e-12)
        assert np.allclose(W[:, 0], 1.0 + 0.1 * v[:, 0], rtol=1e-13)
        assert np.allclose(W[:, 4], 1.0 + 0.05 * z, rtol=1e-13)
    path = write_midplane_slice_vtu(tmp_path / "mid.vtu", op, U)
    pts, conn, offs, types, data = read_vtu(path)
    assert np.all(types == 7) and len(data["pressure"]) == len(pts)
    pvd = write_pvd(tmp_path / "run.pvd", [(0.0, path), (0.5, tmp_path / "mid2.vtu")])
    text = pvd.read_text()
    assert 'timestep="0.5"' in text and 'file="mid.vtu"' in text
