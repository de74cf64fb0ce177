"""The Gmsh reader (hybrid/mesh.py:186-352 of the reference reads the same
files through meshio): hand-written MSH 2.2 and 4.1 files of two tetrahedra
(node tags not contiguous in 4.1), straight and curved second-order cells,
the tag validation of hybrid/mesh.py:334-352, 466-483, round trips of a split
box mesh through all four formats, and, where the reference's meshes are
present, its Taylor-Green (MSH 4.1 ASCII) and SR-71 (MSH 2.2 binary) meshes
against counts taken from the files and the box volume (2 pi)^3."""

import pathlib

import numpy as np
import pytest

from hybrid import box_mesh, read_gmsh, write_gmsh
from hybrid.elements import LOCAL_FACES
from hybrid.topology import Topology

MESHES = pathlib.Path("/root/reference/mesh")

NODES = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)]
# two positively oriented tetrahedra sharing the face (1, 2, 3); the z = 0
# face (0, 1, 2) is "bottom", the other four boundary faces "walls"
TETS = [(0, 1, 2, 3), (1, 2, 3, 4)]
BOUNDARY = {(0, 1, 2): 11, (0, 1, 3): 13, (0, 2, 3): 13, (1, 2, 4): 13, (1, 3, 4): 13,
            (2, 3, 4): 13}
NAMES = '$PhysicalNames\n3\n2 11 "bottom"\n2 13 "walls"\n3 1 "fluid"\n$EndPhysicalNames\n'

MSH41 = ("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n" + NAMES +
         "$Entities\n0 0 2 1\n1 0 0 0 1 1 0 1 11 0\n2 0 0 0 1 1 1 1 13 0\n"
         "1 0 0 0 1 1 1 1 1 0\n$EndEntities\n"
         "$Nodes\n1 5 10 50\n3 1 0 5\n10\n20\n30\n40\n50\n"
         "0 0 0\n1 0 0\n0 1 0\n0 0 1\n1 1 1\n$EndNodes\n"
         "$Elements\n3 8 1 8\n2 1 2 1\n1 10 20 30\n2 2 2 5\n2 10 20 40\n3 10 30 40\n"
         "4 20 30 50\n5 20 40 50\n6 30 40 50\n3 1 4 2\n7 10 20 30 40\n8 20 30 40 50\n"
         "$EndElements\n")


def _msh22(tets=TETS, nodes=NODES, boundary=BOUNDARY, tet_type=4):
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", NAMES.strip(), "$Nodes",
             str(len(nodes))]
    lines += [f"{i + 1} {x} {y} {z}" for i, (x, y, z) in enumerate(nodes)]
    lines += ["$EndNodes", "$Elements", str(len(boundary) + len(tets))]
    k = 1
    for face, tag in boundary.items():
        lines.append(f"{k} 2 2 {tag} {tag} " + " ".join(str(v + 1) for v in face))
        k += 1
    for t in tets:
        lines.append(f"{k} {tet_type} 2 1 1 " + " ".join(str(v + 1) for v in t))
        k += 1
    lines += ["$EndElements"]
    return "\n".join(lines) + "\n"


def _check_two_tets(m):
    assert np.array_equal(m.cells["tetra"], np.array(TETS))
    assert np.allclose(m.points, np.array(NODES, float))
    assert m.physical_names == {"bottom": 11, "walls": 13, "fluid": 1}
    topo = Topology(m)
    f = topo.faces
    bnd = np.flatnonzero(f.neighbor_block < 0)
    assert len(bnd) == 6 and np.count_nonzero(f.neighbor_block >= 0) == 1
    assert sorted(f.tag[bnd].tolist()) == [11] + [13] * 5
    # the bottom face's outward normal is -z
    b = bnd[f.tag[bnd] == 11][0]
    assert np.allclose(f.normal[b], [0, 0, -1])
    assert np.isclose(sum(bl.volume.sum() for bl in topo.blocks), 1 / 6 + 2 / 6)


def test_msh41_ascii(tmp_path):
    p = tmp_path / "two.msh"
    p.write_text(MSH41)
    _check_two_tets(read_gmsh(p))


def test_msh22_ascii(tmp_path):
    p = tmp_path / "two.msh"
    p.write_text(_msh22())
    _check_two_tets(read_gmsh(p))


def _with_midnodes(shift=0.0):
    """The two tetrahedra as tetra10 cells (Gmsh edge order 01 12 20 03 23
    13; any order is accepted), one mid-edge node moved by ``shift``."""
    nodes = list(NODES)
    mid = {}
    tets = []
    for t in TETS:
        row = list(t)
        for a, b in ((0, 1), (1, 2), (2, 0), (0, 3), (2, 3), (1, 3)):
            key = tuple(sorted((t[a], t[b])))
            if key not in mid:
                mid[key] = len(nodes)
                nodes.append(tuple(0.5 * (np.array(NODES[key[0]]) + np.array(NODES[key[1]]))))
            row.append(mid[key])
        tets.append(tuple(row))
    k = mid[(1, 2)]
    nodes[k] = (nodes[k][0], nodes[k][1], nodes[k][2] + shift)
    return nodes, tets


def test_straight_second_order_cells_are_accepted(tmp_path):
    nodes, tets = _with_midnodes()
    p = tmp_path / "p2.msh"
    p.write_text(_msh22(tets, nodes, tet_type=11))
    m = read_gmsh(p)
    assert np.array_equal(m.cells["tetra"], np.array(TETS))


def test_curved_second_order_cells_raise(tmp_path):
    nodes, tets = _with_midnodes(shift=0.05)
    p = tmp_path / "p2.msh"
    p.write_text(_msh22(tets, nodes, tet_type=11))
    with pytest.raises(NotImplementedError, match="2 of 2 second-order tetra cells"):
        read_gmsh(p)


def test_untagged_and_conflicting_faces_raise(tmp_path):
    p = tmp_path / "missing.msh"
    b = dict(BOUNDARY)
    del b[(1, 3, 4)]
    p.write_text(_msh22(boundary=b))
    with pytest.raises(ValueError, match="1 boundary faces lack physical tags"):
        Topology(read_gmsh(p))
    lines = _msh22().split("\n")
    i = lines.index("$Elements")
    lines[i + 1] = str(int(lines[i + 1]) + 1)
    lines.insert(i + 2, "99 2 2 13 13 1 2 3")   # face (0, 1, 2) again, other tag
    p.write_text("\n".join(lines))
    with pytest.raises(ValueError, match="Conflicting physical tags"):
        read_gmsh(p)


@pytest.mark.parametrize("version", ["2.2", "4.1"])
@pytest.mark.parametrize("binary", [False, True])
def test_round_trip_of_a_split_box(tmp_path, version, binary):
    names = {"xmin": "inlet", "xmax": "outlet", "ymin": "walls", "ymax": "walls",
             "zmin": "bottom", "zmax": "top"}
    m = box_mesh((3, 2, 4), lengths=(1.0, 0.5, 2.0), split_fraction=0.5, seed=9,
                 plane_names=names)
    p = tmp_path / "box.msh"
    write_gmsh(p, m, version=version, binary=binary)
    r = read_gmsh(p)
    for t, conn in m.cells.items():
        assert np.array_equal(r.cells[t], conn), t
    assert np.array_equal(r.points, m.points)
    a, b = Topology(m), Topology(r)
    for k in ("face_type", "owner_block", "owner_local", "neighbor_block", "neighbor_local",
              "tag"):
        assert np.array_equal(getattr(a.faces, k), getattr(b.faces, k)), k
    assert np.allclose(a.faces.normal, b.faces.normal, rtol=0, atol=1e-15)
    assert {names.get(k, k) for k in ("xmin", "xmax")} <= set(r.physical_names)


@pytest.mark.skipif(not (MESHES / "taylor_green.msh").exists(), reason="reference meshes absent")
def test_reference_taylor_green_mesh():
    """MSH 4.1 ASCII written by Gmsh: 33026 tetrahedra and 5636 boundary
    triangles (the element blocks of the file), on [0, 2 pi]^3; each
    boundary face lies on the plane its periodic_* name gives."""
    m = read_gmsh(MESHES / "taylor_green.msh")
    assert list(m.cells) == ["tetra"] and len(m.cells["tetra"]) == 33026
    topo = Topology(m)
    L = 2 * np.pi
    assert abs(sum(b.volume.sum() for b in topo.blocks) - L ** 3) < 1e-10 * L ** 3
    f = topo.faces
    bnd = np.flatnonzero(f.neighbor_block < 0)
    assert len(bnd) == 5636
    plane = {21: (0, 0.0), 22: (0, L), 23: (1, 0.0), 24: (1, L), 25: (2, 0.0), 26: (2, L)}
    for tag, (axis, value) in plane.items():
        sel = bnd[f.tag[bnd] == tag]
        assert len(sel), tag
        b = topo.blocks[0]
        for i in sel[:50]:
            fv = LOCAL_FACES["tetra"][f.owner_face[i]]
            x = m.points[b.connectivity[f.owner_local[i]][list(fv)]]
            assert np.allclose(x[:, axis], value, atol=1e-9), tag
        assert np.allclose(np.abs(f.normal[sel, axis]), 1.0)
    names = {v: k for k, v in m.physical_names.items()}
    assert names[21] == "periodic_x_minus" and names[26] == "periodic_z_plus"


@pytest.mark.skipif(not (MESHES / "sr71_pristine.msh").exists(), reason="reference meshes absent")
def test_reference_sr71_binary_mesh():
    """MSH 2.2 binary written by Gmsh: tetrahedra only, every boundary face
    tagged with one of the four named surfaces, positive volumes."""
    m = read_gmsh(MESHES / "sr71_pristine.msh")
    assert set(m.physical_names) == {"inlet", "outlet", "walls", "airfoil_wall", "fluid"}
    topo = Topology(m)
    assert len(topo.blocks) == 1 and topo.blocks[0].element_type == "tetra"
    assert np.all(topo.blocks[0].volume > 0)
    f = topo.faces
    tags = set(f.tag[f.neighbor_block < 0].tolist())
    assert tags == {11, 12, 13, 14}


@pytest.mark.skipif(not (MESHES / "sphere_tunnel.msh").exists(), reason="reference meshes absent")
def test_reference_curved_mesh_is_refused():
    with pytest.raises(NotImplementedError, match="curved"):
        read_gmsh(MESHES / "sphere_tunnel.msh")
