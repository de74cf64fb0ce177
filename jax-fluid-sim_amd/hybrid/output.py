"""Restart files and VTK output (run_hybrid_sphere.py:36-45, 123-131 and
hybrid/output.py:274-441, 743-757 of the reference).

* ``save_restart`` / ``load_restart``: an ``np.savez_compressed`` archive
  with the keys ``time``, ``step`` and ``U_<element type>`` (the reference's
  restart format, readable by either code), written to a temporary file and
  moved into place.
* ``write_tessellated_vtu``: every element's solution nodes as linear
  sub-cells (p^3 sub-hexahedra, sub-prisms or sub-tetrahedra per element),
  point data = primitive fields, in VTK's XML unstructured-grid format with
  base64 binary arrays (meshio is not needed).
* ``write_midplane_slice_vtu``: the planar cut z = plane_z of those linear
  sub-cells, fields linearly interpolated along the cut edges, as polygons.
* ``write_pvd``: the ParaView time-series index.
"""

import base64
import struct
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np
import torch

from .elements import nodes

VTK_TETRA, VTK_HEXAHEDRON, VTK_WEDGE, VTK_POLYGON = 10, 12, 13, 7


# -- restart ---------------------------------------------------------------------
def save_restart(path, states, current_time, step):
    """run_hybrid_sphere.py:36-45: keys time, step, U_<type>."""
    path = Path(path)
    tmp = path.with_name(path.stem + ".tmp.npz")
    arrays = {f"U_{k}": (v.detach().cpu().numpy() if torch.is_tensor(v) else np.asarray(v))
              for k, v in states.items()}
    np.savez_compressed(tmp, time=np.asarray(current_time), step=np.asarray(step), **arrays)
    tmp.replace(path)
    return path


def load_restart(path, solver):
    """run_hybrid_sphere.py:123-131: the solver's blocks from U_<type>, the
    time and the step (0 when absent), as device tensors of the solver
    dtype."""
    with np.load(path, allow_pickle=False) as r:
        states = {}
        for b in solver.mesh.blocks:
            a = r[f"U_{b.element_type}"]
            states[b.element_type] = torch.as_tensor(a, dtype=solver.dtype,
                                                     device=solver.device).contiguous()
        t = float(r["time"])
        step = int(r["step"]) if "step" in r.files else 0
    return states, t, step


# -- linear sub-cells of the solution-node lattice ----------------------------------
def _lattice(element_type, p):
    """{(i, j, k): node} of the reference nodes on the lattice of spacing 1/p."""
    x = nodes(element_type, p)
    idx = np.rint(x * p).astype(int)
    if np.abs(idx - x * p).max() > 1e-9:
        raise ValueError("nodes are not on the 1/p lattice")
    return {tuple(v): n for n, v in enumerate(idx)}


def sub_cells(element_type, p):
    """(vtk cell type, (cells, corners) node indices) of the p^3 linear
    sub-cells that tile the reference element through its solution nodes."""
    L = _lattice(element_type, p)
    cells = []
    if element_type == "hexahedron":
        for k in range(p):
            for j in range(p):
                for i in range(p):
                    c = [(i, j, k), (i + 1, j, k), (i + 1, j + 1, k), (i, j + 1, k)]
                    cells.append([L[v] for v in c] + [L[(a, b, k + 1)] for a, b, _ in c])
        return VTK_HEXAHEDRON, np.asarray(cells, np.int64)
    if element_type == "wedge":
        tris = []
        for j in range(p):
            for i in range(p - j):
                tris.append([(i, j), (i, j + 1), (i + 1, j)])        # base turned away from the top
                if i + j <= p - 2:
                    tris.append([(i + 1, j), (i, j + 1), (i + 1, j + 1)])
        for k in range(p):
            for t in tris:
                cells.append([L[(a, b, k)] for a, b in t] + [L[(a, b, k + 1)] for a, b in t])
        return VTK_WEDGE, np.asarray(cells, np.int64)
    e = np.eye(3, dtype=int)
    for k in range(p):
        for j in range(p - k):
            for i in range(p - k - j):
                v = np.array([i, j, k])
                cells.append([v, v + e[0], v + e[1], v + e[2]])
                if i + j + k <= p - 2:
                    a, b, c = v + e[0], v + e[1], v + e[2]
                    ab, ac, bc = v + e[0] + e[1], v + e[0] + e[2], v + e[1] + e[2]
                    # the octahedron split along its a-bc diagonal
                    cells += [[a, bc, b, ab], [a, bc, ab, ac], [a, bc, ac, c], [a, bc, c, b]]
                if i + j + k <= p - 3:
                    cells.append([v + e[0] + e[1], v + e[0] + e[2], v + e[1] + e[2],
                                  v + e[0] + e[1] + e[2]])
    out = []
    x = nodes(element_type, p)
    for c in cells:
        ids = [L[tuple(v)] for v in c]
        q = x[ids]
        if np.dot(np.cross(q[1] - q[0], q[2] - q[0]), q[3] - q[0]) < 0:
            ids[1], ids[2] = ids[2], ids[1]
        out.append(ids)
    return VTK_TETRA, np.asarray(out, np.int64)


def _points_and_fields(operator, states):
    s = operator.solver
    xs = s.solution_node_coordinates()
    pts, prims, cells, types, off = [], [], [], [], 0
    for b in s.mesh.blocks:
        t = b.element_type
        X = xs[t]
        U = states[t]
        W = s.conserved_to_primitive(U.detach().double().cpu().numpy() if torch.is_tensor(U)
                                     else np.asarray(U, float))
        vt, sub = sub_cells(t, s.polynomial_degree)
        E, N = X.shape[:2]
        cells.append(((off + np.arange(E)[:, None, None] * N + sub[None])
                      .reshape(-1, sub.shape[1]), vt))
        types.append(np.full(E * len(sub), vt, np.uint8))
        pts.append(X.reshape(-1, 3))
        prims.append(W.reshape(-1, W.shape[-1]))
        off += E * N
    return np.concatenate(pts), np.concatenate(prims), cells, np.concatenate(types)


def _point_data(solver, W):
    g = solver.gamma
    rho, vel, p = W[:, 0], W[:, 1:4], W[:, 4]
    c = np.sqrt(np.maximum(g * p / np.maximum(rho, 1e-12), 0.0))
    speed = np.linalg.norm(vel, axis=-1)
    data = {"density": rho, "velocity": vel, "pressure": p,
            "temperature": p / (np.maximum(rho, 1e-300) * solver.gas_constant),
            "speed": speed, "mach": speed / np.maximum(c, 1e-30)}
    for i, name in enumerate(solver.layout.species_names):
        clean = str(name).replace("+", "_plus").replace("-", "_minus")
        data[f"Y_{clean}"] = W[:, 5 + i]
    return data


# -- VTK XML ---------------------------------------------------------------------
def _b64(a):
    raw = np.ascontiguousarray(a).tobytes()
    return base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")


def _data_array(parent, name, a, components=1):
    kinds = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
             np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8"}
    el = ET.SubElement(parent, "DataArray", type=kinds[a.dtype], Name=name,
                       NumberOfComponents=str(components), format="binary")
    el.text = _b64(a)


def _write_vtu(path, points, conn, offsets, types, point_data):
    root = ET.Element("VTKFile", type="UnstructuredGrid", version="1.0",
                      byte_order="LittleEndian", header_type="UInt64")
    piece = ET.SubElement(ET.SubElement(root, "UnstructuredGrid"), "Piece",
                          NumberOfPoints=str(len(points)), NumberOfCells=str(len(types)))
    pd = ET.SubElement(piece, "PointData")
    for k, v in point_data.items():
        v = np.asarray(v, np.float64)
        _data_array(pd, k, v, 1 if v.ndim == 1 else v.shape[1])
    _data_array(ET.SubElement(piece, "Points"), "Points", np.asarray(points, np.float64), 3)
    ce = ET.SubElement(piece, "Cells")
    _data_array(ce, "connectivity", np.asarray(conn, np.int64))
    _data_array(ce, "offsets", np.asarray(offsets, np.int64))
    _data_array(ce, "types", np.asarray(types, np.uint8))
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")
    return path


def read_vtu(path):
    """(points, connectivity, offsets, types, point data) of a file written here."""
    root = ET.parse(path).getroot()
    kinds = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64, "UInt8": np.uint8}

    def arr(el):
        raw = base64.b64decode(el.text)
        n = struct.unpack("<Q", raw[:8])[0]
        a = np.frombuffer(raw[8:8 + n], dtype=kinds[el.get("type")])
        c = int(el.get("NumberOfComponents", "1"))
        return a.reshape(-1, c) if c > 1 else a

    piece = root.find("UnstructuredGrid/Piece")
    pdata = {el.get("Name"): arr(el) for el in piece.find("PointData")}
    pts = arr(piece.find("Points/DataArray"))
    cells = {el.get("Name"): arr(el) for el in piece.find("Cells")}
    return pts, cells["connectivity"], cells["offsets"], cells["types"], pdata


def write_tessellated_vtu(path, operator, states):
    """hybrid/output.py:274-358: each element as linear sub-cells through its
    solution nodes, with the primitive fields at the nodes."""
    pts, W, cells, types = _points_and_fields(operator, states)
    conn = np.concatenate([c.ravel() for c, _ in cells])
    sizes = np.concatenate([np.full(len(c), c.shape[1]) for c, _ in cells])
    return _write_vtu(path, pts, conn, np.cumsum(sizes), types,
                      _point_data(operator.solver, W))


_EDGES = {VTK_TETRA: ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)),
          VTK_WEDGE: ((0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 3), (0, 3), (1, 4), (2, 5)),
          VTK_HEXAHEDRON: ((0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
                           (0, 4), (1, 5), (2, 6), (3, 7))}


def midplane_polygons(operator, states, plane_z=0.0, tol=1e-10):
    """hybrid/output.py:184-261: the cut of every linear sub-cell by the plane
    z = plane_z (a sub-cell face in the plane counted once).  Returns (vertices (n, 3), polygons (list of index arrays),
    primitive fields at the vertices)."""
    pts, W, cells, types = _points_and_fields(operator, states)
    verts, vals, polys = [], [], []
    for c, vt in cells:
        dz = pts[c, 2] - plane_z
        # cells with points below the plane that reach it: a sub-cell face
        # lying in the plane is taken once, from the cell below it
        hit = (dz.min(1) < -tol) & (dz.max(1) >= -tol)
        for cell in c[hit]:
            z = pts[cell, 2] - plane_z
            found = {}
            for a in np.flatnonzero(np.abs(z) <= tol):
                found.setdefault(tuple(np.round(pts[cell[a], :2], 12)),
                                 (pts[cell[a]], W[cell[a]]))
            for a, b in _EDGES[vt]:
                if z[a] * z[b] < -tol * tol:
                    f = -z[a] / (z[b] - z[a])
                    x = pts[cell[a]] + f * (pts[cell[b]] - pts[cell[a]])
                    found.setdefault(tuple(np.round(x[:2], 12)),
                                     (x, W[cell[a]] + f * (W[cell[b]] - W[cell[a]])))
            if len(found) < 3:
                continue
            xy = np.array([v[0] for v in found.values()])
            wv = np.array([v[1] for v in found.values()])
            ctr = xy[:, :2].mean(0)
            order = np.argsort(np.arctan2(xy[:, 1] - ctr[1], xy[:, 0] - ctr[0]))
            polys.append(len(verts) + np.arange(len(order)))
            verts.extend(xy[order])
            vals.extend(wv[order])
    if not polys:
        raise ValueError(f"The mesh does not intersect z={plane_z}")
    return np.asarray(verts), polys, np.asarray(vals)


def write_midplane_slice_vtu(path, operator, states, plane_z=0.0):
    """hybrid/output.py:361-439 (without the mirror image of a half-plane
    sector): the planar cut as polygons with the primitive fields."""
    v, polys, W = midplane_polygons(operator, states, plane_z)
    conn = np.concatenate(polys)
    offsets = np.cumsum([len(q) for q in polys])
    return _write_vtu(path, v, conn, offsets, np.full(len(polys), VTK_POLYGON, np.uint8),
                      _point_data(operator.solver, W))


def write_pvd(path, frames):
    """hybrid/output.py:743-757: ParaView collection of [(time, vtu path)]."""
    path = Path(path)
    root = ET.Element("VTKFile", type="Collection", version="0.1", byte_order="LittleEndian")
    col = ET.SubElement(root, "Collection")
    for t, f in frames:
        ET.SubElement(col, "DataSet", timestep=f"{t:.16g}", group="", part="0",
                      file=Path(f).name)
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")
    return path
