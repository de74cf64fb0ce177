"""Gmsh MSH 2.2 / 4.1 reader (ASCII and binary) for the device path.

Builds the same mesh the reference builds from a Gmsh file through meshio
(hybrid/mesh.py:186-242, 334-352): the volume cells of each type in file
order (blocks ordered hexahedron, wedge, tetra), their corner nodes (the
first 4 / 6 / 8 nodes of a Gmsh cell of any order), the physical names, and
the physical tag of every boundary triangle / quadrilateral keyed by its
sorted corner ids.  A boundary face without a tag, or one face with two
different tags, raises, as the reference's validation does
(hybrid/mesh.py:334-352, 466-483).

Geometry: the device path runs affine cells (constant Jacobian per cell).
Second-order cells (tetra10, wedge15/18, hexahedron20/27) are accepted when
they are straight-sided, i.e. every high-order node sits on the position the
corners' P1 map gives it (then the reference's quadratic geometry,
hybrid/geometry.py:103-109, is the affine one); a curved cell raises
NotImplementedError naming how many cells are curved.
"""

import struct
from pathlib import Path

import numpy as np

from .meshgen import MeshArrays

# Gmsh element type -> (node count, cell kind, corner count)
_GMSH_TYPES = {
    1: (2, None, 0), 8: (3, None, 0), 26: (4, None, 0), 15: (1, None, 0),
    2: (3, "triangle", 3), 9: (6, "triangle", 3), 21: (10, "triangle", 3),
    3: (4, "quad", 4), 16: (8, "quad", 4), 10: (9, "quad", 4),
    4: (4, "tetra", 4), 11: (10, "tetra", 4),
    6: (6, "wedge", 6), 18: (15, "wedge", 6), 13: (18, "wedge", 6),
    5: (8, "hexahedron", 8), 17: (20, "hexahedron", 8), 12: (27, "hexahedron", 8),
}
_VOLUME = ("hexahedron", "wedge", "tetra")

# P2 lattice of each cell in reference coordinates (any order of the nodes:
# a high-order node only has to coincide with one of them)
_P2_POINTS = {
    "tetra": np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, 0, 0], [0, .5, 0],
                       [0, 0, .5], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]], float),
    "wedge": np.array([[x, y, z] for z in (0.0, 1.0, 0.5)
                       for x, y in ((0, 0), (1, 0), (0, 1), (.5, 0), (.5, .5), (0, .5))], float),
    "hexahedron": np.array([[x, y, z] for z in (0.0, 1.0, 0.5) for y in (0.0, 1.0, 0.5)
                            for x in (0.0, 1.0, 0.5)], float),
}


def _p1_basis(kind, xi):
    x, y, z = xi[:, 0], xi[:, 1], xi[:, 2]
    if kind == "tetra":
        return np.stack([1 - x - y - z, x, y, z], -1)
    if kind == "wedge":
        t = (1 - x - y, x, y)
        return np.stack([t[0] * (1 - z), t[1] * (1 - z), t[2] * (1 - z),
                         t[0] * z, t[1] * z, t[2] * z], -1)
    return np.stack([(1 - x) * (1 - y) * (1 - z), x * (1 - y) * (1 - z), x * y * (1 - z),
                     (1 - x) * y * (1 - z), (1 - x) * (1 - y) * z, x * (1 - y) * z, x * y * z,
                     (1 - x) * y * z], -1)


class _Reader:
    def __init__(self, data):
        self.data = data
        self.pos = 0

    def line(self):
        end = self.data.index(b"\n", self.pos)
        out = self.data[self.pos:end].strip()
        self.pos = end + 1
        return out

    def words(self):
        return self.line().split()

    def take(self, fmt, count=1):
        size = struct.calcsize(fmt) * count
        arr = np.frombuffer(self.data, dtype=np.dtype(fmt), count=count, offset=self.pos)
        self.pos += size
        return arr

    def until(self, tag):
        while self.pos < len(self.data):
            if self.line() == tag:
                return
        raise ValueError(f"missing {tag.decode()}")


def _sections(r):
    """Yields the name of each $Section with the reader positioned after it."""
    while r.pos < len(r.data):
        t = r.line()
        if t.startswith(b"$") and not t.startswith(b"$End"):
            yield t[1:].decode()


def read_gmsh(path):
    """Reads a Gmsh MSH 2.2 or 4.1 file (ASCII or binary) into MeshArrays."""
    data = Path(path).read_bytes()
    r = _Reader(data)
    version = binary = None
    size_t = 8
    names = {}
    nodes_tag = nodes_xyz = None
    entity_phys = {}          # (dim, tag) -> first physical tag (4.1)
    blocks = []               # (gmsh type, physical tag per element or None, node-tag array)
    for sec in _sections(r):
        if sec == "MeshFormat":
            w = r.words()
            version, binary, size_t = w[0].decode(), int(w[1]), int(w[2])
            if binary:
                one = r.take("<i")[0]
                if one != 1:
                    raise ValueError("big-endian Gmsh files are not supported")
                r.line()
            if not (version.startswith("2.") or version.startswith("4.")):
                raise ValueError(f"unsupported MSH version {version}")
            r.until(b"$EndMeshFormat")
        elif sec == "PhysicalNames":
            n = int(r.line())
            for _ in range(n):
                w = r.line().split(maxsplit=2)
                names[w[2].decode().strip().strip('"')] = int(w[1])
            r.until(b"$EndPhysicalNames")
        elif sec == "Entities":
            entity_phys = _read_entities(r, binary, size_t)
            r.until(b"$EndEntities")
        elif sec == "Nodes":
            if version.startswith("4"):
                nodes_tag, nodes_xyz = _read_nodes41(r, binary, size_t)
            else:
                nodes_tag, nodes_xyz = _read_nodes22(r, binary)
            r.until(b"$EndNodes")
        elif sec == "Elements":
            if version.startswith("4"):
                blocks = _read_elements41(r, binary, size_t, entity_phys)
            else:
                blocks = _read_elements22(r, binary)
            r.until(b"$EndElements")
        else:
            r.until(f"$End{sec}".encode())
    if nodes_tag is None or not blocks:
        raise ValueError(f"{path}: no $Nodes or $Elements section")
    return _assemble(nodes_tag, nodes_xyz, blocks, names)


def _read_entities(r, binary, size_t):
    st = "<q" if size_t == 8 else "<i"
    out = {}
    if binary:
        counts = r.take(st, 4)
    else:
        counts = [int(x) for x in r.words()]
    for dim in range(4):
        for _ in range(int(counts[dim])):
            if binary:
                tag = int(r.take("<i")[0])
                r.take("<d", 3 if dim == 0 else 6)
                nph = int(r.take(st)[0])
                phys = r.take("<i", nph).tolist()
                if dim > 0:
                    nb = int(r.take(st)[0])
                    r.take("<i", nb)
            else:
                w = r.words()
                tag = int(w[0])
                k = 4 if dim == 0 else 7
                nph = int(w[k])
                phys = [int(x) for x in w[k + 1:k + 1 + nph]]
            if phys:
                out[(dim, tag)] = phys[0]
    return out


def _read_nodes41(r, binary, size_t):
    st = "<q" if size_t == 8 else "<i"
    if binary:
        nblocks, total = (int(x) for x in r.take(st, 4)[:2])
    else:
        w = r.words()
        nblocks, total = int(w[0]), int(w[1])
    tags = np.empty(total, np.int64)
    xyz = np.empty((total, 3))
    k = 0
    for _ in range(nblocks):
        if binary:
            _, _, parametric = (int(x) for x in r.take("<i", 3))
            n = int(r.take(st)[0])
        else:
            w = r.words()
            parametric, n = int(w[2]), int(w[3])
        if parametric:
            raise ValueError("parametric node coordinates are not supported")
        if binary:
            tags[k:k + n] = r.take(st, n)
            xyz[k:k + n] = r.take("<d", 3 * n).reshape(n, 3)
        else:
            for i in range(n):
                tags[k + i] = int(r.line())
            for i in range(n):
                xyz[k + i] = [float(x) for x in r.words()[:3]]
        k += n
    return tags[:k], xyz[:k]


def _read_nodes22(r, binary):
    n = int(r.line())
    if binary:
        rec = np.frombuffer(r.data, dtype=np.dtype([("id", "<i4"), ("x", "<f8", 3)]), count=n,
                            offset=r.pos)
        r.pos += rec.nbytes
        return rec["id"].astype(np.int64), rec["x"].copy()
    tags = np.empty(n, np.int64)
    xyz = np.empty((n, 3))
    for i in range(n):
        w = r.words()
        tags[i] = int(w[0])
        xyz[i] = [float(x) for x in w[1:4]]
    return tags, xyz


def _node_count(gtype):
    if gtype not in _GMSH_TYPES:
        raise ValueError(f"unsupported Gmsh element type {gtype}")
    return _GMSH_TYPES[gtype][0]


def _read_elements41(r, binary, size_t, entity_phys):
    st = "<q" if size_t == 8 else "<i"
    if binary:
        nblocks = int(r.take(st, 4)[0])
    else:
        nblocks = int(r.words()[0])
    out = []
    for _ in range(nblocks):
        if binary:
            dim, etag, gtype = (int(x) for x in r.take("<i", 3))
            n = int(r.take(st)[0])
        else:
            dim, etag, gtype, n = (int(x) for x in r.words())
        k = _node_count(gtype)
        if binary:
            arr = r.take(st, n * (1 + k)).reshape(n, 1 + k)
        else:
            arr = np.array([[int(x) for x in r.words()] for _ in range(n)], np.int64).reshape(n, 1 + k)
        phys = entity_phys.get((dim, etag))
        out.append((gtype, None if phys is None else np.full(n, phys, np.int64),
                    arr[:, 1:].astype(np.int64)))
    return out


def _read_elements22(r, binary):
    total = int(r.line())
    out = []
    if binary:
        done = 0
        while done < total:
            gtype, n, ntags = (int(x) for x in r.take("<i", 3))
            k = _node_count(gtype)
            arr = r.take("<i", n * (1 + ntags + k)).reshape(n, 1 + ntags + k).astype(np.int64)
            phys = arr[:, 1] if ntags > 0 else None
            out.append((gtype, phys, arr[:, 1 + ntags:]))
            done += n
        return out
    rows = {}
    for _ in range(total):
        w = [int(x) for x in r.words()]
        gtype, ntags = w[1], w[2]
        k = _node_count(gtype)
        phys = w[3] if ntags > 0 else -1
        rows.setdefault(gtype, ([], []))
        rows[gtype][0].append(phys)
        rows[gtype][1].append(w[3 + ntags:3 + ntags + k])
    for gtype, (phys, nodes) in rows.items():
        out.append((gtype, np.asarray(phys, np.int64), np.asarray(nodes, np.int64)))
    return out


def _assemble(node_tags, xyz, blocks, names):
    if node_tags.min() < 1:
        raise ValueError("node tags must be positive")
    dense = np.full(int(node_tags.max()) + 1, -1, np.int64)
    dense[node_tags] = np.arange(len(node_tags))

    def lookup(conn):
        if conn.min() < 1 or conn.max() >= len(dense) or np.any(dense[conn] < 0):
            raise ValueError("an element names a node that is not in $Nodes")
        return dense[conn]

    cells = {t: [] for t in _VOLUME}
    faces = {}
    for gtype, phys, conn in blocks:
        k, kind, nc = _GMSH_TYPES[gtype]
        if kind is None or not len(conn):
            continue
        ids = lookup(conn)
        if kind in cells:
            if k > nc:
                _check_straight(kind, ids, xyz, gtype)
            cells[kind].append(ids[:, :nc])
        else:
            for nodes, tag in zip(ids[:, :nc], (phys if phys is not None else
                                                np.full(len(ids), -1))):
                key = tuple(sorted(int(v) for v in nodes))
                if key in faces and faces[key] != int(tag):
                    raise ValueError(f"Conflicting physical tags on face {key}")
                faces[key] = int(tag)
    cells = {t: np.concatenate(v) for t, v in cells.items() if v}
    if not cells:
        raise ValueError("Mesh contains no tetrahedron, wedge, or hexahedron")
    return MeshArrays(xyz, cells, np.arange(len(xyz)), np.zeros(3), names, face_tags=faces)


def _check_straight(kind, ids, xyz, gtype, tol=1e-9):
    """Every node of a second-order cell must coincide with a point of the
    P2 lattice of its corners' P1 map (a straight-sided cell)."""
    nc = {"tetra": 4, "wedge": 6, "hexahedron": 8}[kind]
    X = xyz[ids]                                     # (E, k, 3)
    want = np.einsum("pc,ecd->epd", _p1_basis(kind, _P2_POINTS[kind]), X[:, :nc])
    scale = np.linalg.norm(X[:, :nc] - X[:, :nc].mean(1, keepdims=True), axis=-1).max(1)
    dist = np.linalg.norm(X[:, :, None, :] - want[:, None, :, :], axis=-1).min(-1).max(-1)
    curved = int(np.count_nonzero(dist > tol * np.maximum(scale, 1e-300)))
    if curved:
        raise NotImplementedError(
            f"{curved} of {len(ids)} second-order {kind} cells (Gmsh type {gtype}) are curved: "
            "curved P2 geometry needs per-point metrics, which the device path does not run")


def write_gmsh(path, mesh, version="4.1", binary=False, volume_tag=None):
    """Writes MeshArrays as a Gmsh file with linear cells, every boundary face
    tagged through ``mesh.tag_of`` and the volume under ``volume_tag``
    (default: one above the largest surface tag), named "fluid"."""
    from .elements import LOCAL_FACES
    from .topology import Topology
    topo = Topology(mesh)
    f = topo.faces
    bnd = np.flatnonzero(f.neighbor_block < 0)
    tri, quad = [], []
    for i in bnd:
        b = topo.blocks[f.owner_block[i]]
        fv = LOCAL_FACES[b.element_type][f.owner_face[i]]
        nodes = b.connectivity[f.owner_local[i]][list(fv)]
        if f.sub[i]:
            raise ValueError("quad faces split into triangles cannot be written")
        (tri if len(nodes) == 3 else quad).append((nodes, int(f.tag[i])))
    if volume_tag is None:
        used = [t for _, t in tri + quad] + list(mesh.physical_names.values())
        volume_tag = max(used, default=0) + 1
    pts = np.asarray(mesh.points, float)
    vol = [(t, np.asarray(mesh.cells[t])) for t in _VOLUME if t in mesh.cells]
    gid = {"triangle": 2, "quad": 3, "tetra": 4, "hexahedron": 5, "wedge": 6}
    groups = []    # (dim, gmsh type, phys tag, conn)
    for kind, lst in (("triangle", tri), ("quad", quad)):
        tags = sorted({t for _, t in lst})
        for t in tags:
            groups.append((2, gid[kind], t, np.array([n for n, tt in lst if tt == t], np.int64)))
    for t, conn in vol:
        groups.append((3, gid[t], volume_tag, conn.astype(np.int64)))
    out = bytearray()
    w = lambda s: out.extend(s.encode())
    w(f"$MeshFormat\n{version} {int(binary)} 8\n")
    if binary:
        out.extend(struct.pack("<i", 1))
        w("\n")
    w("$EndMeshFormat\n$PhysicalNames\n")
    pn = [(2, tag, name) for name, tag in mesh.physical_names.items()
          if tag != volume_tag]
    pn.append((3, volume_tag, "fluid"))
    w(f"{len(pn)}\n")
    for dim, tag, name in pn:
        w(f'{dim} {tag} "{name}"\n')
    w("$EndPhysicalNames\n")
    if version.startswith("4"):
        _write41(out, w, pts, groups, binary)
    else:
        _write22(out, w, pts, groups, binary)
    Path(path).write_bytes(bytes(out))


def _write41(out, w, pts, groups, binary):
    # one entity per group: surfaces 1.., volumes 1..
    surf = [g for g in groups if g[0] == 2]
    vols = [g for g in groups if g[0] == 3]
    w("$Entities\n")
    if binary:
        out.extend(struct.pack("<4q", 0, 0, len(surf), len(vols)))
        for i, g in enumerate(surf + vols):
            tag = i + 1 if g[0] == 2 else i + 1 - len(surf)
            out.extend(struct.pack("<i6dqi", tag, 0, 0, 0, 0, 0, 0, 1, g[2]))
            out.extend(struct.pack("<q", 0))
        w("\n")
    else:
        w(f"0 0 {len(surf)} {len(vols)}\n")
        for i, g in enumerate(surf):
            w(f"{i + 1} 0 0 0 0 0 0 1 {g[2]} 0\n")
        for i, g in enumerate(vols):
            w(f"{i + 1} 0 0 0 0 0 0 1 {g[2]} 0\n")
    w("$EndEntities\n$Nodes\n")
    n = len(pts)
    if binary:
        out.extend(struct.pack("<4q", 1, n, 1, n))
        out.extend(struct.pack("<3iq", 3, 1, 0, n))
        out.extend(np.arange(1, n + 1, dtype="<i8").tobytes())
        out.extend(pts.astype("<f8").tobytes())
        w("\n")
    else:
        w(f"1 {n} 1 {n}\n3 1 0 {n}\n")
        w("".join(f"{i + 1}\n" for i in range(n)))
        w("".join(f"{x!r} {y!r} {z!r}\n" for x, y, z in pts.tolist()))
    w("$EndNodes\n$Elements\n")
    total = sum(len(g[3]) for g in groups)
    nxt = 1
    if binary:
        out.extend(struct.pack("<4q", len(groups), total, 1, total))
    else:
        w(f"{len(groups)} {total} 1 {total}\n")
    si = vi = 0
    for g in groups:
        dim, gt, _, conn = g
        if dim == 2:
            si += 1
            etag = si
        else:
            vi += 1
            etag = vi
        m = len(conn)
        ids = np.arange(nxt, nxt + m, dtype=np.int64)
        nxt += m
        rows = np.concatenate([ids[:, None], conn + 1], 1)
        if binary:
            out.extend(struct.pack("<3iq", dim, etag, gt, m))
            out.extend(rows.astype("<i8").tobytes())
        else:
            w(f"{dim} {etag} {gt} {m}\n")
            w("".join(" ".join(str(int(v)) for v in row) + "\n" for row in rows))
    if binary:
        w("\n")
    w("$EndElements\n")


def _write22(out, w, pts, groups, binary):
    n = len(pts)
    w(f"$Nodes\n{n}\n")
    if binary:
        rec = np.zeros(n, dtype=np.dtype([("id", "<i4"), ("x", "<f8", 3)]))
        rec["id"] = np.arange(1, n + 1)
        rec["x"] = pts
        out.extend(rec.tobytes())
        w("\n")
    else:
        w("".join(f"{i + 1} {x!r} {y!r} {z!r}\n" for i, (x, y, z) in enumerate(pts.tolist())))
    total = sum(len(g[3]) for g in groups)
    w(f"$EndNodes\n$Elements\n{total}\n")
    nxt = 1
    for dim, gt, phys, conn in groups:
        m = len(conn)
        ids = np.arange(nxt, nxt + m, dtype=np.int64)
        nxt += m
        rows = np.concatenate([ids[:, None], np.full((m, 2), phys), conn + 1], 1)
        if binary:
            out.extend(struct.pack("<3i", gt, m, 2))
            out.extend(rows.astype("<i4").tobytes())
        else:
            w("".join(f"{int(r[0])} {gt} 2 {int(r[1])} {int(r[2])} " +
                      " ".join(str(int(v)) for v in r[3:]) + "\n" for r in rows))
    if binary:
        w("\n")
    w("$EndElements\n")
