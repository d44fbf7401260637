"""Tecplot FEPolyhedron grid files and the grid transforms.

read_tecplot_grid follows Grid::readTEC (src/grid/grid_reader_tec.cc:36-278):
  - header lines up to the first line that starts with a number or '-': the
    whitespace is stripped and the "a=b" expressions between commas are
    read: Nodes, Elements, Faces, ZONETYPE (FEPolyhedron only), DATAPACKING
    (BLOCK only), TotalNumFaceNodes, NumConnectedBoundaryFaces and
    TotalNumBoundaryConnections (both must be 0);
  - the x, y and z blocks of the node coordinates; then, each block after an
    optional comment line, the node count of every face, the face nodes, the
    left and the right element of every face (1-based; 0 = none);
  - boundary zones: each header holds ZONET (the BC name, quotes stripped;
    BC numbers in order of appearance) and Nodes, then the x, y, z blocks of
    the zone's nodes; each zone node is the nearest node among the nodes of
    the boundary faces (the reference's kd-tree search, kd_nearest3).
The reference builds its grid from this face list directly (create_faces2,
grid_create_elements.cc:444-557: faces in file order, cell nodes as sorted
sets).  This module instead recovers each polyhedral cell as a tetrahedron,
pyramid, prism or hexahedron in CGNS node order from its faces, so the file
enters the same cell-based set-up as every other mesh: the cells, nodes,
BC node sets and geometry are the file's, while the internal face order (and
so the order of the per-cell sums) is that of the cell-based set-up.

translate, scale and rotate restate Grid::translate, Grid::scale and
Grid::rotate (src/grid/grid_transform.cc:24-61) on the raw node coordinates,
operation for operation.
"""
import math
import re

import numpy as np


class TecplotGridError(ValueError):
    pass


def _first_numeric(line):
    s = line.lstrip()
    return bool(s) and (s[0] == "-" or s[0].isdigit())


def _header_pairs(line):
    """strip_white_spaces, StringExplode(",") and StringExplode("=")"""
    s = re.sub(r"[ \t\f\v\n\r]", "", line)
    out = []
    for e in s.split(","):
        t = [x for x in e.split("=") if x != ""]
        if len(t) == 2:
            out.append((t[0], t[1]))
    return out


class _Tokens:
    """whitespace-separated numbers of a line range, read as file >> x"""

    def __init__(self, lines, i):
        self.lines, self.i, self.buf, self.k = lines, i, [], 0

    def _fill(self):
        while self.k >= len(self.buf):
            if self.i >= len(self.lines):
                raise TecplotGridError("unexpected end of file")
            self.buf, self.k = self.lines[self.i].split(), 0
            self.i += 1

    def take(self, n, conv=float):
        out = []
        while len(out) < n:
            self._fill()
            m = min(n - len(out), len(self.buf) - self.k)
            out.extend(conv(x) for x in self.buf[self.k:self.k + m])
            self.k += m
        return out

    def skip_comment(self):
        """the reference's `if (!is_first_numeric(file)) getline(file, line)`"""
        if self.k < len(self.buf):
            return
        while self.i < len(self.lines) and not self.lines[self.i].strip():
            self.i += 1
        if self.i < len(self.lines) and not _first_numeric(self.lines[self.i]):
            self.i += 1
        self.buf, self.k = [], 0

    def line_index(self):
        return self.i if self.k >= len(self.buf) else self.i - 1


def _read_header(lines, i, want):
    vals = {}
    while i < len(lines) and not _first_numeric(lines[i]):
        for a, b in _header_pairs(lines[i]):
            vals[a] = b
        i += 1
    return vals, i


def read_tecplot_grid(path, dimension=3):
    """The raw mesh (the dict fcfd_create takes, as freecfd_amd.cgns returns
    it) of a Tecplot FEPolyhedron grid file, with raw["bc_names"] in BC_n
    order and raw["faces"] = (face_ptr, face_nodes, left, right) as read."""
    with open(path, "r") as f:
        lines = f.read().splitlines()
    head, i = _read_header(lines, 0, None)
    for key in ("Nodes", "Elements", "Faces"):
        if key not in head:
            raise TecplotGridError("header without " + key)
    if head.get("ZONETYPE", "FEPolyhedron") != "FEPolyhedron":
        raise TecplotGridError("ZONETYPE=%s is not supported" % head["ZONETYPE"])
    if head.get("DATAPACKING", "BLOCK") != "BLOCK":
        raise TecplotGridError("DATAPACKING=%s is not supported" % head["DATAPACKING"])
    if int(head.get("NumConnectedBoundaryFaces", "0")) != 0:
        raise TecplotGridError("NumConnectedBoundaryFaces>0 is not supported")
    if int(head.get("TotalNumBoundaryConnections", "0")) != 0:
        raise TecplotGridError("TotalNumBoundaryConnections>0 is not supported")
    nn, nc, nf = int(head["Nodes"]), int(head["Elements"]), int(head["Faces"])
    tk = _Tokens(lines, i)
    xyz = np.array(tk.take(3 * nn)).reshape(3, nn).T.copy()
    tk.skip_comment()
    fcount = np.array(tk.take(nf, int), dtype=np.int64)
    fptr = np.zeros(nf + 1, dtype=np.int64)
    fptr[1:] = np.cumsum(fcount)
    tk.skip_comment()
    fnodes = np.array(tk.take(int(fptr[-1]), int), dtype=np.int64) - 1
    tk.skip_comment()
    left = np.array(tk.take(nf, int), dtype=np.int64) - 1
    tk.skip_comment()
    right = np.array(tk.take(nf, int), dtype=np.int64) - 1
    # boundary nodes (the kd-tree's points): nodes of faces with a missing side
    bface = (left < 0) | (right < 0)
    bnodes = np.unique(np.concatenate([fnodes[fptr[f]:fptr[f + 1]] for f in np.where(bface)[0]])) if bface.any() else \
        np.zeros(0, np.int64)
    names, sets = [], []
    j = tk.line_index()
    while True:
        while j < len(lines) and not lines[j].strip():
            j += 1
        if j >= len(lines):
            break
        zh, j = _read_header(lines, j, None)
        if "ZONET" not in zh:
            raise TecplotGridError("boundary zone without ZONET")
        names.append(zh["ZONET"].replace('"', "").replace("'", ""))
        nbn = int(zh.get("Nodes", "0"))
        tz = _Tokens(lines, j)
        bxyz = np.array(tz.take(3 * nbn)).reshape(3, nbn).T
        sets.append(np.unique(_nearest(xyz[bnodes], bnodes, bxyz)))
        # seek to the next ZONE line or the end of the file
        j = tz.line_index()
        while j < len(lines) and not re.sub(r"\s", "", lines[j]).startswith("ZONE"):
            j += 1
    cell_ptr, cell_nodes = _cells_from_faces(xyz, nc, fptr, fnodes, left, right)
    bc_ptr = np.zeros(len(sets) + 1, dtype=np.int32)
    bc_ptr[1:] = np.cumsum([len(s) for s in sets])
    bc_nodes = np.concatenate(sets).astype(np.int32) if sets else np.zeros(0, np.int32)
    return dict(xyz=xyz.reshape(-1), cell_ptr=cell_ptr, cell_nodes=cell_nodes, bc_ptr=bc_ptr, bc_nodes=bc_nodes,
                dimension=int(dimension), bc_names=names, faces=(fptr, fnodes, left, right))


def _nearest(pts, ids, q):
    """ids of the nearest of pts to each query point (kd_nearest3)"""
    if len(pts) == 0:
        raise TecplotGridError("a boundary zone but no boundary faces")
    try:
        from scipy.spatial import cKDTree
        _, k = cKDTree(pts).query(q)
    except ImportError:   # brute force
        k = np.array([int(np.argmin(((pts - x) ** 2).sum(axis=1))) for x in q])
    return ids[k]


def _cells_from_faces(xyz, nc, fptr, fnodes, left, right):
    """CGNS-ordered TETRA_4 / PYRA_5 / PENTA_6 / HEXA_8 node lists of the
    polyhedral cells: the base face (the quad of a pyramid or a hexahedron,
    either triangle of a prism, any triangle of a tetrahedron) ordered so that
    its normal points into the cell, then the opposite nodes along the
    edges leaving each base node."""
    cf = [[] for _ in range(nc)]
    for f in range(len(left)):
        for c in (left[f], right[f]):
            if c >= 0:
                cf[c].append(f)
    cell_ptr = np.zeros(nc + 1, dtype=np.int32)
    out = []
    for c in range(nc):
        faces = [list(fnodes[fptr[f]:fptr[f + 1]]) for f in cf[c]]
        nodes = sorted({n for fc in faces for n in fc})
        edges = set()
        for fc in faces:
            for a in range(len(fc)):
                e = (fc[a], fc[(a + 1) % len(fc)])
                edges.add((min(e), max(e)))
        ctr = xyz[nodes].mean(axis=0)
        tris = [fc for fc in faces if len(fc) == 3]
        quads = [fc for fc in faces if len(fc) == 4]
        nn = len(nodes)
        if nn == 4 and len(tris) == 4:
            base = tris[0]
        elif nn == 5 and len(quads) == 1 and len(tris) == 4:
            base = quads[0]
        elif nn == 6 and len(tris) == 2 and len(quads) == 3:
            base = tris[0]
        elif nn == 8 and len(quads) == 6:
            base = quads[0]
        else:
            raise TecplotGridError("cell %d (%d nodes, %d faces) is not a tetrahedron, pyramid, prism or hexahedron"
                                   % (c + 1, nn, len(faces)))
        base = list(base)
        p0, p1, p2 = xyz[base[0]], xyz[base[1]], xyz[base[2]]
        nrm = np.cross(p1 - p0, p2 - p0)
        if np.dot(nrm, ctr - p0) < 0:   # CGNS: the base's nodes run counter-clockwise seen from inside
            base = base[::-1]
        rest = [n for n in nodes if n not in base]
        if nn in (4, 5):
            cell = base + rest
        else:
            top = []
            for b in base:
                t = [n for n in rest if (min(b, n), max(b, n)) in edges]
                if len(t) != 1:
                    raise TecplotGridError("cell %d: no unique edge from a base node to the opposite face" % (c + 1))
                top.append(t[0])
            cell = base + top
        out.extend(cell)
        cell_ptr[c + 1] = len(out)
    return cell_ptr, np.array(out, dtype=np.int32)


def write_tecplot_grid(path, raw, metrics, title="Grid_1"):
    """A grid file that read_tecplot_grid (and Grid::readTEC) reads: the set-up's
    faces in reference order (fcfd_metrics: node lists, parent = left element,
    neighbour = right element, 0 on boundaries), one boundary zone per BC
    with the nodes of its faces."""
    xyz = np.asarray(raw["xyz"], dtype=np.float64).reshape(-1, 3)
    fp, fn, fb = metrics["face_parent"], metrics["face_neighbor"], metrics["face_bc"]
    nptr, fnodes = metrics["face_node_ptr"], metrics["face_nodes"]
    # metrics' face nodes are local ids (first touch in cell order): map back to raw ids
    cp = np.asarray(raw["cell_ptr"])
    cn = np.asarray(raw["cell_nodes"])[:cp[-1]]
    _, first = np.unique(cn, return_index=True)
    local = cn[np.sort(first)]
    F, nc = len(fp), len(cp) - 1
    out = ['TITLE = "%s"\n' % title, 'VARIABLES = "x", "y", "z"\n',
           'ZONE T="%s"\n' % title,
           "Nodes=%d, Faces=%d, Elements=%d, ZONETYPE=FEPolyhedron\n" % (len(xyz), F, nc),
           "DATAPACKING=BLOCK\n",
           "TotalNumFaceNodes=%d, NumConnectedBoundaryFaces=0, TotalNumBoundaryConnections=0\n" % nptr[F]]
    for i in range(3):
        out.append("\n".join(repr(float(v)) for v in xyz[:, i]) + "\n")
    out.append("# node count per face\n")
    out.append("".join("%d\n" % (nptr[f + 1] - nptr[f]) for f in range(F)))
    out.append("# face nodes\n")
    out.append("".join(" ".join("%d" % (local[k] + 1) for k in fnodes[nptr[f]:nptr[f + 1]]) + "\n" for f in range(F)))
    out.append("# left elements\n")
    out.append("".join("%d\n" % (c + 1) for c in fp))
    out.append("# right elements\n")
    out.append("".join("%d\n" % (0 if b >= 0 else c + 1) for c, b in zip(fn, fb)))
    names = raw.get("bc_names") or ["BC_%d" % (b + 1) for b in range(len(raw["bc_ptr"]) - 1)]
    for b, name in enumerate(names):
        fs = np.where(fb == b)[0]
        ids = np.unique(np.concatenate([local[fnodes[nptr[f]:nptr[f + 1]]] for f in fs])) if len(fs) else np.zeros(0, int)
        out.append('ZONE T="%s"\nNodes=%d, Elements=%d, ZONETYPE=FEQUADRILATERAL\nDATAPACKING=BLOCK\n' % (name, len(ids), len(fs)))
        for i in range(3):
            out.append("\n".join(repr(float(v)) for v in xyz[ids, i]) + "\n")
    with open(path, "w") as f:
        f.write("".join(out))
    return path


# ---------------------------------------------------------------------------
# grid transforms (grid_transform.cc:24-61), applied to raw["xyz"] in place
# ---------------------------------------------------------------------------
def translate(raw, begin, end):
    """raw.node[n] += end - begin"""
    x = np.asarray(raw["xyz"], dtype=np.float64).reshape(-1, 3)
    diff = np.asarray(end, dtype=np.float64) - np.asarray(begin, dtype=np.float64)
    x += diff
    raw["xyz"] = x.reshape(-1)
    return raw


def scale(raw, anchor, factor):
    """raw.node[n][i] = anchor[i] + factor[i] (raw.node[n][i] - anchor[i])"""
    x = np.asarray(raw["xyz"], dtype=np.float64).reshape(-1, 3)
    a = np.asarray(anchor, dtype=np.float64)
    fa = np.asarray(factor, dtype=np.float64)
    for i in range(3):
        x[:, i] = a[i] + fa[i] * (x[:, i] - a[i])
    raw["xyz"] = x.reshape(-1)
    return raw


def rotate(raw, anchor, axis, angle):
    """rotation by `angle` degrees about `axis` through `anchor` (Rodrigues'
    formula in the reference's arrangement)"""
    x = np.asarray(raw["xyz"], dtype=np.float64).reshape(-1, 3)
    angle = angle * (4. * math.atan(1.) / 180.)
    ax = np.asarray(axis, dtype=np.float64)
    m = math.sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2])
    a0, a1, a2 = ax[0] / m, ax[1] / m, ax[2] / m
    an = np.asarray(anchor, dtype=np.float64)
    p0, p1, p2 = x[:, 0] - an[0], x[:, 1] - an[1], x[:, 2] - an[2]
    c, s = math.cos(angle), math.sin(angle)
    dot = a0 * p0 + a1 * p1 + a2 * p2
    n0 = a0 * dot + (p0 * (1. - a0 * a0) - a0 * (a1 * p1 + a2 * p2)) * c + (-a2 * p1 + a1 * p2) * s
    n1 = a1 * dot + (p1 * (1. - a1 * a1) - a1 * (a0 * p0 + a2 * p2)) * c + (a2 * p0 - a0 * p2) * s
    n2 = a2 * dot + (p2 * (1. - a2 * a2) - a2 * (a0 * p0 + a1 * p1)) * c + (-a1 * p0 + a0 * p1) * s
    x[:, 0] = n0 + an[0]
    x[:, 1] = n1 + an[1]
    x[:, 2] = n2 + an[2]
    raw["xyz"] = x.reshape(-1)
    return raw
