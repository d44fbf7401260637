"""CGNS mesh reader on an in-repo ADF parser, to the reference's reading rules.

The reference reads its meshes through the CGNS 2.x mid-level library
(Grid::readCGNS, src/grid/grid_reader_cgns.cc:28-431); that library is not
here, so this module parses the ADF container itself and walks the CGNS tree
(CGNSBase_t -> Zone_t -> GridCoordinates_t, Elements_t, ZoneBC_t/BC_t) the way
the mid-level calls the reference makes see it.  The result is the raw mesh
dictionary fcfd_create takes (xyz, cell_ptr, cell_nodes, bc_ptr, bc_nodes,
dimension) plus the BC names, in the reference's numbering:

* nodes: zone after zone, a zone's nodes in file order; a node on a boundary
  face of a later zone that lies within 1e-10 (per coordinate) of a boundary
  node of an earlier zone takes that node's id (multi-zone stitching,
  grid_reader_cgns.cc:207-254);
* cells: the volume sections (TETRA_4, PYRA_5, PENTA_6, HEXA_8, and the
  volume members of MIXED) in zone, section and element order
  (grid_reader_cgns.cc:317-352);
* BC regions: numbered by first appearance of their names over the zones
  (grid_reader_cgns.cc:76-103); each holds the sorted set of nodes named by
  its point list or range, or by the nodes of the surface elements its element
  list or range names (grid_reader_cgns.cc:286-303,357-371).

ADF layout read here (the A02011 container of the shipped example meshes):
a disk pointer is 12 hex characters, an 8-digit block number and a 4-digit
byte offset into a 4096-byte block; the file header carries the root node
pointer after its "AdF4" tag; a node record is the "NoDe" tag, its name and
label (32 characters each), the child count and child-table capacity (8 hex
each), the child-table pointer, the data type (32 characters), the rank
(2 hex), 12 dimensions (8 hex each), the data-chunk count (4 hex) and pointer,
and the "TaiL" tag; a child table is "SNTb", its end pointer, then (name,
pointer) entries; a data chunk is "DaTa", its end pointer (to the chunk's end
tag), then the values in the file's byte order; data in several chunks is
reached through a "DCtb" table of (start, end) pointers.  Links to other
files are not followed.
"""
import numpy as np

_BLOCK = 4096
_DTYPES = {"I4": "i4", "I8": "i8", "U4": "u4", "U8": "u8", "R4": "f4", "R8": "f8", "C1": "S1", "B1": "u1"}

# CGNS ElementType_t values and their node counts (cgnslib.h of CGNS 2.x)
TRI_3, QUAD_4, TETRA_4, PYRA_5, PENTA_6, HEXA_8, MIXED = 5, 7, 10, 12, 14, 17, 20
_NPE = {3: 2, 4: 3, 5: 3, 6: 6, 7: 4, 8: 8, 9: 9, 10: 4, 11: 10, 12: 5, 13: 14, 14: 6, 15: 15, 16: 18,
        17: 8, 18: 20, 19: 27}
_VOLUME = (TETRA_4, PYRA_5, PENTA_6, HEXA_8)
STITCH_TOLERANCE = 1.e-10      # grid_reader_cgns.cc:26


class AdfError(ValueError):
    pass


class AdfNode:
    __slots__ = ("file", "name", "label", "dtype", "dims", "children", "_nchunks", "_chunk")

    def __getitem__(self, name):
        for c in self.children:
            if c.name == name:
                return c
        raise KeyError("%s has no child %r" % (self.name, name))

    def get(self, name):
        for c in self.children:
            if c.name == name:
                return c
        return None

    def by_label(self, label):
        return [c for c in self.children if c.label == label]

    def data(self):
        """The node's values: a 1-D array in file order, a str for C1."""
        if self.dtype == "MT":
            return None
        if self.dtype == "LK":
            raise AdfError("node %s is a link to another file; links are not followed" % self.name)
        if self.dtype not in _DTYPES:
            raise AdfError("node %s has unsupported data type %s" % (self.name, self.dtype))
        n = int(np.prod(self.dims)) if self.dims else 0
        dt = np.dtype(_DTYPES[self.dtype]).newbyteorder(self.file.order)
        if n == 0:
            return "" if self.dtype == "C1" else np.zeros(0, dt)
        f, buf = self.file, self.file.buf
        if self._nchunks == 1:
            spans = [(self._chunk, None)]
        else:
            t = self._chunk
            if buf[t:t + 4] != b"DCtb":
                raise AdfError("bad data-chunk table at %d" % t)
            spans = [(f._ptr(t + 16 + 24 * i), f._ptr(t + 28 + 24 * i)) for i in range(self._nchunks)]
        parts = []
        for start, end in spans:
            if buf[start:start + 4] != b"DaTa":
                raise AdfError("bad data chunk tag at %d" % start)
            end = f._ptr(start + 4) if end is None else end
            if end > len(buf) or end < start + 16:
                raise AdfError("data chunk of node %s runs past the end of the file" % self.name)
            parts.append(buf[start + 16:end])
        raw = parts[0] if len(parts) == 1 else b"".join(parts)
        if n * dt.itemsize > len(raw):
            raise AdfError("node %s holds %d bytes of data, its dimensions ask for %d" % (self.name, len(raw), n * dt.itemsize))
        a = np.frombuffer(raw, dtype=dt, count=n)
        if self.dtype == "C1":
            return a.tobytes().decode("latin-1").rstrip("\0 ")
        return a.astype(dt.newbyteorder("="))

    def ints(self):
        return self.data().astype(np.int64)


class AdfFile:
    """An ADF file read whole into memory; `root` is its top node."""

    def __init__(self, path):
        with open(path, "rb") as f:
            self.buf = f.read()
        b = self.buf
        if b[4:27] != b"ADF Database Version A0" or b[32:36] != b"AdF0":
            raise AdfError("%s is not an ADF file" % path)
        k = b.find(b"AdF2", 0, 186)
        k4 = b.find(b"AdF4", 0, 186)
        if k < 0 or k4 < 0:
            raise AdfError("%s: ADF header tags missing" % path)
        fmt = chr(b[k + 4])
        if fmt not in "LB":
            raise AdfError("%s: numeric format %r is not IEEE" % (path, fmt))
        self.order = "<" if fmt == "L" else ">"
        self.root = self._node(self._ptr(k4 + 4))

    def _ptr(self, p):
        s = self.buf[p:p + 12].decode("ascii")
        return int(s[:8], 16) * _BLOCK + int(s[8:], 16)

    def _str(self, p, n):
        return self.buf[p:p + n].decode("latin-1").rstrip("\0 ")

    def _node(self, p, depth=0):
        b = self.buf
        if depth > 64 or b[p:p + 4] != b"NoDe" or b[p + 242:p + 246] != b"TaiL":
            raise AdfError("bad node record at %d" % p)
        nd = AdfNode()
        nd.file = self
        nd.name = self._str(p + 4, 32)
        nd.label = self._str(p + 36, 32)
        nsub = int(b[p + 68:p + 76], 16)
        nd.dtype = self._str(p + 96, 32)
        rank = int(b[p + 128:p + 130], 16)
        nd.dims = [int(b[p + 130 + 8 * i:p + 138 + 8 * i], 16) for i in range(rank)]
        nd._nchunks = int(b[p + 226:p + 230], 16)
        nd._chunk = self._ptr(p + 230)
        nd.children = []
        if nsub:
            t = self._ptr(p + 84)
            if b[t:t + 4] != b"SNTb":
                raise AdfError("bad child table at %d" % t)
            for i in range(nsub):
                e = t + 16 + 44 * i
                nd.children.append(self._node(self._ptr(e + 32), depth + 1))
        return nd


def _zone_sections(zone):
    """(element type, start, end, connectivity) of each Elements_t child in
    child order (cg_section_read / cg_elements_read)."""
    out = []
    for s in zone.by_label("Elements_t"):
        et = int(s.ints()[0])
        rng = s["ElementRange"].ints()
        out.append((et, int(rng[0]), int(rng[1]), s["ElementConnectivity"].ints()))
    return out


def _zone_bocos(zone):
    """(name, point-set type, values) of each BC_t child of ZoneBC, in child
    order (cg_boco_info / cg_boco_read of CGNS 2.x: the point-set type is
    the name of the child that holds it)."""
    zbc = zone.get("ZoneBC")
    out = []
    if zbc is None:
        return out
    for bc in zbc.by_label("BC_t"):
        for kind in ("PointList", "PointRange", "ElementList", "ElementRange"):
            c = bc.get(kind)
            if c is not None:
                out.append((bc.name, kind, c.ints()))
                break
        else:
            raise AdfError("boundary condition %s has no point or element set" % bc.name)
    return out


def _surface_nodes(sections):
    """The 0-based node numbers on the zone's TRI_3 and QUAD_4 sections
    (grid_reader_cgns.cc:107-131)."""
    s = set()
    for et, a, b, conn in sections:
        if et in (TRI_3, QUAD_4):
            s.update((conn[:(b - a + 1) * _NPE[et]] - 1).tolist())
    return s


def read_cgns(path, dimension=None):
    """Read a CGNS (ADF) mesh the way Grid::readCGNS does; returns the raw
    mesh dict (xyz, cell_ptr, cell_nodes, bc_ptr, bc_nodes, dimension,
    bc_names).  `dimension` is the grid's `dimension` input key (2 for the
    one-cell-thick meshes of the shipped examples); by default the base's
    cell dimension."""
    f = AdfFile(path)
    bases = f.root.by_label("CGNSBase_t")
    if not bases:
        raise AdfError("%s holds no CGNSBase_t" % path)
    base = bases[0]                                   # the reference reads base 1 only
    cell_dim = int(base.ints()[0])
    # the CGNS library lists zones sorted by name; the shipped two-zone
    # cylinder mesh (examples/cylinder) needs that order for its BC numbers
    # to meet the BC_n types of its input file
    zones = sorted(base.by_label("Zone_t"), key=lambda z: z.name)
    bc_names = []                                     # bocoNameMap, by first appearance
    for z in zones:
        for name, _, _ in _zone_bocos(z):
            if name not in bc_names:
                bc_names.append(name)
    bc_sets = [set() for _ in bc_names]
    coords, zmaps, boxes, zbnd = [], [], [], []
    cell_ptr, cell_nodes = [0], []
    n_global = 0
    for zi, z in enumerate(zones):
        if z["ZoneType"].data() != "Unstructured":
            raise AdfError("zone %s is not unstructured" % z.name)
        nn = int(z.ints()[0])
        gc = z["GridCoordinates"]
        xyz = np.stack([gc["Coordinate" + a].data().astype(np.float64)[:nn] for a in "XYZ"], axis=1)
        coords.append(xyz)
        boxes.append((xyz.min(axis=0), xyz.max(axis=0)))
        sections = _zone_sections(z)
        zbnd.append(sorted(_surface_nodes(sections)))
        zmap = np.full(nn, -1, np.int64)
        if zi == 0:
            zmap[:] = np.arange(nn)
            n_global = nn
        else:
            tol = STITCH_TOLERANCE
            for n in zbnd[zi]:
                x = xyz[n]
                for zj in range(zi):
                    lo, hi = boxes[zj]
                    if (x < lo - tol).any() or (x > hi + tol).any():
                        continue
                    cand = np.asarray(zbnd[zj], np.int64)
                    if len(cand) == 0:
                        continue
                    hit = np.nonzero((np.abs(coords[zj][cand] - x) < tol).all(axis=1))[0]
                    if len(hit):
                        zmap[n] = zmaps[zj][cand[hit[0]]]
                        break
            for n in range(nn):
                if zmap[n] == -1:
                    zmap[n] = n_global
                    n_global += 1
        zmaps.append(zmap)
        # boundary conditions of this zone
        elem_lists = {}
        for name, kind, vals in _zone_bocos(z):
            b = bc_names.index(name)
            if kind == "PointList":
                bc_sets[b].update(zmap[vals - 1].tolist())
            elif kind == "PointRange":
                bc_sets[b].update(zmap[np.arange(vals[0], vals[1] + 1) - 1].tolist())
            elif kind == "ElementList":
                elem_lists.setdefault(b, set()).update(vals.tolist())
            else:
                elem_lists.setdefault(b, set()).update(range(int(vals[0]), int(vals[1]) + 1))
        # sections: cells, and the nodes of boundary elements
        for et, a, e, conn in sections:
            ne = e - a + 1
            if et == MIXED:
                k = 0
                for _ in range(ne):
                    t = int(conn[k])
                    if t not in _NPE:
                        raise AdfError("unknown element type %d in a MIXED section" % t)
                    npe = _NPE[t]
                    cell_nodes.extend(zmap[conn[k + 1:k + 1 + npe] - 1].tolist())
                    cell_ptr.append(len(cell_nodes))
                    k += npe + 1
            elif et in _VOLUME:
                npe = _NPE[et]
                nodes = zmap[conn[:ne * npe] - 1]
                cell_nodes.extend(nodes.tolist())
                cell_ptr.extend((cell_ptr[-1] + npe * np.arange(1, ne + 1)).tolist())
            elif et in _NPE:
                # The reference walks the section once per element-list BC and
                # reads the k-th member element's nodes at k * npe from where the
                # previous match left off, with one cursor for all BCs of the
                # section (grid_reader_cgns.cc:357-371).  That is each member's
                # own nodes when a section holds exactly the elements of one BC,
                # as the shipped meshes do; other layouts keep its reading.
                npe = _NPE[et]
                ids = a + np.arange(ne)
                ci = 0
                for b in sorted(elem_lists):
                    k = int(np.isin(ids, np.fromiter(elem_lists[b], np.int64)).sum())
                    if ci + k * npe > len(conn):
                        raise AdfError("element list of BC %s overruns section connectivity" % bc_names[b])
                    bc_sets[b].update(zmap[conn[ci:ci + k * npe] - 1].tolist())
                    ci += k * npe
            else:
                raise AdfError("unknown element type %d" % et)
    xyz = np.zeros((n_global, 3))
    for c, zmap in zip(coords, zmaps):
        xyz[zmap] = c
    bc_ptr, bc_nodes = [0], []
    for s in bc_sets:
        bc_nodes.extend(sorted(s))
        bc_ptr.append(len(bc_nodes))
    return {"xyz": xyz.reshape(-1), "cell_ptr": np.asarray(cell_ptr, np.int32),
            "cell_nodes": np.asarray(cell_nodes, np.int32), "bc_ptr": np.asarray(bc_ptr, np.int32),
            "bc_nodes": np.asarray(bc_nodes, np.int32), "dimension": cell_dim if dimension is None else int(dimension),
            "bc_names": bc_names}
