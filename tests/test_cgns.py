"""CGNS (ADF) mesh reader (freecfd_amd/cgns.py) to the reference's reading
rules (src/grid/grid_reader_cgns.cc:28-431).

Two kinds of input:
* the reference's shipped example meshes, read where the reference tree is
  present: node and cell counts against the Zone_t sizes, BC numbers against
  the BC_n types of each example's input file, every boundary face matched to
  a BC (the set-up fails otherwise) with as many faces as its element range
  names, the two-zone cylinder stitched along its interface;
* small ADF files written here (single and multi-chunk data), whose expected
  raw mesh is known from their construction: point lists, point ranges,
  element lists and ranges, a MIXED section, zones listed out of name order,
  a BC name shared by two zones, stitching, and the reference's element-list
  cursor."""
import os

import numpy as np
import pytest

from freecfd_amd import cases
from freecfd_amd import cgns as CG

EX = "/root/reference/examples"


# ---------------------------------------------------------------- ADF writer

class _N:
    def __init__(self, name, label, dtype="MT", data=None, children=(), chunks=1):
        self.name, self.label, self.dtype, self.children, self.chunks = name, label, dtype, list(children), chunks
        if dtype == "C1":
            self.raw, self.dims = data.encode(), [len(data)]
        elif dtype == "MT":
            self.raw, self.dims = b"", []
        else:
            a = np.ascontiguousarray(data, dtype={"I4": "<i4", "R8": "<f8"}[dtype])
            self.raw, self.dims = a.tobytes(), list(a.shape[::-1])


def _ptr(pos):
    return b"%08X%04X" % (pos // 4096, pos % 4096)


def _write_adf(path, root_children):
    out = bytearray(b"\xc0\xa8\xa3\xa9ADF Database Version A02011>")
    out += b"AdF0" + b"Thu Dec  3 00:00:00 2026".ljust(28)
    out += b"AdF1" + b"Thu Dec  3 00:00:00 2026".ljust(28)
    out += b"AdF2LLAdF3" + b"010204040408040404040404" + b"AdF4" + b"\0" * 48 + b"AdF5"
    assert len(out) == 186

    def emit(nd):
        kids = [emit(c) for c in nd.children]
        chunk, nch = _ptr(0), 0
        if nd.raw:
            pieces = [nd.raw] if nd.chunks == 1 else [nd.raw[:len(nd.raw) // 2 // 4 * 4], nd.raw[len(nd.raw) // 2 // 4 * 4:]]
            spans = []
            for piece in pieces:
                start = len(out)
                out.extend(b"DaTa" + _ptr(start + 16 + len(piece)) + piece + b"dEnD")
                spans.append((start, start + 16 + len(piece)))
            if nd.chunks == 1:
                chunk, nch = _ptr(spans[0][0]), 1
            else:
                t = len(out)
                out.extend(b"DCtb" + _ptr(t + 16 + 24 * len(spans)) + b"".join(_ptr(a) + _ptr(b) for a, b in spans) + b"dcTE")
                chunk, nch = _ptr(t), len(spans)
        table = _ptr(0)
        if kids:
            t = len(out)
            out.extend(b"SNTb" + _ptr(t + 16 + 44 * len(kids)) +
                       b"".join(c.name.encode().ljust(32) + _ptr(p) for c, p in zip(nd.children, kids)) + b"snTb")
            table = _ptr(t)
        pos = len(out)
        rec = (b"NoDe" + nd.name.encode().ljust(32) + nd.label.encode().ljust(32) + b"%08X" % len(kids) +
               b"%08X" % len(kids) + table + nd.dtype.encode().ljust(32) + b"%02X" % len(nd.dims) +
               b"".join(b"%08X" % d for d in nd.dims) + b"00000000" * (12 - len(nd.dims)) +
               b"%04X" % nch + chunk + b"TaiL")
        assert len(rec) == 246
        out.extend(rec)
        return pos

    root = emit(_N("ADF MotherNode", "Root Node of ADF File", children=root_children))
    out[134:146] = _ptr(root)
    with open(path, "wb") as f:
        f.write(bytes(out))


# -------------------------------------------------------- synthetic meshes

def _box_zone(x0, order):
    """A 2 x 2 x 2 hex block on [x0, x0+1] x [0,1] x [0,1]; `order(i,j,k)` is
    the 0-based node number; returns (xyz, hexes, quads on x=x0+1 (i=2),
    wall quads on y and z = 0 or 1), all 1-based."""
    xyz = np.zeros((27, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                xyz[order(i, j, k)] = (x0 + 0.5 * i, 0.5 * j, 0.5 * k)
    n = lambda i, j, k: order(i, j, k) + 1
    hexes = [[n(i, j, k), n(i + 1, j, k), n(i + 1, j + 1, k), n(i, j + 1, k),
              n(i, j, k + 1), n(i + 1, j, k + 1), n(i + 1, j + 1, k + 1), n(i, j + 1, k + 1)]
             for k in range(2) for j in range(2) for i in range(2)]
    quads_x0 = [[n(0, j, k), n(0, j + 1, k), n(0, j + 1, k + 1), n(0, j, k + 1)] for j in range(2) for k in range(2)]
    quads_x2 = [[n(2, j, k), n(2, j + 1, k), n(2, j + 1, k + 1), n(2, j, k + 1)] for j in range(2) for k in range(2)]
    walls = []
    for j in (0, 2):
        walls += [[n(i, j, k), n(i + 1, j, k), n(i + 1, j, k + 1), n(i, j, k + 1)] for i in range(2) for k in range(2)]
    for k in (0, 2):
        walls += [[n(i, j, k), n(i + 1, j, k), n(i + 1, j + 1, k), n(i, j + 1, k)] for i in range(2) for j in range(2)]
    return xyz, hexes, quads_x0, quads_x2, walls


def _zone(name, xyz, sections, bcs, chunks=1):
    kids = [_N("ZoneType", "ZoneType_t", "C1", "Unstructured"),
            _N("GridCoordinates", "GridCoordinates_t", children=[
                _N("Coordinate" + a, "DataArray_t", "R8", xyz[:, d], chunks=chunks) for d, a in enumerate("XYZ")])]
    ncell = 0
    for sname, et, start, conn in sections:
        ne = len(conn)
        flat = [v for c in conn for v in c]
        kids.append(_N(sname, "Elements_t", "I4", [et, 0], [
            _N("ElementRange", "IndexRange_t", "I4", [start, start + ne - 1]),
            _N("ElementConnectivity", "DataArray_t", "I4", flat, chunks=chunks)]))
        if et in CG._VOLUME or et == CG.MIXED:
            ncell += ne
    kids.append(_N("ZoneBC", "ZoneBC_t", children=[
        _N(bname, "BC_t", "C1", "Null", [_N(kind, "IndexRange_t" if "Range" in kind else "IndexArray_t", "I4",
                                            np.asarray(vals).reshape(1, -1))])
        for bname, kind, vals in bcs]))
    return _N(name, "Zone_t", "I4", np.array([[len(xyz), ncell, 0]]), kids)


def _two_box_file(path, chunks=1):
    """Zone blk-1 on x in [0,1] (node (i,j,k) -> i+3j+9k) and zone blk-2 on
    [1,2] (node (i,j,k) -> j+3k+9i, so the x=2 nodes are 19..27), written
    blk-2 first.  blk-1: BCs walls (ElementRange) and left (PointList of the
    x=0 nodes); blk-2: walls again (ElementList, shuffled), right
    (PointRange 19..27), and its hexes in a MIXED section."""
    o1 = lambda i, j, k: i + 3 * j + 9 * k
    o2 = lambda i, j, k: j + 3 * k + 9 * i
    x1, h1, _, q1, w1 = _box_zone(0.0, o1)
    x2, h2, q2, _, w2 = _box_zone(1.0, o2)
    left = [o1(0, j, k) + 1 for j in range(3) for k in range(3)][::-1]
    z1 = _zone("blk-1", x1, [("hex", CG.HEXA_8, 1, h1), ("iface", CG.QUAD_4, 9, q1), ("walls", CG.QUAD_4, 13, w1)],
               [("walls", "ElementRange", [13, 28]), ("left", "PointList", left)], chunks)
    mixed = [[CG.HEXA_8] + h for h in h2]
    rng = np.random.default_rng(0)
    z2 = _zone("blk-2", x2, [("mixed", CG.MIXED, 1, mixed), ("iface", CG.QUAD_4, 9, q2), ("walls", CG.QUAD_4, 13, w2)],
               [("walls", "ElementList", rng.permutation(np.arange(13, 29))), ("right", "PointRange", [19, 27])], chunks)
    base = _N("Base", "CGNSBase_t", "I4", [3, 3], [z2, z1])
    _write_adf(path, [_N("CGNSLibraryVersion", "CGNSLibraryVersion_t"), base])
    return o1, o2, x1, x2, h1, h2


@pytest.mark.parametrize("chunks", [1, 2])
def test_two_zone_box(tmp_path, chunks):
    path = str(tmp_path / "box.cgns")
    o1, o2, x1, x2, h1, h2 = _two_box_file(path, chunks)
    raw = CG.read_cgns(path)
    # global nodes: blk-1 in its order, then blk-2's nodes off the interface
    # (i > 0) in its order; its i = 0 nodes take blk-1's i = 2 ids
    g2 = {}
    nxt = 27
    for n2 in range(27):
        i, r = divmod(n2, 9)
        k, j = divmod(r, 3)
        if i == 0:
            g2[n2] = o1(2, j, k)
        else:
            g2[n2] = nxt
            nxt += 1
    xyz = raw["xyz"].reshape(-1, 3)
    assert xyz.shape == (45, 3) and raw["dimension"] == 3
    assert np.array_equal(xyz[:27], x1)
    assert all(np.array_equal(xyz[g2[n]], x2[n]) for n in range(27))
    want_cells = [[v - 1 for v in h] for h in h1] + [[g2[v - 1] for v in h] for h in h2]
    cp, cn = raw["cell_ptr"], raw["cell_nodes"]
    assert cp.tolist() == list(range(0, 8 * 16 + 1, 8))
    assert [cn[cp[c]:cp[c + 1]].tolist() for c in range(16)] == want_cells
    # BCs: first appearance over the zones in name order; node sets by geometry
    assert raw["bc_names"] == ["walls", "left", "right"]
    sets = [raw["bc_nodes"][raw["bc_ptr"][b]:raw["bc_ptr"][b + 1]].tolist() for b in range(3)]
    on = lambda v, t: np.abs(v - t) < 1e-12
    walls = np.nonzero(on(xyz[:, 1], 0) | on(xyz[:, 1], 1) | on(xyz[:, 2], 0) | on(xyz[:, 2], 1))[0]
    assert sets == [walls.tolist(), np.nonzero(on(xyz[:, 0], 0))[0].tolist(), np.nonzero(on(xyz[:, 0], 2))[0].tolist()]
    # the set-up matches every boundary face to a BC; the interface is interior
    from freecfd_amd.ns import Grid
    g = Grid(raw, cases.make_options([cases.bc_symmetry()] * 3))
    m = g.metrics()
    assert np.bincount(m["face_bc"][m["face_bc"] >= 0]).tolist() == [32, 4, 4]
    assert g.info.n_faces == 40 + 2 * 12 + 4               # boundary, per-zone interior, interface
    assert abs(m["cell_volume"][:16].sum() - 2.0) < 1e-12
    g.close()


def test_element_list_cursor(tmp_path):
    """One hex, its six faces in one QUAD_4 section (elements 2..7), and a BC
    whose element list names elements 4 and 6: the reference's cursor reads
    the first two faces' nodes (grid_reader_cgns.cc:357-371)."""
    o = lambda i, j, k: i + 2 * j + 4 * k
    xyz = np.array([(i, j, k) for k in range(2) for j in range(2) for i in range(2)], float)
    n = lambda i, j, k: o(i, j, k) + 1
    hexa = [n(0, 0, 0), n(1, 0, 0), n(1, 1, 0), n(0, 1, 0), n(0, 0, 1), n(1, 0, 1), n(1, 1, 1), n(0, 1, 1)]
    faces = [[n(0, 0, 0), n(0, 1, 0), n(1, 1, 0), n(1, 0, 0)], [n(0, 0, 1), n(1, 0, 1), n(1, 1, 1), n(0, 1, 1)],
             [n(0, 0, 0), n(1, 0, 0), n(1, 0, 1), n(0, 0, 1)], [n(0, 1, 0), n(0, 1, 1), n(1, 1, 1), n(1, 1, 0)],
             [n(0, 0, 0), n(0, 0, 1), n(0, 1, 1), n(0, 1, 0)], [n(1, 0, 0), n(1, 1, 0), n(1, 1, 1), n(1, 0, 1)]]
    z = _zone("z", xyz, [("hex", CG.HEXA_8, 1, [hexa]), ("faces", CG.QUAD_4, 2, faces)],
              [("all", "PointRange", [1, 8]), ("two", "ElementList", [4, 6])])
    path = str(tmp_path / "one.cgns")
    _write_adf(path, [_N("Base", "CGNSBase_t", "I4", [3, 3], [z])])
    raw = CG.read_cgns(path, dimension=2)
    assert raw["dimension"] == 2 and raw["bc_names"] == ["all", "two"]
    assert raw["bc_nodes"][:raw["bc_ptr"][1]].tolist() == list(range(8))
    assert raw["bc_nodes"][raw["bc_ptr"][1]:].tolist() == sorted({v - 1 for v in faces[0] + faces[1]})
    # the cursor is shared by the section's element-list BCs: a second list
    # over the same six faces would read past the connectivity
    z = _zone("z", xyz, [("hex", CG.HEXA_8, 1, [hexa]), ("faces", CG.QUAD_4, 2, faces)],
              [("all", "ElementRange", [2, 7]), ("two", "ElementList", [4, 6])])
    _write_adf(path, [_N("Base", "CGNSBase_t", "I4", [3, 3], [z])])
    with pytest.raises(CG.AdfError, match="overruns"):
        CG.read_cgns(path)


def test_rejects_what_it_does_not_read(tmp_path):
    bad = tmp_path / "bad.cgns"
    bad.write_bytes(b"\0" * 400)
    with pytest.raises(CG.AdfError):
        CG.read_cgns(str(bad))
    z = _zone("z", np.zeros((4, 3)), [], [])
    z.children[0] = _N("ZoneType", "ZoneType_t", "C1", "Structured")
    path = str(tmp_path / "s.cgns")
    _write_adf(path, [_N("Base", "CGNSBase_t", "I4", [3, 3], [z])])
    with pytest.raises(CG.AdfError):
        CG.read_cgns(path)


# ------------------------------------------------------ the shipped meshes

_EXAMPLES = {
    # file: (zone sizes [nodes, cells], stitched nodes, BC names in BC_n order,
    #        the BC_n types of the example's input file)
    "wedge/wedge.cgns": ([(24442, 12000)], 0, ["inlet", "outlet", "symmetry", "wall"],
                         ["inlet", "outlet", "symmetry", "wall"]),
    "cylinder/cylinder.cgns": ([(5148, 2475), (26332, 26077)], 198,
                               ["Unspecified", "wall", "top-bottom", "inlet", "bc-4"],
                               ["symmetry", "wall", "symmetry", "inlet", "outlet"]),
    "hypersonic_cylinder/cylinder.cgns": ([(72762, 36000)], 0, ["inlet", "outlet", "symmetry", "wall"], None),
}


@pytest.mark.skipif(not os.path.isdir(EX), reason="the reference's example meshes are not present")
@pytest.mark.parametrize("name", sorted(_EXAMPLES))
def test_shipped_example_mesh(name):
    sizes, stitched, names, types = _EXAMPLES[name]
    raw = CG.read_cgns(os.path.join(EX, name), dimension=2)
    assert len(raw["xyz"]) // 3 == sum(s[0] for s in sizes) - stitched
    assert len(raw["cell_ptr"]) - 1 == sum(s[1] for s in sizes)
    assert raw["bc_names"] == names
    if types is not None:
        text = open(os.path.join(EX, os.path.dirname(name), os.path.basename(name).replace(".cgns", ".in"))).read()
        for b, t in enumerate(types):
            assert "BC_%d (type=%s;" % (b + 1, t) in text
    # BC face counts against the element ranges of the BC_t nodes
    f = CG.AdfFile(os.path.join(EX, name))
    zones = f.root["Base"].by_label("Zone_t")
    want = [0] * len(names)
    for z in zones:
        for bname, kind, vals in CG._zone_bocos(z):
            assert kind == "ElementRange"
            want[names.index(bname)] += int(vals[1] - vals[0] + 1)
    from freecfd_amd.ns import Grid
    g = Grid(raw, cases.make_options([cases.bc_symmetry()] * len(names)))
    m = g.metrics()
    fb = m["face_bc"]
    assert np.bincount(fb[fb >= 0], minlength=len(names)).tolist() == want
    assert (m["cell_volume"][:g.info.n_cells] > 0).all()
    g.close()


def test_wedge_example_fixture():
    """tests/golden/wedge_example_mesh.npz (the mesh of the shipped wedge
    example for its GPU test) has the example's sizes and BC names, and,
    where the reference tree is present, equals a fresh read of wedge.cgns"""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "wedge_example_mesh.npz")
    d = np.load(path, allow_pickle=False)
    assert len(d["xyz"]) // 3 == 24442 and len(d["cell_ptr"]) - 1 == 12000 and int(d["dimension"]) == 2
    assert [str(x) for x in d["bc_names"]] == ["inlet", "outlet", "symmetry", "wall"]
    if os.path.isdir(EX):
        raw = CG.read_cgns(os.path.join(EX, "wedge/wedge.cgns"), dimension=2)
        for k in ("xyz", "cell_ptr", "cell_nodes", "bc_ptr", "bc_nodes"):
            assert np.array_equal(np.asarray(raw[k]), d[k]), k
