"""Synthetic unstructured meshes in the reference's raw cell-based input form.

Every generator returns a dict with
  xyz        (n_nodes, 3) float64 node coordinates
  cell_ptr   (n_cells+1,) int32 offsets into cell_nodes
  cell_nodes int32 node ids, CGNS element node order (TETRA_4, PENTA_6, HEXA_8)
  bc_names   list of region names
  bc_ptr     (n_bc+1,) int32 offsets into bc_nodes
  bc_nodes   int32 node ids of each boundary region (a node set, like a CGNS
             boundary condition given by point list; src/grid/grid.h:61-78)
  dimension  2 for one-cell-thick extruded meshes, else 3
The element node order is what the face tables of
src/grid/grid_create_elements.cc:187-214 expect: with it, every face normal
computed by src/grid/grid.cc:129-149 points out of its parent cell.
"""
import numpy as np


def _bc_pack(names_masks):
    names, lists = [], []
    for name, ids in names_masks:
        names.append(name)
        lists.append(np.unique(np.asarray(ids, dtype=np.int64)).astype(np.int32))
    ptr = np.zeros(len(lists) + 1, np.int32)
    ptr[1:] = np.cumsum([len(a) for a in lists])
    return names, ptr, (np.concatenate(lists) if lists else np.zeros(0, np.int32))


def _hex_cells(nid, ni, nj, nk):
    """Hex connectivity (CGNS HEXA_8) for cells (i,j,k), i fastest.  nid(i,j,k) maps
    (broadcast) integer index arrays to node ids."""
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    c = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                  nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)], axis=1)
    return c.astype(np.int32)


def _pack_cells(cell_lists):
    """cell_lists: list of (n_cells_i, nodes_per_cell_i) arrays in output order."""
    cells = []
    for arr in cell_lists:
        cells.extend(list(arr))
    ptr = np.zeros(len(cells) + 1, np.int32)
    ptr[1:] = np.cumsum([len(c) for c in cells])
    nodes = np.concatenate(cells).astype(np.int32) if cells else np.zeros(0, np.int32)
    return ptr, nodes


def structured_hex(X, bcs, dimension=3, periodic_i=False):
    """X: (NI, NJ, NK, 3) node coordinates (NI = cells_i + 1 unless periodic_i).
    bcs: list of (name, predicate(i, j, k) -> bool array) over node index arrays."""
    NI, NJ, NK = X.shape[:3]
    ni = NI if periodic_i else NI - 1
    nj, nk = NJ - 1, NK - 1

    def nid(i, j, k):
        ii = np.mod(i, NI) if periodic_i else i
        return ii + NI * (j + NJ * k)

    cells = _hex_cells(nid, ni, nj, nk)
    xyz = X.transpose(2, 1, 0, 3).reshape(-1, 3).copy()   # node id = i + NI*(j + NJ*k)
    k, j, i = np.meshgrid(np.arange(NK), np.arange(NJ), np.arange(NI), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    ids = i + NI * (j + NJ * k)
    names, bptr, bnodes = _bc_pack([(n, ids[pred(i, j, k)]) for n, pred in bcs])
    cptr = np.arange(0, 8 * len(cells) + 1, 8, dtype=np.int32)
    return dict(xyz=xyz, cell_ptr=cptr, cell_nodes=cells.reshape(-1), bc_names=names, bc_ptr=bptr,
                bc_nodes=bnodes, dimension=dimension, shape=(ni, nj, nk))


def pie(nr=6, nt=12, nz=3, radius=1.0, height=0.5, jitter=0.0, seed=0):
    """A cylinder of nr x nt x nz hexahedra in (r, theta, z) whose innermost
    ring touches the axis: those cells are hexahedra with two repeated nodes
    (the axis node at the bottom and at the top, positions 0 = 3 and 4 = 7 of
    the CGNS HEXA_8), the collapsed-edge cells a mesh generator writes for
    prisms.  BCs: "outer" (r = radius), "bottom" (z = 0), "top" (z = height)."""
    rng = np.random.default_rng(seed)
    rs = np.linspace(0.0, radius, nr + 1)
    ts = np.linspace(0.0, 2 * np.pi, nt + 1)[:-1]
    zs = np.linspace(0.0, height, nz + 1)
    # nodes: per z level the axis node, then rings i = 1..nr, nt nodes each
    per = 1 + nr * nt
    xyz = np.zeros(((nz + 1) * per, 3))
    for k, z in enumerate(zs):
        xyz[k * per] = (0.0, 0.0, z)
        for i in range(1, nr + 1):
            for j, t in enumerate(ts):
                r = rs[i]
                if jitter and i < nr:
                    r = r + jitter * (rs[1] - rs[0]) * rng.uniform(-1, 1)
                xyz[k * per + 1 + (i - 1) * nt + j] = (r * np.cos(t), r * np.sin(t), z)

    def nid(i, j, k):
        return k * per if i == 0 else k * per + 1 + (i - 1) * nt + (j % nt)

    cells = []
    for k in range(nz):
        for i in range(nr):
            for j in range(nt):
                cells.append([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                              nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)])
    cells = np.array(cells, dtype=np.int32)
    outer = [nid(nr, j, k) for k in range(nz + 1) for j in range(nt)]
    bottom = [nid(i, j, 0) for i in range(nr + 1) for j in range(nt)]
    top = [nid(i, j, nz) for i in range(nr + 1) for j in range(nt)]
    names, bptr, bnodes = _bc_pack([("outer", np.unique(outer)), ("bottom", np.unique(bottom)), ("top", np.unique(top))])
    cptr = np.arange(0, 8 * len(cells) + 1, 8, dtype=np.int32)
    return dict(xyz=xyz, cell_ptr=cptr, cell_nodes=cells.reshape(-1), bc_names=names, bc_ptr=bptr, bc_nodes=bnodes,
                dimension=3)


# ---------------------------------------------------------------------------
# benchmark meshes
# ---------------------------------------------------------------------------

def wedge(ni=96, nj=48, length=1.0, height=0.5, ramp_x=0.2, angle_deg=15.0, dz=0.01):
    """One-cell-thick 2-D compression ramp: bottom y_w(x) = 0 for x < ramp_x else
    (x - ramp_x) tan(angle).  Regions: inlet, outlet, bottom, top, front, back."""
    x = np.linspace(0.0, length, ni + 1)
    yw = np.where(x < ramp_x, 0.0, (x - ramp_x) * np.tan(np.radians(angle_deg)))
    X = np.zeros((ni + 1, nj + 1, 2, 3))
    s = np.linspace(0.0, 1.0, nj + 1)
    for k in range(2):
        X[:, :, k, 0] = x[:, None]
        X[:, :, k, 1] = yw[:, None] + (height - yw[:, None]) * s[None, :]
        X[:, :, k, 2] = k * dz
    bcs = [("inlet", lambda i, j, k: i == 0), ("outlet", lambda i, j, k: i == ni),
           ("bottom", lambda i, j, k: j == 0), ("top", lambda i, j, k: j == nj),
           ("front", lambda i, j, k: k == 0), ("back", lambda i, j, k: k == 1)]
    return structured_hex(X, bcs, dimension=2)


def cylinder_ogrid(nt=256, nr=128, r0=0.5, r1=30.0, ratio=1.05, jitter=0.05, dz=0.05, seed=2):
    """O-grid around a unit-diameter cylinder, geometric radial stretching, interior
    nodes jittered by +-jitter of the local spacing (seeded).  Index a = radial,
    b = circumferential (periodic); node (a, b) sits at radius r_a, angle 2 pi b / nt.
    Regions: wall (r=r0), inlet (outer, x<=0), outlet (outer, x>=0), front, back."""
    dr0 = (r1 - r0) * (ratio - 1.0) / (ratio ** nr - 1.0)
    r = r0 + dr0 * (ratio ** np.arange(nr + 1) - 1.0) / (ratio - 1.0)
    r[-1] = r1
    th = 2.0 * np.pi * np.arange(nt) / nt
    rng = np.random.default_rng(seed)
    R = np.repeat(r[:, None], nt, axis=1)
    TH = np.repeat(th[None, :], nr + 1, axis=0)
    dr = np.gradient(r)[:, None] * np.ones((1, nt))
    dth = (2.0 * np.pi / nt) * np.ones((nr + 1, nt))
    jr = rng.uniform(-jitter, jitter, size=R.shape) * dr
    jt = rng.uniform(-jitter, jitter, size=R.shape) * dth
    jr[0, :] = jr[-1, :] = 0.0
    jt[0, :] = jt[-1, :] = 0.0
    R = R + jr
    TH = TH + jt
    # radial index a is the first (i) direction, circumferential b the second (j, periodic)
    NA, NB = nr + 1, nt
    xyz = np.zeros((NA * NB * 2, 3))

    def nid(a, b, k):
        return a + NA * (np.mod(b, NB) + NB * k)

    for k in range(2):
        for b in range(NB):
            ids = nid(np.arange(NA), b, k)
            xyz[ids, 0] = R[:, b] * np.cos(TH[:, b])
            xyz[ids, 1] = R[:, b] * np.sin(TH[:, b])
            xyz[ids, 2] = k * dz
    cells = _hex_cells(nid, nr, nt, 1)
    k, b, a = np.meshgrid(np.arange(2), np.arange(NB), np.arange(NA), indexing="ij")
    a, b, k = a.ravel(), b.ravel(), k.ravel()
    ids = nid(a, b, k)
    q = nt // 4
    bcs = [("wall", ids[a == 0]),
           ("inlet", ids[(a == nr) & (b >= q) & (b <= 3 * q)]),
           ("outlet", ids[(a == nr) & ((b <= q) | (b >= 3 * q))]),
           ("front", ids[k == 0]), ("back", ids[k == 1])]
    names, bptr, bnodes = _bc_pack(bcs)
    cptr = np.arange(0, 8 * len(cells) + 1, 8, dtype=np.int32)
    return dict(xyz=xyz, cell_ptr=cptr, cell_nodes=cells.reshape(-1), bc_names=names, bc_ptr=bptr,
                bc_nodes=bnodes, dimension=2, shape=(nr, nt, 1))


# Kuhn split of a hexahedron into six positively oriented tetrahedra along the
# 0-6 diagonal (CGNS TETRA_4 order).  Conforming between neighbouring hexes.
_KUHN = np.array([[0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6], [0, 4, 5, 6], [0, 5, 1, 6]])


def tet_box(n=64, L=1.0):
    """[0,L]^3 with n^3 hexes, each split into 6 tetrahedra.  Regions:
    xmin, xmax, ymin, ymax, zmin, zmax."""
    g = np.linspace(0.0, L, n + 1)
    X = np.zeros((n + 1, n + 1, n + 1, 3))
    X[..., 0] = g[:, None, None]
    X[..., 1] = g[None, :, None]
    X[..., 2] = g[None, None, :]
    NI = n + 1

    def nid(i, j, k):
        return i + NI * (j + NI * k)

    hexes = _hex_cells(nid, n, n, n)
    tets = hexes[:, _KUHN].reshape(-1, 4)
    xyz = X.transpose(2, 1, 0, 3).reshape(-1, 3).copy()
    tets = _orient_tets(xyz, tets)
    k, j, i = np.meshgrid(np.arange(NI), np.arange(NI), np.arange(NI), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    ids = nid(i, j, k)
    bcs = [("xmin", ids[i == 0]), ("xmax", ids[i == n]), ("ymin", ids[j == 0]), ("ymax", ids[j == n]),
           ("zmin", ids[k == 0]), ("zmax", ids[k == n])]
    names, bptr, bnodes = _bc_pack(bcs)
    cptr = np.arange(0, 4 * len(tets) + 1, 4, dtype=np.int32)
    return dict(xyz=xyz, cell_ptr=cptr, cell_nodes=tets.reshape(-1), bc_names=names, bc_ptr=bptr,
                bc_nodes=bnodes, dimension=3, shape=(n, n, n))


def _orient_tets(xyz, tets):
    a, b, c, d = (xyz[tets[:, q]] for q in range(4))
    vol = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
    flip = vol < 0
    t = tets.copy()
    t[flip, 1], t[flip, 2] = tets[flip, 2], tets[flip, 1]
    return t


def _stretch(n, ratio, first=None, total=1.0):
    """n+1 points from 0 with geometric cell growth `ratio`; scaled to `total`
    unless `first` (first spacing) is given."""
    if ratio == 1.0:
        s = np.linspace(0.0, 1.0, n + 1)
    else:
        s = (ratio ** np.arange(n + 1) - 1.0) / (ratio ** n - 1.0)
    if first is not None:
        return s * first / s[1]
    return s * total


def bump_channel(ni=256, nj=96, nk=64, length=3.0, height=1.0, width=0.5, thickness=0.10, ratio=1.03):
    """3-D channel with a 10%-thick circular-arc bump on the lower wall between
    x = 1 and x = 2, wall-normal stretching `ratio`.  Regions: inlet, outlet,
    lower, upper, zmin, zmax."""
    x = np.linspace(0.0, length, ni + 1)
    chord = 1.0
    Rb = (0.25 * chord ** 2 + thickness ** 2) / (2.0 * thickness)
    on = (x >= 1.0) & (x <= 2.0)
    yb = np.zeros_like(x)
    yb[on] = np.sqrt(Rb ** 2 - (x[on] - 1.5) ** 2) - (Rb - thickness)
    s = _stretch(nj, ratio)
    z = np.linspace(0.0, width, nk + 1)
    X = np.zeros((ni + 1, nj + 1, nk + 1, 3))
    X[..., 0] = x[:, None, None]
    X[..., 1] = yb[:, None, None] + (height - yb[:, None, None]) * s[None, :, None]
    X[..., 2] = z[None, None, :]
    bcs = [("inlet", lambda i, j, k: i == 0), ("outlet", lambda i, j, k: i == ni),
           ("lower", lambda i, j, k: j == 0), ("upper", lambda i, j, k: j == nj),
           ("zmin", lambda i, j, k: k == 0), ("zmax", lambda i, j, k: k == nk)]
    return structured_hex(X, bcs, dimension=3)


def flat_plate(ni=512, nj=128, nk=64, length=1.0, width=0.25, first=5.5e-6, ratio=1.1, prism_frac=0.2):
    """Flat plate y=0, x in [0,L], geometric wall-normal spacing from `first`;
    the top `prism_frac` of the nj layers are split into two PENTA_6 prisms per
    hex (triangles in the x-y plane, extruded along z).  Regions: inlet,
    outlet, wall, top, zmin, zmax."""
    x = np.linspace(0.0, length, ni + 1)
    y = _stretch(nj, ratio, first=first)
    z = np.linspace(0.0, width, nk + 1)
    NI, NJ, NK = ni + 1, nj + 1, nk + 1

    def nid(i, j, k):
        return i + NI * (j + NJ * k)

    jp = nj - int(round(prism_frac * nj))       # first prism layer
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    hx = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                   nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)], axis=1)
    is_prism = j >= jp
    # cells in (k, j, i) order; a split hex contributes prism A then prism B
    pa = hx[:, [0, 1, 2, 4, 5, 6]]
    pb = hx[:, [0, 2, 3, 4, 6, 7]]
    sizes = np.where(is_prism, 12, 8)
    flat = np.empty(int(sizes.sum()), np.int32)
    off = np.concatenate([[0], np.cumsum(sizes)])
    hexmask = ~is_prism
    idx_hex = off[:-1][hexmask]
    for q in range(8):
        flat[idx_hex + q] = hx[hexmask, q]
    idx_pr = off[:-1][is_prism]
    for q in range(6):
        flat[idx_pr + q] = pa[is_prism, q]
        flat[idx_pr + 6 + q] = pb[is_prism, q]
    npc = np.repeat(np.where(is_prism, 6, 8), np.where(is_prism, 2, 1))
    cptr = np.zeros(len(npc) + 1, np.int32)
    cptr[1:] = np.cumsum(npc)
    X = np.zeros((NI, NJ, NK, 3))
    X[..., 0] = x[:, None, None]
    X[..., 1] = y[None, :, None]
    X[..., 2] = z[None, None, :]
    xyz = X.transpose(2, 1, 0, 3).reshape(-1, 3).copy()
    kk, jj, ii = np.meshgrid(np.arange(NK), np.arange(NJ), np.arange(NI), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    ids = nid(ii, jj, kk)
    bcs = [("inlet", ids[ii == 0]), ("outlet", ids[ii == ni]), ("wall", ids[jj == 0]), ("top", ids[jj == nj]),
           ("zmin", ids[kk == 0]), ("zmax", ids[kk == nk])]
    names, bptr, bnodes = _bc_pack(bcs)
    return dict(xyz=xyz, cell_ptr=cptr, cell_nodes=flat, bc_names=names, bc_ptr=bptr, bc_nodes=bnodes,
                dimension=3, shape=(ni, nj, nk))


def cell_centroids(raw):
    """Node-average cell centroids (src/grid/grid.cc:161-163) in raw cell order."""
    ptr = raw["cell_ptr"]
    nodes = raw["cell_nodes"]
    xyz = raw["xyz"]
    counts = np.diff(ptr)
    sums = np.add.reduceat(xyz[nodes], ptr[:-1], axis=0)
    return sums / counts[:, None]
