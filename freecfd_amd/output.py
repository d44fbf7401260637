"""Volume output in the reference's layouts, one rank: Tecplot FEPolyhedron
(below), VTK XML (write_vtk_volume_arrays), legacy VTK
(write_vtk_legacy_arrays); the loads history files (write_loads_arrays); and
Tecplot surface output (write_tecplot_surface_arrays).

Tecplot FEPolyhedron
(write_volume_output.cc:87-139 and write_tec_header :162-204,
write_tec_nodes :206-228, write_tec_var :230-432, write_tec_face_node_counts
:434-459, write_tec_face_nodes :461-490, write_tec_left :492-514,
write_tec_right :516-542).

./volume_output/volume_<step>_<g>.dat holds, in this order:
- the header: variable names (vectors as name_x, name_y, name_z), zone,
  counts, DATAPACKING=BLOCK, one DOUBLE per variable, the cell-centred
  variable range;
- x, y, z of every node (output id order: local node ids, first touch in
  cell order, grid_create_elements.cc:35-59 and :774-812), each block ending
  with a newline;
- each cell variable (vector components one block each), with no newline
  after the block;
- "# node count per face", one count per face; "# face nodes", the face's
  node output ids + 1 followed by a space each, one face per line;
  "# left elements", parent + 1; "# right elements", neighbour + 1, or 0 on
  a boundary face.
Values are written as `file << scientific << setprecision(8)` writes them
(printf's %.8e), ten to a line, TAB-separated.  A symbolic link
./volume_latest_<g>.dat points at the file, as the reference's `ln -sf`.

Variable names as in the reference's `volumevariables` list: p, T, rho, V,
Mach, volume, dt, rank, gradp, gradu, gradv, gradw, gradT, limiterp, limiterV,
limiterT, null; and nut for the build-only SA variable.
"""
import os

import numpy as np

_VEC = {"V", "gradp", "gradu", "gradv", "gradw", "gradT", "limiterV"}


def local_nodes(raw):
    """Raw ids of the local nodes in local (first touch in cell order) order."""
    cp = np.asarray(raw["cell_ptr"])
    cn = np.asarray(raw["cell_nodes"])[:cp[-1]]
    _, first = np.unique(cn, return_index=True)
    return cn[np.sort(first)]


def _block(vals):
    """The values, %.8e, a newline after every tenth and a TAB otherwise."""
    out = []
    for k, v in enumerate(vals):
        out.append("%.8e" % v)
        out.append("\n" if (k + 1) % 10 == 0 else "\t")
    return "".join(out)


def cell_fields(W, G=None, PHI=None, gamma=1.4, Mw=28.97, Pref=0.0, Tref=0.0, volume=None, dt=None):
    """The named cell variables of the state W (n_cells x 5 or 6: p, u, v, w,
    T[, nu~]), the gradients G (n x 3 nv) and limiters PHI (n x nv), in
    reference cell order (material.cc:108-153 for rho and the sound speed)."""
    n = W.shape[0]
    R = 8314.47215 / Mw
    p, V, T = W[:, 0], W[:, 1:4], W[:, 4]
    rho = (p + Pref) / (R * (T + Tref))
    a = np.sqrt(gamma * (p + Pref) / rho)
    f = {"p": p, "T": T, "rho": rho, "V": V, "Mach": np.sqrt((V * V).sum(axis=1)) / a, "rank": np.zeros(n, dtype=int)}
    if W.shape[1] == 6:
        f["nut"] = W[:, 5]
    if volume is not None:
        f["volume"] = np.asarray(volume)[:n]
    if dt is not None:
        f["dt"] = np.asarray(dt)[:n]
    if G is not None:
        for k, name in enumerate(("gradp", "gradu", "gradv", "gradw", "gradT")):
            f[name] = G[:n, 3 * k:3 * k + 3]
    if PHI is not None:
        f["limiterp"], f["limiterV"], f["limiterT"] = PHI[:n, 0], PHI[:n, 1:4], PHI[:n, 4]
    return f


def write_tecplot_volume_arrays(root, step, raw, metrics, fields, variables, gid=0):
    """Write the file from the raw mesh, fcfd_metrics output (reference
    numbering) and a dict of cell fields (cell_fields)."""
    nodes = local_nodes(raw)
    xyz = np.asarray(raw["xyz"], dtype=np.float64).reshape(-1, 3)[nodes]
    fp, fn, fb = metrics["face_parent"], metrics["face_neighbor"], metrics["face_bc"]
    nptr, fnodes = metrics["face_node_ptr"], metrics["face_nodes"]
    F = len(fp)
    n = len(raw["cell_ptr"]) - 1
    names, nvars = [], 3
    for v in variables:
        if v == "null":
            continue
        if v not in fields:
            raise ValueError("volume variable %r is not available" % v)
        if v in _VEC:
            names += [v + "_x", v + "_y", v + "_z"]
            nvars += 3
        else:
            names.append(v)
            nvars += 1
    out = ['VARIABLES = "x", "y", "z" ', "".join(',"%s" ' % s for s in names), "\n",
           'ZONE T="Grid_%d"\n' % (gid + 1),
           "Nodes=%d, Faces=%d, Elements=%d, ZONETYPE=FEPolyhedron\n" % (len(nodes), F, n),
           "DATAPACKING=BLOCK\n",
           "TotalNumFaceNodes=%d, NumConnectedBoundaryFaces=0, TotalNumBoundaryConnections=0 \n" % nptr[F],
           "DT=(" + "DOUBLE " * nvars + ")\n",
           "VARLOCATION=([4]=CELLCENTERED)\n" if nvars == 4 else "VARLOCATION=([4-%d]=CELLCENTERED)\n" % nvars]
    for i in range(3):
        out.append(_block(xyz[:, i]) + "\n")
    for v in variables:
        if v == "null":
            continue
        a = np.asarray(fields[v])
        if v in _VEC:
            for i in range(3):
                out.append(_block(a[:n, i]))
        elif v == "rank":
            out.append("".join("%d%s" % (x, "\n" if (k + 1) % 10 == 0 else "\t") for k, x in enumerate(a[:n])))
        else:
            out.append(_block(a[:n]))
    out.append("\n# node count per face\n")
    out.append("".join("%d\n" % (nptr[f + 1] - nptr[f]) for f in range(F)))
    out.append("# face nodes\n")
    out.append("".join("".join("%d " % (k + 1) for k in fnodes[nptr[f]:nptr[f + 1]]) + "\n" for f in range(F)))
    out.append("# left elements\n")
    out.append("".join("%d\n" % (c + 1) for c in fp))
    out.append("# right elements\n")
    out.append("".join("%d\n" % (0 if b >= 0 else c + 1) for c, b in zip(fn, fb)))
    os.makedirs(os.path.join(root, "volume_output"), exist_ok=True)
    name = "./volume_output/volume_%d_%d.dat" % (step, gid + 1)
    path = os.path.join(root, name)
    with open(path, "w") as f:
        f.write("".join(out))
    link = os.path.join(root, "volume_latest_%d.dat" % (gid + 1))
    if os.path.lexists(link):
        os.remove(link)
    os.symlink(name, link)
    return path


def write_tecplot_volume(ns, raw, step, variables=("p", "V", "T"), root=".", gid=0, dt=None):
    """write_volume_output(gid, step) with format tecplot for a NavierStokes
    context: the current state and, when requested, the gradients and
    limiters of the last residual evaluation."""
    W = ns.state_host(include_ghosts=False)
    G = PHI = None
    if any(v.startswith("grad") or v.startswith("limiter") for v in variables):
        G, PHI = ns.gradients_limiters_host()
    m = ns.metrics()
    o = ns._opt
    f = cell_fields(W, G, PHI, gamma=o.gamma, Mw=o.Mw, Pref=o.Pref, Tref=o.Tref, volume=m["cell_volume"], dt=dt)
    return write_tecplot_volume_arrays(root, step, raw, m, f, variables, gid)


def _vtk_var_open(name, vec):
    return '<DataArray Name="%s" %stype="Float32" format="ascii" >\n' % (name, 'Number of Components="3" ' if vec else "")


def write_vtk_volume_arrays(root, step, raw, fields, variables, gid=0):
    """Format `vtk` of write_volume_output (write_vtk_parallel
    write_volume_output.cc:685-719 and write_vtk :555-683), one rank: the
    .pvtu index and the .vtu piece.  Quirks kept as the reference writes
    them: vectors carry the attribute `Number of Components="3"`, a `dt`
    variable writes the cells' T, and "null" writes an empty array.  Points
    are `setw(16) << setprecision(8) << scientific` (%16.8e), one coordinate
    per line; cell values then keep the stream's scientific precision 8
    (%.8e)."""
    nodes = local_nodes(raw)
    xyz = np.asarray(raw["xyz"], dtype=np.float64).reshape(-1, 3)[nodes]
    g2l = np.full(int(np.asarray(raw["cell_nodes"]).max()) + 1, -1, dtype=np.int64)
    g2l[nodes] = np.arange(len(nodes))
    cp = np.asarray(raw["cell_ptr"])
    cn = np.asarray(raw["cell_nodes"])
    n = len(cp) - 1
    d = os.path.join(root, "volume_output", "%d" % step)
    os.makedirs(d, exist_ok=True)
    pv = ['<?xml version="1.0"?>\n', '<VTKFile type="PUnstructuredGrid">\n', '<PUnstructuredGrid GhostLevel="0">\n', "<PPoints>\n",
          '<DataArray NumberOfComponents="3" type="Float32" format="ascii" />\n', "</PPoints>\n", '<PCellData format="ascii">\n']
    pv += [_vtk_var_open(v, v in _VEC) for v in variables]
    pv += ["</PCellData>\n", '<Piece Source="grid_%d_proc_0.vtu" />\n' % (gid + 1), "</PUnstructuredGrid>\n", "</VTKFile>\n"]
    with open(os.path.join(d, "grid_%d_volume_%d.pvtu" % (gid + 1, step)), "w") as f:
        f.write("".join(pv))
    out = ['<?xml version="1.0"?>\n', '<VTKFile type="UnstructuredGrid">\n', "<UnstructuredGrid>\n",
           '<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (len(nodes), n), "<Points>\n",
           '<DataArray NumberOfComponents="3" type="Float32" format="ascii" >\n']
    out.append("".join("%16.8e\n" % x for x in xyz.reshape(-1)))
    out += ["</DataArray>\n", "</Points>\n", "<Cells>\n", '<DataArray Name="connectivity" type="Int32" format="ascii" >\n']
    out.append("".join("".join("%d " % g2l[k] for k in cn[cp[c]:cp[c + 1]]) + "\n" for c in range(n)))
    out += ["</DataArray>\n", '<DataArray Name="offsets" type="Int32" format="ascii" >\n']
    out.append("".join("%d\n" % o for o in cp[1:]))
    out += ["</DataArray>\n", '<DataArray Name="types" type="UInt8" format="ascii" >\n']
    vtk_type = {4: "10", 8: "12", 6: "13", 5: "14"}
    out.append("".join(vtk_type[int(k)] + "\n" for k in np.diff(cp)))
    out += ["</DataArray>\n", "</Cells>\n", '<CellData format="ascii">\n']
    for v in variables:
        vec = v in _VEC
        out.append(_vtk_var_open(v, vec))
        if v == "null":
            pass
        elif v == "rank":
            out.append("0\n" * n)
        else:
            src = "T" if v == "dt" else v
            if src not in fields:
                raise ValueError("volume variable %r is not available" % v)
            a = np.asarray(fields[src])
            if vec:
                out.append("".join("%.8e %.8e %.8e \n" % tuple(a[c]) for c in range(n)))
            else:
                out.append("".join("%.8e\n" % x for x in a[:n]))
        out.append("</DataArray>\n")
    out += ["</CellData>\n", "</Piece>\n", "</UnstructuredGrid>\n", "</VTKFile>\n"]
    path = os.path.join(d, "grid_%d_proc_0.vtu" % (gid + 1))
    with open(path, "w") as f:
        f.write("".join(out))
    return path


def write_vtk_volume(ns, raw, step, variables=("p", "V", "T"), root=".", gid=0):
    """write_volume_output(gid, step) with format vtk for a NavierStokes context."""
    W = ns.state_host(include_ghosts=False)
    G = PHI = None
    if any(v.startswith("grad") or v.startswith("limiter") for v in variables):
        G, PHI = ns.gradients_limiters_host()
    m = ns.metrics()
    o = ns._opt
    f = cell_fields(W, G, PHI, gamma=o.gamma, Mw=o.Mw, Pref=o.Pref, Tref=o.Tref, volume=m["cell_volume"])
    return write_vtk_volume_arrays(root, step, raw, f, variables, gid)


# write_vtk_legacy's variable chain (write_volume_output.cc:742-786): the
# names it writes values for; any other name gets its SCALARS header and no
# values.  Of these, the ones without a source here are refused.
_LEGACY_VALUES = {"p", "T", "rho", "dt", "rank", "V", "gradp", "gradu", "gradv", "gradw", "gradT", "limiterp",
                  "limiterV", "limiterT", "Mach"}
_LEGACY_ABSENT = {"mu", "lambda", "Cp", "resp", "resT", "resV", "k", "omega", "mu_t", "gradrho", "gradk", "gradomega"}


def write_vtk_legacy_arrays(root, step, raw, fields, variables, gid=0):
    """Format `vtklegacy` of write_volume_output, one rank: the .visit index
    (write_visit_parallel, write_volume_output.cc:789-809, with its
    ./volume_output/volume_latest_<g>.visit link) and the legacy ASCII piece
    (write_vtk_legacy :701-787).  As the reference writes them: points one
    coordinate a line as `setw(16) << setprecision(8) << scientific` (%16.8e),
    later doubles in the stream's scientific precision 8 (%.8e); vector
    variables (V, grad*, limiterV) as three SCALARS blocks name_x, name_y,
    name_z; `dt` writes the cells' T; a name outside its variable chain
    (null, volume, ...) writes its header with no values.  nu~ of the
    build-only SA model is written under `nut`."""
    nodes = local_nodes(raw)
    xyz = np.asarray(raw["xyz"], dtype=np.float64).reshape(-1, 3)[nodes]
    g2l = np.full(int(np.asarray(raw["cell_nodes"]).max()) + 1, -1, dtype=np.int64)
    g2l[nodes] = np.arange(len(nodes))
    cp = np.asarray(raw["cell_ptr"])
    cn = np.asarray(raw["cell_nodes"])
    n = len(cp) - 1
    g = gid + 1
    vo = os.path.join(root, "volume_output")
    d = os.path.join(vo, "%d" % step)
    os.makedirs(d, exist_ok=True)
    visit = "volume_%d_%d.visit" % (step, g)
    with open(os.path.join(vo, visit), "w") as f:
        f.write("!NBLOCKS 1\n./%d/grid_%d_proc_0.vtk\n" % (step, g))
    link = os.path.join(vo, "volume_latest_%d.visit" % g)
    if os.path.lexists(link):
        os.remove(link)
    os.symlink("./" + visit, link)
    out = ["# vtk DataFile Version 2.0\n", "Grid_%d\n" % g, "ASCII\n", "DATASET UNSTRUCTURED_GRID\n",
           "POINTS %d double\n" % len(nodes), "".join("%16.8e\n" % x for x in xyz.reshape(-1)),
           "CELLS %d %d\n" % (n, cp[n] + n)]
    out.append("".join("%d" % (cp[c + 1] - cp[c]) + "".join(" %d" % g2l[k] for k in cn[cp[c]:cp[c + 1]]) + "\n"
                       for c in range(n)))
    out.append("CELL_TYPES %d\n" % n)
    vtk_type = {4: "10\n", 8: "12\n", 6: "13\n", 5: "14\n"}
    out.append("".join(vtk_type.get(int(k), "") for k in np.diff(cp)))
    out.append("CELL_DATA %d\n" % n)
    for v in variables:
        if v in _LEGACY_ABSENT:
            raise ValueError("volume variable %r is not available" % v)
        vec = v == "V" or v.startswith("grad") or v == "limiterV"
        src = "T" if v == "dt" else v
        has = v in _LEGACY_VALUES or v == "nut"
        if has and src not in fields:
            raise ValueError("volume variable %r is not available" % v)
        for i in range(3 if vec else 1):
            out.append("SCALARS %s double\nLOOKUP_TABLE default\n" % (v + ("_x", "_y", "_z")[i] if vec else v))
            if not has:
                continue
            a = np.asarray(fields[src])
            if v == "rank":
                out.append("0\n" * n)
            else:
                out.append("".join("%.8e\n" % x for x in (a[:n, i] if vec else a[:n])))
    path = os.path.join(d, "grid_%d_proc_0.vtk" % g)
    with open(path, "w") as f:
        f.write("".join(out))
    return path


def write_vtk_legacy(ns, raw, step, variables=("p", "V", "T"), root=".", gid=0):
    """write_volume_output(gid, step) with format vtklegacy for a NavierStokes context."""
    W = ns.state_host(include_ghosts=False)
    G = PHI = None
    if any(v.startswith("grad") or v.startswith("limiter") for v in variables):
        G, PHI = ns.gradients_limiters_host()
    o = ns._opt
    f = cell_fields(W, G, PHI, gamma=o.gamma, Mw=o.Mw, Pref=o.Pref, Tref=o.Tref)
    return write_vtk_legacy_arrays(root, step, raw, f, variables, gid)


def write_loads_arrays(root, step, time, loads, include_bcs, gid=0):
    """write_loads (write_volume_output.cc:65-85): for each BC index b in
    include_bcs, append `step<TAB>time<TAB>Fx<TAB>Fy<TAB>Fz<TAB>Mx<TAB>My<TAB>Mz`
    to ./volume_output/loads_grid_<g>_BC_<b+1>.dat, the doubles as
    `scientific << setprecision(8)` writes them (%.8e).  loads is (n_bc, 6):
    forces then moments per BC."""
    L = np.asarray(loads, dtype=np.float64)
    vo = os.path.join(root, "volume_output")
    os.makedirs(vo, exist_ok=True)
    paths = []
    for b in include_bcs:
        path = os.path.join(vo, "loads_grid_%d_BC_%d.dat" % (gid + 1, b + 1))
        with open(path, "a") as f:
            f.write("%d\t%.8e\t" % (step, time) + "\t".join("%.8e" % x for x in L[b]) + "\n")
        paths.append(path)
    return paths


def write_loads(ns, step, time, root=".", gid=0):
    """write_loads(gid, step, time) for a NavierStokes context: the BCs with
    include_in_loads, the loads of its current state (fcfd_loads)."""
    inc = [b for b in range(int(ns._opt.n_bc)) if ns._opt.bc[b].include_in_loads]
    return write_loads_arrays(root, step, time, ns.loads(), inc, gid)


def face_values(metrics, Wg, bcs, gamma=1.4, Mw=28.97, Pref=0.0, Tref=0.0):
    """p, T, rho, V and Mach on every face, as Variable::face gives them
    (variable/variable.h:135-148): on a face of a BC where the variable is
    fixed (fixedonBC), the BC's value; otherwise the face-average stencil
    over cells and ghosts (fcfd_metrics avg_ptr/avg_idx/avg_w) applied to the
    cell values, summed in stencil order.  Wg is the state with ghosts
    ((N+B) x nv, reference order); bcs the options' BC records; Mach is
    |V_f| / a(p_f, T_f) (write_surface_output.cc:327-332, material.cc)."""
    ap, ai, aw = metrics["avg_ptr"], metrics["avg_idx"], metrics["avg_w"]
    fb = metrics["face_bc"]
    F = len(fb)
    R = 8314.47215 / Mw
    Wg = np.asarray(Wg, dtype=np.float64)
    rho_c = (Wg[:, 0] + Pref) / (R * (Wg[:, 4] + Tref))
    X = np.concatenate([Wg[:, :5], rho_c[:, None]], axis=1)       # p u v w T rho
    out = np.zeros((F, 6))
    cnt = np.diff(ap)
    for k in range(int(cnt.max()) if F else 0):
        sel = np.nonzero(cnt > k)[0]
        out[sel] += aw[ap[sel] + k][:, None] * X[ai[ap[sel] + k]]
    for b, bc in enumerate(bcs):
        on = fb == b
        if bc.p_fixed:
            out[on, 0] = bc.p
        if bc.V_fixed:
            out[on, 1:4] = np.asarray(bc.V[:3])
        if bc.T_fixed:
            out[on, 4] = bc.T
        if bc.rho_fixed:
            out[on, 5] = bc.rho
    p, V, T = out[:, 0], out[:, 1:4], out[:, 4]
    a = np.sqrt(gamma * (p + Pref) / ((p + Pref) / (R * (T + Tref))))
    return {"p": p, "V": V, "T": T, "rho": out[:, 5], "Mach": np.sqrt((V * V).sum(axis=1)) / a,
            "rank": np.zeros(F, dtype=int)}


def write_tecplot_surface_arrays(root, step, raw, metrics, fvals, variables, gid=0):
    """write_surface_output with format tecplot (write_surface_output.cc:57-433),
    one rank: ./surface_output/surface_<step>_<g>.dat and the
    ./surface_latest_<g>.dat link.  One FEQUADRILATERAL zone per BC region;
    the first zone carries the boundary nodes (local node ids that lie on a
    boundary face, in increasing order: get_bc_output_ids,
    grid_create_elements.cc:883-931), later zones share them
    (VARSHARELIST).  Per zone, the face values of the BC's faces in reference
    face order, one block per variable (vector components V, tau: one block
    each), ten to a line and with no newline after a block; then one line
    per face of its output node ids + 1, TAB after each, a triangle's third
    node written twice.  fvals: face_values (p, T, rho, V, Mach, rank), and
    the face loop's mdot, qdot, tau (per face, reference order) when given."""
    fb = metrics["face_bc"]
    nptr, fnodes = metrics["face_node_ptr"], metrics["face_nodes"]
    n_bc = int(np.asarray(raw["bc_ptr"]).size - 1)
    xyz = np.asarray(raw["xyz"], dtype=np.float64).reshape(-1, 3)[local_nodes(raw)]
    bface = np.nonzero(fb >= 0)[0]
    bnodes = np.unique(np.concatenate([fnodes[nptr[f]:nptr[f + 1]] for f in bface])) if len(bface) else np.zeros(0, int)
    out_id = np.full(len(xyz), -2, dtype=np.int64)
    out_id[bnodes] = np.arange(len(bnodes))
    vec = {"V", "tau"}
    nvars = 3
    for v in variables:
        if v != "null" and v not in fvals:
            raise ValueError("surface variable %r is not available" % v)
        nvars += 3 if v in vec else (0 if v == "null" else 1)
    out = []
    for b in range(n_bc):
        faces = np.nonzero(fb == b)[0]
        if b == 0:
            out.append('VARIABLES = "x", "y", "z" ' + "".join(
                ',"%s_x" ,"%s_y" ,"%s_z" ' % (v, v, v) if v in vec else ('' if v == "null" else ',"%s" ' % v)
                for v in variables) + "\n")
        out.append('ZONE, T="BC_%d", ZONETYPE=FEQUADRILATERAL, DATAPACKING=BLOCK\n' % (b + 1))
        if b == 0:
            out.append("NODES=%d, " % len(bnodes))
        out.append("ELEMENTS=%d\n" % len(faces))
        out.append("VARLOCATION=([4]=CELLCENTERED)\n" if nvars == 4 else "VARLOCATION=([4-%d]=CELLCENTERED)\n" % nvars)
        if b != 0:
            out.append("VARSHARELIST=([1-3]=1)\n")
        if b == 0:
            for i in range(3):
                out.append(_block(xyz[bnodes, i]) + "\n")
        for v in variables:
            if v == "null":
                continue
            a = np.asarray(fvals[v])
            if v == "rank":
                out.append("".join("%d%s" % (x, "\n" if (k + 1) % 10 == 0 else "\t") for k, x in enumerate(a[faces])))
            elif v in vec:
                for i in range(3):
                    out.append(_block(a[faces, i]))
            else:
                out.append(_block(a[faces]))
        for f in faces:
            ids = out_id[fnodes[nptr[f]:nptr[f + 1]]] + 1
            row = list(ids[:3]) + ([ids[3]] if len(ids) == 4 else [ids[2]])
            out.append("".join("%d\t" % x for x in row) + "\n")
    os.makedirs(os.path.join(root, "surface_output"), exist_ok=True)
    name = "./surface_output/surface_%d_%d.dat" % (step, gid + 1)
    path = os.path.join(root, name)
    with open(path, "w") as f:
        f.write("".join(out))
    link = os.path.join(root, "surface_latest_%d.dat" % (gid + 1))
    if os.path.lexists(link):
        os.remove(link)
    os.symlink(name, link)
    return path


def write_tecplot_surface(ns, raw, step, variables=("p", "V", "T"), root=".", gid=0):
    """write_surface_output(gid, step) with format tecplot for a NavierStokes
    context: face values from its current state and the ghosts of its last
    residual evaluation (or gradients_limiters call); mdot, qdot and tau
    (write_surface_output.cc:266-290, 390) from the face loop's surface
    quantities of the current state (fcfd_surface: mdot.face, qdot.bc,
    tau.bc, ns_assemble_linear_system.cc:122-127)."""
    m = ns.metrics()
    o = ns._opt
    fv = face_values(m, ns.state_host(include_ghosts=True), [o.bc[b] for b in range(int(o.n_bc))],
                     gamma=o.gamma, Mw=o.Mw, Pref=o.Pref, Tref=o.Tref)
    if {"mdot", "qdot", "tau"} & set(variables):
        mdot, qdot, tau = ns.surface()
        fv.update(mdot=mdot, qdot=qdot, tau=tau)
    return write_tecplot_surface_arrays(root, step, raw, m, fv, variables, gid)
