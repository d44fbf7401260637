"""Tecplot FEPolyhedron volume output in the reference's layout
(freecfd_amd/output.py; write_volume_output.cc:87-542): the file parsed back
against the mesh metrics (node order pinned by the face centroids, which the
reference computes as node averages, grid.cc:123-178), the header, the face
connectivity and the cell variables; the number format against a compiled
C++ ostream."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from freecfd_amd import cases
from freecfd_amd import output as O


def _parse(path, n_nodes, n_faces):
    text = open(path).read()
    head, tail = text.split("\n# node count per face\n")
    lines = head.split("\n")
    header, nums = lines[:7], np.array(" ".join(lines[7:]).split(), dtype=np.float64)
    counts_txt, rest = tail.split("# face nodes\n")
    nodes_txt, rest = rest.split("# left elements\n")
    left_txt, right_txt = rest.split("# right elements\n")
    counts = np.array(counts_txt.split(), dtype=int)
    faces = [[int(t) for t in ln.split()] for ln in nodes_txt.rstrip("\n").split("\n")]
    return (header, nums[:3 * n_nodes].reshape(3, n_nodes).T, nums[3 * n_nodes:], counts, faces,
            np.array(left_txt.split(), dtype=int), np.array(right_txt.split(), dtype=int), text)


@pytest.mark.parametrize("name", ["plate", "tetbox"])
def test_volume_file_against_the_mesh(tmp_path, name):
    from freecfd_amd.ns import Grid
    cs = cases.plate_case(6, 4, 2, sa=True) if name == "plate" else cases.tetbox_case(2)
    g = Grid(cs.raw, cs.opt)
    m = g.metrics()
    N, F = g.info.n_cells, g.info.n_faces
    nv = cs.W0.shape[1]
    rng = np.random.default_rng(1)
    G, PHI = rng.standard_normal((N, 3 * nv)), rng.random((N, nv))
    f = O.cell_fields(cs.W0, G, PHI, gamma=cs.opt.gamma, Mw=cs.opt.Mw, Pref=cs.opt.Pref, Tref=cs.opt.Tref,
                      volume=m["cell_volume"])
    variables = ["p", "V", "null", "T", "rho", "Mach", "volume", "gradu", "limiterT", "rank"] + (["nut"] if nv == 6 else [])
    path = O.write_tecplot_volume_arrays(str(tmp_path), 3, cs.raw, m, f, variables)
    n_nodes = len(O.local_nodes(cs.raw))
    header, xyz, vals, counts, faces, left, right, text = _parse(path, n_nodes, F)
    names = ["p", "V_x", "V_y", "V_z", "T", "rho", "Mach", "volume", "gradu_x", "gradu_y", "gradu_z", "limiterT", "rank"] + \
        (["nut"] if nv == 6 else [])
    nvars = 3 + len(names)
    assert header[0] == 'VARIABLES = "x", "y", "z" ' + "".join(',"%s" ' % s for s in names)
    assert header[1] == 'ZONE T="Grid_1"'
    assert header[2] == "Nodes=%d, Faces=%d, Elements=%d, ZONETYPE=FEPolyhedron" % (n_nodes, F, N)
    assert header[3] == "DATAPACKING=BLOCK"
    assert header[4] == "TotalNumFaceNodes=%d, NumConnectedBoundaryFaces=0, TotalNumBoundaryConnections=0 " % counts.sum()
    assert header[5] == "DT=(" + "DOUBLE " * nvars + ")"
    assert header[6] == "VARLOCATION=([4-%d]=CELLCENTERED)" % nvars
    # nodes in output-id order: every face centroid is the average of its nodes
    assert counts.tolist() == np.diff(m["face_node_ptr"]).tolist()
    fc = m["face_centroid"].reshape(-1, 3)
    cen = np.array([xyz[np.array(fn) - 1].mean(axis=0) for fn in faces])
    scale = np.abs(fc).max()
    assert np.abs(cen - fc).max() <= 1e-7 * scale
    assert left.tolist() == (m["face_parent"] + 1).tolist()
    assert right.tolist() == np.where(m["face_bc"] >= 0, 0, m["face_neighbor"] + 1).tolist()
    # the cell variables, block after block
    want = [f["p"], f["V"][:, 0], f["V"][:, 1], f["V"][:, 2], f["T"], f["rho"], f["Mach"], f["volume"][:N],
            G[:, 3], G[:, 4], G[:, 5], PHI[:, 4], np.zeros(N)] + ([cs.W0[:, 5]] if nv == 6 else [])
    got = vals.reshape(len(want), N)
    for w, x in zip(want, got):
        assert np.allclose(x, w, rtol=1e-8, atol=0.0)
    assert "\t" in text and text.count("\n") > 10


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_numbers_match_cpp_scientific_8(tmp_path):
    vals = [101325.0, -3.5e-7, 0.0, 1.0 / 3.0, 6.02214076e23, 300.0, -0.0]
    src = tmp_path / "fmt.cc"
    src.write_text("#include <fstream>\n#include <iomanip>\nint main(int, char** argv) {\n  std::ofstream f(argv[1]);\n"
                   "  f << std::scientific << std::setprecision(8);\n"
                   + "".join("  f << %r << \"\\n\";\n" % v for v in vals) + "}\n")
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    out = tmp_path / "out.txt"
    subprocess.run([str(exe), str(out)], check=True)
    assert out.read_text() == "".join("%.8e\n" % v for v in vals)


@pytest.mark.gpu
def test_volume_output_of_a_context(tmp_path):
    """Written from a live context: the p and gradient blocks parse back to the
    state and the gradients of the last residual."""
    from freecfd_amd.ns import NavierStokes
    cs = cases.cylinder_case(32, 16, noise=0.01)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    ns.assemble_residual()
    path = O.write_tecplot_volume(ns, cs.raw, 10, variables=("p", "gradp"), root=str(tmp_path))
    N, F = ns.info.n_cells, ns.info.n_faces
    G, _ = ns.gradients_limiters_host()
    _, _, vals, *_ = _parse(path, len(O.local_nodes(cs.raw)), F)
    got = vals.reshape(4, N)
    assert np.allclose(got[0], cs.W0[:, 0], rtol=1e-8, atol=0)
    for i in range(3):
        assert np.allclose(got[1 + i], G[:N, i], rtol=1e-8, atol=1e-300)
    assert (tmp_path / "volume_latest_1.dat").is_symlink()
    ns.close()


def _vtk_arrays(text):
    """The DataArray bodies of a .vtu file in order: (opening tag, lines)."""
    out, cur, tag = [], None, None
    for ln in text.split("\n"):
        if ln.startswith("<DataArray"):
            tag, cur = ln, []
        elif ln.startswith("</DataArray>"):
            out.append((tag, cur))
            cur = None
        elif cur is not None:
            cur.append(ln)
    return out


@pytest.mark.parametrize("name", ["plate", "tetbox"])
def test_vtk_file_against_the_mesh(tmp_path, name):
    from freecfd_amd.ns import Grid
    cs = cases.plate_case(6, 4, 2) if name == "plate" else cases.tetbox_case(2)
    g = Grid(cs.raw, cs.opt)
    m = g.metrics()
    N = g.info.n_cells
    f = O.cell_fields(cs.W0, gamma=cs.opt.gamma, Mw=cs.opt.Mw, Pref=cs.opt.Pref, Tref=cs.opt.Tref)
    variables = ["p", "V", "dt", "null", "rank"]
    path = O.write_vtk_volume_arrays(str(tmp_path), 4, cs.raw, f, variables)
    text = open(path).read()
    nodes = O.local_nodes(cs.raw)
    assert text.startswith('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid">\n<UnstructuredGrid>\n'
                           '<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (len(nodes), N))
    arrs = _vtk_arrays(text)
    pts = np.array([float(x) for x in arrs[0][1]]).reshape(-1, 3)
    assert np.allclose(pts, np.asarray(cs.raw["xyz"]).reshape(-1, 3)[nodes], rtol=1e-8, atol=1e-300)
    conn = [[int(t) for t in ln.split()] for ln in arrs[1][1]]
    cp, cn = np.asarray(cs.raw["cell_ptr"]), np.asarray(cs.raw["cell_nodes"])
    assert [nodes[c].tolist() for c in conn] == [cn[cp[c]:cp[c + 1]].tolist() for c in range(N)]
    assert [int(x) for x in arrs[2][1]] == cp[1:].tolist()
    assert [int(x) for x in arrs[3][1]] == [{4: 10, 8: 12, 6: 13, 5: 14}[k] for k in np.diff(cp)]
    tags = [t for t, _ in arrs[4:]]
    assert tags == ['<DataArray Name="p" type="Float32" format="ascii" >',
                    '<DataArray Name="V" Number of Components="3" type="Float32" format="ascii" >',
                    '<DataArray Name="dt" type="Float32" format="ascii" >',
                    '<DataArray Name="null" type="Float32" format="ascii" >',
                    '<DataArray Name="rank" type="Float32" format="ascii" >']
    p = np.array([float(x) for x in arrs[4][1]])
    V = np.array([[float(x) for x in ln.split()] for ln in arrs[5][1]])
    dt = np.array([float(x) for x in arrs[6][1]])
    assert np.allclose(p, cs.W0[:, 0], rtol=1e-8) and np.allclose(V, cs.W0[:, 1:4], rtol=1e-8, atol=1e-300)
    assert np.allclose(dt, cs.W0[:, 4], rtol=1e-8)          # the reference's dt array holds T
    assert arrs[7][1] == [] and arrs[8][1] == ["0"] * N
    pv = open(str(tmp_path / "volume_output" / "4" / "grid_1_volume_4.pvtu")).read()
    assert '<Piece Source="grid_1_proc_0.vtu" />' in pv and pv.count("<DataArray") == 6


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_vtk_numbers_match_cpp_setw16(tmp_path):
    """Points as `setw(16) << setprecision(8) << scientific << x`; later
    values keep the stream's scientific precision 8."""
    pts, vals = [0.5, -12.25, 1e-9], [101325.0, 0.1]
    src = tmp_path / "fmt.cc"
    src.write_text("#include <fstream>\n#include <iomanip>\nint main(int, char** argv) {\n  std::ofstream f(argv[1]);\n"
                   + "".join("  f << std::setw(16) << std::setprecision(8) << std::scientific << %r << std::endl;\n" % v for v in pts)
                   + "".join("  f << %r << std::endl;\n" % v for v in vals) + "}\n")
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    out = tmp_path / "out.txt"
    subprocess.run([str(exe), str(out)], check=True)
    assert out.read_text() == "".join("%16.8e\n" % v for v in pts) + "".join("%.8e\n" % v for v in vals)


def _legacy_blocks(lines, start):
    """SCALARS blocks from `start`: [(name, values as strings)]."""
    out, i = [], start
    while i < len(lines):
        assert lines[i].startswith("SCALARS ") and lines[i].endswith(" double") and lines[i + 1] == "LOOKUP_TABLE default"
        j = i + 2
        while j < len(lines) and not lines[j].startswith("SCALARS "):
            j += 1
        out.append((lines[i].split()[1], lines[i + 2:j]))
        i = j
    return out


@pytest.mark.parametrize("name", ["plate", "tetbox"])
def test_vtk_legacy_file_against_the_mesh(tmp_path, name):
    """Format vtklegacy (write_volume_output.cc:701-809) parsed back: points
    in local order, cells in local node ids with their counts, cell types,
    the SCALARS blocks (vectors by component, dt holding T, null and volume
    with no values), and the .visit index with its latest link."""
    cs = cases.plate_case(6, 4, 2, sa=True) if name == "plate" else cases.tetbox_case(2)
    N = len(cs.raw["cell_ptr"]) - 1
    f = O.cell_fields(cs.W0, gamma=cs.opt.gamma, Mw=cs.opt.Mw, Pref=cs.opt.Pref, Tref=cs.opt.Tref)
    variables = ["p", "V", "dt", "null", "Mach", "volume", "rank"] + (["nut"] if cs.W0.shape[1] == 6 else [])
    path = O.write_vtk_legacy_arrays(str(tmp_path), 8, cs.raw, f, variables)
    assert path == str(tmp_path / "volume_output" / "8" / "grid_1_proc_0.vtk")
    lines = open(path).read().split("\n")
    assert lines[-1] == ""
    lines = lines[:-1]
    nodes = O.local_nodes(cs.raw)
    P = len(nodes)
    assert lines[:5] == ["# vtk DataFile Version 2.0", "Grid_1", "ASCII", "DATASET UNSTRUCTURED_GRID", "POINTS %d double" % P]
    assert all(len(x) == 16 for x in lines[5:5 + 3 * P])
    pts = np.array([float(x) for x in lines[5:5 + 3 * P]]).reshape(-1, 3)
    assert np.allclose(pts, np.asarray(cs.raw["xyz"]).reshape(-1, 3)[nodes], rtol=1e-8, atol=1e-300)
    cp, cn = np.asarray(cs.raw["cell_ptr"]), np.asarray(cs.raw["cell_nodes"])
    k = 5 + 3 * P
    assert lines[k] == "CELLS %d %d" % (N, cp[N] + N)
    cells = [[int(t) for t in ln.split(" ")] for ln in lines[k + 1:k + 1 + N]]
    assert [c[0] for c in cells] == np.diff(cp).tolist()
    assert [nodes[c[1:]].tolist() for c in cells] == [cn[cp[c]:cp[c + 1]].tolist() for c in range(N)]
    k += 1 + N
    assert lines[k] == "CELL_TYPES %d" % N
    assert [int(x) for x in lines[k + 1:k + 1 + N]] == [{4: 10, 8: 12, 6: 13, 5: 14}[n] for n in np.diff(cp)]
    k += 1 + N
    assert lines[k] == "CELL_DATA %d" % N
    blocks = _legacy_blocks(lines, k + 1)
    names = ["p", "V_x", "V_y", "V_z", "dt", "null", "Mach", "volume", "rank"] + (["nut"] if cs.W0.shape[1] == 6 else [])
    assert [b[0] for b in blocks] == names
    vals = dict(blocks)
    want = {"p": cs.W0[:, 0], "V_x": cs.W0[:, 1], "V_y": cs.W0[:, 2], "V_z": cs.W0[:, 3], "dt": cs.W0[:, 4],
            "Mach": f["Mach"]}
    if "nut" in vals:
        want["nut"] = cs.W0[:, 5]
    for key, w in want.items():
        assert all(x == "%.8e" % y for x, y in zip(vals[key], w)) and len(vals[key]) == N, key
    assert vals["null"] == [] and vals["volume"] == [] and vals["rank"] == ["0"] * N
    vo = tmp_path / "volume_output"
    assert (vo / "volume_8_1.visit").read_text() == "!NBLOCKS 1\n./8/grid_1_proc_0.vtk\n"
    assert os.readlink(str(vo / "volume_latest_1.visit")) == "./volume_8_1.visit"
    with pytest.raises(ValueError):
        O.write_vtk_legacy_arrays(str(tmp_path), 9, cs.raw, f, ["mu_t"])


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_loads_lines_match_a_cpp_ostream(tmp_path):
    """write_loads appends `step \\t time \\t Fx .. Mz` per included BC in
    `scientific << setprecision(8)`: the same statements compiled and run."""
    L = np.array([[1.5, -2.25e-3, 0.0, 7.0, 1e-12, -3e5], [0.1, 0.2, 0.3, 0.4, 0.5, 0.6], [9.0, 8.0, 7.0, 6.0, 5.0, 4.0]])
    rows = [(12, 0.125, 0), (12, 0.125, 2), (13, 2.5, 0)]
    src = tmp_path / "fmt.cc"
    body = ""
    for step, time, b in rows:
        body += ('  { std::ofstream f(std::string(argv[1]) + "/loads_grid_1_BC_%d.dat", std::ios::app);\n'
                 '    f << std::scientific << std::setprecision(8);\n    f << %d << "\\t" << %r << "\\t";\n'
                 '    f << %s << std::endl; }\n') % (b + 1, step, time, ' << "\\t" << '.join(repr(float(x)) for x in L[b]))
    src.write_text("#include <fstream>\n#include <iomanip>\n#include <string>\nint main(int, char** argv) {\n" + body + "}\n")
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    cdir = tmp_path / "cpp"
    cdir.mkdir()
    subprocess.run([str(exe), str(cdir)], check=True)
    O.write_loads_arrays(str(tmp_path), 12, 0.125, L, [0, 2])
    O.write_loads_arrays(str(tmp_path), 13, 2.5, L, [0])
    for b in (1, 3):
        name = "loads_grid_1_BC_%d.dat" % b
        assert (tmp_path / "volume_output" / name).read_text() == (cdir / name).read_text()
    assert not (tmp_path / "volume_output" / "loads_grid_1_BC_2.dat").exists()


@pytest.mark.gpu
def test_loads_file_of_a_context(tmp_path):
    """The loads file written from a live context holds fcfd_loads of the
    wall (the only BC with include_in_loads), one line per call."""
    from freecfd_amd.ns import NavierStokes
    cs = cases.cylinder_case(32, 16, noise=0.01)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    F = ns.loads()
    paths = O.write_loads(ns, 4, 0.5, root=str(tmp_path))
    O.write_loads(ns, 5, 0.75, root=str(tmp_path))
    assert paths == [str(tmp_path / "volume_output" / "loads_grid_1_BC_1.dat")]
    rows = [ln.split("\t") for ln in open(paths[0]).read().rstrip("\n").split("\n")]
    assert [r[:2] for r in rows] == [["4", "%.8e" % 0.5], ["5", "%.8e" % 0.75]]
    for r in rows:
        assert np.allclose([float(x) for x in r[2:]], F[0], rtol=1e-8, atol=1e-300)
    assert abs(F[0, 0]) > 0
    ns.close()


def _parse_surface(text, n_bc, nvals):
    """Zones of a surface file: [(header lines, coordinates or None, values
    (nvals x faces), connectivity rows)]."""
    chunks = text.split("ZONE, ")
    assert chunks[0].startswith('VARIABLES = "x", "y", "z" ')
    zones = []
    for b, ch in enumerate(chunks[1:]):
        lines = ch.split("\n")
        nh = 3 if b == 0 else 4
        head = lines[:nh]
        toks = "\n".join(lines[nh:]).split()
        nf = int(head[1].split("ELEMENTS=")[1])
        k, xyz = 0, None
        if b == 0:
            nn = int(head[1].split("NODES=")[1].split(",")[0])
            xyz = np.array(toks[:3 * nn], dtype=float).reshape(3, nn).T
            k = 3 * nn
        vals = toks[k:k + nvals * nf]
        conn = np.array(toks[k + nvals * nf:], dtype=int).reshape(nf, 4) if nf else np.zeros((0, 4), int)
        zones.append((head, xyz, vals, conn))
    assert len(zones) == n_bc
    return zones


def test_surface_file_against_the_mesh(tmp_path):
    """Tecplot surface output (write_surface_output.cc:57-433) parsed back:
    zone headers, the boundary nodes (sorted local ids; every face centroid is
    the average of its distinct nodes), per-zone face counts in reference
    face order, the values of each block, and a triangle's repeated node."""
    from freecfd_amd.ns import Grid
    cs = cases.cylinder_case(16, 8)
    g = Grid(cs.raw, cs.opt)
    m = g.metrics()
    N, B = g.info.n_cells, g.info.n_ghosts
    rng = np.random.default_rng(4)
    Wg = np.concatenate([cs.W0, np.tile(cs.W0[0], (B, 1))], axis=0) * (1 + 0.05 * rng.random((N + B, 5)))
    bcs = [cs.opt.bc[b] for b in range(cs.opt.n_bc)]
    fv = O.face_values(m, Wg, bcs, gamma=cs.opt.gamma, Mw=cs.opt.Mw)
    variables = ["p", "V", "null", "T", "Mach", "rank"]
    path = O.write_tecplot_surface_arrays(str(tmp_path), 6, cs.raw, m, fv, variables)
    assert os.readlink(str(tmp_path / "surface_latest_1.dat")) == "./surface_output/surface_6_1.dat"
    text = open(path).read()
    assert text.split("\n")[0] == 'VARIABLES = "x", "y", "z" ,"p" ,"V_x" ,"V_y" ,"V_z" ,"T" ,"Mach" ,"rank" '
    zones = _parse_surface(text, cs.opt.n_bc, 7)
    fb, nptr, fnodes = m["face_bc"], m["face_node_ptr"], m["face_nodes"]
    bnodes = np.unique(np.concatenate([fnodes[nptr[f]:nptr[f + 1]] for f in np.nonzero(fb >= 0)[0]]))
    xyz_local = np.asarray(cs.raw["xyz"]).reshape(-1, 3)[O.local_nodes(cs.raw)]
    head0, xyz, _, _ = zones[0]
    assert head0 == ['T="BC_1", ZONETYPE=FEQUADRILATERAL, DATAPACKING=BLOCK',
                     "NODES=%d, ELEMENTS=%d" % (len(bnodes), (fb == 0).sum()), "VARLOCATION=([4-10]=CELLCENTERED)"]
    assert np.allclose(xyz, xyz_local[bnodes], rtol=1e-8, atol=1e-300)
    fc = m["face_centroid"].reshape(-1, 3)
    for b, (head, _, vals, conn) in enumerate(zones):
        faces = np.nonzero(fb == b)[0]
        if b:
            assert head == ['T="BC_%d", ZONETYPE=FEQUADRILATERAL, DATAPACKING=BLOCK' % (b + 1),
                            "ELEMENTS=%d" % len(faces), "VARLOCATION=([4-10]=CELLCENTERED)", "VARSHARELIST=([1-3]=1)"]
        want = [fv["p"][faces], fv["V"][faces, 0], fv["V"][faces, 1], fv["V"][faces, 2], fv["T"][faces], fv["Mach"][faces]]
        nf = len(faces)
        for j, w in enumerate(want):
            assert vals[j * nf:(j + 1) * nf] == ["%.8e" % x for x in w]
        assert vals[6 * nf:] == ["0"] * nf
        for f, row in zip(faces, conn):
            k = nptr[f + 1] - nptr[f]
            if k == 3:
                assert row[3] == row[2]
            cen = xyz[row[:k] - 1].mean(axis=0)
            assert np.abs(cen - fc[f]).max() <= 1e-7 * np.abs(fc).max()      # 9 printed digits
    g.close()


def test_face_values_follow_fixed_bcs_and_stencils():
    """Variable::face on a uniform state is that state off the fixed BCs and
    the BC value on them; on any state a boundary face's unfixed value lies
    between its parent's and its ghost's (a convex stencil)."""
    from freecfd_amd.ns import Grid
    cs = cases.cylinder_case(16, 8)
    g = Grid(cs.raw, cs.opt)
    m = g.metrics()
    N, B = g.info.n_cells, g.info.n_ghosts
    bcs = [cs.opt.bc[b] for b in range(cs.opt.n_bc)]
    W = np.tile(np.array([9.0e4, 3.0, -2.0, 0.5, 280.0]), (N + B, 1))
    fv = O.face_values(m, W, bcs)
    fb = m["face_bc"]
    assert np.allclose(fv["p"][fb != 2], 9.0e4, rtol=1e-14) and (fv["p"][fb == 2] == 101325.0).all()
    assert np.allclose(fv["T"][fb != 1], 280.0, rtol=1e-14) and (fv["T"][fb == 1] == 300.0).all()
    assert (fv["V"][fb == 0] == 0.0).all() and (fv["V"][fb == 1] == np.array(list(bcs[1].V))).all()
    assert np.allclose(fv["V"][fb >= 3], [3.0, -2.0, 0.5], rtol=1e-14)
    # rho is not fixed on the inlet: its face value averages the cells rho, not rho(p_f, T_f)
    assert np.allclose(fv["rho"], 9.0e4 / (8314.47215 / 28.97 * 280.0), rtol=1e-14)
    rng = np.random.default_rng(2)
    W = W * (1 + 0.1 * rng.random(W.shape))
    fv = O.face_values(m, W, bcs)
    par = m["face_parent"]
    bf = np.nonzero(fb == 3)[0]
    ghost = N + np.array([m["avg_idx"][m["avg_ptr"][f]:m["avg_ptr"][f + 1]].max() - N for f in bf])
    lo, hi = np.minimum(W[par[bf], 0], W[ghost, 0]), np.maximum(W[par[bf], 0], W[ghost, 0])
    assert ((fv["p"][bf] >= lo * (1 - 1e-15)) & (fv["p"][bf] <= hi * (1 + 1e-15))).all()
    g.close()


@pytest.mark.gpu
def test_surface_output_of_a_context(tmp_path):
    """Written from a live context after a residual: the fixed inlet T and
    outlet p, and the no-slip wall's zero velocity, in their zones."""
    from freecfd_amd.ns import NavierStokes
    cs = cases.cylinder_case(32, 16, noise=0.01)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    ns.assemble_residual()
    path = O.write_tecplot_surface(ns, cs.raw, 3, variables=("p", "V", "T"), root=str(tmp_path))
    zones = _parse_surface(open(path).read(), cs.opt.n_bc, 5)
    m = ns.metrics()
    fb = m["face_bc"]
    for b, (_, _, vals, _) in enumerate(zones):
        nf = int((fb == b).sum())
        v = np.array(vals, dtype=float).reshape(5, nf)
        if b == 0:
            assert (v[1:4] == 0).all()
        if b == 1:
            assert (v[4] == 300.0).all()
        if b == 2:
            assert (v[0] == 101325.0).all()
        assert (v[0] > 0).all() and (v[4] > 0).all()
    ns.close()
