"""Restart files in the reference's layout, one rank.

Writer: write_restart.cc:44-99 (partition map, time file, dt),
ns_write_restart.cc:25-35 (p, V, T), Variable::dump_cell_data
(variable/variable.h:229-245: the cell values as raw binary, cell by cell).
Reader: read_restart.cc:44-103, ns_read_restart.cc:25-43,
Variable::read_cell_data (variable/variable.h:248-271).  With one rank the
global cell ids are the cell indices and the files hold the cells in the
reference's numbering.

    restart/partitionMap_<g>.dat  text: the rank count; per rank
                                  "nodeCount<TAB>cellCount", then one global
                                  cell id per line
    restart/<step>/time.dat       text: the time; the three NS first residuals
                                  (TAB-separated); the two RANS ones, or
                                  "-1<TAB>-1"; the heat one, or "-1"
    restart/<step>/dt.<g>         one double per cell
    restart/<step>/p.<g>, T.<g>   one double per cell
    restart/<step>/V.<g>          three doubles (x, y, z: Vec3D) per cell
    restart/<step>/nut.<g>        one double per cell: nu~ of the build-only
                                  SA model (the reference reads no such file)

<g> is the grid index + 1.  Numbers in the text files are written as a
default C++ ostream writes them (6 significant digits, printf's %g).  The
reader recomputes nothing: ghosts, gradients and limiters follow from the
state on the next residual (read_restart's update_boundaries,
calc_cell_grads and calc_limiter).
"""
import os

import numpy as np


def _g(x):
    return "%g" % x


def write_restart_arrays(root, step, W, n_nodes, time=0.0, first_residuals=(1.0, 1.0, 1.0), dt=None, gid=0):
    """Write the state W (n_cells x 5: p, u, v, w, T; or x 6 with nu~) in
    reference cell order.  dt (n_cells) is written as given, zeros when None
    (the reference's reader does not read it back)."""
    W = np.ascontiguousarray(W, dtype=np.float64)
    n = W.shape[0]
    gs = ".%d" % (gid + 1)
    rdir = os.path.join(root, "restart")
    sdir = os.path.join(rdir, "%d" % step)
    os.makedirs(sdir, exist_ok=True)
    with open(os.path.join(rdir, "partitionMap_%d.dat" % (gid + 1)), "w") as f:
        f.write("1\n%d\t%d\n" % (n_nodes, n))
        f.write("".join("%d\n" % c for c in range(n)))
    r = first_residuals
    with open(os.path.join(sdir, "time.dat"), "w") as f:
        f.write("%s\n%s\t%s\t%s\n-1\t-1\n-1\n" % (_g(time), _g(r[0]), _g(r[1]), _g(r[2])))
    d = np.zeros(n) if dt is None else np.ascontiguousarray(dt, dtype=np.float64).reshape(n)
    d.tofile(os.path.join(sdir, "dt" + gs))
    W[:, 0].copy().tofile(os.path.join(sdir, "p" + gs))
    W[:, 1:4].copy().tofile(os.path.join(sdir, "V" + gs))
    W[:, 4].copy().tofile(os.path.join(sdir, "T" + gs))
    if W.shape[1] == 6:
        W[:, 5].copy().tofile(os.path.join(sdir, "nut" + gs))


def read_restart_arrays(root, step, gid=0):
    """Read a restart written by the reference (one or more ranks) or by
    write_restart_arrays.  Returns (W, time, first_residuals) with W in the
    order of the global cell ids (n_cells x 5, or x 6 when a nut file is
    present)."""
    gs = ".%d" % (gid + 1)
    rdir = os.path.join(root, "restart")
    sdir = os.path.join(rdir, "%d" % step)
    with open(os.path.join(rdir, "partitionMap_%d.dat" % (gid + 1))) as f:
        tok = f.read().split()
    nprocs, k, ids = int(tok[0]), 1, []
    for _ in range(nprocs):
        ncells = int(tok[k + 1])
        ids.extend(int(t) for t in tok[k + 2:k + 2 + ncells])
        k += 2 + ncells
    ids = np.asarray(ids)
    n = len(ids)
    with open(os.path.join(sdir, "time.dat")) as f:
        t = f.read().split()
    time, first = float(t[0]), (float(t[1]), float(t[2]), float(t[3]))
    cols = {"p": np.fromfile(os.path.join(sdir, "p" + gs), dtype=np.float64, count=n),
            "V": np.fromfile(os.path.join(sdir, "V" + gs), dtype=np.float64, count=3 * n).reshape(n, 3),
            "T": np.fromfile(os.path.join(sdir, "T" + gs), dtype=np.float64, count=n)}
    nut_path = os.path.join(sdir, "nut" + gs)
    nv = 6 if os.path.exists(nut_path) else 5
    W = np.empty((n, nv))
    # the file position of a cell is its rank in the partition map; its row
    # is its global id (read_cell_data)
    W[ids, 0] = cols["p"]
    W[ids, 1:4] = cols["V"]
    W[ids, 4] = cols["T"]
    if nv == 6:
        W[ids, 5] = np.fromfile(nut_path, dtype=np.float64, count=n)
    return W, time, first


def write_restart(ns, step, time=0.0, first_residuals=(1.0, 1.0, 1.0), dt=None, root=".", gid=0):
    """write_restart(gid, step, time) of a NavierStokes context's current
    state (dt: host array or device tensor of the cells' time steps)."""
    if dt is not None and hasattr(dt, "cpu"):
        dt = ns.to_host(dt.reshape(1, -1), 1)[:, 0]
    write_restart_arrays(root, step, ns.state_host(include_ghosts=False), ns.info.n_nodes, time, first_residuals, dt, gid)


def read_restart(ns, step, root=".", gid=0):
    """read_restart(gid, step, time) into a NavierStokes context; returns
    (time, first_residuals)."""
    W, time, first = read_restart_arrays(root, step, gid)
    if W.shape != (ns.info.n_cells, ns.nv):
        raise ValueError("restart holds %s values, the context %s" % (W.shape, (ns.info.n_cells, ns.nv)))
    ns.set_state(W)
    return time, first
