"""Restart files in the reference's layout (freecfd_amd/restart.py): the exact
bytes of each file as the reference's writer produces them
(write_restart.cc:44-99, ns_write_restart.cc:25-35, variable/variable.h:229-245),
the text formatting checked against a compiled C++ ostream, multi-rank
partition maps on the read side (read_restart.cc:48-68,
variable/variable.h:248-271), and round trips."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from freecfd_amd import restart as RS


def _state(n, nv, seed=0):
    rng = np.random.default_rng(seed)
    W = np.empty((n, nv))
    W[:, 0] = 1e5 * (1 + 0.1 * rng.random(n))
    W[:, 1:4] = 100 * rng.standard_normal((n, 3))
    W[:, 4] = 300 * (1 + 0.1 * rng.random(n))
    if nv == 6:
        W[:, 5] = 1e-5 * rng.random(n)
    return W


def test_files_have_the_reference_bytes(tmp_path):
    W = _state(7, 5)
    dt = np.arange(7) * 1e-3
    RS.write_restart_arrays(str(tmp_path), 40, W, 12, time=0.25, first_residuals=(1.5e-3, 2.0, 3e5), dt=dt)
    r = tmp_path / "restart"
    assert (r / "partitionMap_1.dat").read_text() == "1\n12\t7\n" + "".join("%d\n" % c for c in range(7))
    assert (r / "40" / "time.dat").read_text() == "0.25\n0.0015\t2\t300000\n-1\t-1\n-1\n"
    assert (r / "40" / "p.1").read_bytes() == W[:, 0].tobytes()
    assert (r / "40" / "V.1").read_bytes() == np.ascontiguousarray(W[:, 1:4]).tobytes()   # x, y, z per cell
    assert (r / "40" / "T.1").read_bytes() == W[:, 4].tobytes()
    assert (r / "40" / "dt.1").read_bytes() == dt.tobytes()
    assert not (r / "40" / "nut.1").exists()


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_text_numbers_match_a_cpp_ostream(tmp_path):
    """The text files' numbers are written the way `file << x` writes a
    double with a default-constructed ofstream: compile that and compare."""
    vals = [0.25, 1.5e-3, 2.0, 3e5, 1e-5, 123456789.0, -1.0, 0.1 + 0.2, 7e-300, 0.0]
    src = tmp_path / "fmt.cc"
    src.write_text("#include <fstream>\nint main(int, char** argv) {\n  std::ofstream f(argv[1]);\n"
                   + "".join("  f << %r << std::endl;\n" % v for v in vals) + "}\n")
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    out = tmp_path / "out.txt"
    subprocess.run([str(exe), str(out)], check=True)
    assert out.read_text() == "".join(RS._g(v) + "\n" for v in vals)


def test_reads_a_multi_rank_restart(tmp_path):
    """Two ranks with their own cell order: every value lands at its global
    id (read_cell_data reads the file in partition-map order)."""
    n = 9
    W = _state(n, 5, seed=3)
    part = [[4, 0, 7, 2], [8, 1, 6, 3, 5]]
    r = tmp_path / "restart"
    (r / "3").mkdir(parents=True)
    txt = "2\n" + "".join("10\t%d\n" % len(p) + "".join("%d\n" % c for c in p) for p in part)
    (r / "partitionMap_1.dat").write_text(txt)
    (r / "3" / "time.dat").write_text("1.5\n1\t1\t1\n-1\t-1\n-1\n")
    order = [c for p in part for c in p]
    W[order, 0].tofile(str(r / "3" / "p.1"))
    np.ascontiguousarray(W[order, 1:4]).tofile(str(r / "3" / "V.1"))
    W[order, 4].tofile(str(r / "3" / "T.1"))
    W2, time, first = RS.read_restart_arrays(str(tmp_path), 3)
    assert time == 1.5 and first == (1.0, 1.0, 1.0)
    assert np.array_equal(W2, W)


@pytest.mark.parametrize("nv", [5, 6])
def test_round_trip(tmp_path, nv):
    W = _state(11, nv, seed=nv)
    RS.write_restart_arrays(str(tmp_path), 7, W, 30, time=3.0, first_residuals=(0.5, 0.25, 0.125))
    W2, time, first = RS.read_restart_arrays(str(tmp_path), 7)
    assert np.array_equal(W2, W)
    assert time == 3.0 and first == (0.5, 0.25, 0.125)
    assert os.path.exists(str(tmp_path / "restart" / "7" / "nut.1")) == (nv == 6)


@pytest.mark.gpu
def test_restart_reproduces_the_residual(tmp_path):
    """A context's state written and read into a fresh context gives a
    bitwise-equal residual (SA plate with prisms: six variables)."""
    from freecfd_amd import cases
    from freecfd_amd.ns import NavierStokes
    cs = cases.smooth_state(cases.plate_case(16, 10, 4, sa=True))
    a = NavierStokes(cs.raw, cs.opt)
    a.set_state(cs.W0)
    Ra = a.assemble_residual().clone()
    RS.write_restart(a, 5, time=0.5, root=str(tmp_path))
    b = NavierStokes(cs.raw, cs.opt)
    time, _ = RS.read_restart(b, 5, root=str(tmp_path))
    Rb = b.assemble_residual().clone()
    b.status()
    assert time == 0.5
    assert bool((Ra == Rb).all())
    a.close()
    b.close()
