"""Host set-up of the product (hash/bitmask connectivity) against the CPU
restatement (the reference's linear searches): connectivity, face order, BC
tags, ghosts, neighbour lists, stencil maps bit-exact; geometry and weights
bitwise equal as well (same formulas, no FMA contraction)."""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases


def _check(cs):
    from freecfd_amd.ns import Grid
    from oracle.oracle import Oracle
    mo = Oracle(cs.raw, cs.opt).metrics()
    mg = Grid(cs.raw, cs.opt).metrics()
    for k, a in mo.items():
        if k == "cell_perm":
            continue
        b = mg[k]
        assert a.shape == b.shape, k
        assert np.array_equal(a, b), "%s differs (%d entries)" % (k, int((a != b).sum()))


@pytest.mark.parametrize("name,make", [
    ("wedge", lambda: cases.wedge_case(24, 12)),
    ("wedge_full", lambda: cases.wedge_case(96, 48)),
    ("cylinder", lambda: cases.cylinder_case(32, 16)),
    ("tetbox", lambda: cases.tetbox_case(5)),
    ("bump", lambda: cases.bump_case(16, 8, 4)),
    ("plate_prisms", lambda: cases.plate_case(16, 10, 4)),
])
def test_setup_bit_exact(name, make, oracle_lib):
    _check(make())


def test_setup_idw_and_simple(oracle_lib):
    for method in (A.INTERP_IDW, A.INTERP_SIMPLE):
        cs = cases.bump_case(6, 4, 3)
        cs.opt.interp_method = method
        _check(cs)


def test_setup_lsqr_hex_and_curvilinear_tets(oracle_lib):
    cs = cases.bump_case(6, 4, 3)
    cs.opt.grad_hex = A.GRAD_LSQR
    _check(cs)
    cs = cases.tetbox_case(3)
    cs.opt.grad_other = A.GRAD_CURVILINEAR
    _check(cs)
