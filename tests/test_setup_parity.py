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
    ("pie_collapsed_hexahedra", lambda: cases.pie_case()),
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


@pytest.mark.parametrize("flux", [A.FLUX_ROE, A.FLUX_SW])
def test_oracle_bucketed_wall_distance_equals_the_scan(flux, oracle_lib):
    """The oracle's exact bucketed wall search gives the reference's
    brute-force distances bit for bit, and (Steger-Warming) the same first
    closest wall face, whose normal sets the face dissipation factor: the
    residual is bitwise equal."""
    from oracle import oracle as O
    cs = cases.plate_case(16, 10, 4, noise=0.01, first=1e-3, sa=True)
    cs.opt.convective_flux = flux
    try:
        O.set_wall_brute_force(True)
        a = O.Oracle(cs.raw, cs.opt)
        O.set_wall_brute_force(False)
        b = O.Oracle(cs.raw, cs.opt)
    finally:
        O.set_wall_brute_force(False)
    wa, wb = a.metrics()["wall_distance"], b.metrics()["wall_distance"]
    assert np.array_equal(wa, wb) and np.all(wa < 1e20)
    Ra, _ = a.residual(cs.W0)
    Rb, _ = b.residual(cs.W0)
    assert np.array_equal(Ra, Rb)
