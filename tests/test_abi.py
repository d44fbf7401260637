"""The C-ABI library loads on the CPU and exports every entry point that
include/fcfd_gpu.h declares (no compute calls without a GPU)."""
import ctypes

from freecfd_amd import _lib


def test_library_exports_header_symbols():
    L = _lib.lib()
    names = _lib.exported_symbols()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_create_host_reports_errors_without_exit():
    from freecfd_amd import _abi as A, cases
    cs = cases.wedge_case(4, 2)
    cs.opt.bc[1].specified = A.SPEC_STATE      # outlet with a complete state: input error
    m = A.make_mesh(cs.raw)
    h = ctypes.c_void_p()
    rc = _lib.lib().fcfd_create_host(ctypes.byref(m), ctypes.byref(cs.opt), ctypes.byref(h))
    assert rc == A.EBC
    assert b"outlet" in _lib.lib().fcfd_last_error(None)


def test_set_bc_validates_on_a_host_context():
    """fcfd_set_bc replaces BC values in a live (host-only) context; it refuses
    a change of BC type and an out-of-range index."""
    from freecfd_amd import _abi as A, cases
    cs = cases.wedge_case(4, 2)
    m = A.make_mesh(cs.raw)
    h = ctypes.c_void_p()
    L = _lib.lib()
    assert L.fcfd_create_host(ctypes.byref(m), ctypes.byref(cs.opt), ctypes.byref(h)) == 0
    try:
        bc = A.BC.from_buffer_copy(cs.opt.bc[0])
        bc.V[0] *= 1.01
        assert L.fcfd_set_bc(h, 0, ctypes.byref(bc)) == 0
        bad = A.BC.from_buffer_copy(cs.opt.bc[0])
        bad.type = A.BC_WALL
        assert L.fcfd_set_bc(h, 0, ctypes.byref(bad)) == A.EBC
        assert b"type" in L.fcfd_last_error(h)
        assert L.fcfd_set_bc(h, 99, ctypes.byref(bc)) == A.EINVAL
        assert L.fcfd_set_bc(h, -1, ctypes.byref(bc)) == A.EINVAL
    finally:
        L.fcfd_destroy(h)


def test_set_bc_refuses_switching_the_diffusive_flux():
    """On an inviscid case a FIXED_Q wall with qdot = 0 leaves the diffusive
    flux off; a new qdot != 0 would switch it on, which changes the kernels
    and the adjoint workspace chosen at create: fcfd_set_bc refuses it and
    keeps the old values."""
    from freecfd_amd import _abi as A, cases
    cs = cases.wedge_case(4, 2)
    wall = [b for b in range(cs.opt.n_bc) if cs.opt.bc[b].type == A.BC_WALL][0]
    cs.opt.bc[wall].thermal = A.THERMAL_FIXED_Q
    cs.opt.bc[wall].qdot = 0.0
    m = A.make_mesh(cs.raw)
    h = ctypes.c_void_p()
    L = _lib.lib()
    assert L.fcfd_create_host(ctypes.byref(m), ctypes.byref(cs.opt), ctypes.byref(h)) == 0
    try:
        bc = A.BC.from_buffer_copy(cs.opt.bc[wall])
        bc.qdot = 250.0
        assert L.fcfd_set_bc(h, wall, ctypes.byref(bc)) == A.EBC
        assert b"diffusive" in L.fcfd_last_error(h)
        bc.qdot = 0.0     # the unchanged flag is accepted
        assert L.fcfd_set_bc(h, wall, ctypes.byref(bc)) == 0
    finally:
        L.fcfd_destroy(h)


def test_two_equation_rans_validation_on_a_host_context():
    """FCFD_TURB_KW contexts are created on the host; an unknown model or a
    zero molecular viscosity is refused at create, and the model's entry
    points refuse a context without device data or without the model."""
    from freecfd_amd import _abi as A, cases
    cs = cases.kw_case(model=A.RANS_SST)
    L = _lib.lib()
    m = A.make_mesh(cs.raw)
    h = ctypes.c_void_p()
    assert L.fcfd_create_host(ctypes.byref(m), ctypes.byref(cs.opt), ctypes.byref(h)) == 0
    try:
        assert L.fcfd_rans_init(h, None, 0.05, 10.0, None, None) == A.EINVAL
        assert L.fcfd_rans_get(h, None, None, None, None, None) == A.EINVAL
        assert b"device" in L.fcfd_last_error(h)
    finally:
        L.fcfd_destroy(h)
    bad = cases.kw_case(model=A.RANS_SST)
    bad.opt.rans_model = 7
    assert L.fcfd_create_host(ctypes.byref(A.make_mesh(bad.raw)), ctypes.byref(bad.opt), ctypes.byref(h)) == A.EINVAL
    assert b"RANS" in L.fcfd_last_error(None)
    inv = cases.kw_case(model=A.RANS_SST)
    inv.opt.visc_model, inv.opt.mu = A.VISC_CONSTANT, 0.0
    assert L.fcfd_create_host(ctypes.byref(A.make_mesh(inv.raw)), ctypes.byref(inv.opt), ctypes.byref(h)) == A.EINVAL
    lam = cases.wedge_case(4, 2)
    hl = ctypes.c_void_p()
    assert L.fcfd_create_host(ctypes.byref(A.make_mesh(lam.raw)), ctypes.byref(lam.opt), ctypes.byref(hl)) == 0
    try:
        assert L.fcfd_rans_get(hl, None, None, None, None, None) == A.EINVAL
    finally:
        L.fcfd_destroy(hl)


def test_convective_flux_validates_on_a_host_context():
    """fcfd_convective_flux needs device data: a host-only context refuses it
    with FCFD_EINVAL (and a NULL context too), before any launch."""
    from freecfd_amd import _abi as A, cases
    cs = cases.wedge_case(4, 2)
    m = A.make_mesh(cs.raw)
    h = ctypes.c_void_p()
    L = _lib.lib()
    assert L.fcfd_create_host(ctypes.byref(m), ctypes.byref(cs.opt), ctypes.byref(h)) == 0
    try:
        assert L.fcfd_convective_flux(h, A.FLUX_ROE, 0, None, None, None, None, None, None, None, None, None, None) == A.EINVAL
        assert b"device" in L.fcfd_last_error(h)
        assert L.fcfd_convective_flux(None, A.FLUX_ROE, 0, None, None, None, None, None, None, None, None, None, None) == A.EINVAL
    finally:
        L.fcfd_destroy(h)
