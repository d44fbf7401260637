"""Loader for the in-tree libfcfd.so (fails loudly when it is missing)."""
import ctypes as C
import os

from . import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfcfd.so")
_LIB = None

_vp = C.c_void_p
_dpp = C.POINTER(C.POINTER(C.c_double))


def lib():
    """Return the loaded library; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libfcfd.so is not built (run `python -m freecfd_amd.build`): " + LIB_PATH)
    L = C.CDLL(LIB_PATH)
    sig = {
        "fcfd_create": (C.c_int, [C.POINTER(A.Mesh), C.POINTER(A.Options), C.POINTER(_vp)]),
        "fcfd_create_host": (C.c_int, [C.POINTER(A.Mesh), C.POINTER(A.Options), C.POINTER(_vp)]),
        "fcfd_destroy": (None, [_vp]),
        "fcfd_last_error": (C.c_char_p, [_vp]),
        "fcfd_get_info": (C.c_int, [_vp, C.POINTER(A.Info)]),
        "fcfd_metrics": (C.c_int, [_vp, C.POINTER(A.MetricsOut)]),
        "fcfd_state_to_device": (C.c_int, [_vp, _vp, _vp]),
        "fcfd_cells_to_host": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp]),
        "fcfd_residual": (C.c_int, [_vp, _vp, _vp, _vp]),
        "fcfd_internal_buffers": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
        "fcfd_gradients_limiters": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "fcfd_status": (C.c_int, [_vp, C.c_int]),
        "fcfd_profile": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
        "fcfd_fd_jacobian": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "fcfd_jacobian_buffers": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
        "fcfd_jacobian_to_host": (C.c_int, [_vp, _vp, _vp, _vp]),
        "fcfd_jvp": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
        "fcfd_newton_step": (C.c_int, [_vp, _vp, C.POINTER(A.NewtonOpts), C.POINTER(A.NewtonStats), _vp]),
        "fcfd_time_step": (C.c_int, [_vp, _vp, C.c_int, C.c_double, _vp, C.POINTER(C.c_double), _vp]),
        "fcfd_time_step_noalloc": (C.c_int, [_vp, _vp, C.c_int, C.c_double, _vp, _vp, _vp]),
        "fcfd_vjp": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "fcfd_gradients_to_host": (C.c_int, [_vp, _vp, _vp]),
        "fcfd_convective_flux": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "fcfd_loads": (C.c_int, [_vp, _vp, _vp, _vp]),
        "fcfd_adjoint_step": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(A.NewtonOpts), C.POINTER(A.NewtonStats), _vp]),
        "fcfd_set_bc": (C.c_int, [_vp, C.c_int, C.POINTER(A.BC)]),
        "fcfd_surface": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "fcfd_pseudo_step": (C.c_int, [_vp, _vp, C.POINTER(A.NewtonOpts), C.POINTER(A.PseudoOpts), C.POINTER(A.NewtonStats), _vp]),
        "fcfd_bc_tangent": (C.c_int, [_vp, _vp, C.c_int, C.POINTER(A.BC), _vp, _vp, _vp]),
        "fcfd_vjp_params": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "fcfd_rans_init": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.POINTER(A.RansOpts), _vp]),
        "fcfd_rans_set": (C.c_int, [_vp, _vp, _vp, _vp]),
        "fcfd_rans_get": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "fcfd_rans_terms": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "fcfd_rans_step": (C.c_int, [_vp, _vp, _vp, C.POINTER(A.RansOpts), C.POINTER(A.RansStats), _vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def exported_symbols():
    """Names declared in include/fcfd_gpu.h (extern "C" entry points)."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "fcfd_gpu.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(fcfd_\w+)\s*\(", text, re.M)))
