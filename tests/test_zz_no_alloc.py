"""fcfd_time_step_noalloc (the time step from an entry point that allocates
nothing, its scratch a static device array) against fcfd_time_step, bitwise,
for every time step type: two calls in a row on one GPU, the second on a
smaller mesh, each after the other entry point has left its own values."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _compare(ns):
    import torch
    from freecfd_amd import _abi as A
    from freecfd_amd._lib import lib
    L = lib()
    ld = ns.info.ld_cell
    N = ns.info.n_cells
    for dt_type, cfl in ((A.DT_CFL_LOCAL, 7.5), (A.DT_CFL_MAX, 3.0), (A.DT_FIXED, 1.25e-4)):
        dt_ref = torch.full((ld,), -1.0, dtype=torch.float64, device="cuda")
        dt_new = torch.full((ld,), -2.0, dtype=torch.float64, device="cuda")
        mc_new = torch.full((1,), -3.0, dtype=torch.float64, device="cuda")
        mc_ref = C.c_double(0.0)
        s = torch.cuda.current_stream().cuda_stream
        rc = L.fcfd_time_step(ns.h, ns.W.data_ptr(), dt_type, cfl, dt_ref.data_ptr(), C.byref(mc_ref), C.c_void_p(s))
        assert rc == 0, L.fcfd_last_error(ns.h)
        rc = L.fcfd_time_step_noalloc(ns.h, ns.W.data_ptr(), dt_type, cfl, dt_new.data_ptr(), mc_new.data_ptr(), C.c_void_p(s))
        assert rc == 0, L.fcfd_last_error(ns.h)
        torch.cuda.synchronize()
        a, b = dt_ref[:N].cpu().numpy(), dt_new[:N].cpu().numpy()
        assert np.array_equal(a, b), (dt_type, np.abs(a - b).max())
        assert mc_new.item() == mc_ref.value, (dt_type, mc_new.item(), mc_ref.value)
        # a NULL max_cfl_dev is accepted
        rc = L.fcfd_time_step_noalloc(ns.h, ns.W.data_ptr(), dt_type, cfl, dt_new.data_ptr(), None, C.c_void_p(s))
        assert rc == 0
        torch.cuda.synchronize()
        assert np.array_equal(a, dt_new[:N].cpu().numpy())
    # bad arguments are refused before any launch
    assert L.fcfd_time_step_noalloc(ns.h, ns.W.data_ptr(), 99, 1.0, dt_new.data_ptr(), None, None) == A.EINVAL
    assert L.fcfd_time_step_noalloc(ns.h, ns.W.data_ptr(), A.DT_CFL_MAX, -1.0, dt_new.data_ptr(), None, None) == A.EINVAL


def test_time_step_noalloc_matches_time_step_twice():
    from freecfd_amd import cases
    from freecfd_amd.ns import NavierStokes
    print("no-alloc time step: cylinder 64x32 then wedge 24x12", flush=True)
    big = cases.cylinder_case(64, 32, noise=0.01)
    small = cases.wedge_case(24, 12)
    for cs in (big, small):
        ns = NavierStokes(cs.raw, cs.opt)
        ns.set_state(cs.W0)
        ns.assemble_residual()   # ghosts filled as in a solve
        _compare(ns)
        ns.close()
