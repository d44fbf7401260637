"""GPU surface quantities of the face loop (mdot.face, qdot.bc, tau.bc,
ns_assemble_linear_system.cc:122-127) against the CPU restatement, and the
surface writer's mdot/qdot/tau blocks (write_surface_output.cc:266-290, 390)
from a live context."""
import numpy as np
import pytest

from freecfd_amd import cases
from freecfd_amd import output as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,make", [
    ("cylinder_viscous", lambda: cases.cylinder_case(32, 16, noise=0.01)),
    ("plate_sa_prisms", lambda: cases.plate_case(12, 8, 3, first=1e-3, noise=0.01, sa=True)),
    ("channel_fixed_q", lambda: cases.bc_variant_case("vel_rho__out_rho__wall_fixed_q")),
    ("channel_fixed_t", lambda: cases.bc_variant_case("stagnation__out_t_no_reverse__wall_fixed_t")),
])
def test_surface_quantities_match_oracle(name, make):
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    cs = make()
    print("case", name, flush=True)
    o = Oracle(cs.raw, cs.opt)
    o.residual(cs.W0)
    mo, qo, to = o.surface()
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    mg, qg, tg = ns.surface()
    ns.status()
    for label, a, b in (("mdot", mg, mo), ("qdot", qg, qo), ("tau", tg, to)):
        err = np.abs(a - b).max() / (np.abs(b).max() + 1e-300)
        print("  %s: max |gpu - oracle| / max|oracle| = %.2e" % (label, err), flush=True)
        assert err <= 1e-12
    ns.close()


def test_surface_writer_mdot_qdot_tau(tmp_path):
    from freecfd_amd.ns import NavierStokes
    from test_output import _parse_surface
    cs = cases.bc_variant_case("vel_rho__out_rho__wall_fixed_q")
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    ns.assemble_residual()
    path = O.write_tecplot_surface(ns, cs.raw, 2, variables=("p", "mdot", "qdot", "tau"), root=str(tmp_path))
    text = open(path).read()
    assert text.split("\n")[0] == 'VARIABLES = "x", "y", "z" ,"p" ,"mdot" ,"qdot" ,"tau_x" ,"tau_y" ,"tau_z" '
    zones = _parse_surface(text, cs.opt.n_bc, 6)
    m = ns.metrics()
    mdot, qdot, tau = ns.surface()
    fb = m["face_bc"]
    for b, (_, _, vals, _) in enumerate(zones):
        faces = np.nonzero(fb == b)[0]
        nf = len(faces)
        v = np.array(vals, dtype=float).reshape(6, nf)
        assert np.allclose(v[1], mdot[faces], rtol=1e-8, atol=1e-300)
        assert np.allclose(v[2], qdot[faces], rtol=1e-8, atol=1e-300)
        assert np.allclose(v[3:6], tau[faces].T, rtol=1e-8, atol=1e-300)
        if b == 2:   # the FIXED_Q wall
            assert (v[2] == 5.0e3).all()
    ns.close()
