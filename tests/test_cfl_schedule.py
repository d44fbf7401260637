"""The CFL schedule file of update_time_step (time_step.cc:63-77, 160-175),
host-side: the CFL at time step k is the value of the last (step, CFL) pair
with step < k (pairs read in file order up to the first at or beyond k),
and the ramp stops once a pair applies.  Checked on the schedule the
reference ships with its hypersonic cylinder example (cfl.sch), restated
here as data, and on a ramp that hands over to the schedule."""
import os

from freecfd_amd.ns import read_cfl_schedule, scheduled_cfl

SHIPPED = [(1, 1.0), (10, 50.0), (20, 200.0), (30, 300.0), (40, 400.0), (50, 500.0), (60, 600.0), (70, 700.0),
           (80, 800.0), (90, 900.0), (100, 1000.0), (200, 5000.0)]


def test_read(tmp_path):
    p = tmp_path / "cfl.sch"
    p.write_text("".join("%d\t%g\n" % x for x in SHIPPED))
    assert read_cfl_schedule(str(p)) == SHIPPED
    ref = "/root/reference/examples/hypersonic_cylinder/cfl.sch"
    if os.path.exists(ref):
        assert read_cfl_schedule(ref) == SHIPPED


def test_schedule_values():
    # step 1: no pair below 1, the ramp's value stays
    assert scheduled_cfl(SHIPPED, 1, 7.0) == (7.0, False)
    # steps 2..10: 1 < k  -> 1.0 ; k = 11..20 -> 50
    assert scheduled_cfl(SHIPPED, 2, 7.0) == (1.0, True)
    assert scheduled_cfl(SHIPPED, 10, 7.0) == (1.0, True)
    assert scheduled_cfl(SHIPPED, 11, 7.0) == (50.0, True)
    assert scheduled_cfl(SHIPPED, 101, 7.0) == (1000.0, True)
    assert scheduled_cfl(SHIPPED, 200, 7.0) == (1000.0, True)
    assert scheduled_cfl(SHIPPED, 201, 7.0) == (5000.0, True)
    assert scheduled_cfl(SHIPPED, 10 ** 6, 7.0) == (5000.0, True)
