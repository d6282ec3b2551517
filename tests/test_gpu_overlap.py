"""GPU: the step with the collide in z-slabs overlapped with the IBM pass (d3q19_options.overlap_slabs)
gives bitwise the results of the step with one collide launch and one IBM launch.

Each IBM launch of the overlapped step runs the spheres whose stencils end in one slab, on a second
stream, after that slab has collided; the spreading is integer and every sphere's arithmetic is the
same, so populations, velocities and sphere states must be bitwise equal after graph-replayed steps
and after directly launched (event-timed) steps.  A wall-bounded box (stencils clipped at the walls)
and a fully periodic one (stencils that wrap in z go to the last slab) are both run.
"""
import numpy as np
import pytest

import d3q19
from workloads import deck

pytestmark = pytest.mark.gpu


def casson_box(walls):
    eta = (0.7 - 0.5) / 3.0
    return deck(size=(96, 80, 64), walls=walls, tau=0.7, uw=0.02, np_=40, diam=10.0, markers=128, dist=4, seed=31, gap=1.0,
                coll=1, ibm=1, visc=4, vparams=(1e-5, eta, 1e4, 0), local=1, clamp=(0.505, 5.0), zbuf=3.0)


def run(walls, slabs):
    c = d3q19.Case(lines=casson_box(walls))
    c.opt.overlap_slabs = slabs
    s = d3q19.Sim(c)
    s.write_particles(c.parKin)
    s.run(1, 40, d3q19.STEP_WRITE_U)       # graph replay
    s.time_steps(41, 6)                     # direct launches with per-kernel events
    s.run(47, 3, d3q19.STEP_WRITE_U)
    s.sync()
    out = (s.read_f(), s.read_u(), s.read_tau(), s.read_particles_f64())
    s.close()
    return out


@pytest.mark.parametrize("walls", [(0, 0, 1), (0, 0, 0)], ids=["z-walls", "periodic"])
def test_overlapped_step_is_bitwise_the_plain_step(walls):
    print("overlap: walls %s, 1 slab vs 4 and 7 slabs" % (walls,), flush=True)
    f1, u1, t1, p1 = run(walls, 1)
    for k in (4, 7):
        fk, uk, tk, pk = run(walls, k)
        print("  %d slabs: max|df| %.3e max|du| %.3e" % (k, np.abs(fk - f1).max(), np.abs(uk - u1).max()), flush=True)
        assert np.array_equal(fk, f1) and np.array_equal(uk, u1) and np.array_equal(tk, t1)
        for a, b in zip(pk, p1):
            assert np.array_equal(a, b)
