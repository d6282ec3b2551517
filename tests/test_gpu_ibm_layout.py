"""GPU: the IBM pass's layouts of Peskin-4 markers give bitwise the same results.

d3q19_options.ibm_layout maps markers onto the hardware in one of three ways: one marker per
wavefront (a lane per stencil node) in blocks of 8-marker groups, four markers per wavefront (a
16-lane row each, a lane per stencil row of four nodes) in the same groups, or four per wavefront in
the largest groups (up to 64 markers) whose LDS boxes fit.  Both layouts sum a marker's 64 weighted velocities with the same association,
and the spreading is integer, so populations, velocities, relaxation times and sphere states must
be bitwise equal.  The LDS-staged kernel runs on a wall-bounded Casson box; the directly spreading
kernel (a periodic extent too short for the LDS boxes) on a narrow periodic box.
"""
import numpy as np
import pytest

import d3q19
from workloads import deck

pytestmark = pytest.mark.gpu

LAYOUTS = (d3q19.IBM_LAYOUT_WAVE, d3q19.IBM_LAYOUT_ROWS, d3q19.IBM_LAYOUT_ROWS_LARGE)


def casson_box(size, np_, walls=(0, 0, 1), storage=0):
    eta = (0.7 - 0.5) / 3.0
    return deck(size=size, walls=walls, tau=0.7, uw=0.02, np_=np_, diam=10.0, markers=128, dist=4, seed=17, gap=1.0,
                coll=1, ibm=1, visc=4, vparams=(1e-5, eta, 1e4, 0), local=1, clamp=(0.505, 5.0), zbuf=3.0, storage=storage)


def run(lines, layout):
    c = d3q19.Case(lines=lines)
    c.opt.ibm_layout = layout
    s = d3q19.Sim(c)
    s.write_particles(c.parKin)
    s.run(1, 30, d3q19.STEP_WRITE_U)   # graph replay
    s.time_steps(31, 4)                 # direct launches
    s.sync()
    out = (s.read_f(), s.read_u(), s.read_tau(), s.read_particles_f64())
    s.close()
    return out


def check_layouts(lines, what):
    ref = run(lines, LAYOUTS[0])
    for lay in LAYOUTS[1:]:
        got = run(lines, lay)
        print("%s: layout %d vs %d: max|df| %.3e max|du| %.3e" % (what, lay, LAYOUTS[0], np.abs(got[0] - ref[0]).max(),
                                                               np.abs(got[1] - ref[1]).max()), flush=True)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
        for a, b in zip(got[3], ref[3]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("storage", [0, 3], ids=["fp32", "fp64"])
def test_layouts_bitwise_equal_lds_staged(storage):
    print("ibm layouts: 72x64x48 Casson box, 24 spheres, storage %d" % storage, flush=True)
    check_layouts(casson_box((72, 64, 48), 24, storage=storage), "lds-staged")


def test_layouts_bitwise_equal_direct_spreading():
    # periodic x extent 18 < 2 (r + 3) + 4: stencils of one sphere wrap onto each other, so no LDS box
    print("ibm layouts: 18x64x48 periodic-x box, 6 spheres (direct spreading)", flush=True)
    check_layouts(casson_box((18, 64, 48), 6), "direct")
