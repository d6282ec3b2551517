"""GPU: a lattice whose population buffers exceed 4 GiB.

The collide addresses a population as a scalar base (its run in the node's z-plane, 64-bit) plus a
32-bit lane offset within the plane, so a lattice's depth is not limited by the 32-bit offsets.
512 x 512 x 256 interior nodes in fp32 hold 19 x 4 B x 67 M = 5.1 GB per population buffer.

Check, independent of size: a sphere-free shear cell, periodic in x and y with walls normal to z,
started at rest, stays a function of z alone, node for node, so its per-plane mean velocity must
equal bitwise that of an 8 x 8 x 256 cell with the same parameters (the plane means sum 2^18 and
2^6 identical values in FP64, exactly).  The two lattices store their populations in runs of one
row each, of 512 and of 8 nodes.
"""
import numpy as np
import pytest

import d3q19
from cases import case

pytestmark = pytest.mark.gpu


def profile(size, steps):
    c = case(size=size, walls=(0, 0, 1), tau=0.7, uw=0.03, coll=1)
    s = d3q19.Sim(c)
    s.run(1, steps, d3q19.STEP_WRITE_U)  # u is stored on every node in the last step
    s.sync()
    assert s.check_finite() == 0
    u = s.plane_mean_u()
    s.close()
    return u


def test_population_buffers_beyond_4gib_match_a_narrow_cell():
    big, small = (512, 512, 256), (8, 8, 256)
    pop_bytes = 19 * 4 * big[0] * big[1] * big[2]
    assert pop_bytes > (4 << 30)
    ub = profile(big, 24)
    us = profile(small, 24)
    dev = np.abs(ub - us).max()
    print("population buffer %.2f GB; max |u_big - u_small| over the plane means %.3e; u_x at the upper wall plane %.6e" %
          (pop_bytes / 1e9, dev, ub[0, -3]), flush=True)
    assert np.abs(us[0]).max() > 1e-4  # the walls have set the flow going
    assert np.array_equal(ub, us)
