"""GPU: the spread force is consumed exactly where the spreading put it.

The IBM pass sets a per-node tag wherever it adds force; the next collide reads the force at tagged
nodes only and clears force and tag together.  After a full step (which ends with an IBM pass) the
pending force is nonzero at the spheres' stencil nodes, and one collide must leave the force field
zero everywhere:
- in periodic and walled boxes, with Peskin-4 and with the reference's smoothed trilinear kernel
  (whose spreading reaches a node further);
- after the spheres were moved by the host between the IBM pass and the collide (positions written
  two lattice units away): the pending force of the old positions is still consumed where it was
  spread, which marks built from positions cannot guarantee.
"""
import numpy as np
import pytest

import d3q19
from cases import case

pytestmark = pytest.mark.gpu


def pending_force_then_collide(c, steps, shift=None):
    s = d3q19.Sim(c)
    s.upload_case()
    s.run(1, steps)
    s.sync()
    before = s.read_force()
    if shift is not None:
        k = c.parKin[: 16 * c.np].reshape(4, c.np, 4).copy()
        k[0, :, :3] += np.asarray(shift, dtype=np.float32)
        s.write_particles(k.ravel())
    s.collide_stream()
    s.sync()
    after = s.read_force()
    s.close()
    return before, after


@pytest.mark.parametrize("walls,ibm", [((0, 0, 1), 1), ((0, 0, 0), 1), ((0, 0, 1), 0)])
def test_collide_consumes_every_spread_node(walls, ibm):
    c = case(size=(28, 26, 24), walls=walls, coll=1, ibm=ibm, np_=3, diam=6.0, markers=128, dist=4, seed=21, tau=0.8, uw=0.02)
    before, after = pending_force_then_collide(c, 6)
    nz = int(np.count_nonzero(before))
    print("walls %s ibm %d: %d nonzero force components pending, %d after the collide" % (walls, ibm, nz, np.count_nonzero(after)),
          flush=True)
    assert nz > 100
    assert not np.any(after)


def test_pending_force_consumed_after_the_spheres_moved():
    c = case(size=(28, 26, 24), walls=(0, 0, 1), coll=1, ibm=1, np_=2, diam=6.0, markers=128, dist=4, seed=22, tau=0.8, uw=0.02)
    before, after = pending_force_then_collide(c, 4, shift=(2.0, 0.0, 0.0))
    print("moved spheres: %d pending, %d after the collide" % (np.count_nonzero(before), np.count_nonzero(after)), flush=True)
    assert np.count_nonzero(before) > 100
    assert not np.any(after)
