"""GPU: the allocation-free plane-mean entry point (d3q19_plane_mean_u_noalloc) against
d3q19_plane_mean_u, through the C-ABI.

Both run the same reduction kernel; the variant stages its output in the library's file-scope
device array instead of a buffer it allocates, so the results are bitwise equal.  Two calls in a
row, the second on a smaller lattice (the shared array keeps the first call's contents beyond the
second's planes), then a lattice above the array's 2048 planes, which is refused.
"""
import numpy as np
import pytest

import d3q19
from cases import case

pytestmark = pytest.mark.gpu


def sheared(size, np_=0, seed=1):
    c = case(size=size, walls=(0, 0, 1), tau=0.8, uw=0.02, np_=np_, diam=6.0, markers=128, dist=4, seed=seed, coll=1, ibm=1)
    s = d3q19.Sim(c)
    if c.np:
        s.write_particles(c.parKin)
    s.run(1, 30, d3q19.STEP_WRITE_U)
    s.sync()
    return s


def test_plane_mean_noalloc_matches_allocating_entry_point():
    big = sheared((24, 20, 40), np_=2, seed=3)
    small = sheared((16, 16, 12), np_=1, seed=4)
    for s in (big, small, big):
        ref = s.plane_mean_u()
        got = s.plane_mean_u_noalloc()
        print("plane means: Nz=%d max|u_x| %.3e, equal %s" % (s.case.N[2], np.abs(ref[0]).max(), np.array_equal(ref, got)),
              flush=True)
        assert np.abs(ref[0]).max() > 0
        assert np.array_equal(ref, got)
        assert got[:, 0].tolist() == [0.0, 0.0, 0.0] and got[:, -1].tolist() == [0.0, 0.0, 0.0]
    big.close()
    small.close()


def test_plane_mean_noalloc_refuses_lattices_above_its_scratch():
    c = case(size=(6, 6, 2100), walls=(0, 0, 1), tau=0.8, uw=0.02)
    s = d3q19.Sim(c)
    with pytest.raises(d3q19.D3Q19Error) as e:
        s.plane_mean_u_noalloc()
    assert e.value.code == d3q19.E_CAPACITY
    assert s.plane_mean_u().shape == (3, c.N[2])
    s.close()
