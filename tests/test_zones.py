"""Contact-pair enumeration through the reference's neighbour-zone lists (CPU, oracle only).

Pins of the oracle's zone mode:
- Its zone assignment, membership lists and neighbour lists equal those that the reference's own
  initialize_particle_zones (D3Q19-OpenCL.c:344-463, compiled from the reference tree into
  oracle/_ref) builds from the same float positions, quirks included: neighbour ids formed with
  NumZones[2] where NumZones[1] is meant, and zones listed more than once when NumZones < 3.
- Its zone-mode contact forces equal the sum over pairs pi > pj of the pair's contact force
  times the number of times pj's zone appears in the neighbour list of pi's zone, with those
  lists taken from the reference's function.
- Closed form: with one zone per axis in a fully periodic box every offset maps to the zone
  itself, so each pair is visited 3 x 3 x 3 = 27 times.
"""
import ctypes as C

import numpy as np
import pytest

import d3q19
import oracle_py
from cases import case
from test_host import RefHostParams, needs_ref, ref

# (box, walls, spheres, diameter, rebuild frequency): zone grids (2, 4, 2), (5, 4, 3), (2, 2, 2),
# (3, 3, 3) periodic, (4, 3, 5) with NumZones[2] > NumZones[1]
GRIDS = [
    ((20, 30, 16), (0, 0, 1), 10, 6.0, 10),
    ((40, 28, 24), (0, 0, 1), 24, 6.0, 10),
    ((14, 14, 14), (0, 0, 1), 4, 6.0, 10),
    ((21, 21, 21), (0, 0, 0), 9, 6.0, 10),
    ((28, 21, 36), (0, 0, 1), 20, 6.0, 10),
]


def zone_case(box, walls, np_, diam, rebuild, seed=3, gap=0.3, kcol=0.05, storage=0, coll=1):
    """Seeded spheres, then every odd sphere moved next to its even partner (surface gap in
    [-0.2, 0.25], random direction) so that contacts and squeeze pairs straddle zone faces."""
    c = case(size=box, walls=walls, np_=np_, diam=diam, markers=128, dist=4, seed=seed, gap=gap, kcol=kcol,
             tau=0.8, uw=0.02, zbuf=3.0, contacts=d3q19.CONTACT_ZONES, rebuild=rebuild, storage=storage, coll=coll)
    rng = np.random.default_rng(seed)
    k = c.parKin[: 16 * c.np].reshape(4, c.np, 4)
    L = np.array(list(c.ip.SystemSize), dtype=np.float64)
    for p in range(1, c.np, 2):
        for _ in range(100):
            e = rng.normal(size=3)
            e /= np.linalg.norm(e)
            x = k[0, p - 1, :3].astype(np.float64) + e * (diam + rng.uniform(-0.2, 0.25))
            x = np.mod(x, L)
            if walls[2] == 0 or 3.5 < x[2] < L[2] - 3.5:
                break
        k[0, p, :3] = x.astype(np.float32)
    c.parKin[: 16 * c.np] = k.ravel()
    return c


def reference_zone_lists(c):
    """The reference's initialize_particle_zones on the case's float positions."""
    np_ = c.np
    nzt = int(np.prod(list(c.ip.NumZones)))
    rh = RefHostParams()
    rh.RebuildFreq = c.hp.RebuildFreq
    for d in range(3):
        rh.DomainDecomp[d] = c.hp.DomainDecomp[d]
    nthreads = int(np.prod(list(rh.DomainDecomp)))
    ri, rf = d3q19.IntParams(), d3q19.FlpParams()
    C.memmove(C.byref(ri), C.byref(c.ip), C.sizeof(ri))
    C.memmove(C.byref(rf), C.byref(c.fp), C.sizeof(rf))
    rk = np.ascontiguousarray(c.parKin[: 16 * np_], dtype=np.float32)
    pz, tm, npt = np.zeros(np_, np.int32), np.zeros(nthreads * np_, np.int32), np.zeros(nthreads, np.int32)
    zm, nz, zn = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    ref().initialize_particle_zones(C.byref(rh), C.byref(ri), C.byref(rf), rk.ctypes.data_as(C.c_void_p),
                                    pz.ctypes.data_as(C.c_void_p), C.byref(zm), C.byref(nz),
                                    tm.ctypes.data_as(C.c_void_p), npt.ctypes.data_as(C.c_void_p), C.byref(zn))
    assert list(ri.NumZones) == list(c.ip.NumZones) and list(rf.ZoneWidth) == list(c.fp.ZoneWidth)
    num = np.ctypeslib.as_array(nz, shape=(nzt,)).copy()
    mem = np.ctypeslib.as_array(zm, shape=(nzt * np_,)).copy().reshape(nzt, np_)
    neigh = np.ctypeslib.as_array(zn, shape=(28 * nzt,)).copy().reshape(nzt, 28)
    libc = C.CDLL("libc.so.6")
    for p in (zm, nz, zn):
        libc.free(C.cast(p, C.c_void_p))
    return pz, num, mem, neigh


def multiplicity(pz, neigh, pi, pj):
    lst = neigh[pz[pi], 1:1 + neigh[pz[pi], 0]]
    return int(np.count_nonzero(lst == pz[pj]))


@needs_ref
@pytest.mark.parametrize("grid", range(len(GRIDS)))
def test_oracle_zone_lists_match_reference(grid):
    c = zone_case(*GRIDS[grid])
    o = oracle_py.Oracle(c, "f64").load_case()
    pz, num, mem, neigh = o.zone_lists()
    rpz, rnum, rmem, rneigh = reference_zone_lists(c)
    print("zones %s, spheres %d, neighbour entries per zone %s" % (list(c.ip.NumZones), c.np, sorted(set(rneigh[:, 0]))))
    assert np.array_equal(pz, rpz)
    assert np.array_equal(num, rnum)
    for z in range(len(num)):
        assert np.array_equal(mem[z, : num[z]], rmem[z, : rnum[z]])
    assert np.array_equal(neigh, rneigh)
    o.close()


@needs_ref
@pytest.mark.parametrize("grid", range(len(GRIDS)))
def test_zone_contacts_are_pair_forces_times_reference_multiplicity(grid):
    c = zone_case(*GRIDS[grid])
    rpz, _, _, rneigh = reference_zone_lists(c)
    o = oracle_py.Oracle(c, "f64").load_case()
    pf = o.arr(oracle_py.PARFORCE)
    pf[:] = 0
    o.call("particle_forces")
    got = pf.reshape(2, c.np, 4)[0, :, :3].copy()
    want = np.zeros((c.np, 3))
    pairs = 0
    for pi in range(c.np):
        for pj in range(pi):
            pf[:] = 0
            o.fn("pair_force_of")(o.h, pi, pj)
            f = pf.reshape(2, c.np, 4)[0, :, :3]
            if np.any(f):
                pairs += 1
            want += multiplicity(rpz, rneigh, pi, pj) * f
    # the z-wall repulsion does not depend on the zones
    k = float(np.float32(c.fp.ParForceParams[0]))
    z = o.arr(oracle_py.PARKIN).reshape(4, c.np, 4)[0, :, 2]
    lo = float(np.float32(c.fp.ParticleZBuffer)) - z
    hi = z - (c.ip.SystemSize[2] - float(np.float32(c.fp.ParticleZBuffer)))
    want[:, 2] += np.where(lo > 0, k * lo, 0.0) - np.where(hi > 0, k * hi, 0.0)
    print("zones %s: %d pairs in contact range" % (list(c.ip.NumZones), pairs))
    assert pairs > 0
    assert np.abs(got - want).max() <= 1e-12 * max(np.abs(want).max(), 1e-30)
    o.close()


def test_single_zone_periodic_box_visits_each_pair_27_times():
    # 12^3 periodic box, d = 6, rebuild 10: minimum zone width 2 * 0.05 * 10 + 6 = 7 > 12 / 2, so one
    # zone per axis; its 27 offsets all wrap to zone 0, which is listed 27 times.  The two spheres
    # overlap by 0.1 (harmonic force) and sit mid-height (no z-wall repulsion).
    out = []
    for mode in (d3q19.CONTACT_ZONES, d3q19.CONTACT_ALL_PAIRS):
        c = case(size=12, walls=(0, 0, 0), np_=2, diam=6.0, markers=128, dist=4, seed=2, kcol=0.05, tau=0.8, zbuf=1.0,
                 contacts=mode)
        assert list(c.ip.NumZones) == [1, 1, 1]
        k = c.parKin[: 16 * c.np].reshape(4, c.np, 4)
        k[0, 0, :3] = (3.0, 6.0, 6.0)
        k[0, 1, :3] = (8.9, 6.5, 6.2)
        c.parKin[: 16 * c.np] = k.ravel()
        o = oracle_py.Oracle(c, "f64").load_case()
        pf = o.arr(oracle_py.PARFORCE)
        pf[:] = 0
        o.call("particle_forces")
        out.append(pf.reshape(2, c.np, 4)[0, :, :3].copy())
        o.close()
    assert np.abs(out[1]).max() > 1e-3
    assert np.abs(out[0] - 27 * out[1]).max() <= 1e-12 * np.abs(27 * out[1]).max()


def test_zone_lists_rebuilt_every_rebuild_freq_steps():
    c = zone_case((20, 30, 16), (0, 0, 1), 10, 6.0, 3)
    o = oracle_py.Oracle(c, "f64").load_case()
    o.step(3)
    assert o.fn("steps_done")(o.h) == 3
    pz_after = o.zone_lists()[0]
    # the lists now hold the zones of the positions after step 3
    pos = o.arr(oracle_py.PARKIN).reshape(4, c.np, 4)[0, :, :3]
    zw = np.array(list(c.fp.ZoneWidth), dtype=np.float32)
    nz = list(c.ip.NumZones)
    idx = (pos.astype(np.float32) / zw).astype(np.int32)
    assert np.array_equal(pz_after, idx[:, 0] + nz[0] * (idx[:, 1] + nz[1] * idx[:, 2]))
    o.close()
