"""GPU parity: the HIP library through the C-ABI against the CPU oracle on identical seeded inputs.

Two bars, both stated per test:
- fp64 storage: the GPU's float64 build against the float64 oracle, populations within 1e-10
  relative L-inf and sphere states within 1e-9.  This pins every term, sign and index of the
  algorithm, free of float32 rounding.
- fp32 storage: populations within 1e-5 relative L-inf of the float32 oracle (the reference's
  arithmetic).  Velocity, tau and sphere states are small differences of populations, so float32
  rounding alone moves them by more than 1e-5 relative within tens of steps (the float32 oracle
  differs from the float64 oracle by 7e-5 in u after the 100-step configuration-1 run).  For them
  the bar is the rounding envelope of the reference arithmetic: the GPU's error against the
  float64 oracle must stay within max(1e-5, 3x the float32 oracle's own error against it).
Relative L-inf = max|a - b| / max|b| over interior nodes.
"""
import numpy as np
import pytest

import d3q19
import oracle_py
from cases import case, interior, perturb, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-5
ENVELOPE = 3.0


def run_triplet(c, steps, kernel_level=False):
    """Run `steps` steps on the GPU and on the float32 and float64 oracles from the same state."""
    print("case N=%s np=%d coll=%d ibm=%d vis=%d st=%d steps=%d" % (c.N, c.np, c.opt.collision, c.opt.ibm_kernel,
                                                                     c.opt.variable_viscosity, c.opt.storage, steps), flush=True)
    s = d3q19.Sim(c)
    s.upload_case()
    if kernel_level:
        for _ in range(steps):
            s.collide_stream()
            if c.np:
                s.particle_dynamics()
                s.particle_fluid_forces()
                s.particle_particle_forces()
    else:
        s.run(1, steps, d3q19.STEP_WRITE_U)
    s.sync()
    o32 = oracle_py.Oracle(c, "f32").load_case()
    o64 = oracle_py.Oracle(c, "f64").load_case()
    o32.step(steps)
    o64.step(steps)
    return s, o32, o64


def envelope_check(name, g, o32, o64):
    eg, er = rel_linf(g, o64), rel_linf(o32, o64)
    bar = max(TOL, ENVELOPE * er)
    print("%s: gpu-vs-f64 %.3e  f32oracle-vs-f64 %.3e  gpu-vs-f32oracle %.3e  bar %.3e" % (name, eg, er, rel_linf(g, o32), bar),
          flush=True)
    assert eg <= bar, "%s outside the float32 envelope: %.3e > %.3e" % (name, eg, bar)


def compare_fluid(s, o32, o64, c):
    N = c.N
    if c.opt.storage == 3:
        fg = interior(s.read_f64(), N, 19)
        ef = rel_linf(fg, interior(o64.arr(oracle_py.F_C), N, 19))
        eu = rel_linf(interior(s.read_u(), N, 3), interior(o64.arr(oracle_py.U), N, 3))
        print("fp64: rel Linf f=%.3e u(float32 export)=%.3e" % (ef, eu), flush=True)
        assert ef < 1e-10
        assert eu < 2e-7
        return
    fg = interior(s.read_f(), N, 19)
    ef = rel_linf(fg, interior(o32.arr(oracle_py.F_C), N, 19))
    print("fp32: rel Linf f=%.3e" % ef, flush=True)
    assert ef < TOL, "population field differs: %.3e" % ef
    envelope_check("u", interior(s.read_u(), N, 3), interior(o32.arr(oracle_py.U), N, 3), interior(o64.arr(oracle_py.U), N, 3))
    if c.opt.variable_viscosity:
        envelope_check("tau", interior(s.read_tau(), N, 1), interior(o32.arr(oracle_py.TAU), N, 1), interior(o64.arr(oracle_py.TAU), N, 1))


def compare_particles(s, o32, o64, c):
    if c.np == 0:
        return
    pk, pf, pff = s.read_particles()
    pos, vel, om, ff, tf = s.read_particles_f64()
    k32 = o32.arr(oracle_py.PARKIN).reshape(4, c.np, 4)
    k64 = o64.arr(oracle_py.PARKIN).reshape(4, c.np, 4)
    nfa = max(c.ip.WorkGroupsPerParticle, 1)
    f32 = o32.arr(oracle_py.PARFLUID).reshape(2, c.np, nfa, 4).sum(2)
    f64 = o64.arr(oracle_py.PARFLUID).reshape(2, c.np, nfa, 4).sum(2)
    L = np.array(list(c.ip.SystemSize), dtype=np.float64)

    def wrapdiff(a, b):
        d = np.abs(a - b)
        return np.minimum(d, L - d).max() / L.max()

    gp = pos.reshape(-1, 3)[: c.np]
    if c.opt.storage == 3:
        e = [wrapdiff(gp, k64[0, :, :3]), rel_linf(vel.reshape(-1, 3)[: c.np], k64[1, :, :3]),
             rel_linf(ff.reshape(-1, 3)[: c.np], f64[0, :, :3])]
        print("fp64 particles: pos %.3e vel %.3e force %.3e" % tuple(e), flush=True)
        assert max(e) < 1e-9
        return
    eg, er = wrapdiff(gp, k64[0, :, :3]), wrapdiff(k32[0, :, :3], k64[0, :, :3])
    print("pos: gpu %.3e f32oracle %.3e" % (eg, er), flush=True)
    assert eg <= max(TOL, ENVELOPE * er)
    envelope_check("vel", vel.reshape(-1, 3)[: c.np], k32[1, :, :3], k64[1, :, :3])
    envelope_check("fluid force", ff.reshape(-1, 3)[: c.np], f32[0, :, :3], f64[0, :, :3])
    if np.abs(k64[3, :, :3]).max() > 0:
        envelope_check("omega", om.reshape(-1, 3)[: c.np], k32[3, :, :3], k64[3, :, :3])
    # the float4 read-back carries the same state
    assert np.abs(pk[: 16 * c.np].reshape(4, c.np, 4)[0, :, :3] - gp).max() < 1e-4


def run_pair(c, steps, kernel_level=False):
    s, o32, o64 = run_triplet(c, steps, kernel_level)
    compare_fluid(s, o32, o64, c)
    compare_particles(s, o32, o64, c)
    s.close()


@pytest.mark.parametrize("coll", [0, 1])
def test_periodic_box_single_collide(coll):
    c = perturb(case(size=(12, 10, 9), walls=(0, 0, 0), coll=coll, tau=0.9), seed=3)
    run_pair(c, 1, kernel_level=True)


@pytest.mark.parametrize("coll", [0, 1])
def test_wall_box_first_steps(coll):
    # the first step uses the raw uploaded wall unknowns, later steps rebuild them
    c = perturb(case(size=(16, 12, 10), walls=(0, 0, 1), coll=coll, tau=0.7, uw=0.02), seed=4)
    run_pair(c, 5)


def test_wall_axis_x_and_tangential():
    c = perturb(case(size=(10, 12, 14), walls=(1, 0, 0), coll=0, tau=0.8, uw=0.02, tangential=(0, 1, 0)), seed=5)
    run_pair(c, 4)


@pytest.mark.parametrize("visc,vp", [(2, (0.02, 0.7, 0, 0)), (4, (1e-4, 0.05, 0, 0)), (4, (1e-4, 0.05, 1e4, 0)), (3, (1e-4, 0.02, 0.8, 0))])
@pytest.mark.parametrize("coll", [0, 1])
def test_variable_viscosity(visc, vp, coll):
    c = perturb(case(size=(14, 12, 12), coll=coll, visc=visc, vparams=vp, local=1, clamp=(0.51, 5.0), tau=0.9, uw=0.03), seed=6)
    run_pair(c, 6)


def test_reference_mode_sphere_linear_ibm():
    # reference build: MRT, constant tau, trilinear IBM with 27-point smoothing
    c = case(size=(24, 24, 24), coll=0, ibm=0, np_=1, diam=8.0, dist=4, seed=11, tau=0.8, uw=0.02)
    run_pair(c, 30)


def test_config_mode_sphere_peskin():
    c = case(size=(24, 24, 24), coll=1, ibm=1, np_=1, diam=8.0, markers=256, dist=4, seed=12, tau=0.8, uw=0.02)
    run_pair(c, 30)


def test_kernel_level_entry_points_match_fused_step():
    c = case(size=(20, 20, 20), coll=0, ibm=0, np_=2, diam=6.0, markers=128, dist=4, seed=13, tau=0.8, uw=0.02)
    run_pair(c, 10, kernel_level=True)


def test_contacts_many_spheres():
    # close-packed spheres with a stiff spring so that harmonic and squeeze forces fire
    c = case(size=(32, 32, 32), coll=1, ibm=1, np_=12, diam=6.0, markers=128, dist=4, seed=14, gap=0.05, kcol=0.05,
             tau=0.8, uw=0.03, zbuf=4.0)
    run_pair(c, 20)


@pytest.mark.parametrize("grid", [0, 4])
def test_zone_contacts_fp64(grid):
    # contacts through the reference's zone lists (zone grids (2, 4, 2) and (4, 3, 5): pairs
    # visited 0, 1 or 2 times), lists rebuilt every 3 steps; fp64 build against the f64 oracle
    from test_zones import GRIDS, zone_case
    box, walls, np_, diam, _ = GRIDS[grid]
    c = zone_case(box, walls, np_, diam, 3, storage=3)
    run_pair(c, 13)


def test_zone_contacts_fp32_and_kernel_level():
    from test_zones import GRIDS, zone_case
    box, walls, np_, diam, _ = GRIDS[0]
    run_pair(zone_case(box, walls, np_, diam, 4), 17)
    run_pair(zone_case(box, walls, np_, diam, 2, coll=0), 7, kernel_level=True)


FP64_CASES = {
    "mrt-linear-ibm-walls": dict(size=(16, 14, 12), coll=0, ibm=0, np_=1, diam=6.0, markers=128, tau=0.8, uw=0.02),
    "bgk-peskin-walls": dict(size=(16, 14, 12), coll=1, ibm=1, np_=1, diam=6.0, markers=128, tau=0.8, uw=0.02),
    "periodic": dict(size=(12, 10, 9), walls=(0, 0, 0), coll=0, tau=0.9),
    "wall-x-tangential": dict(size=(10, 12, 14), walls=(1, 0, 0), coll=0, tau=0.8, uw=0.02, tangential=(0, 1, 0)),
    "power-law-mrt": dict(size=(14, 12, 12), coll=0, visc=2, vparams=(0.02, 0.7, 0, 0), local=1, clamp=(0.51, 5.0), tau=0.9, uw=0.03),
    "casson-pap-bgk": dict(size=(14, 12, 12), coll=1, visc=4, vparams=(1e-4, 0.05, 1e4, 0), local=1, clamp=(0.51, 5.0), tau=0.9, uw=0.03),
    "hb-bgk": dict(size=(14, 12, 12), coll=1, visc=3, vparams=(1e-4, 0.02, 0.8, 0), local=1, clamp=(0.51, 5.0), tau=0.9, uw=0.03),
    "contacts": dict(size=(32, 32, 32), coll=1, ibm=1, np_=12, diam=6.0, markers=128, seed=14, gap=0.05, kcol=0.05, tau=0.8, uw=0.03, zbuf=4.0),
}


@pytest.mark.parametrize("name", sorted(FP64_CASES))
def test_fp64_build_matches_f64_oracle(name):
    c = perturb(case(storage=3, **FP64_CASES[name]), seed=7)
    run_pair(c, 12)


def test_c1_geometry_100_steps():
    # configuration 1 geometry: 64^3 box, walls +-0.01, BGK tau 0.8, one d=16 sphere with 256
    # markers at the centre, Peskin 4-point; 100 steps; fields and sphere state within 1e-5.
    c = case(size=64, coll=1, ibm=1, np_=1, diam=16.0, markers=256, dist=3, buffer=32.0, tau=0.8, uw=0.01)
    run_pair(c, 100)


@pytest.mark.parametrize("ibm", [0, 1])
def test_marked_bricks_cover_every_stencil_across_periodic_faces(ibm):
    """Brick marks decide where the collide writes u and consumes the spread force; on boxes whose
    periodic extents resolve the nearest image they are cut to the bricks a marker stencil can
    reach.  36x36 periodic (> 2 (r + 3) + 16 for d = 8), z walls, a seeded random f_eq(rho, u)
    start (|u| <= 0.02) so that the spheres move, four spheres placed across the x face, the y face,
    an xy corner and inside; 30 fused steps without STEP_WRITE_U (so u is written only in marked
    bricks), fp64 build: populations within 1e-10 of the float64 oracle, positions within 1e-9
    lattice units, velocities within 1e-9 of the wall speed.  A stencil node in an unmarked brick
    would read a stale u or leave its force unconsumed, far outside these bars."""
    c = perturb(case(size=(36, 36, 32), coll=1, ibm=ibm, np_=4, diam=8.0, markers=128, dist=4, seed=21, tau=0.8,
                     uw=0.03, storage=3), seed=5)
    kin = np.asarray(c.parKin)
    cen = ((0.6, 18.2, 15.5), (17.9, 35.5, 16.4), (35.3, 0.4, 13.8), (18.5, 9.0, 19.2))
    for p, xyz in enumerate(cen):
        kin[4 * p: 4 * p + 3] = xyz  # parKin[0]: positions, one float4 per sphere
    s = d3q19.Sim(c)
    s.upload_case()
    s.run(1, 30)
    s.sync()
    o64 = oracle_py.Oracle(c, "f64").load_case()
    o64.step(30)
    N = c.N
    ef = rel_linf(interior(s.read_f64(), N, 19), interior(o64.arr(oracle_py.F_C), N, 19))
    pos, vel, om, ff, tf = s.read_particles_f64()
    k64 = o64.arr(oracle_py.PARKIN).reshape(4, c.np, 4)
    L = np.array(list(c.ip.SystemSize), dtype=np.float64)
    d = np.abs(pos.reshape(-1, 3)[: c.np] - k64[0, :, :3])
    ep = np.minimum(d, L - d).max()
    ev = np.abs(vel.reshape(-1, 3)[: c.np] - k64[1, :, :3]).max() / 0.03
    print("ibm %d: f %.3e  pos %.3e  vel/U %.3e (|v| up to %.2e U)" % (ibm, ef, ep, ev, np.abs(k64[1, :, :3]).max() / 0.03),
          flush=True)
    assert s.check_finite() == 0
    s.close()
    assert ef < 1e-10
    assert ep < 1e-9 and ev < 1e-9


@pytest.mark.parametrize("size,coll,storage,visc", [((96, 24, 24), 0, 0, 1), ((96, 24, 24), 1, 0, 2), ((96, 24, 24), 1, 1, 2),
                                                    ((96, 24, 24), 0, 3, 1), ((300, 20, 20), 0, 0, 1), ((300, 20, 20), 1, 1, 2)])
def test_tile_shape_changes_no_result_bit(size, coll, storage, visc):
    """No result may depend on which nodes share a wavefront or a block (the sphere sums of the fused
    step run over 256 virtual threads whatever the block size): 40 steps of a sphere workload
    (forced and unforced nodes in the same rows) with 64x4, 128x2, 256x1 and the default tiles
    (rows of 96 nodes: 128x2; rows of 300 nodes: 256 consecutive nodes of a plane in node order,
    across rows) give the same populations, velocities and sphere state bit for bit."""
    vp = (0, 0, 0, 0) if visc == 1 else (2.0e-2, 0.7, 0, 0)
    runs = []
    for block in ((64, 4), (128, 2), (256, 1), (0, 0)):
        c = case(size=size, coll=coll, ibm=1, np_=6, diam=8.0, markers=128, dist=4, seed=31, tau=0.8, uw=0.03,
                 storage=storage, visc=visc, vparams=vp, local=1 if visc != 1 else 0, clamp=(0.51, 5.0) if visc != 1 else (0, 0))
        c.opt.block[0], c.opt.block[1] = block
        s = d3q19.Sim(c)
        s.upload_case()
        s.run(1, 40)
        runs.append((s.read_f(), s.read_u(), s.read_particles_f64()[0].copy()))
        assert s.check_finite() == 0
        s.close()
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            assert np.array_equal(a, b)
