"""GPU parity: the HIP library through the C-ABI against the CPU oracle on identical seeded inputs.

Tolerances (stated per test): the fluid fields must match within 1e-5 relative L-inf (relative to
the oracle field's maximum magnitude) after the given number of steps; particle positions,
velocities and hydrodynamic forces within 1e-5 relative.  The oracle runs in float32 with the
reference's arithmetic (parity mode) or in float64 (fp64 storage); the GPU computes rho, u and the
collision in float32 (float64 for fp64 storage) and the particle state in float64.
"""
import numpy as np
import pytest

import d3q19
import oracle_py
from cases import case, interior, perturb, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-5


def run_pair(c, steps, precision="f32", kernel_level=False):
    """Run `steps` steps on the GPU and on the oracle from the same state; return both."""
    print("case N=%s np=%d coll=%d ibm=%d vis=%d st=%d steps=%d" % (c.N, c.np, c.opt.collision, c.opt.ibm_kernel,
                                                                     c.opt.variable_viscosity, c.opt.storage, steps), flush=True)
    s = d3q19.Sim(c)
    s.upload_case()
    o = oracle_py.Oracle(c, precision).load_case()
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
    o.step(steps)
    return s, o


def compare_fluid(s, o, c, tol=TOL):
    N = c.N
    fg = interior(s.read_f() if o.dtype == np.float32 else s.read_f64(), N, 19)
    fo = interior(o.arr(oracle_py.F_C), N, 19)
    ug = interior(s.read_u(), N, 3)
    uo = interior(o.arr(oracle_py.U), N, 3)
    ef, eu = rel_linf(fg, fo), rel_linf(ug, uo)
    print("rel Linf f=%.3e u=%.3e" % (ef, eu), flush=True)
    assert ef < tol, "population field differs: %.3e" % ef
    assert eu < tol, "velocity field differs: %.3e" % eu
    return ef, eu


def compare_particles(s, o, c, tol=TOL):
    if c.np == 0:
        return
    pk, pf, pff = s.read_particles()
    ok = o.arr(oracle_py.PARKIN).reshape(4, c.np, 4)
    gk = pk[: 16 * c.np].reshape(4, c.np, 4)
    sys_ = np.array(list(c.ip.SystemSize), dtype=np.float64)
    dpos = np.abs(gk[0, :, :3] - ok[0, :, :3])
    dpos = np.minimum(dpos, sys_ - dpos)  # periodic wrap
    e_pos = dpos.max() / sys_.max()
    e_vel = rel_linf(gk[1, :, :3], ok[1, :, :3]) if np.abs(ok[1]).max() > 0 else 0.0
    e_om = rel_linf(gk[3, :, :3], ok[3, :, :3]) if np.abs(ok[3]).max() > 1e-12 else 0.0
    nfa = max(c.ip.WorkGroupsPerParticle, 1)
    of = o.arr(oracle_py.PARFLUID).reshape(2, c.np, nfa, 4).sum(2)
    gf = pff[: 8 * c.np * nfa].reshape(2, c.np, nfa, 4).sum(2)
    e_f = rel_linf(gf[0, :, :3], of[0, :, :3])
    print("particles: pos %.3e vel %.3e omega %.3e force %.3e" % (e_pos, e_vel, e_om, e_f), flush=True)
    assert e_pos < tol and e_vel < tol and e_om < 10 * tol and e_f < 10 * tol


@pytest.mark.parametrize("coll", [0, 1])
def test_periodic_box_single_collide(coll):
    c = perturb(case(size=(12, 10, 9), walls=(0, 0, 0), coll=coll, tau=0.9), seed=3)
    s, o = run_pair(c, 1, kernel_level=True)
    compare_fluid(s, o, c)


@pytest.mark.parametrize("coll", [0, 1])
def test_wall_box_first_steps(coll):
    # the first step uses the raw uploaded wall unknowns, later steps rebuild them
    c = perturb(case(size=(16, 12, 10), walls=(0, 0, 1), coll=coll, tau=0.7, uw=0.02), seed=4)
    s, o = run_pair(c, 5)
    compare_fluid(s, o, c)


def test_wall_axis_x_and_tangential():
    c = perturb(case(size=(10, 12, 14), walls=(1, 0, 0), coll=0, tau=0.8, uw=0.02, tangential=(0, 1, 0)), seed=5)
    s, o = run_pair(c, 4)
    compare_fluid(s, o, c)


@pytest.mark.parametrize("visc,vp", [(2, (0.02, 0.7, 0, 0)), (4, (1e-4, 0.05, 0, 0)), (4, (1e-4, 0.05, 1e4, 0)), (3, (1e-4, 0.02, 0.8, 0))])
@pytest.mark.parametrize("coll", [0, 1])
def test_variable_viscosity(visc, vp, coll):
    c = perturb(case(size=(14, 12, 12), coll=coll, visc=visc, vparams=vp, local=1, clamp=(0.51, 5.0), tau=0.9, uw=0.03), seed=6)
    s, o = run_pair(c, 6)
    compare_fluid(s, o, c, tol=2e-5)
    tg = interior(s.read_tau(), c.N, 1)
    to = interior(o.arr(oracle_py.TAU), c.N, 1)
    assert rel_linf(tg, to) < 2e-5


def test_reference_mode_sphere_linear_ibm():
    # reference build: MRT, constant tau, trilinear IBM with 27-point smoothing
    c = case(size=(24, 24, 24), coll=0, ibm=0, np_=1, diam=8.0, dist=4, seed=11, tau=0.8, uw=0.02)
    s, o = run_pair(c, 30)
    compare_fluid(s, o, c)
    compare_particles(s, o, c)


def test_config_mode_sphere_peskin():
    c = case(size=(24, 24, 24), coll=1, ibm=1, np_=1, diam=8.0, markers=256, dist=4, seed=12, tau=0.8, uw=0.02)
    s, o = run_pair(c, 30)
    compare_fluid(s, o, c)
    compare_particles(s, o, c)


def test_kernel_level_entry_points_match_fused_step():
    c = case(size=(20, 20, 20), coll=0, ibm=0, np_=2, diam=6.0, markers=128, dist=4, seed=13, tau=0.8, uw=0.02)
    s, o = run_pair(c, 10, kernel_level=True)
    compare_fluid(s, o, c)
    compare_particles(s, o, c)


def test_contacts_many_spheres():
    # close-packed spheres with a stiff spring so that harmonic and squeeze forces fire
    c = case(size=(32, 32, 32), coll=1, ibm=1, np_=12, diam=6.0, markers=128, dist=4, seed=14, gap=0.05, kcol=0.05,
             tau=0.8, uw=0.03, zbuf=4.0)
    s, o = run_pair(c, 20)
    compare_fluid(s, o, c)
    compare_particles(s, o, c)


def test_fp64_storage_matches_f64_oracle():
    c = perturb(case(size=(16, 14, 12), coll=0, storage=3, tau=0.8, uw=0.02, np_=1, diam=6.0, markers=128, ibm=1), seed=7)
    s, o = run_pair(c, 8, precision="f64")
    compare_fluid(s, o, c, tol=1e-10)


def test_c1_geometry_100_steps():
    # configuration 1 geometry: 64^3 box, walls +-0.01, BGK tau 0.8, one d=16 sphere with 256
    # markers at the centre, Peskin 4-point; 100 steps; fields and sphere state within 1e-5.
    c = case(size=64, coll=1, ibm=1, np_=1, diam=16.0, markers=256, dist=3, buffer=32.0, tau=0.8, uw=0.01)
    s, o = run_pair(c, 100)
    compare_fluid(s, o, c)
    compare_particles(s, o, c)
