"""GPU: long-run agreement with the float64 oracle on the 64^3 single-sphere workload.

Workload: 64^3 box, walls at +-0.01, BGK tau 0.8, one d=16 sphere (256 markers) at the centre,
Peskin-4.  2000 steps of the GPU, fp32 and fp16-shifted storage, against 2000 steps of the float64
oracle from the same state.

Bars, 1e-3 for fp32 storage:
- centreline profile u_x(z) at x = y = centre: relative L-inf;
- sphere velocity: absolute, in units of the wall speed (the sphere sits on the zero-velocity
  plane);
- sphere angular velocity: relative to the oracle's;
- sphere position: absolute, in lattice units;
- wall stress: relative;
- drag coefficient: absolute.
fp16-shifted storage (stochastic rounding, the default) keeps the same bars (measured: profile
1.9e-4, velocity 1.3e-5 U, angular velocity 1.3e-4, position 9.9e-5, wall stress 8.5e-6) except
the drag coefficient.  The net fluid force on the centred sphere vanishes by symmetry (the
oracle's C_d is 1.7e-5), so C_d = |F| / (rho U^2 pi a^2 / 2) reads the force's noise floor: fp32
1.0e-4, fp16 1.8e-3, held to 5e-3.  (With round-to-nearest stores, whose errors are frozen in
space and cancel in the symmetric sum, the drag reads 4.7e-4, but the wall stress is off by 1.3e-2
and the profile by 1.7e-3: the stress lives in the non-equilibrium part of the populations, about
7e-4 of w_q here, and every node of a wall plane then makes the same rounding error.)
"""
import numpy as np
import pytest

import d3q19
import oracle_py
from test_gpu_observables import numpy_wall_stress
from workloads import c1_deck

pytestmark = pytest.mark.gpu

STEPS = 2000
U_WALL = float(np.float32(0.01))


@pytest.fixture(scope="module")
def oracle_state():
    c = d3q19.Case(lines=c1_deck(storage=0))
    o = oracle_py.Oracle(c, "f64").load_case()
    print("oracle: %d steps of %s" % (STEPS, c.N), flush=True)
    o.step(STEPS)
    N = c.N
    u = o.arr(oracle_py.U).reshape(3, N[2], N[1], N[0])
    kin = o.arr(oracle_py.PARKIN).reshape(4, c.np, 4)
    nfa = max(c.ip.WorkGroupsPerParticle, 1)
    ff = o.arr(oracle_py.PARFLUID).reshape(2, c.np, nfa, 4).sum(2)[0, :, :3]
    lo, hi = numpy_wall_stress(o.arr(oracle_py.F_C).astype(np.float64), None, c, float(np.float32(0.8)))
    a = 0.5 * c.fp.ParticleDiam
    state = {"ux_centre": u[0, 1:-1, N[1] // 2, N[0] // 2].astype(np.float64).copy(),
             "pos": kin[0, :, :3].astype(np.float64).copy(), "vel": kin[1, :, :3].astype(np.float64).copy(),
             "omega": kin[3, :, :3].astype(np.float64).copy(), "wall_stress": 0.5 * (lo + hi),
             "drag": np.linalg.norm(ff, axis=1).mean() / (0.5 * U_WALL ** 2 * np.pi * a * a)}
    o.close()
    return state


@pytest.mark.parametrize("storage", [0, 1])
def test_c1_2000_steps_match_f64_oracle(oracle_state, storage):
    ref = oracle_state
    c = d3q19.Case(lines=c1_deck(storage=storage))
    s = d3q19.Sim(c)
    s.upload_case()
    s.run(1, STEPS, d3q19.STEP_WRITE_U)
    N = c.N
    u = s.read_u().reshape(3, N[2], N[1], N[0])
    ux = u[0, 1:-1, N[1] // 2, N[0] // 2].astype(np.float64)
    pos, vel, om, ff, tf = s.read_particles_f64()
    o = s.observables()
    assert s.check_finite() == 0
    s.close()
    e_prof = np.abs(ux - ref["ux_centre"]).max() / np.abs(ref["ux_centre"]).max()
    e_vel = np.abs(vel.reshape(-1, 3)[: c.np] - ref["vel"]).max() / U_WALL
    e_om = np.abs(om.reshape(-1, 3)[: c.np] - ref["omega"]).max() / np.abs(ref["omega"]).max()
    e_pos = np.abs(pos.reshape(-1, 3)[: c.np] - ref["pos"]).max()
    e_ws = abs(o.wall_stress - ref["wall_stress"]) / abs(ref["wall_stress"])
    e_cd = abs(o.drag_coefficient - ref["drag"])
    print("storage %d: profile %.2e vel/U %.2e omega %.2e pos %.2e wall stress %.2e drag %.2e (C_d %.6f vs %.6f)" % (
        storage, e_prof, e_vel, e_om, e_pos, e_ws, e_cd, o.drag_coefficient, ref["drag"]), flush=True)
    assert e_vel < 1e-3
    assert e_om < 1e-3
    assert e_pos < 1e-3
    assert e_cd < (1e-3 if storage == 0 else 5e-3)
    assert e_prof < 1e-3
    assert e_ws < 1e-3


@pytest.fixture(scope="module")
def c2_runs():
    """5000 steps of the 128^3 power-law workload per storage format, run on first use."""
    from workloads import c2_deck
    res = {}

    def get(st):
        if st not in res:
            c = d3q19.Case(lines=c2_deck(storage=st))
            s = d3q19.Sim(c)
            s.upload_case()
            s.run(1, C2_STEPS)
            o = s.observables()
            pos = s.read_particles_f64()[0].reshape(-1, 3)[: c.np].copy()
            assert s.check_finite() == 0
            s.close()
            res[st] = (o.relative_viscosity, o.drag_coefficient, pos)
        return res[st]
    return get


C2_STEPS = 5000


def c2_errors(c2_runs, st):
    L = 128.0
    ref = c2_runs(3)
    rv, cd, pos = c2_runs(st)
    d = np.abs(pos - ref[2])
    d = np.minimum(d, L - d).max()
    e_rv, e_cd = abs(rv / ref[0] - 1.0), abs(cd / ref[1] - 1.0)
    print("storage %d: relative viscosity %.6f (fp64 %.6f, err %.2e)  C_d %.6f (fp64 %.6f, err %.2e)  pos %.2e" % (
        st, rv, ref[0], e_rv, cd, ref[1], e_cd, d), flush=True)
    return e_rv, e_cd, d


def test_c2_5000_steps_fp32_against_fp64_build(c2_runs):
    """The 128^3 power-law workload (32 spheres, 192 markers each, n = 0.7, tau clamp [0.51, 5]):
    5000 steps in fp32 storage against the fp64 build of the same kernels (the GPU's FP64 oracle,
    itself within 1e-9 of the float64 oracle in test_gpu_parity.py).  Bars: the relative suspension
    viscosity and the drag coefficient within 1e-3 relative (measured 9.2e-5, 1.4e-4); every
    sphere's position within 2.5e-2 lattice units (measured 5.9e-3, 0.05% of the diameter: contacts
    and shear amplify rounding along the trajectories)."""
    e_rv, e_cd, d = c2_errors(c2_runs, 0)
    assert e_rv < 1e-3
    assert e_cd < 1e-3
    assert d < 2.5e-2


def test_c2_5000_steps_fp16_against_fp64_build(c2_runs):
    """As above for fp16-shifted storage with stochastic rounding (the default), with the same bars
    (measured: relative viscosity 7.3e-5, drag 2.0e-4, positions 4.9e-3).  With round-to-nearest
    stores the same run drifts: 1.8e-2, 1.2e-2 and 5.9e-2 after 5000 steps, growing with time."""
    e_rv, e_cd, d = c2_errors(c2_runs, 1)
    assert e_rv < 1e-3
    assert e_cd < 1e-3
    assert d < 2.5e-2
