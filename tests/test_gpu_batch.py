"""GPU: batched ensembles and the serving scheduler through the C-ABI.

- A cell stepped in a batch matches the same cell stepped as a single simulation.  Populations
  agree within 1e-5 relative L-inf in fp32; the batched and single kernels share their per-node
  code.
- A cell's results do not depend on the rest of the batch.  The same cell alone and among other
  cells (added, removed, autotuned in between) gives bitwise equal populations and observables.
- The server's results match the plain sequential solve (d3q19_solve): same status and step
  count, observables within 1e-5.  Each request's result is bitwise the same alone and in a mix
  that arrives staggered under a memory cap that forces queueing.
"""
import numpy as np
import pytest

import d3q19
from cases import case, rel_linf

pytestmark = pytest.mark.gpu


def shear_cell(edge, np_=0, diam=8.0, markers=128, uw=0.02, visc=1, vp=(0, 0, 0, 0), tau=0.8, seed=1, storage=0):
    local = 0 if visc == 1 else 1
    return case(size=edge, walls=(0, 0, 1), tau=tau, uw=uw, np_=np_, diam=diam, markers=markers, dist=4, seed=seed,
                gap=2.0, coll=1, ibm=1, visc=visc, vparams=vp, local=local, clamp=(0.51, 5.0) if local else (0.0, 0.0),
                storage=storage)


def mix_cells():
    """Uneven cells: Newtonian, power-law and Casson (Papanastasiou) rheology, 0-3 spheres."""
    return [
        shear_cell(16, np_=1, diam=6.0, markers=128, uw=0.02, seed=3),
        shear_cell(24, np_=3, diam=6.0, markers=160, uw=0.01, visc=2, vp=(0.02, 0.7, 0, 0), tau=0.9, seed=4),
        shear_cell(12, uw=0.03, visc=4, vp=(1e-5, 0.0667, 1e4, 0), tau=0.7),
        shear_cell((20, 16, 18), np_=2, diam=5.0, markers=128, uw=0.015, visc=2, vp=(0.03, 1.2, 0, 0), tau=0.9, seed=5),
    ]


def single_run(c, steps):
    s = d3q19.Sim(c)
    if c.np:
        s.write_particles(c.parKin)
    s.run(1, steps)
    s.sync()
    f = s.read_f()
    o = s.observables()
    s.close()
    return f, o


def test_batch_matches_single_simulations():
    cells = mix_cells()
    b = d3q19.Batch(storage=0, collision=d3q19.COLL_BGK, ibm=d3q19.IBM_PESKIN4, arena_bytes=1 << 30)
    ids = [b.add(c) for c in cells]
    b.run(31)  # odd: exercises the direct steps around the graph pairs
    b.sync()
    obs = b.observables(ids)
    for i, c in zip(ids, cells):
        fb = b.read_f(i)
        fs, os_ = single_run(c, 31)
        e = rel_linf(fb, fs)
        print("cell %d N=%s np=%d: f rel Linf %.3e  stress batch %.9e single %.9e" % (i, c.N, c.np, e, obs[i].wall_stress,
                                                                                 os_.wall_stress), flush=True)
        assert e < 1e-5
        assert abs(obs[i].wall_stress - os_.wall_stress) <= 1e-5 * abs(os_.wall_stress)
    b.close()


def test_batch_population_layouts_match_single_simulations():
    """Batch members take one of three population layouts by their row length: chunks walked in
    node order (rows of 64: runs that are whole wavefronts), chunks walked in padded row slots
    (rows of 50 and 100: padded to 64 and 128 slots, the padding lanes idle) and planes (rows of
    20); rows of 32 go two to a chunk.  Each matches the same cell as a single simulation (one row per chunk, walked in row tiles)."""
    cells = [
        shear_cell((50, 50, 24), np_=1, diam=6.0, markers=128, uw=0.02, seed=6),
        shear_cell((64, 32, 16), uw=0.01, visc=2, vp=(0.02, 0.7, 0, 0), tau=0.9),
        shear_cell((20, 20, 20), np_=1, diam=5.0, markers=128, uw=0.03, seed=7),
        shear_cell((100, 40, 16), np_=2, diam=6.0, markers=160, uw=0.015, visc=4, vp=(1e-5, 0.0667, 1e4, 0), tau=0.7, seed=8),
        shear_cell((32, 32, 24), np_=1, diam=6.0, markers=128, uw=0.02, seed=9),
    ]
    b = d3q19.Batch(storage=0, collision=d3q19.COLL_BGK, ibm=d3q19.IBM_PESKIN4, arena_bytes=1 << 30)
    ids = [b.add(c) for c in cells]
    b.run(21)
    b.sync()
    obs = b.observables(ids)
    for i, c in zip(ids, cells):
        fb = b.read_f(i)
        fs, os_ = single_run(c, 21)
        e = rel_linf(fb, fs)
        print("cell %d N=%s np=%d: f rel Linf %.3e  stress batch %.9e single %.9e" % (i, c.N, c.np, e, obs[i].wall_stress,
                                                                                 os_.wall_stress), flush=True)
        assert e < 1e-5
        assert abs(obs[i].wall_stress - os_.wall_stress) <= 1e-5 * abs(os_.wall_stress)
    b.close()


def test_cell_alone_equals_cell_in_mix_bitwise():
    cells = mix_cells()
    x = cells[1]
    alone = d3q19.Batch(storage=1, arena_bytes=1 << 29)
    xa = alone.add(x)
    alone.run(40)
    mix = d3q19.Batch(storage=1, arena_bytes=1 << 30)
    first = mix.add(cells[0])
    mix.run(7)                      # x joins a batch already under way (odd phase)
    xm = mix.add(x)
    others = [mix.add(cells[2])]
    mix.run(15)
    mix.autotune()                  # trial launches must leave every cell unchanged
    mix.remove(first)
    others.append(mix.add(cells[3]))
    mix.run(25)
    fa, fm = alone.read_f(xa), mix.read_f(xm)
    oa, om = alone.observables([xa])[0], mix.observables([xm])[0]
    pa, pm = alone.read_particles_f64(xa), mix.read_particles_f64(xm)
    print("alone vs mix: max |df| %.3e  stress %.17e %.17e" % (np.abs(fa - fm).max(), oa.wall_stress, om.wall_stress), flush=True)
    assert np.array_equal(fa, fm)
    assert oa.wall_stress == om.wall_stress and oa.drag_coefficient == om.drag_coefficient
    for a, m in zip(pa, pm):
        assert np.array_equal(a, m)
    alone.close()
    mix.close()


def test_autotune_choice_and_state_unchanged():
    cells = mix_cells()
    b1 = d3q19.Batch(storage=1, arena_bytes=1 << 30)
    b2 = d3q19.Batch(storage=1, arena_bytes=1 << 30)
    i1 = [b1.add(c) for c in cells]
    i2 = [b2.add(c) for c in cells]
    b1.run(10)
    b2.run(10)
    t = b1.autotune()
    print("autotuned:", t.as_dict(), flush=True)
    assert t.tile in (64, 128, 256) and t.span in (1, 2, 4, 8) and t.grid_collide >= 0 and t.collide_ms > 0
    assert b1.tuning().tile == t.tile
    # the trial launches write neither u (the velocity the next IBM pass reads) nor tau nor force
    for a, c in zip(i1, i2):
        assert np.array_equal(b1.read_u(a), b2.read_u(c))
    b1.run(10)
    b2.run(10)
    for a, c in zip(i1, i2):
        assert np.array_equal(b1.read_f(a), b2.read_f(c))
    b1.close()
    b2.close()


def test_batch_write_u_flag_and_memory_accounting():
    c = shear_cell(16, np_=1, diam=6.0)
    b = d3q19.Batch(storage=0, arena_bytes=1 << 28)
    a0, u0, n0 = b.memory()
    i = b.add(c)
    a1, u1, n1 = b.memory()
    assert n1 == 1 and u1 - u0 == b.cell_bytes(c) and a1 == a0 == 1 << 28
    b.run(9, d3q19.STEP_WRITE_U)
    u = b.read_u(i)
    s = d3q19.Sim(c)
    s.write_particles(c.parKin)
    s.run(1, 9, d3q19.STEP_WRITE_U)
    us = s.read_u()
    s.close()
    assert rel_linf(u, us) < 1e-5
    b.remove(i)
    assert b.memory()[1] == u0
    with pytest.raises(d3q19.D3Q19Error):
        b.read_u(i)
    # a cell larger than the arena is refused with the capacity code
    big = shear_cell(128)
    with pytest.raises(d3q19.D3Q19Error) as ei:
        b.add(big)
    assert ei.value.code == -7
    b.close()


def requests():
    return [
        shear_cell(12, uw=0.02),
        shear_cell(16, np_=1, diam=6.0, markers=128, uw=0.02, seed=8),
        shear_cell(14, uw=0.01, visc=2, vp=(0.02, 0.8, 0, 0), tau=0.9),
        shear_cell(10, uw=0.03, visc=4, vp=(1e-5, 0.0667, 1e4, 0), tau=0.7),
        shear_cell(18, np_=2, diam=5.0, markers=128, uw=0.015, seed=9),
    ]


SP = dict(tolerance=1e-4, check_interval=50, min_steps=200, max_steps=6000)


def test_server_matches_sequential_solves():
    reqs = requests()
    sp = d3q19.solve_params(**SP)
    srv = d3q19.Server(storage=0, arena_bytes=1 << 30)
    ids = [srv.submit(c, sp) for c in reqs]
    done = {}
    for _ in range(1000):
        for r in srv.progress():
            done[r.id] = r
        if len(done) == len(reqs):
            break
    assert len(done) == len(reqs)
    for rid, c in zip(ids, reqs):
        r = done[rid]
        s = d3q19.solve(c, sp)
        print("req %d N=%s: server status %d steps %d stress %.9e | solve status %d steps %d stress %.9e" % (
            rid, c.N, r.status, r.steps, r.obs.wall_stress, s.status, s.steps, s.obs.wall_stress), flush=True)
        assert r.status == s.status and r.steps == s.steps
        assert abs(r.obs.wall_stress - s.obs.wall_stress) <= 1e-5 * abs(s.obs.wall_stress)
        assert r.t_submit <= r.t_admit <= r.t_finish
    srv.close()


def run_server(reqs, sp, arena, stagger):
    """Submit reqs[i] after stagger[i] scheduling rounds; returns results by request index."""
    srv = d3q19.Server(storage=1, arena_bytes=arena)
    rid_of = {}
    done = {}
    max_queued = 0
    rounds = 0
    while len(done) < len(reqs):
        for i, st in enumerate(stagger):
            if st == rounds and i not in rid_of.values():
                rid_of[srv.submit(reqs[i], sp)] = i
        for r in srv.progress():
            done[rid_of[r.id]] = r
        q, res, used = srv.status()
        max_queued = max(max_queued, q)
        rounds += 1
        assert rounds < 2000
    srv.close()
    return done, max_queued


def test_server_result_alone_equals_result_in_mix():
    reqs = requests()
    sp = d3q19.solve_params(**SP)
    alone = {}
    for i, c in enumerate(reqs):
        d, _ = run_server([c], sp, 1 << 29, [0])
        alone[i] = d[0]
    # staggered arrivals into an arena that holds the largest cell plus one small one
    probe = d3q19.Batch(storage=1, arena_bytes=1 << 20)
    sizes = sorted(probe.cell_bytes(c) for c in reqs)
    probe.close()
    arena = ((sizes[-1] + sizes[1]) + 255) // 256 * 256
    mix, max_queued = run_server(reqs, sp, arena, [0, 0, 1, 3, 3])
    print("max queued %d" % max_queued, flush=True)
    assert max_queued >= 1
    for i in range(len(reqs)):
        a, m = alone[i], mix[i]
        print("req %d: alone steps %d stress %.17e | mix steps %d stress %.17e" % (i, a.steps, a.obs.wall_stress, m.steps,
                                                                                   m.obs.wall_stress), flush=True)
        assert a.status == m.status and a.steps == m.steps
        assert a.obs.wall_stress == m.obs.wall_stress
        assert a.obs.drag_coefficient == m.obs.drag_coefficient


def kicked_cell():
    """A shear cell whose one sphere starts at 1.5 lattice units per step: it crosses more than one
    lattice unit in its first step, beyond what its dilated brick marks cover."""
    c = shear_cell(24, np_=1, diam=6.0, markers=128, uw=0.02, seed=2)
    k = c.parKin[: 16 * c.np].reshape(4, c.np, 4)
    k[1, 0, :3] = (1.5, 0.0, 0.0)
    c.parKin[: 16 * c.np] = k.ravel()
    return c


def test_displacement_watchdog_ends_requests_with_an_error():
    """A sphere that moves a lattice unit or more in one step makes d3q19_check_finite, the
    sequential solve and the server report D3Q19_E_STATE instead of a converged result; the flag
    travels with a checkpoint."""
    c = kicked_cell()
    s = d3q19.Sim(c)
    s.write_particles(c.parKin)
    s.run(1, 4)
    s.sync()
    assert s.check_finite() == d3q19.E_STATE
    img = s.checkpoint()
    s.close()
    s2 = d3q19.Sim(c)
    assert s2.check_finite() == 0
    s2.restore(img)
    assert s2.check_finite() == d3q19.E_STATE
    s2.close()
    sp = d3q19.solve_params(tolerance=1e-4, check_interval=20, min_steps=0, max_steps=200)
    r = d3q19.solve(c, sp)
    print("solve: status %d after %d steps" % (r.status, r.steps), flush=True)
    assert r.status == d3q19.E_STATE and r.steps == 20
    srv = d3q19.Server(storage=0, arena_bytes=1 << 28)
    calm = shear_cell(16, np_=1, diam=6.0, seed=3)
    ids = [srv.submit(calm, sp), srv.submit(c, sp)]
    done = {}
    for _ in range(50):
        for res in srv.progress():
            done[res.id] = res
        if len(done) == 2:
            break
    srv.close()
    print("server: statuses %s" % {k: (v.status, v.steps) for k, v in done.items()}, flush=True)
    assert done[ids[1]].status == d3q19.E_STATE and done[ids[1]].steps == 20
    assert done[ids[0]].status in (0, 1)


def test_nonfinite_density_ends_requests_with_an_error():
    """A request whose lattice turns non-finite (a wall speed that is not a number: wall speeds are
    not validated, as in the reference) ends with D3Q19_E_NONFINITE at the first check in the
    sequential solve and in the server, whose one-launch watchdog scans every resident cell; a calm
    request beside it in the server is unaffected."""
    wild = case(size=16, walls=(0, 0, 1), tau=0.8, uw=float("nan"), coll=1, ibm=1)
    sp = d3q19.solve_params(tolerance=1e-4, check_interval=20, min_steps=0, max_steps=2000)
    r = d3q19.solve(wild, sp)
    print("solve: status %d after %d steps" % (r.status, r.steps), flush=True)
    assert r.status == d3q19.E_NONFINITE and r.steps == 20
    calm = shear_cell(16, np_=1, diam=6.0, seed=3)
    srv = d3q19.Server(storage=0, arena_bytes=1 << 28)
    ids = [srv.submit(calm, sp), srv.submit(wild, sp)]
    done = {}
    for _ in range(200):
        for res in srv.progress():
            done[res.id] = res
        if len(done) == 2:
            break
    srv.close()
    print("server: statuses %s" % {k: (v.status, v.steps) for k, v in done.items()}, flush=True)
    assert done[ids[1]].status == d3q19.E_NONFINITE and done[ids[1]].steps == r.steps
    seq_calm = d3q19.solve(calm, sp)
    assert done[ids[0]].status == seq_calm.status and done[ids[0]].steps == seq_calm.steps
