"""Benchmark workloads of the D3Q19 shear-cell library, written as input decks in the reference's
input-file format (read by the C host's parser, host/params.c) plus the seeded generators of the
batched ensemble and of the serving request stream.

Conventions of the decks: `size` is the interior box edge; a walled axis gets LatticeSize =
size + 3 (SystemSize = size, wet wall nodes at interior 0 and size), a periodic axis
LatticeSize = size + 2.  Walls move at +uw (upper) and -uw (lower) along x.
"""
import math

import numpy as np


def deck(size, walls=(0, 0, 1), tau=0.8, uw=0.01, np_=0, diam=8.0, markers=0, dist=4, seed=1, gap=2.0,
         coll=1, ibm=1, visc=1, vparams=(0, 0, 0, 0), local=0, clamp=(0.0, 0.0), storage=0, zbuf=None,
         body=(0.0, 0.0, 0.0), buffer=None, kcol=1.0, mass_model=0, init="zero", u0=(0.0, 0.0, 0.0), tangential=(0, 0, 0),
         iterations=100, contacts=1, rebuild=10, rounding=0):
    """Lines of an input file.  size = interior box edge(s) (SystemSize for periodic axes).
    contacts: 1 all sphere pairs (the benchmark workloads), 0 the reference's neighbour-zone lists.
    rounding: 16-bit population stores, 0 stochastic, 1 round to nearest."""
    if np.isscalar(size):
        size = (size, size, size)
    N = [size[d] + 2 + (1 if walls[d] else 0) for d in range(3)]
    lines = [
        "iterations %d" % iterations,
        "total_lattice_size %d %d %d" % tuple(N),
        "boundary_conditions_xyz %d %d %d" % tuple(walls),
        "tangential_vel_bcs %d %d %d" % tuple(tangential),
        "newtonian_tau %r" % tau,
        "viscosity_model %d" % visc,
        "viscosity_params %r %r %r %r" % tuple(vparams),
        "velocity_bc_upper %r 0.0 0.0" % uw,
        "velocity_bc_lower %r 0.0 0.0" % (-uw),
        "constant_body_force %r %r %r" % tuple(body),
        "num_particles %d" % np_,
        "particle_diameter %r" % diam,
        "marker_count %d" % markers,
        "initial_particle_distribution %d" % dist,
        "placement_seed %d" % seed,
        "placement_gap %r" % gap,
        "collision_operator %d" % coll,
        "ibm_kernel %d" % ibm,
        "local_relaxation %d" % local,
        "tau_clamp %r %r" % tuple(clamp),
        "population_storage %d" % storage,
        "particle_collision_params %r" % kcol,
        "mass_model %d" % mass_model,
        "initial_f %s" % init,
        "initial_vel %r %r %r" % tuple(u0),
        "contact_detection %d" % contacts,
        "rebuild_neigh_list_freq %d" % rebuild,
        "population_rounding %d" % rounding,
    ]
    if zbuf is not None:
        lines.append("z_wall_particle_buffer %r" % zbuf)
    if buffer is not None:
        lines.append("initial_particle_buffer %r" % buffer)
    return lines


def power_law_K(n, tau_wall, U, H):
    """Consistency K that gives tau = tau_wall at the nominal shear rate 2U/H."""
    nu = (tau_wall - 0.5) / 3.0
    return nu * (2.0 * U / H) ** (1.0 - n)


def c1_deck(storage=0):
    """64^3, walls +-0.01, BGK tau 0.8, one d=16 sphere (256 markers) at the centre, Peskin-4."""
    return deck(size=64, tau=0.8, uw=0.01, np_=1, diam=16.0, markers=256, dist=3, buffer=32.0, coll=1, ibm=1,
                storage=storage)


def c2_deck(storage=0):
    """128^3 power-law (n = 0.7, tau_wall ~ 0.9, clamp [0.51, 5]), walls +-0.02, 32 spheres d=12 x 192
    markers at seeded centres with surface gaps >= 2, BGK, Peskin-4."""
    n, tau, U, H = 0.7, 0.9, 0.02, 128.0
    return deck(size=128, tau=tau, uw=U, np_=32, diam=12.0, markers=192, dist=4, seed=2024, gap=2.0, coll=1, ibm=1,
                visc=2, vparams=(power_law_K(n, tau, U, H), n, 0, 0), local=1, clamp=(0.51, 5.0), storage=storage)


C3_SPHERES = 500


def c3_deck(storage=0, size=(256, 256, 128), spheres=C3_SPHERES):
    """256x256x128 Casson shear cell: yield stress 1e-5, plastic viscosity (0.7 - 0.5)/3, Papanastasiou
    m = 1e4, tau clamp [0.505, 5]; walls +-0.02; 500 spheres d=10 x 128 markers at seeded centres;
    BGK, Peskin-4.  (500 spheres of radius 5 fill 3.1% of the box.)"""
    eta = (0.7 - 0.5) / 3.0
    return deck(size=size, tau=0.7, uw=0.02, np_=spheres, diam=10.0, markers=128, dist=4, seed=77, gap=1.0, coll=1,
                ibm=1, visc=4, vparams=(1e-5, eta, 1e4, 0), local=1, clamp=(0.505, 5.0), storage=storage)


EDGES_C4 = (32, 48, 64, 80, 96)
MARKER_SETS = (128, 160, 192, 256)


def _max_spheres(edge, diam, phi=0.12):
    """Sphere count that keeps the solid fraction of an edge^3 box at most phi."""
    v = math.pi * diam ** 3 / 6.0
    return max(1, int(phi * edge ** 3 / v))


def rheology(rng, kind, U, H, storage=1):
    """Deck keywords of one random rheology: 0 Newtonian (tau in [0.6, 1.0]), 1 power law (n in
    [0.5, 1.3], tau_wall 0.8 at the nominal rate), 2 Casson (yield stress 1e-5, plastic viscosity of
    tau 0.7, Papanastasiou m = 1e4)."""
    if kind == 0:
        return dict(tau=float(rng.uniform(0.6, 1.0)), visc=1, vparams=(0, 0, 0, 0), local=0, clamp=(0.0, 0.0))
    if kind == 1:
        n = float(rng.uniform(0.5, 1.3))
        return dict(tau=0.8, visc=2, vparams=(power_law_K(n, 0.8, U, H), n, 0, 0), local=1, clamp=(0.51, 5.0))
    return dict(tau=0.7, visc=4, vparams=(1e-5, (0.7 - 0.5) / 3.0, 1e4, 0), local=1, clamp=(0.505, 5.0))


def c4_decks(seed=4, cells=64, storage=1):
    """The batched ensemble: 64 independent cubic shear cells, edges uniform over {32,...,96}, 1-16
    spheres of radius 4-8 (capped so the solid fraction stays <= 12%) with 128/160/192/256 markers,
    wall speed uniform in [0.005, 0.03], rheology 1/3 Newtonian, 1/3 power law, 1/3 Casson."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(cells):
        edge = int(rng.choice(EDGES_C4))
        diam = float(2.0 * rng.uniform(4.0, 8.0))
        nsph = min(int(rng.integers(1, 17)), _max_spheres(edge, diam))
        markers = int(rng.choice(MARKER_SETS))
        U = float(rng.uniform(0.005, 0.03))
        rh = rheology(rng, i % 3, U, float(edge), storage)
        out.append(deck(size=edge, uw=U, np_=nsph, diam=diam, markers=markers, dist=4, seed=1000 + i, gap=1.0, coll=1,
                        ibm=1, storage=storage, **rh))
    return out


def c5_requests(seed=5, count=24, edge_range=(32, 160), max_spheres=64, storage=0):
    """The serving stream: box edge log-uniform over edge_range, 0-max_spheres spheres (radius 4-8,
    capped to a 12% solid fraction), wall speed and rheology as the batched ensemble."""
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = math.log(edge_range[0]), math.log(edge_range[1])
    for i in range(count):
        edge = int(round(math.exp(rng.uniform(lo, hi))))
        diam = float(2.0 * rng.uniform(4.0, 8.0))
        nsph = min(int(rng.integers(0, max_spheres + 1)), _max_spheres(edge, diam))
        markers = int(rng.choice(MARKER_SETS))
        U = float(rng.uniform(0.005, 0.03))
        rh = rheology(rng, int(rng.integers(0, 3)), U, float(edge), storage)
        out.append(deck(size=edge, uw=U, np_=nsph, diam=diam, markers=markers, dist=4, seed=5000 + i, gap=1.0, coll=1,
                        ibm=1, storage=storage, **rh))
    return out


def poisson_arrivals(rate, count, seed=6):
    """Arrival offsets (seconds) of a Poisson process of the given rate (requests per second)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


def precision_study(deck_fn, steps, storages=(0, 1, 2), reference=3, nearest=(1, 2)):
    """Run the same workload with each population storage and with the fp64 build; report L2 and
    L-inf errors of u (relative to the fp64 run, over interior nodes), and the errors of the drag
    coefficient, wall stress and sphere positions.  The 16-bit storages round stochastically (the
    default); those listed in `nearest` are also run with round-to-nearest stores ("*_nearest")."""
    import d3q19
    runs = {}
    variants = [(reference, 0)] + [(st, 0) for st in storages] + [(st, 1) for st in nearest]
    for st, rnd in variants:
        c = d3q19.Case(lines=deck_fn(storage=st) + ["population_rounding %d" % rnd])
        s = d3q19.Sim(c)
        if c.np:
            s.write_particles(c.parKin)
        s.run(1, steps, d3q19.STEP_WRITE_U)
        s.sync()
        u = s.read_u().reshape(3, c.N[2], c.N[1], c.N[0])[:, 1:-1, 1:-1, 1:-1].astype(np.float64)
        o = s.observables()
        pos = s.read_particles_f64()[0].reshape(-1, 3)[: c.np] if c.np else np.zeros((0, 3))
        s.close()
        runs[(st, rnd)] = (u, o, pos, c)
        del c
    u64, o64, p64, c = runs[(reference, 0)]
    L = np.array(list(c.ip.SystemSize), dtype=np.float64)
    out = {"reference": "fp64 build", "steps": steps,
           "fp64": {"drag_coefficient": o64.drag_coefficient, "wall_stress": o64.wall_stress,
                    "relative_viscosity": o64.relative_viscosity}}
    names = {0: "fp32", 1: "fp16_shift", 2: "bf16_shift", 3: "fp64"}
    for st, rnd in variants[1:]:
        u, o, pos, _ = runs[(st, rnd)]
        d = u - u64
        dp = np.abs(pos - p64)
        dp = np.minimum(dp, L - dp) if len(pos) else dp
        out[names[st] + ("_nearest" if rnd else "")] = {
            "u_l2_rel": float(np.sqrt((d * d).sum() / (u64 * u64).sum())),
            "u_linf_rel": float(np.abs(d).max() / np.abs(u64).max()),
            "drag_coefficient": o.drag_coefficient,
            "drag_abs_err": abs(o.drag_coefficient - o64.drag_coefficient),
            "drag_rel_err": abs(o.drag_coefficient - o64.drag_coefficient) / max(abs(o64.drag_coefficient), 1e-300),
            "wall_stress_rel_err": abs(o.wall_stress - o64.wall_stress) / max(abs(o64.wall_stress), 1e-300),
            "position_max_err": float(dp.max()) if len(pos) else 0.0,
        }
    return out
