"""Input decks for the tests and the benchmark, written in the reference's input-file format and
read by the C host's own parser (host/params.c)."""
import numpy as np

import d3q19


def deck(size, walls=(0, 0, 1), tau=0.8, uw=0.01, np_=0, diam=8.0, markers=0, dist=4, seed=1, gap=2.0,
         coll=1, ibm=1, visc=1, vparams=(0, 0, 0, 0), local=0, clamp=(0.0, 0.0), storage=0, zbuf=None,
         body=(0.0, 0.0, 0.0), buffer=None, kcol=1.0, mass_model=0, init="zero", u0=(0.0, 0.0, 0.0), tangential=(0, 0, 0)):
    """Lines of an input file.  size = interior box edge(s) (SystemSize for periodic axes)."""
    if np.isscalar(size):
        size = (size, size, size)
    N = [size[d] + 2 + (1 if walls[d] else 0) for d in range(3)]
    lines = [
        "iterations 100",
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
    ]
    if zbuf is not None:
        lines.append("z_wall_particle_buffer %r" % zbuf)
    if buffer is not None:
        lines.append("initial_particle_buffer %r" % buffer)
    return lines


def case(**kw):
    return d3q19.Case(lines=deck(**kw))


def perturb(c, seed=0, amp_u=0.02, amp_rho=0.01):
    """Replace the case's initial populations by f_eq(rho, u) with seeded random rho and u per node
    (every node, shell included), so that every population and boundary path is exercised."""
    rng = np.random.default_rng(seed)
    nn = c.nn
    rho = 1.0 + amp_rho * (2 * rng.random(nn) - 1)
    u = amp_u * (2 * rng.random((3, nn)) - 1)
    cvec = np.array(c.ip.BasisVel, dtype=np.float64)
    w = np.array(c.fp.EqWeights, dtype=np.float64)
    cu = cvec @ u
    usq = (u * u).sum(0)
    f = w[:, None] * rho[None, :] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq[None, :])
    c.f[:] = f.astype(np.float32).ravel()
    return c


def interior(a, N, ncomp):
    """View of a reference-layout field restricted to interior nodes: shape (ncomp, z, y, x)."""
    return a.reshape(ncomp, N[2], N[1], N[0])[:, 1:-1, 1:-1, 1:-1]


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))
