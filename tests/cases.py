"""Input decks for the tests and the benchmark, written in the reference's input-file format and
read by the C host's own parser (host/params.c)."""
import numpy as np

import d3q19
from workloads import deck  # noqa: F401  (the decks of the tests and of the benchmark)


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
