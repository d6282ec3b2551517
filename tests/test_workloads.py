"""CPU: the benchmark workload generators build valid cases through the C host (parser, sphere
discretisation, seeded placement) and follow their stated distributions."""
import numpy as np

import d3q19
from workloads import EDGES_C4, MARKER_SETS, c2_deck, c3_deck, c4_decks, c5_requests, poisson_arrivals


def test_c2_and_c3_decks_build():
    c2 = d3q19.Case(lines=c2_deck())
    assert c2.N == (130, 130, 131) and c2.np == 32 and c2.ip.PointsPerParticle == 192
    assert c2.ip.ViscosityModel == 2 and c2.opt.variable_viscosity == 1
    c3 = d3q19.Case(lines=c3_deck(storage=1))
    assert c3.N == (258, 258, 131) and c3.np == 500 and c3.ip.PointsPerParticle == 128
    assert c3.ip.ViscosityModel == 4 and abs(c3.fp.ViscosityParams[2] - 1e4) < 1e-3 and c3.opt.storage == 1
    pos = c3.parKin[: 4 * c3.np].reshape(-1, 4)[:, :3]
    assert (pos[:, 2] >= 5.0).all() and (pos[:, 2] <= 128 - 5.0).all()


def test_c4_ensemble_cells():
    decks = c4_decks()
    assert len(decks) == 64
    kinds = []
    for d in decks:
        c = d3q19.Case(lines=d)
        edge = c.N[0] - 2
        assert edge in EDGES_C4 and c.N == (edge + 2, edge + 2, edge + 3)
        assert 1 <= c.np <= 16
        assert c.ip.PointsPerParticle in MARKER_SETS
        assert 8.0 <= c.fp.ParticleDiam <= 16.0
        assert 0.005 <= c.fp.VelUpper[0] <= 0.03 and c.fp.VelLower[0] == -c.fp.VelUpper[0]
        assert c.opt.storage == 1
        kinds.append(c.ip.ViscosityModel)
    assert abs(kinds.count(1) - 64 / 3) < 1.5 and abs(kinds.count(2) - 64 / 3) < 1.5 and abs(kinds.count(4) - 64 / 3) < 1.5


def test_c5_request_stream():
    decks = c5_requests(count=24)
    edges = []
    for d in decks:
        c = d3q19.Case(lines=d)
        edges.append(c.N[0] - 2)
        assert 0 <= c.np <= 64
    assert min(edges) >= 32 and max(edges) <= 160
    arr = poisson_arrivals(4.0, 2000)
    assert (np.diff(arr) > 0).all()
    assert abs(arr[-1] / 2000 - 0.25) < 0.02
