"""Pins of the CPU oracle (oracle/d3q19_oracle.c) -- CPU only.

Each check is chosen so that a plausible restatement mistake fails it: the reference's own shipped
results (input_file.txt, 216 spheres, 100 steps), the reference's moment-transform coefficient
tables, closed-form moments, the MRT == BGK identity, conservation, pure-permutation streaming,
steady Couette flow, power-law channel flow, Peskin kernel identities and the IBM momentum balance.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import d3q19
import oracle_py
from cases import case, interior, perturb

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def velocity_set():
    c = case(size=8)
    return np.array(c.ip.BasisVel, dtype=np.float64), np.array(c.fp.EqWeights, dtype=np.float64)


def test_moment_matrix_matches_reference_tables():
    t = json.load(open(os.path.join(GOLD, "mrt_tables.json")))
    Mref, Minvref = np.array(t["M"]), np.array(t["Minv"])
    o = oracle_py.Oracle(case(size=8), "f64")
    M, Minv = o.moment_matrix()
    # the reference sets the density moment of f_neq to zero instead of computing row 0
    assert np.array_equal(M[1:], Mref[1:])
    assert np.all(M[0] == 1)
    # the reference's inverse is hard-coded to 8 significant digits
    assert np.abs(Minv - Minvref).max() < 5e-9
    assert np.abs(Minv @ M - np.eye(19)).max() < 1e-14


def test_equilibrium_moments():
    c, w = velocity_set()
    rng = np.random.default_rng(0)
    for _ in range(5):
        rho = 1 + 0.1 * rng.random()
        u = 0.1 * (2 * rng.random(3) - 1)
        f = oracle_py.feq(rho, u)
        assert abs(f.sum() - rho) < 1e-14
        assert np.abs(c.T @ f - rho * u).max() < 1e-14
        P = np.einsum("qa,qb,q->ab", c, c, f)
        assert np.abs(P - (rho * np.outer(u, u) + rho / 3 * np.eye(3))).max() < 1e-14


def test_guo_term_moments():
    c, w = velocity_set()
    rng = np.random.default_rng(1)
    for _ in range(5):
        u = 0.1 * (2 * rng.random(3) - 1)
        g = 1e-3 * (2 * rng.random(3) - 1)
        F = oracle_py.guo(u, g)
        assert abs(F.sum()) < 1e-16
        assert np.abs(c.T @ F - g).max() < 1e-16
        P = np.einsum("qa,qb,q->ab", c, c, F)
        assert np.abs(P - (np.outer(u, g) + np.outer(g, u))).max() < 1e-16


def test_mrt_with_uniform_rates_equals_bgk():
    base = dict(size=(10, 8, 9), walls=(0, 0, 1), tau=0.73, uw=0.02, body=(1e-5, 0, 0))
    cm = perturb(case(coll=0, **base), seed=2)
    cb = perturb(case(coll=1, **base), seed=2)
    om = oracle_py.Oracle(cm, "f64").load_case()
    om.set_s_override(np.full(19, 1 / float(np.float32(0.73))))  # NewtonianTau is float32
    ob = oracle_py.Oracle(cb, "f64").load_case()
    om.step(3)
    ob.step(3)
    fm, fb = interior(om.arr(oracle_py.F_C), cm.N, 19), interior(ob.arr(oracle_py.F_C), cb.N, 19)
    assert np.abs(fm - fb).max() < 1e-14


def test_rest_state_is_fixed_point():
    c = case(size=(8, 7, 6), walls=(0, 0, 1), uw=0.0, coll=0)
    o = oracle_py.Oracle(c, "f64").load_case()
    w = oracle_py.feq(1.0, np.zeros(3))  # exact f64 weights
    o.arr(oracle_py.F_C)[:] = np.repeat(w, c.nn)
    o.arr(oracle_py.F_S)[:] = np.repeat(w, c.nn)
    f0 = o.arr(oracle_py.F_C).copy()
    o.step(5)
    assert np.abs(o.arr(oracle_py.F_C) - f0).max() < 1e-15


def test_periodic_streaming_is_a_permutation():
    # with every relaxation rate zero and no force, collide is the identity, so one step is pure
    # streaming: f'[q](r) = f[q](r - c_q) with periodic wrap over the interior
    c = perturb(case(size=(7, 6, 5), walls=(0, 0, 0), coll=0), seed=3)
    o = oracle_py.Oracle(c, "f64").load_case()
    o.set_s_override(np.zeros(19))
    f0 = interior(o.arr(oracle_py.F_C).copy(), c.N, 19)
    o.step(1)
    f1 = interior(o.arr(oracle_py.F_C), c.N, 19)
    cv, _ = velocity_set()
    for q in range(19):
        exp = np.roll(f0[q], shift=(int(cv[q, 2]), int(cv[q, 1]), int(cv[q, 0])), axis=(0, 1, 2))
        assert np.array_equal(f1[q], exp), q


def test_mass_and_momentum_conserved_periodic():
    c = perturb(case(size=(9, 8, 7), walls=(0, 0, 0), coll=0, tau=0.6), seed=4)
    o = oracle_py.Oracle(c, "f64").load_case()
    cv, _ = velocity_set()

    def totals():
        f = interior(o.arr(oracle_py.F_C), c.N, 19).reshape(19, -1)
        return f.sum(), (cv.T @ f).sum(1)

    m0, j0 = totals()
    o.step(10)
    m1, j1 = totals()
    assert abs(m1 - m0) / m0 < 1e-13
    assert np.abs(j1 - j0).max() < 1e-12


def test_couette_steady_state_is_linear():
    # walls at z lattice layers 1 and N-2 move at -U and +U; steady state is a linear profile
    U = 0.01
    c = case(size=(4, 4, 16), tau=1.0, uw=U, coll=1)
    o = oracle_py.Oracle(c, "f64").load_case()
    o.step(3000)
    u = interior(o.arr(oracle_py.U), c.N, 3)[0].mean(axis=(1, 2))
    z = np.arange(len(u))
    a, b = np.polyfit(z, u, 1)
    assert np.abs(u - (a * z + b)).max() < 1e-6
    assert abs(u.mean()) < 1e-7
    # the wet-node wall (rho fixed to 1, no transverse correction) slips slightly; the slope stays
    # within 10% of 2U/(Nz-3)
    assert abs(a / (2 * U / (len(u) - 1)) - 1) < 0.1


@pytest.mark.parametrize("n,K", [(1.0, 0.1), (1.2, 0.2)])
def test_power_law_channel_matches_closed_form(n, K):
    # body-force channel between resting walls (the reference's validation case shape,
    # input_file_power_law.txt), local tau from the strain rate; steady profile against
    # u(s) = n/(n+1) (g/K)^(1/n) (h^(1+1/n) - |s|^(1+1/n)), s from the channel centre, walls at the
    # wet wall layers (h = (nz-1)/2).  n = 1 is the Newtonian parabola with nu = K.
    g = 2e-6
    c = case(size=(3, 3, 24), walls=(0, 0, 1), tau=0.9, uw=0.0, coll=1, visc=2, vparams=(K, n, 0, 0),
             local=1, body=(g, 0, 0), clamp=(0.505, 50.0))
    o = oracle_py.Oracle(c, "f64").load_case()
    o.step(30000)
    u = interior(o.arr(oracle_py.U), c.N, 3)[0].mean(axis=(1, 2))
    nz = len(u)
    s = np.arange(nz) - (nz - 1) / 2.0
    h = (nz - 1) / 2.0
    Kf, nf, gf = [float(np.float32(x)) for x in (K, n, g)]
    ua = nf / (nf + 1) * (gf / Kf) ** (1 / nf) * (h ** (1 + 1 / nf) - np.abs(s) ** (1 + 1 / nf))
    assert np.abs(u - ua).max() / ua.max() < 0.005


def test_tau_laws():
    f = lambda x: float(np.float32(x))  # parameters travel as float32 in flp_param_struct
    c = case(size=8, visc=2, vparams=(0.02, 1.0, 0, 0), local=1)
    assert abs(oracle_py.Oracle(c, "f64").tau_of(1e-3) - (3 * f(0.02) + 0.5)) < 1e-15    # n = 1 is Newtonian
    c = case(size=8, visc=2, vparams=(0.02, 0.5, 0, 0), local=1)
    assert abs(oracle_py.Oracle(c, "f64").tau_of(4e-4) - (3 * f(0.02) / 0.02 + 0.5)) < 1e-12  # k / sqrt(g)
    c = case(size=8, visc=2, vparams=(0.02, 0.7, 0, 0), local=1)
    pl = 3 * f(0.02) * 4e-4 ** (f(0.7) - 1) + 0.5
    assert abs(oracle_py.Oracle(c, "f64").tau_of(4e-4) - pl) < 1e-12                        # k g^(n-1)
    c = case(size=8, visc=2, vparams=(1e-6, 0.5, 0, 0), local=1)
    assert abs(oracle_py.Oracle(c, "f64").tau_of(1.0) - 0.505) < 1e-15                      # reference floor
    c = case(size=8, visc=2, vparams=(10.0, 0.5, 0, 0), local=1, clamp=(0.51, 5.0))
    assert oracle_py.Oracle(c, "f64").tau_of(1e-6) == pytest.approx(5.0)                  # configuration ceiling
    ty, eta = f(1e-5), f(0.05)
    c = case(size=8, visc=4, vparams=(1e-5, 0.05, 0, 0), local=1)
    gdot = 1e-3
    casson = 3 * (np.sqrt(ty / gdot) + np.sqrt(eta)) ** 2 + 0.5
    assert abs(oracle_py.Oracle(c, "f64").tau_of(gdot) - casson) < 1e-12
    c = case(size=8, visc=4, vparams=(1e-5, 0.05, 1e12, 0), local=1)                        # Papanastasiou m -> inf
    assert abs(oracle_py.Oracle(c, "f64").tau_of(gdot) - casson) < 1e-12
    c = case(size=8, visc=4, vparams=(1e-5, 0.05, 1.0, 0), local=1)
    pap = 3 * (np.sqrt(eta) + np.sqrt(ty / gdot) * (1 - np.exp(-np.sqrt(1.0 * gdot)))) ** 2 + 0.5
    assert abs(oracle_py.Oracle(c, "f64").tau_of(gdot) - pap) < 1e-12
    c = case(size=8, visc=3, vparams=(1e-5, 0.02, 0.7, 0), local=1)
    hb = 3 * (ty / gdot + f(0.02) * gdot ** (f(0.7) - 1)) + 0.5
    assert abs(oracle_py.Oracle(c, "f64").tau_of(gdot) - hb) < 1e-12


def test_peskin_kernel_identities():
    for r in np.linspace(0, 1, 17):
        ks = np.arange(-3, 4)
        phi = np.array([oracle_py.peskin(r - k) for k in ks])
        assert abs(phi.sum() - 1) < 1e-14
        assert abs(((r - ks) * phi).sum()) < 1e-14
        assert abs((phi ** 2).sum() - 3 / 8) < 1e-14
    assert oracle_py.peskin(2.0) == 0.0 and oracle_py.peskin(2.5) == 0.0


@pytest.mark.parametrize("ibm", [0, 1])
def test_ibm_momentum_balance(ibm):
    c = case(size=(20, 20, 20), coll=1, ibm=ibm, np_=2, diam=6.0, markers=128, dist=4, seed=5, uw=0.02)
    o = oracle_py.Oracle(c, "f64").load_case()
    o.step(3)
    g = interior(o.arr(oracle_py.GPF), c.N, 3).reshape(3, -1).sum(1)
    F = o.arr(oracle_py.PARFLUID).reshape(2, c.np, -1, 4).sum(2)[0, :, :3].sum(0)
    assert np.abs(F).max() > 0
    assert np.abs(g + F).max() < 1e-12 * max(1.0, np.abs(F).max())


@pytest.mark.slow
def test_reference_shipped_outputs():
    """The reference's own results of input_file.txt (velocity_profile_z.txt, 94 z-plane means at
    t = 100, and the 94x94 x-z plane at y = 48 of matlab_postproc/velocity_field_centerline.txt)
    against the oracle on the same input; profile through the C host's compute_shear_stress."""
    gold = np.load(os.path.join(GOLD, "reference_outputs.npz"))
    c = d3q19.Case(path=os.path.join(GOLD, "inputs", "input_file.txt"))
    o = oracle_py.Oracle(c, "f32").load_case()
    o.step(100)
    u = o.arr(oracle_py.U).astype(np.float32).copy()
    od = d3q19.OutputData()
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        d3q19.host_lib().host_compute_shear_stress(C.byref(od), C.byref(c.hp), C.byref(c.ip), C.byref(c.fp),
                                                   u.ctypes.data_as(C.c_void_p), d.encode(), 1)
        prof = np.loadtxt(os.path.join(d, "velocity_profile_z.txt"))
    assert np.abs(prof - gold["profile"]).max() <= 1.01e-6          # printed with %8.6f
    N = c.N
    uu = u.reshape(3, N[2], N[1], N[0])
    idx = gold["centerline_idx"].astype(int)
    mine = uu[:, idx[:, 2], idx[:, 1], idx[:, 0]].T
    assert np.abs(mine - gold["centerline_u"]).max() < 5e-7           # printed with %9.7f
