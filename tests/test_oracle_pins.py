"""Independent pins of the oracle's contact law and immersed-boundary kernels (CPU only).

Contacts (particle_particle_forces + compute_squeeze_force, CPU_program.cl:76-184, 319-343), held
to closed forms computed here from the sphere positions alone:
- two spheres overlapping by d (minimum image across a periodic face too) repel each other with
  |F| = k d along the line of centres, equal and opposite, with no torque;
- two spheres at a surface gap 0 < h < 0.1 r approaching at speed v feel Reynolds' squeeze
  force between two spheres of radii a1, a2, 6 pi mu v (a1 a2 / (a1 + a2))^2 / h, with the gap
  floored at 0.1 lattice units and mu = (tau - 1/2) / 3 the Newtonian lattice viscosity (rho = 1);
  it vanishes beyond 0.1 r and for non-Newtonian fluids;
- a sphere closer than ZBuffer to either z wall is pushed back by k times its depth.

Interpolation and spreading (GPU_program.cl:81-256 and the Peskin 4-point variant): in a linear
velocity field u(x) = a + B x both kernels interpolate exactly (partition of unity and vanishing
first moment), so every marker's force equals (a + B x_m) dA, its torque r0 x F; and the spread
force has total -sum F_m and first moment -sum x_m F_m.

Each check is also run against copies of the oracle with one plausible mistake planted in the
source (a sign, a factor, a radius, an index shift, a dropped minimum image): every planted
mistake must fail at least one check, so the checks pin what they claim to.
"""
import os
import subprocess

import numpy as np
import pytest

import d3q19
import oracle_py
from workloads import deck

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SRC = os.path.join(os.path.dirname(HERE), "oracle", "d3q19_oracle.c")


# ---- contacts ---------------------------------------------------------------------------------
def contact_case(visc=1):
    vp = (0.0, 0.0, 0.0, 0.0) if visc == 1 else (0.01, 0.7, 0.0, 0.0)
    return d3q19.Case(lines=deck(size=(40, 40, 40), walls=(0, 0, 1), np_=2, diam=8.0, markers=128, dist=4, seed=1,
                                 kcol=0.7, tau=0.8, zbuf=5.0, contacts=1, visc=visc, vparams=vp))


def contact_forces(c, pos, vel, prec="f64", lib_path=None):
    """Forces and torques (np x 3 each) of the oracle's particle_particle_forces for the given
    sphere positions and velocities (np x 3)."""
    o = oracle_py.Oracle(c, prec, lib_path=lib_path).load_case()
    k = o.arr(oracle_py.PARKIN).reshape(4, c.np, 4)
    k[:] = 0
    k[0, :, :3] = pos
    k[1, :, :3] = vel
    k[2, :, 3] = 1
    o.arr(oracle_py.PARFORCE)[:] = 0
    o.call("particle_forces")
    pf = o.arr(oracle_py.PARFORCE).reshape(2, c.np, 4).astype(np.float64).copy()
    o.close()
    return pf[0, :, :3], pf[1, :, :3]


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def check_contacts(prec="f64", lib_path=None, rtol=1e-12):
    """All contact closed forms; raises AssertionError on the first that fails."""
    c = contact_case()
    L = np.array(list(c.ip.SystemSize), dtype=np.float64)
    r = float(np.float32(8.0)) / 2
    k = float(np.float32(0.7))
    mu = (float(np.float32(0.8)) - 0.5) / 3.0
    zb = float(np.float32(5.0))
    rng = np.random.default_rng(7)

    def close(a, b, what):
        scale = max(np.abs(b).max(), 1e-30)
        assert np.abs(np.asarray(a) - np.asarray(b)).max() <= rtol * scale + 1e-300, (what, a, b)

    # harmonic overlap along a random direction, in the bulk
    for overlap in (0.3, 1.1):
        e = unit(rng.normal(size=3))
        xi = np.array([20.2, 19.7, 20.4])
        xj = xi + e * (2 * r - overlap)
        F, T = contact_forces(c, np.array([xi, xj]), np.zeros((2, 3)), prec, lib_path)
        dist = np.linalg.norm(xj - xi)
        d = 2 * r - dist
        e = (xj - xi) / dist
        close(F[0], -k * d * e, "overlap force on i")
        close(F[1], k * d * e, "overlap force on j")
        assert np.abs(T).max() == 0.0, "contact torque"
    # overlap across the periodic x face: the minimum image decides the direction
    xi = np.array([1.0, 20.0, 20.0])
    xj = np.array([L[0] - 6.5, 20.0, 20.0])
    F, _ = contact_forces(c, np.array([xi, xj]), np.zeros((2, 3)), prec, lib_path)
    d = 2 * r - 7.5
    close(F[0], np.array([k * d, 0, 0]), "periodic overlap force on i")
    close(F[1], np.array([-k * d, 0, 0]), "periodic overlap force on j")
    # squeeze: Reynolds' force between two spheres with reduced radius a1 a2 / (a1 + a2) = r / 2
    v = 1e-3
    for h in (0.05, 0.2, 0.35):
        for sgn in (1, -1):  # approaching, separating
            e = unit(rng.normal(size=3))
            xi = np.array([20.0, 20.0, 20.0])
            xj = xi + e * (2 * r + h)
            vel = np.array([0.5 * sgn * v * e, -0.5 * sgn * v * e])
            F, T = contact_forces(c, np.array([xi, xj]), vel, prec, lib_path)
            reduced = r * r / (r + r)
            mag = 6 * np.pi * mu * v * reduced ** 2 / max(h, 0.1)
            e = (xj - xi) / np.linalg.norm(xj - xi)
            close(F[0], -sgn * mag * e, "squeeze force on i (h=%g)" % h)
            close(F[1], sgn * mag * e, "squeeze force on j (h=%g)" % h)
            assert np.abs(T).max() == 0.0
    # no squeeze at h >= 0.1 r, and none for a non-Newtonian fluid
    e = unit([1.0, 2.0, -0.5])
    xi = np.array([20.0, 20.0, 20.0])
    F, _ = contact_forces(c, np.array([xi, xi + e * (2 * r + 0.45)]), np.array([0.5 * v * e, -0.5 * v * e]), prec, lib_path)
    assert np.abs(F).max() == 0.0, "squeeze beyond 0.1 r"
    cpl = contact_case(visc=2)
    F, _ = contact_forces(cpl, np.array([xi, xi + e * (2 * r + 0.2)]), np.array([0.5 * v * e, -0.5 * v * e]), prec, lib_path)
    assert np.abs(F).max() == 0.0, "squeeze in a power-law fluid"
    # z walls: a spring of stiffness k pushes a sphere back out of either buffer layer
    for z, fz in ((2.0, k * (zb - 2.0)), (L[2] - 1.5, -k * (zb - 1.5))):
        F, _ = contact_forces(c, np.array([[10.0, 10.0, z], [30.0, 30.0, 20.0]]), np.zeros((2, 3)), prec, lib_path)
        close(F[0], np.array([0, 0, fz]), "z-wall force at z=%g" % z)
        assert np.abs(F[1]).max() == 0.0


# ---- interpolation and spreading --------------------------------------------------------------
def ibm_case(ibm):
    # one work item per work group, so that the per-work-group partials are the markers' own forces
    return d3q19.Case(lines=deck(size=(24, 24, 24), walls=(0, 0, 1), np_=1, diam=6.0, markers=128, dist=3,
                                 ibm=ibm, zbuf=3.0), max_wg=1)


def check_ibm(ibm, prec="f64", lib_path=None, rtol=1e-12):
    c = ibm_case(ibm)
    N = c.N
    nn = c.nn
    rng = np.random.default_rng(11 + ibm)
    a = rng.uniform(-0.02, 0.02, 3)
    B = rng.uniform(-2e-3, 2e-3, (3, 3))
    o = oracle_py.Oracle(c, prec, lib_path=lib_path).load_case()
    # particle coordinate of a lattice node = lattice index - BufferSize (1)
    X = np.stack(np.meshgrid(np.arange(N[0]), np.arange(N[1]), np.arange(N[2]), indexing="ij"), 0).astype(np.float64) - 1.0
    X = X.transpose(0, 3, 2, 1).reshape(3, nn)  # node i = x + Nx (y + Ny z)
    u = a[:, None] + B @ X
    o.arr(oracle_py.U)[:] = u.ravel()
    o.arr(oracle_py.GPF)[:] = 0
    k = o.arr(oracle_py.PARKIN).reshape(4, 1, 4)
    k[:] = 0
    centre = np.array([11.3, 12.7, 10.4])
    k[0, 0, :3] = centre
    k[2, 0, 3] = 1
    o.call("ibm")
    ppp = c.ip.PointsPerParticle
    pts = o.arr(oracle_py.POINTS).reshape(ppp, 4)[:, :3].astype(np.float64)
    pff = o.arr(oracle_py.PARFLUID).reshape(2, ppp, 4).astype(np.float64)
    dA = float(np.float32(c.fp.PointArea))
    xm = centre + pts
    F_exp = (a[None, :] + xm @ B.T) * dA
    F = pff[0, :, :3]
    assert np.abs(F - F_exp).max() <= rtol * np.abs(F_exp).max(), ("interpolated force", np.abs(F - F_exp).max())
    T_exp = np.cross(pts, F_exp)
    assert np.abs(pff[1, :, :3] - T_exp).max() <= rtol * np.abs(F_exp).max() * 4, "marker torque"
    g = o.arr(oracle_py.GPF).reshape(3, nn).astype(np.float64)
    tot = g.sum(1)
    assert np.abs(tot + F_exp.sum(0)).max() <= rtol * 10 * np.abs(F_exp).sum(), ("spread total", tot, -F_exp.sum(0))
    mom = g @ X.T  # [component, axis]: sum over nodes of x_axis * g_component
    mom_exp = -(F_exp.T @ xm)
    assert np.abs(mom - mom_exp).max() <= rtol * 100 * np.abs(mom_exp).max(), ("spread first moment", mom, mom_exp)
    o.close()


def test_contact_law_closed_forms_f64():
    check_contacts("f64")


def test_contact_law_closed_forms_f32():
    # float32 positions near 20 lu carry ~2e-6 of rounding into a 0.2 lu gap
    check_contacts("f32", rtol=5e-5)


@pytest.mark.parametrize("ibm", [0, 1], ids=["trilinear", "peskin4"])
def test_ibm_exact_on_linear_fields(ibm):
    check_ibm(ibm)
    check_ibm(ibm, "f32", rtol=2e-6)


# Planted mistakes: (name, text in d3q19_oracle.c, replacement, which check must fail)
MUTANTS = [
    ("overlap sign", "pf[4 * pi + d] -= e[d] * fm; pf[4 * pj + d] += e[d] * fm;",
     "pf[4 * pi + d] += e[d] * fm; pf[4 * pj + d] -= e[d] * fm;", "contacts"),
    ("overlap factor 1/2", "real fm = k * overlap;", "real fm = (real)0.5 * k * overlap;", "contacts"),
    ("no minimum image", "rij[d] = (r_fabs(rij[d]) < w[d] / (real)2.0) ? rij[d] : (rij[d] - (real)sg * w[d]);",
     "rij[d] = rij[d] + (real)0 * (real)sg;", "contacts"),
    ("squeeze rstar = r", "real rstar = rp / (real)2.0;", "real rstar = rp;", "contacts"),
    ("squeeze gap not floored", "real ms = minsep < (real)0.1 ? (real)0.1 : minsep;", "real ms = minsep;", "contacts"),
    ("squeeze viscosity (tau-1/2)/6", "real eta0 = ((real)s->fp.NewtonianTau - (real)0.5) / (real)3.0;",
     "real eta0 = ((real)s->fp.NewtonianTau - (real)0.5) / (real)6.0;", "contacts"),
    ("squeeze range r", "if (minsep > 0 && minsep < (real)0.1 * rp) {", "if (minsep > 0 && minsep < rp) {", "contacts"),
    ("z-wall sign", "if (lo > 0) pf[4 * pi + 2] += k * lo;", "if (lo > 0) pf[4 * pi + 2] -= k * lo;", "contacts"),
    ("peskin stencil shifted", "int node = base[d] - 1 + k; /* particle coordinate of the node */",
     "int node = base[d] - 2 + k; /* particle coordinate of the node */", "peskin4"),
    ("peskin offset sign", "wd[d][k] = peskin4(rpp[d] - (real)node);", "wd[d][k] = peskin4((real)node - rpp[d] + (real)0.25);",
     "peskin4"),
    ("peskin outer branch", "if (r < 2) return ((real)5 - (real)2 * r - r_sqrt((real)-7 + (real)12 * r - (real)4 * r * r)) / (real)8;",
     "if (r < 2) return ((real)5 - (real)2 * r - r_sqrt((real)-7 + (real)12 * r - (real)4 * r * r)) / (real)9;", "peskin4"),
    ("trilinear weight", "w[d] = (real)1.0 - (rpp[d] - (real)fl[d]);", "w[d] = (rpp[d] - (real)fl[d]);", "trilinear"),
    ("marker velocity omits omega x r", "cross3(av, r0, wr);", "cross3(av, r0, wr); wr[0] = wr[1] = wr[2] = 0;", None),
    ("spread sign", "gpf[nodes[n]] -= wts[n] * F[0];", "gpf[nodes[n]] += wts[n] * F[0];", "trilinear"),
]


def build_mutant(tmp, name, old, new):
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, "mutation site not found once: " + name
    path = os.path.join(tmp, "mut.c")
    with open(path, "w") as f:
        f.write(src.replace(old, new))
    so = os.path.join(tmp, "liboracle_mut_%d.so" % abs(hash(name)))
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off", "-I" + os.path.dirname(ORACLE_SRC),
                    "-DORC_REAL=double", "-DORC_SFX=_f64", "-o", so, path, "-lm"], check=True, capture_output=True)
    return so


@pytest.mark.parametrize("mutant", [m for m in MUTANTS if m[3]], ids=[m[0] for m in MUTANTS if m[3]])
def test_planted_mistakes_fail_the_pins(mutant, tmp_path):
    name, old, new, which = mutant
    so = build_mutant(str(tmp_path), name, old, new)
    with pytest.raises(AssertionError):
        if which == "contacts":
            check_contacts("f64", lib_path=so)
        else:
            check_ibm(0 if which == "trilinear" else 1, "f64", lib_path=so)


def test_marker_velocity_includes_rotation():
    """Marker velocity v + omega x r0 (GPU_program.cl:200): a sphere spinning in fluid at rest gets
    marker forces -(omega x r0) dA; checked with the oracle's own planted-mistake copy failing."""
    def forces(lib_path=None):
        c = ibm_case(1)
        o = oracle_py.Oracle(c, "f64", lib_path=lib_path).load_case()
        o.arr(oracle_py.U)[:] = 0
        k = o.arr(oracle_py.PARKIN).reshape(4, 1, 4)
        k[:] = 0
        k[0, 0, :3] = [11.3, 12.7, 10.4]
        k[2, 0, 3] = 1
        w = np.array([1e-3, -2e-3, 5e-4])
        k[3, 0, :3] = w
        o.call("ibm")
        ppp = c.ip.PointsPerParticle
        pts = o.arr(oracle_py.POINTS).reshape(ppp, 4)[:, :3].astype(np.float64)
        F = o.arr(oracle_py.PARFLUID).reshape(2, ppp, 4)[0, :, :3].astype(np.float64).copy()
        exp = -np.cross(w, pts) * float(np.float32(c.fp.PointArea))
        o.close()
        return np.abs(F - exp).max() / np.abs(exp).max()
    assert forces() < 1e-12
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        m = [x for x in MUTANTS if x[0] == "marker velocity omits omega x r"][0]
        assert forces(build_mutant(tmp, m[0], m[1], m[2])) > 0.5
