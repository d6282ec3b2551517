"""GPU: the FP64 shear-cell observables (wall stress, apparent viscosity, drag coefficient).

Pins:
- Closed form.  In steady plane Couette flow between wet-node walls at interior z = 0 and
  z = n_z - 1 moving at -U and +U, the shear rate is uniform, 2U/(n_z - 1).  So the wall stress
  of a Newtonian fluid is nu * 2U/(n_z - 1) and its apparent viscosity is nu (to 1e-6 with fp64
  populations).  For a power-law
  fluid the local viscosity is K * gdot^(n - 1) everywhere, so the relative viscosity is 1.
- Restatement.  The kernel's plane sums equal a numpy evaluation of the stress formula on the
  same population field (read back from the GPU in float64), and the drag coefficient equals
  the numpy mean of |F_fluid| over the spheres, normalised by rho/2 U^2 pi a^2.
"""
import numpy as np
import pytest

import d3q19
from cases import case, perturb

pytestmark = pytest.mark.gpu


def couette(coll, visc=1, vparams=(0, 0, 0, 0), local=0, tau=0.8, uw=0.01, nz=20, storage=0):
    return case(size=(4, 4, nz), coll=coll, tau=tau, uw=uw, visc=visc, vparams=vparams, local=local,
                clamp=(0.505, 0.0) if local else (0.0, 0.0), storage=storage)


@pytest.mark.parametrize("storage", [3, 0])
@pytest.mark.parametrize("coll", [d3q19.COLL_MRT, d3q19.COLL_BGK])
def test_newtonian_couette_apparent_viscosity_is_nu(coll, storage):
    # fp64 storage: the steady profile is linear and the apparent viscosity is nu to ~1e-7 (nu
    # itself comes from the float32 NewtonianTau).  fp32 storage: float32 rounding of the
    # populations shifts the steady state; the stress moves by ~1.2e-4 relative and the profile
    # bends by ~3e-5 of U.
    tol = 1e-6 if storage == 3 else 2e-4
    c = couette(coll, storage=storage)
    s = d3q19.Sim(c)
    s.run(1, 40000, d3q19.STEP_WRITE_U)
    o = s.observables()
    um = s.plane_mean_u()[0]  # reference z index; interior planes 1 .. Nz-2
    s.close()
    nu = (float(np.float32(0.8)) - 0.5) / 3.0
    U = float(np.float32(0.01))
    prof = um[1:-1]
    z = np.arange(len(prof), dtype=np.float64)
    slope, icpt = np.polyfit(z, prof, 1)
    lin = np.abs(prof - (slope * z + icpt)).max()
    print("coll %d: wall stress %.9e (lo %.9e hi %.9e) nu_app %.9e nu %.9e  slope %.9e nominal %.9e  lin %.2e" % (
        coll, o.wall_stress, o.wall_stress_lower, o.wall_stress_upper, o.apparent_viscosity, nu, slope, 2 * U / 20, lin),
        flush=True)
    assert o.wall_axis == 2 and o.flow_axis == 0
    assert abs(o.shear_rate - 2 * U / 20.0) < 1e-15
    assert lin < (1e-6 if storage == 3 else 1e-4) * U
    assert abs(o.wall_stress - nu * slope) < tol * nu * slope
    assert abs(o.wall_stress_lower - o.wall_stress_upper) < tol * nu * slope
    assert abs(o.apparent_viscosity / nu - 1.0) < tol
    assert abs(o.fluid_viscosity - nu) < 1e-12


@pytest.mark.parametrize("coll", [d3q19.COLL_MRT, d3q19.COLL_BGK])
def test_power_law_couette_relative_viscosity_is_one(coll):
    n, U, H = 0.7, 0.01, 20.0
    nu0 = (0.9 - 0.5) / 3.0
    K = nu0 * (2 * U / H) ** (1 - n)
    c = couette(coll, visc=2, vparams=(K, n, 0, 0), local=1, tau=0.9, uw=U)
    s = d3q19.Sim(c)
    s.run(1, 40000)
    o = s.observables()
    s.close()
    print("coll %d: nu_app %.9e nu_fluid %.9e rel %.9e" % (coll, o.apparent_viscosity, o.fluid_viscosity,
                                                           o.relative_viscosity), flush=True)
    assert abs(o.fluid_viscosity - float(np.float32(K)) * (2 * U / H) ** (n - 1)) < 1e-6 * o.fluid_viscosity
    assert abs(o.relative_viscosity - 1.0) < 1e-3


def numpy_wall_stress(f, tau_field, c, tau0):
    """sigma_xz summed over the planes next to the walls, from a reference-layout f (float64)."""
    Nx, Ny, Nz = c.N
    cv = np.array(c.ip.BasisVel, dtype=np.float64)
    F = f.reshape(19, Nz, Ny, Nx)
    g = np.array(c.fp.ConstBodyForce, dtype=np.float64)
    out = []
    for zl in (2, Nz - 3):  # interior planes 1 and n_z - 2
        fp = F[:, zl, 1:-1, 1:-1].reshape(19, -1)
        rho = fp.sum(0)
        u = (cv.T @ fp + 0.5 * g[:, None]) / rho
        pi = (cv[:, 0] * cv[:, 2]) @ fp
        tau = tau0 if tau_field is None else tau_field.reshape(Nz, Ny, Nx)[zl, 1:-1, 1:-1].ravel().astype(np.float64)
        sig = -(1 - 0.5 / tau) * (pi - rho * u[0] * u[2] + 0.5 * (u[0] * g[2] + u[2] * g[0]))
        out.append(sig.mean())
    return out


@pytest.mark.parametrize("visc", [1, 2])
def test_wall_stress_matches_numpy_restatement_fp64(visc):
    vp = (0.02, 0.7, 0, 0) if visc == 2 else (0, 0, 0, 0)
    c = perturb(case(size=(10, 8, 12), coll=1, visc=visc, vparams=vp, local=1 if visc == 2 else 0, clamp=(0.51, 5.0),
                     tau=0.9, uw=0.03, storage=3, body=(1e-5, 0.0, 2e-6)), seed=21)
    s = d3q19.Sim(c)
    s.upload_case()
    s.run(1, 7)
    s.sync()
    f = s.read_f64()
    tau = s.read_tau() if visc == 2 else None
    o = s.observables()
    s.close()
    lo, hi = numpy_wall_stress(f, tau, c, float(np.float32(0.9)))
    print("gpu lo %.15e hi %.15e | numpy lo %.15e hi %.15e" % (o.wall_stress_lower, o.wall_stress_upper, lo, hi), flush=True)
    # tau is read back as float32; the stress scales with (1 - 1/(2 tau))
    tol = 1e-12 if visc == 1 else 2e-6
    assert abs(o.wall_stress_lower - lo) <= tol * max(abs(lo), 1e-12)
    assert abs(o.wall_stress_upper - hi) <= tol * max(abs(hi), 1e-12)


def test_drag_coefficient_matches_numpy():
    c = case(size=(24, 24, 24), coll=1, ibm=1, np_=3, diam=6.0, markers=128, dist=4, seed=5, tau=0.8, uw=0.02, storage=3)
    s = d3q19.Sim(c)
    s.write_particles(c.parKin)
    s.run(1, 40)
    o = s.observables()
    pos, vel, om, ff, tf = s.read_particles_f64()
    s.close()
    F = ff.reshape(-1, 3)[: c.np]
    mean = np.linalg.norm(F, axis=1).mean()
    U, a = float(np.float32(0.02)), 3.0  # wall speeds are float32 parameters
    cd = mean / (0.5 * U * U * np.pi * a * a)
    print("drag gpu %.15e numpy %.15e (mean |F| %.6e)" % (o.drag_coefficient, cd, mean), flush=True)
    assert mean > 0
    assert abs(o.mean_fluid_force - mean) <= 1e-12 * mean
    assert abs(o.drag_coefficient - cd) <= 1e-12 * cd
