// particles.hip -- rigid-sphere kernels, FP64, one thread per sphere.
//
// dynamics  : particle_dynamics, CPU_program.cl:187-277 -- explicit Newton-Euler with dt = 1,
//             positions wrapped with fmod over SystemSize, the quaternion slot overwritten with
//             dq/dt / 2 as the reference does (nothing reads it), contact force/torque reset.
// contacts  : particle_particle_forces + compute_squeeze_force, CPU_program.cl:76-184, 319-343.
//             Each thread sums the forces on its own sphere over all partners, so there is no
//             shared read-modify-write (the reference races on parForce[pj], CPU_program.cl:154,167).
#include "internal.h"

namespace d3q19 {

struct PartArgs {
    int np;
    double sys[3];
    double gconst[3];
    double mass, inertia;
    double diam;
    double kcol;
    int force_model;
    int visc_model;
    double tau0;
    double zbuf;
    double* pos;
    double* vel;
    double* quat;
    double* omega;
    const double* ffluid;
    const double* tfluid;
    double* fcont;
    double* tcont;
};

__global__ void dynamics_kernel(PartArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.np) return;
    double acc[3], aacc[3];
    for (int k = 0; k < 3; k++) {
        acc[k] = (a.ffluid[3 * p + k] + a.fcont[3 * p + k]) / a.mass + a.gconst[k];
        aacc[k] = (a.tfluid[3 * p + k] + a.tcont[3 * p + k]) / a.inertia;
    }
    for (int k = 0; k < 3; k++) {
        const double rt = a.pos[3 * p + k] + a.vel[3 * p + k] + 0.5 * acc[k];
        a.pos[3 * p + k] = fmod(rt + a.sys[k], a.sys[k]);
        a.vel[3 * p + k] += acc[k];
    }
    double q[4], av[3];
    for (int k = 0; k < 4; k++) q[k] = a.quat[4 * p + k];
    for (int k = 0; k < 3; k++) av[k] = a.omega[3 * p + k];
    const double dqw = -(av[0] * q[0] + av[1] * q[1] + av[2] * q[2]);
    const double cr[3] = {av[1] * q[2] - av[2] * q[1], av[2] * q[0] - av[0] * q[2], av[0] * q[1] - av[1] * q[0]};
    for (int k = 0; k < 3; k++) a.quat[4 * p + k] = 0.5 * (q[3] * av[k] + cr[k]);
    a.quat[4 * p + 3] = 0.5 * dqw;
    for (int k = 0; k < 3; k++) {
        a.omega[3 * p + k] = av[k] + aacc[k];
        a.fcont[3 * p + k] = 0.0;
        a.tcont[3 * p + k] = 0.0;
    }
}

__global__ void contacts_kernel(PartArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.np) return;
    const double rp = 0.5 * a.diam;
    double F[3] = {0, 0, 0};
    const double xi[3] = {a.pos[3 * i], a.pos[3 * i + 1], a.pos[3 * i + 2]};
    const double vi[3] = {a.vel[3 * i], a.vel[3 * i + 1], a.vel[3 * i + 2]};
    for (int j = 0; j < a.np; j++) {
        if (j == i) continue;
        // pair (pi, pj) = (max, min) as in the reference; r = x_pj - x_pi
        const int pi = i > j ? i : j, pj = i > j ? j : i;
        double r[3], v[3];
        for (int d = 0; d < 3; d++) {
            double rd = a.pos[3 * pj + d] - a.pos[3 * pi + d];
            const double sg = rd < 0 ? -1.0 : (rd > 0 ? 1.0 : 0.0);
            rd = (fabs(rd) < a.sys[d] / 2.0) ? rd : (rd - sg * a.sys[d]);
            r[d] = rd;
            v[d] = a.vel[3 * pj + d] - a.vel[3 * pi + d];
        }
        const double rsep = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double minsep = rsep - 2.0 * rp;
        const double e[3] = {r[0] / rsep, r[1] / rsep, r[2] / rsep};
        const double vrel = -(e[0] * v[0] + e[1] * v[1] + e[2] * v[2]);
        double fmag = 0.0;
        if (-minsep > 0 && a.force_model == 1) fmag += a.kcol * (-minsep);
        if (minsep > 0 && minsep < 0.1 * rp && a.visc_model == D3Q19_VISC_NEWTONIAN) {
            const double rstar = rp / 2.0;
            const double ms = minsep < 0.1 ? 0.1 : minsep;
            const double eta0 = (a.tau0 - 0.5) / 3.0;
            fmag += 6.0 * 3.14159265358979323846 * eta0 * rstar * rstar * vrel / ms;
        }
        // the reference subtracts e*f from pi and adds it to pj
        const double sgn = (i == pi) ? -1.0 : 1.0;
        for (int d = 0; d < 3; d++) F[d] += sgn * e[d] * fmag;
    }
    (void)vi;
    const double lo = a.zbuf - xi[2];
    if (lo > 0) F[2] += a.kcol * lo;
    const double hi = xi[2] - (a.sys[2] - a.zbuf);
    if (hi > 0) F[2] -= a.kcol * hi;
    for (int d = 0; d < 3; d++) a.fcont[3 * i + d] += F[d];
}

static PartArgs part_args(Sim* s) {
    PartArgs a;
    a.np = s->np;
    for (int d = 0; d < 3; d++) {
        a.sys[d] = (double)s->ip.SystemSize[d];
        a.gconst[d] = (double)s->fp.ConstBodyForce[d];
    }
    a.mass = (double)s->fp.ParticleMass;
    a.inertia = (double)s->fp.ParticleMomInertia;
    a.diam = (double)s->fp.ParticleDiam;
    a.kcol = (double)s->fp.ParForceParams[0];
    a.force_model = s->ip.ParForceModel;
    a.visc_model = s->ip.ViscosityModel;
    a.tau0 = (double)s->fp.NewtonianTau;
    a.zbuf = (double)s->fp.ParticleZBuffer;
    a.pos = s->pos;
    a.vel = s->vel;
    a.quat = s->quat;
    a.omega = s->omega;
    a.ffluid = s->ffluid;
    a.tfluid = s->tfluid;
    a.fcont = s->fcont;
    a.tcont = s->tcont;
    return a;
}

int launch_dynamics(Sim* s, hipStream_t st) {
    if (s->np == 0) return 0;
    PartArgs a = part_args(s);
    hipLaunchKernelGGL(dynamics_kernel, dim3((s->np + 63) / 64), dim3(64), 0, st, a);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_contacts(Sim* s, hipStream_t st) {
    if (s->np == 0) return 0;
    PartArgs a = part_args(s);
    hipLaunchKernelGGL(contacts_kernel, dim3((s->np + 63) / 64), dim3(64), 0, st, a);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace d3q19
