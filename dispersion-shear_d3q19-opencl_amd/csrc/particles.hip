// particles.hip -- rigid-sphere kernels, FP64, one thread block per sphere.
//
// sphere_pre  (start of a step): particle_dynamics (CPU_program.cl:187-277) -- explicit
//             Newton-Euler with dt = 1, positions wrapped with fmod over SystemSize, the quaternion
//             slot overwritten with dq/dt / 2 as the reference does (nothing reads it), contact
//             force/torque consumed and reset -- then the brick marking of the new position.
// sphere_post (end of a step): the sphere's force and torque summed over its markers in a fixed
//             order (the reference's work-group partials + per-particle sum, GPU_program.cl:229-255,
//             CPU_program.cl:218-223), then particle_particle_forces + compute_squeeze_force
//             (CPU_program.cl:76-184, 319-343) over all partners, summed in a fixed order by the
//             block (the reference races on parForce[pj], CPU_program.cl:154,167).
//
// Contacts depend only on positions and velocities, which do not change between the collide and
// the end of a step, so computing them after the spreading pass (as the reference does) or before
// it gives the same forces for the next step.
//
// Brick epochs: every kernel of a step uses e = epochs[0] + 1; the spreading kernel commits it
// (epochs[0] = epochs[1] = e), so no kernel writes an epoch that another block of it reads.
#include "internal.h"

namespace d3q19 {

// ---- brick marking ------------------------------------------------------------------------------------
// Stamps epoch e on every brick that a marker stencil of a sphere centred at c3 can touch.  Marker
// floor coordinates along an axis span [floor(c-R), floor(c+R)]; each is wrapped like the markers
// (fmod over SystemSize, GPU_program.cl:144) and its stencil covers nodes -1..+2 around it (taken
// as -2..+3); periodic axes wrap node indices, wall axes clip them.  The per-axis brick sets are
// built in LDS, then the block stamps their product.
__device__ void mark_sphere(const double* c3, const Geom& g, double radius, uint32_t* mk, uint32_t e) {
    __shared__ uint8_t fl[3][1024];
    __shared__ int lists[3][64];
    __shared__ int counts[3];
    const int nd[3] = {g.nx, g.ny, g.nz};
    const int bn[3] = {g.bnx, g.bny, g.bnz};
    for (int d = 0; d < 3; d++)
        for (int i = threadIdx.x; i < bn[d]; i += blockDim.x) fl[d][i] = 0;
    __syncthreads();
    bool ok = true;
    int lo[3], span[3];
    for (int d = 0; d < 3; d++)
        if (!(c3[d] > -1e7 && c3[d] < 1e7)) ok = false;  // non-finite sphere: nothing to mark
    for (int d = 0; d < 3; d++) {
        lo[d] = ok ? (int)floor(c3[d] - radius) : 0;
        span[d] = ok ? (int)floor(c3[d] + radius) - lo[d] + 1 : 0;
    }
    for (int d = 0; d < 3; d++) {
        const int L = g.sys[d];
        for (int i = threadIdx.x; i < span[d] * 6; i += blockDim.x) {
            const int v = lo[d] + i / 6;
            int node = ((v % L) + L) % L + (i % 6) - 2;
            if (g.per[d]) node = ((node % nd[d]) + nd[d]) % nd[d];
            else if (node < 0 || node >= nd[d]) continue;
            fl[d][node >> 2] = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int d = threadIdx.x;
        int cnt = 0;
        for (int b = 0; b < bn[d] && cnt < 64; b++)
            if (fl[d][b]) lists[d][cnt++] = b;
        counts[d] = cnt;
    }
    __syncthreads();
    const int total = counts[0] * counts[1] * counts[2];
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int a0 = i % counts[0], a1 = (i / counts[0]) % counts[1], a2 = i / (counts[0] * counts[1]);
        mk[lists[0][a0] + g.bnx * (lists[1][a1] + g.bny * lists[2][a2])] = e;
    }
}

__device__ __forceinline__ void dynamics_one(const PartArgs& a, int p) {
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

// dynamics of sphere p (thread 0), then marking of its bricks for this step's epoch; whole block
__device__ __forceinline__ void sphere_pre(const PartArgs& a, int p, int do_dyn, int do_mark) {
    __shared__ double c3[3];
    if (threadIdx.x == 0) {
        if (do_dyn) dynamics_one(a, p);
        for (int k = 0; k < 3; k++) c3[k] = a.pos[3 * p + k];
    }
    __syncthreads();
    if (do_mark) {
        const uint32_t e = a.epochs[0] + 1u;
        mark_sphere(c3, a.g, a.radius, (e & 1u) ? a.mask1 : a.mask0, e);
    }
}

__global__ void __launch_bounds__(256) sphere_pre_kernel(PartArgs a, int do_dyn, int do_mark) {
    sphere_pre(a, blockIdx.x, do_dyn, do_mark);
}

// fixed-order sums over the block of K values per thread (wave butterfly, then the waves in
// order): deterministic
template <int K>
__device__ __forceinline__ void block_sums(double* v, double (*sh)[K]) {
    for (int k = 0; k < K; k++)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < K; k++) sh[w][k] = v[k];
    __syncthreads();
    for (int k = 0; k < K; k++) {
        double t = 0;
        for (int i = 0; i < nw; i++) t += sh[i][k];
        v[k] = t;
    }
    __syncthreads();
}

// FP64 sums of sphere p's marker forces/torques, then its contact forces; whole block
__device__ __forceinline__ void sphere_post(const PartArgs& a, int p, int do_reduce, int do_contacts) {
    __shared__ double sh6[4][6];
    __shared__ double sh3[4][3];
    if (do_reduce) {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int m = threadIdx.x; m < a.ppp; m += blockDim.x) {
            const double* mf = a.mforce + 6 * ((long long)p * a.ppp + m);
            for (int k = 0; k < 6; k++) acc[k] += mf[k];
        }
        block_sums<6>(acc, sh6);
        if (threadIdx.x == 0)
            for (int k = 0; k < 3; k++) {
                a.ffluid[3 * p + k] = acc[k];
                a.tfluid[3 * p + k] = acc[3 + k];
            }
    }
    if (!do_contacts) return;
    const double rp = 0.5 * a.diam;
    const double xi[3] = {a.pos[3 * p], a.pos[3 * p + 1], a.pos[3 * p + 2]};
    const double vi[3] = {a.vel[3 * p], a.vel[3 * p + 1], a.vel[3 * p + 2]};
    double F[3] = {0, 0, 0};
    for (int j = threadIdx.x; j < a.np; j += blockDim.x) {
        if (j == p) continue;
        // pair (pi, pj) = (max, min) as in the reference: r = x_pj - x_pi, v = v_pj - v_pi
        const bool ipi = p > j;
        double r[3], v[3];
        for (int d = 0; d < 3; d++) {
            const double xj = a.pos[3 * j + d], vj = a.vel[3 * j + d];
            double rd = ipi ? xj - xi[d] : xi[d] - xj;
            const double sg = rd < 0 ? -1.0 : (rd > 0 ? 1.0 : 0.0);
            rd = (fabs(rd) < a.sys[d] / 2.0) ? rd : (rd - sg * a.sys[d]);
            r[d] = rd;
            v[d] = ipi ? vj - vi[d] : vi[d] - vj;
        }
        const double rsep = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double minsep = rsep - 2.0 * rp;
        if (minsep >= 0.1 * rp) continue;  // neither overlap nor squeeze range
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
        const double sgn = ipi ? -1.0 : 1.0;
        for (int d = 0; d < 3; d++) F[d] += sgn * e[d] * fmag;
    }
    block_sums<3>(F, sh3);
    if (threadIdx.x == 0) {
        // z-wall repulsion (CPU_program.cl:175-182)
        const double lo = a.zbuf - xi[2];
        if (lo > 0) F[2] += a.kcol * lo;
        const double hi = xi[2] - (a.sys[2] - a.zbuf);
        if (hi > 0) F[2] -= a.kcol * hi;
        for (int d = 0; d < 3; d++) a.fcont[3 * p + d] += F[d];
    }
}

__global__ void __launch_bounds__(256) sphere_post_kernel(PartArgs a, int do_reduce, int do_contacts) {
    sphere_post(a, blockIdx.x, do_reduce, do_contacts);
}

// Batched: a persistent grid of blocks walks the sphere list of all cells.
__global__ void __launch_bounds__(256) sphere_pre_batch_kernel(BatchTables t) {
    const int nsph = t.cnt->nsph;
    for (int k = blockIdx.x; k < nsph; k += gridDim.x) {
        const int4 d = t.sph[k];
        const CellDesc& cd = t.cells[d.x];
        sphere_pre(cd.p, d.y, 1, 1);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) sphere_post_batch_kernel(BatchTables t) {
    const int nsph = t.cnt->nsph;
    for (int k = blockIdx.x; k < nsph; k += gridDim.x) {
        const int4 d = t.sph[k];
        const CellDesc& cd = t.cells[d.x];
        sphere_post(cd.p, d.y, 1, 1);
        __syncthreads();
    }
}

PartArgs make_part_args(Sim* s) {
    PartArgs a;
    a.np = s->np;
    a.ppp = s->ppp;
    a.g = s->g;
    for (int d = 0; d < 3; d++) {
        a.sys[d] = (double)s->ip.SystemSize[d];
        a.gconst[d] = (double)s->fp.ConstBodyForce[d];
    }
    a.mass = (double)s->fp.ParticleMass;
    a.inertia = (double)s->fp.ParticleMomInertia;
    a.diam = (double)s->fp.ParticleDiam;
    // marker offsets are scaled unit vectors (|r0|^2 within 1% of 1, D3Q19-OpenCL.c:597-601)
    a.radius = 0.5 * (double)s->fp.ParticleDiam * 1.01 + 1e-3;
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
    a.mforce = s->mforce;
    a.mask0 = s->mask[0];
    a.mask1 = s->mask[1];
    a.epochs = s->epochs;
    return a;
}

int launch_sphere_pre(Sim* s, int do_dyn, int do_mark, hipStream_t st) {
    if (s->np == 0) return 0;
    hipLaunchKernelGGL(sphere_pre_kernel, dim3(s->np), dim3(256), 0, st, make_part_args(s), do_dyn, do_mark);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_sphere_post(Sim* s, int do_reduce, int do_contacts, hipStream_t st) {
    if (s->np == 0) return 0;
    hipLaunchKernelGGL(sphere_post_kernel, dim3(s->np), dim3(256), 0, st, make_part_args(s), do_reduce, do_contacts);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_sphere_pre_batch(const BatchLaunch& L, hipStream_t st) {
    hipLaunchKernelGGL(sphere_pre_batch_kernel, dim3(L.grid_sphere), dim3(256), 0, st, L.t);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_sphere_post_batch(const BatchLaunch& L, hipStream_t st) {
    hipLaunchKernelGGL(sphere_post_batch_kernel, dim3(L.grid_sphere), dim3(256), 0, st, L.t);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace d3q19
