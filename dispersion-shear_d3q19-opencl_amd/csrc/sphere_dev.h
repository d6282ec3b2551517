// sphere_dev.h -- device code of the rigid spheres (FP64), shared by the sphere kernels
// (particles.hip) and the fused step kernels (lbm.hip), which run the spheres' turn from one
// step to the next in blocks of the collide launch.
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
// sphere_turn (fused step): sphere_post of step t, then the dynamics of step t+1 written to the
//             other position/velocity buffer (partners' contacts still read x(t), v(t)), then the
//             marks of step t+2 for x(t+1) dilated by one lattice unit.
//
// Contacts depend only on positions and velocities, which do not change between the collide and
// the end of a step, so computing them after the spreading pass (as the reference does) or before
// it gives the same forces for the next step.
#pragma once
#include "internal.h"

namespace d3q19 {

// brick stamps of epoch e live in array e mod 3
template <class P>
__host__ __device__ inline P mask_of(P m0, P m1, P m2, uint32_t e) {
    const uint32_t r = e % 3u;
    return r == 0u ? m0 : (r == 1u ? m1 : m2);
}

// ---- brick marking ------------------------------------------------------------------------------------
// Stamps epoch e on every brick that a marker stencil of a sphere centred at c3 can touch.  Marker
// floor coordinates along an axis span [floor(c-R), floor(c+R)]; each is wrapped like the markers
// (fmod over SystemSize, GPU_program.cl:144) and its stencil covers nodes -1..+2 around it, taken
// as -pad..+pad+1 (pad 2: exact marks; pad 3 and more: marks of x(t) that also cover the stencils at
// x(t+1) while the sphere moves less than pad - 2 lattice units per step and axis).  Periodic axes
// wrap node indices, wall axes clip them.  The per-axis brick sets are built in LDS, then the block
// stamps their product.
template <class MP>
__device__ void mark_sphere(const double* c3, const Geom& g, double radius, double rin, double rout, MP mk, uint32_t e, int pad) {
    __shared__ uint8_t fl[3][1024];
    __shared__ int lists[3][MARK_BRICKS];  // sim_init refuses spheres that need more bricks per axis
    __shared__ float dn[3][MARK_BRICKS], df[3][MARK_BRICKS];
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
        const int w = 2 * pad + 2;
        for (int i = threadIdx.x; i < span[d] * w; i += blockDim.x) {
            const int v = lo[d] + i / w;
            int node = ((v % L) + L) % L + (i % w) - pad;
            if (g.per[d]) node = ((node % nd[d]) + nd[d]) % nd[d];
            else if (node < 0 || node >= nd[d]) continue;
            fl[d][node >> 2] = 1;
        }
    }
    __syncthreads();
    // Shell test: a brick is needed only if its nodes, widened by pad on each axis, come as close
    // to the centre as the longest marker offset and as far as the shortest one (a stencil node
    // lies within 2 nodes of its marker on each axis, plus 1 for a step's motion when pad = 3).
    // Per axis, the distance from the centre to the widened brick (nearest periodic image) is
    // kept as its nearest (dn) and farthest (df) extent.  On a periodic axis shorter than the
    // marked span plus margins, images are ambiguous and every brick of the box is kept.
    const bool shell = ok && rout > 0.0;
    if (threadIdx.x < 3) {
        const int d = threadIdx.x;
        int cnt = 0;
        for (int b = 0; b < bn[d] && cnt < MARK_BRICKS; b++)
            if (fl[d][b]) {
                double lo_r = 4.0 * b - pad - c3[d], hi_r = 4.0 * b + 3 + pad - c3[d];
                if (g.per[d]) {
                    const double k = rint(0.5 * (lo_r + hi_r) / nd[d]);
                    lo_r -= k * nd[d];
                    hi_r -= k * nd[d];
                }
                const bool amb = g.per[d] && nd[d] < 2.0 * (rout + pad) + 16.0;
                dn[d][cnt] = amb ? 0.0f : (float)(lo_r > 0 ? lo_r : (hi_r < 0 ? -hi_r : 0.0));
                df[d][cnt] = amb ? 3.0e38f : (float)fmax(fabs(lo_r), fabs(hi_r));
                lists[d][cnt++] = b;
            }
        counts[d] = cnt;
    }
    __syncthreads();
    const int total = counts[0] * counts[1] * counts[2];
    const float rin2 = (float)(rin * rin) * 0.999f, rout2 = (float)(rout * rout) * 1.001f;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int a0 = i % counts[0], a1 = (i / counts[0]) % counts[1], a2 = i / (counts[0] * counts[1]);
        if (shell) {
            const float dmin2 = dn[0][a0] * dn[0][a0] + dn[1][a1] * dn[1][a1] + dn[2][a2] * dn[2][a2];
            const float dmax2 = fminf(df[0][a0] * df[0][a0] + df[1][a1] * df[1][a1] + df[2][a2] * df[2][a2], 3.0e38f);
            if (dmin2 > rout2 || dmax2 < rin2) continue;
        }
        mk[lists[0][a0] + g.bnx * (lists[1][a1] + g.bny * lists[2][a2])] = e;
    }
}

// ---- contact zones ----------------------------------------------------------------------------------
// Zone of sphere p from its current position: (int)(x / ZoneWidth) per axis in float, as the
// reference computes it on its float positions (D3Q19-OpenCL.c:401-404, CPU_program.cl:299-302);
// an id outside the zone grid (out of bounds in the reference) is -1 and meets no sphere.
__device__ __forceinline__ int zone_of(const PartArgs& a, int p) {
    int id[3];
    for (int d = 0; d < 3; d++) id[d] = (int)((float)a.pos[3 * p + d] / a.zw[d]);
    const long long z = id[0] + (long long)a.nz[0] * (id[1] + (long long)a.nz[1] * id[2]);
    return (z >= 0 && z < (long long)a.nz[0] * a.nz[1] * a.nz[2]) ? (int)z : -1;
}

// Number of times zone zj appears in the neighbour list of zone zi as the reference builds it
// (D3Q19-OpenCL.c:419-462): offsets -1..1 per axis, none past a velocity-boundary face, indices
// wrapped (s < 0 -> NumZones - 1, else s % NumZones), the id formed with NumZones[2] where
// NumZones[1] is meant (line 450).  particle_particle_forces visits pair (pi, pj), pi > pj, once
// per appearance of pj's zone in pi's list (CPU_program.cl:106-121).
__device__ __forceinline__ int zone_multiplicity(const PartArgs& a, int zi, int zj) {
    if (zi < 0 || zj < 0) return 0;
    const int n0 = a.nz[0], n1 = a.nz[1], n2 = a.nz[2];
    const int c[3] = {zi % n0, (zi / n0) % n1, zi / (n0 * n1)};
    const int nd[3] = {n0, n1, n2};
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = (!a.g.per[d] && c[d] == 0) ? 0 : -1;
        hi[d] = (!a.g.per[d] && c[d] == nd[d] - 1) ? 0 : 1;
    }
    int cnt = 0;
    for (int n = lo[2]; n <= hi[2]; n++)
        for (int m = lo[1]; m <= hi[1]; m++)
            for (int l = lo[0]; l <= hi[0]; l++) {
                const int ls = c[0] + l, ms = c[1] + m, ns = c[2] + n;
                const int lw = ls < 0 ? n0 - 1 : ls % n0;
                const int mw = ms < 0 ? n1 - 1 : ms % n1;
                const int nw = ns < 0 ? n2 - 1 : ns % n2;
                cnt += (lw + n0 * (mw + n2 * nw) == zj) ? 1 : 0;
            }
    return cnt;
}

// generation of the zone lists that the contacts of step t read: the lists of the rebuild after
// step rebuild * ((t - 1) / rebuild), or the initial lists
__device__ __forceinline__ int zone_generation(const PartArgs& a, long long t) { return t >= 1 ? (int)(((t - 1) / a.rebuild) & 1) : 0; }

// reads pos/vel, writes pos_out/vel_out (the same arrays except in the fused turn)
__device__ __forceinline__ void dynamics_one(const PartArgs& a, int p) {
    double acc[3], aacc[3];
    for (int k = 0; k < 3; k++) {
        acc[k] = (a.ffluid[3 * p + k] + a.fcont[3 * p + k]) / a.mass + a.gconst[k];
        aacc[k] = (a.tfluid[3 * p + k] + a.tcont[3 * p + k]) / a.inertia;
    }
    for (int k = 0; k < 3; k++) {
        const double x0 = a.pos[3 * p + k];
        const double rt = x0 + a.vel[3 * p + k] + 0.5 * acc[k];
        const double x1 = fmod(rt + a.sys[k], a.sys[k]);
        a.pos_out[3 * p + k] = x1;
        a.vel_out[3 * p + k] = a.vel[3 * p + k] + acc[k];
        // a displacement that the dilated marks cannot cover is reported, never run silently
        double dx = fabs(x1 - x0);
        dx = fmin(dx, a.sys[k] - dx);
        if (!(dx < 1.0) && a.flag) __hip_atomic_fetch_or(a.flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        mark_sphere(c3, a.g, a.radius, a.rin, a.rout, mask_of(a.mask0, a.mask1, a.mask2, e), e, 2);
    }
}

// Fixed-order sums over 256 virtual threads, whatever the block size (a multiple of 64, at most
// 256): virtual thread vt (threadIdx.x, threadIdx.x + blockDim.x, ...) accumulates items vt,
// vt + 256, ... in order, each virtual wave sums by butterfly (wave_store), then every thread adds
// the four waves in order (wave_total).  The bits do not depend on the launch shape, so the fused
// step (collide-tile blocks) and the separate sphere kernels (256 threads) agree exactly.
constexpr int VTHREADS = 256;

template <int K>
__device__ __forceinline__ void wave_store(double* v, double (*sh)[K], int vt) {
    for (int k = 0; k < K; k++)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    if ((vt & 63) == 0)
        for (int k = 0; k < K; k++) sh[vt >> 6][k] = v[k];
}

template <int K>
__device__ __forceinline__ void wave_total(double* v, double (*sh)[K]) {
    __syncthreads();
    for (int k = 0; k < K; k++) {
        double t = 0;
        for (int i = 0; i < VTHREADS / 64; i++) t += sh[i][k];
        v[k] = t;
    }
    __syncthreads();
}

// FP64 sums of sphere p's marker forces/torques, then its contact forces; whole block
__device__ __forceinline__ void sphere_post(const PartArgs& a, int p, int do_reduce, int do_contacts) {
    __shared__ double sh6[4][6];
    __shared__ double sh3[4][3];
    if (do_reduce) {
        double acc[6];
        for (int vt = threadIdx.x; vt < VTHREADS; vt += blockDim.x) {  // wave-uniform trip count
            for (int k = 0; k < 6; k++) acc[k] = 0;
            for (int m = vt; m < a.ppp; m += VTHREADS) {
                const auto mf = a.mforce + 6 * ((long long)p * a.ppp + m);
                for (int k = 0; k < 6; k++) acc[k] += mf[k];
            }
            wave_store<6>(acc, sh6, vt);
        }
        wave_total<6>(acc, sh6);
        if (threadIdx.x == 0)
            for (int k = 0; k < 3; k++) {
                a.ffluid[3 * p + k] = acc[k];
                a.tfluid[3 * p + k] = acc[3 + k];
            }
    }
    if (!do_contacts) return;
    // step number of these contacts: the step whose spreading committed the current epoch
    const long long t = (long long)a.epochs[0] + a.tofs;
    const int zgen = zone_generation(a, t);
    const int zp = a.zone_mode ? a.zone[zgen * a.np + p] : 0;
    const double rp = 0.5 * a.diam;
    const double xi[3] = {a.pos[3 * p], a.pos[3 * p + 1], a.pos[3 * p + 2]};
    const double vi[3] = {a.vel[3 * p], a.vel[3 * p + 1], a.vel[3 * p + 2]};
    double F[3];
    for (int vt = threadIdx.x; vt < VTHREADS; vt += blockDim.x) {  // wave-uniform trip count
        F[0] = F[1] = F[2] = 0;
        for (int j = vt; j < a.np; j += VTHREADS) {
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
            double mult = 1.0;
            if (a.zone_mode) {
                const int zj = a.zone[zgen * a.np + j];
                mult = (double)(ipi ? zone_multiplicity(a, zp, zj) : zone_multiplicity(a, zj, zp));
                if (mult == 0.0) continue;
            }
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
            for (int d = 0; d < 3; d++) F[d] += mult * (sgn * e[d] * fmag);
        }
        wave_store<3>(F, sh3, vt);
    }
    wave_total<3>(F, sh3);
    if (threadIdx.x == 0) {
        // z-wall repulsion (CPU_program.cl:175-182)
        const double lo = a.zbuf - xi[2];
        if (lo > 0) F[2] += a.kcol * lo;
        const double hi = xi[2] - (a.sys[2] - a.zbuf);
        if (hi > 0) F[2] -= a.kcol * hi;
        for (int d = 0; d < 3; d++) a.fcont[3 * p + d] += F[d];
        // zone lists rebuilt after the contacts of steps t % RebuildFreq == 0 (sim_main.c:428-444),
        // into the other generation, which no contact of this step reads
        if (a.zone_mode && t >= 1 && t % a.rebuild == 0) a.zone[(int)((t / a.rebuild) & 1) * a.np + p] = zone_of(a, p);
    }
}

// end of step t and start of step t+1 for sphere p (whole block); marks epoch e_mark (= t+2)
__device__ __forceinline__ void sphere_turn(const PartArgs& a, int p, int do_post, uint32_t e_mark) {
    if (do_post) sphere_post(a, p, 1, 1);
    __shared__ double c3n[3];
    __syncthreads();
    if (threadIdx.x == 0) {
        dynamics_one(a, p);
        for (int k = 0; k < 3; k++) c3n[k] = a.pos_out[3 * p + k];
    }
    __syncthreads();
    mark_sphere(c3n, a.g, a.radius, a.rin, a.rout, mask_of(a.mask0, a.mask1, a.mask2, e_mark), e_mark, 3);
}

// marks of epoch e for the current positions, dilated (the first step of a fused sequence)
__device__ __forceinline__ void sphere_prime(const PartArgs& a, int p, uint32_t e) {
    __shared__ double c3p[3];
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) c3p[k] = a.pos[3 * p + k];
    __syncthreads();
    mark_sphere(c3p, a.g, a.radius, a.rin, a.rout, mask_of(a.mask0, a.mask1, a.mask2, e), e, 3);
}

}  // namespace d3q19
