// particles.hip -- rigid-sphere kernels, FP64, one thread block per sphere (the device code is in
// sphere_dev.h): the pre- and post-step kernels of the kernel-level entry points and of the first
// and last steps of a run, the dilated marking that starts a fused sequence, and their batched
// forms over the sphere list of all cells.
#include "sphere_dev.h"

namespace d3q19 {

__global__ void __launch_bounds__(256) sphere_pre_kernel(PartArgs a, int do_dyn, int do_mark) {
    sphere_pre(a, blockIdx.x, do_dyn, do_mark);
}

__global__ void __launch_bounds__(256) sphere_post_kernel(PartArgs a, int do_reduce, int do_contacts) {
    sphere_post(a, blockIdx.x, do_reduce, do_contacts);
}

__global__ void __launch_bounds__(256) zone_init_kernel(PartArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.np) return;
    const int z = zone_of(a, p);
    a.zone[p] = z;
    a.zone[a.np + p] = z;
}

__global__ void __launch_bounds__(256) sphere_prime_kernel(PartArgs a) {
    sphere_prime(a, blockIdx.x, a.epochs[0] + 1u);
}

// Batched: a persistent grid of blocks walks the sphere list of all cells.  sp is the batch's step
// parity: a cell's current positions are posb[par ^ sp], and p[par ^ sp] are its sphere arguments
// for them.  The tables come in as restrict arguments: nothing in these kernels writes them, so the
// descriptor fields are scalar loads that stores to the sphere arrays do not invalidate.
__global__ void __launch_bounds__(256) sphere_prime_batch_kernel(BatchTables t, const CellDesc* __restrict__ cells,
                                                                 const int4* __restrict__ sph, int sp) {
    const int nsph = t.cnt->nsph;
    for (int k = blockIdx.x; k < nsph; k += gridDim.x) {
        const int4 d = sph[k];
        const CellDesc& cd = cells[d.x];
        sphere_prime(cd.p[cd.par ^ sp], d.y, t.epochs[0] + 1u);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) sphere_post_batch_kernel(BatchTables t, const CellDesc* __restrict__ cells,
                                                                const int4* __restrict__ sph, int sp) {
    const int nsph = t.cnt->nsph;
    for (int k = blockIdx.x; k < nsph; k += gridDim.x) {
        const int4 d = sph[k];
        const CellDesc& cd = cells[d.x];
        sphere_post(cd.p[cd.par ^ sp], d.y, 1, 1);
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
    a.radius = std::max(0.5 * (double)s->fp.ParticleDiam * 1.01 + 1e-3, s->mark_rout);
    a.rin = s->mark_rin;
    a.rout = s->mark_rout;
    a.kcol = (double)s->fp.ParForceParams[0];
    a.force_model = s->ip.ParForceModel;
    a.visc_model = s->ip.ViscosityModel;
    a.tau0 = (double)s->fp.NewtonianTau;
    a.zbuf = (double)s->fp.ParticleZBuffer;
    a.pos = gp(s->pos);
    a.vel = gp(s->vel);
    a.pos_out = gp(s->pos);
    a.vel_out = gp(s->vel);
    a.quat = gp(s->quat);
    a.omega = gp(s->omega);
    a.ffluid = gp(s->ffluid);
    a.tfluid = gp(s->tfluid);
    a.fcont = gp(s->fcont);
    a.tcont = gp(s->tcont);
    a.mforce = gp(s->mforce);
    a.mask0 = gp(s->mask[0]);
    a.mask1 = gp(s->mask[1]);
    a.mask2 = gp(s->mask[2]);
    a.flag = gp(s->d_flag);
    a.epochs = gp(s->epochs);
    a.zone_mode = s->opt.contact_detection == D3Q19_CONTACT_ZONES ? 1 : 0;
    a.rebuild = s->opt.rebuild_freq > 0 ? s->opt.rebuild_freq : 10;
    a.tofs = s->tofs;
    for (int d = 0; d < 3; d++) {
        a.nz[d] = s->ip.NumZones[d];
        a.zw[d] = s->fp.ZoneWidth[d];
    }
    a.zone = gp(s->zone);
    return a;
}

int launch_zone_init(Sim* s, hipStream_t st) {
    if (s->np == 0 || s->opt.contact_detection != D3Q19_CONTACT_ZONES) return 0;
    hipLaunchKernelGGL(zone_init_kernel, dim3((s->np + 255) / 256), dim3(256), 0, st, make_part_args(s));
    HIP_OK(hipGetLastError());
    return 0;
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

int launch_sphere_prime(Sim* s, hipStream_t st) {
    if (s->np == 0) return 0;
    hipLaunchKernelGGL(sphere_prime_kernel, dim3(s->np), dim3(256), 0, st, make_part_args(s));
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_sphere_prime_batch(const BatchLaunch& L, int sp, hipStream_t st) {
    hipLaunchKernelGGL(sphere_prime_batch_kernel, dim3(L.grid_sphere), dim3(256), 0, st, L.t, L.t.cells, L.t.sph, sp);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_sphere_post_batch(const BatchLaunch& L, int sp, hipStream_t st) {
    hipLaunchKernelGGL(sphere_post_batch_kernel, dim3(L.grid_sphere), dim3(256), 0, st, L.t, L.t.cells, L.t.sph, sp);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace d3q19
