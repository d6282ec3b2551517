// ibm.hip -- immersed-boundary kernels: marker interpolation, penalty force, deterministic
// spreading and the per-sphere force/torque sums (particle_fluid_forces_linear_stencil,
// sum_particle_fluid_forces and reset_particle_fluid_forces, GPU_program.cl:81-313).
//
// Spreading is order independent: every contribution is rounded to a 2^-40 fixed-point integer
// and added with 64-bit integer atomics, whose sum does not depend on the order of arrival, so the
// force field is bitwise reproducible (the reference's slot counter + float read-modify-write,
// GPU_program.cl:217-224, is not).  The reference's 27-point smoothing of the spread force at
// the next collide (GPU_program.cl:356-377, buffer shell contributing zero and no periodic wrap)
// is folded into the spread: each marker spreads directly with the convolution of its trilinear
// weights and the (1/4,1/2,1/4)^3 kernel, restricted to interior target nodes.
#include "internal.h"

namespace d3q19 {

struct IbmArgs {
    Geom g;
    int np, ppp, mode;
    double dA;
    const double* pos;
    const double* vel;
    const double* omega;
    const float* pts;
    const void* u;
    long long* facc;
    double* mforce;
};

__device__ __forceinline__ double peskin4(double r) {
    r = fabs(r);
    if (r < 1.0) return (3.0 - 2.0 * r + sqrt(1.0 + 4.0 * r - 4.0 * r * r)) * 0.125;
    if (r < 2.0) return (5.0 - 2.0 * r - sqrt(-7.0 + 12.0 * r - 4.0 * r * r)) * 0.125;
    return 0.0;
}

__device__ __forceinline__ void spread_add(long long* facc, long long nint, int node, double w, const double* F) {
    for (int c = 0; c < 3; c++) {
        const double v = -w * F[c];
        const long long iv = __double2ll_rn(v * FIX_SCALE);
        if (iv != 0) atomicAdd((unsigned long long*)(facc + c * nint + node), (unsigned long long)iv);
    }
}

template <class R, int MODE>
__global__ void __launch_bounds__(256) ibm_marker_kernel(IbmArgs a) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= a.np * a.ppp) return;
    const Geom& g = a.g;
    const int p = gid / a.ppp, m = gid - p * a.ppp;
    const double r0[3] = {(double)a.pts[4 * m], (double)a.pts[4 * m + 1], (double)a.pts[4 * m + 2]};
    const int nd[3] = {g.nx, g.ny, g.nz};
    double rpp[3];
    bool finite = true;
    for (int d = 0; d < 3; d++) {
        const double L = (double)g.sys[d];
        rpp[d] = fmod((a.pos[3 * p + d] + r0[d]) + L, L);  // GPU_program.cl:141-144
        if (!(rpp[d] >= 0.0 && rpp[d] < L)) finite = false;  // NaN or out of range
    }
    if (!finite) {
        // a sphere whose state is no longer finite contributes nothing (the watchdog reports it)
        double* mf = a.mforce + 6 * (long long)gid;
        for (int k = 0; k < 6; k++) mf[k] = 0.0 / 0.0;
        return;
    }
    const R* u = (const R*)a.u;
    const long long nint = g.nint;
    double up[3] = {0, 0, 0};
    double F[3], T[3];
    if (MODE == D3Q19_IBM_LINEAR_SMOOTHED) {
        int ia[3], ib[3];
        double wa[3];
        for (int d = 0; d < 3; d++) {
            const double fl = floor(rpp[d]);
            ia[d] = (int)fl;
            ib[d] = (ia[d] == nd[d] - 1) ? 0 : ia[d] + 1;  // GPU_program.cl:160-162
            wa[d] = 1.0 - (rpp[d] - fl);
        }
        // trilinear interpolation, corner order of GPU_program.cl:164, 171-178
        for (int c = 0; c < 8; c++) {
            const int bx = (c == 1 || c == 4 || c == 5 || c == 7), by = (c == 2 || c == 4 || c == 6 || c == 7),
                      bz = (c == 3 || c == 5 || c == 6 || c == 7);
            const double w = (bx ? 1.0 - wa[0] : wa[0]) * (by ? 1.0 - wa[1] : wa[1]) * (bz ? 1.0 - wa[2] : wa[2]);
            const int node = (bx ? ib[0] : ia[0]) + g.nx * ((by ? ib[1] : ia[1]) + g.ny * (bz ? ib[2] : ia[2]));
            for (int k = 0; k < 3; k++) up[k] += w * (double)u[k * nint + node];
        }
        double wr[3];
        const double* om = a.omega + 3 * p;
        wr[0] = om[1] * r0[2] - om[2] * r0[1];
        wr[1] = om[2] * r0[0] - om[0] * r0[2];
        wr[2] = om[0] * r0[1] - om[1] * r0[0];
        for (int k = 0; k < 3; k++) F[k] = (up[k] - (a.vel[3 * p + k] + wr[k])) * a.dA;
        // spread targets per axis: trilinear pair convolved with (1/4,1/2,1/4), interior only
        int ti[3][4];
        double tw[3][4];
        int tn[3];
        for (int d = 0; d < 3; d++) {
            const double A = wa[d], B = 1.0 - wa[d];
            int cnt = 0;
            if (ib[d] == ia[d] + 1) {
                const int base = ia[d] - 1;
                const double w4[4] = {0.25 * A, 0.5 * A + 0.25 * B, 0.25 * A + 0.5 * B, 0.25 * B};
                for (int k = 0; k < 4; k++) {
                    const int node = base + k;
                    if (node >= 0 && node < nd[d]) { ti[d][cnt] = node; tw[d][cnt] = w4[k]; cnt++; }
                }
            } else {
                const int nodes[4] = {ia[d] - 1, ia[d], ib[d], ib[d] + 1};
                const double w4[4] = {0.25 * A, 0.5 * A, 0.5 * B, 0.25 * B};
                for (int k = 0; k < 4; k++)
                    if (nodes[k] >= 0 && nodes[k] < nd[d]) { ti[d][cnt] = nodes[k]; tw[d][cnt] = w4[k]; cnt++; }
            }
            tn[d] = cnt;
        }
        for (int kz = 0; kz < tn[2]; kz++)
            for (int ky = 0; ky < tn[1]; ky++)
                for (int kx = 0; kx < tn[0]; kx++) {
                    const int node = ti[0][kx] + g.nx * (ti[1][ky] + g.ny * ti[2][kz]);
                    spread_add(a.facc, nint, node, tw[0][kx] * tw[1][ky] * tw[2][kz], F);
                }
    } else {
        int ti[3][4], ok[3][4];
        double tw[3][4];
        for (int d = 0; d < 3; d++) {
            const int base = (int)floor(rpp[d]);
            for (int k = 0; k < 4; k++) {
                int node = base - 1 + k;
                tw[d][k] = peskin4(rpp[d] - (double)node);
                if (g.per[d]) {
                    node = ((node % nd[d]) + nd[d]) % nd[d];
                    ok[d][k] = 1;
                } else {
                    ok[d][k] = (node >= 0 && node < nd[d]);
                }
                ti[d][k] = node;
            }
        }
        for (int kz = 0; kz < 4; kz++)
            for (int ky = 0; ky < 4; ky++)
                for (int kx = 0; kx < 4; kx++) {
                    if (!(ok[0][kx] && ok[1][ky] && ok[2][kz])) continue;
                    const double w = tw[0][kx] * tw[1][ky] * tw[2][kz];
                    const int node = ti[0][kx] + g.nx * (ti[1][ky] + g.ny * ti[2][kz]);
                    for (int c = 0; c < 3; c++) up[c] += w * (double)u[c * nint + node];
                }
        double wr[3];
        const double* om = a.omega + 3 * p;
        wr[0] = om[1] * r0[2] - om[2] * r0[1];
        wr[1] = om[2] * r0[0] - om[0] * r0[2];
        wr[2] = om[0] * r0[1] - om[1] * r0[0];
        for (int k = 0; k < 3; k++) F[k] = (up[k] - (a.vel[3 * p + k] + wr[k])) * a.dA;
        for (int kz = 0; kz < 4; kz++)
            for (int ky = 0; ky < 4; ky++)
                for (int kx = 0; kx < 4; kx++) {
                    if (!(ok[0][kx] && ok[1][ky] && ok[2][kz])) continue;
                    const int node = ti[0][kx] + g.nx * (ti[1][ky] + g.ny * ti[2][kz]);
                    spread_add(a.facc, nint, node, tw[0][kx] * tw[1][ky] * tw[2][kz], F);
                }
    }
    T[0] = r0[1] * F[2] - r0[2] * F[1];
    T[1] = r0[2] * F[0] - r0[0] * F[2];
    T[2] = r0[0] * F[1] - r0[1] * F[0];
    double* mf = a.mforce + 6 * (long long)gid;
    for (int k = 0; k < 3; k++) {
        mf[k] = F[k];
        mf[3 + k] = T[k];
    }
}

// per-sphere FP64 sums of the marker forces and torques in a fixed order
__global__ void __launch_bounds__(256) ibm_reduce_kernel(const double* __restrict__ mforce, int ppp, double* ff, double* tf) {
    const int p = blockIdx.x;
    __shared__ double sh[6][256];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int m = threadIdx.x; m < ppp; m += blockDim.x) {
        const double* mf = mforce + 6 * ((long long)p * ppp + m);
        for (int k = 0; k < 6; k++) acc[k] += mf[k];
    }
    for (int k = 0; k < 6; k++) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int k = 0; k < 6; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) {
            ff[3 * p + k] = sh[k][0];
            tf[3 * p + k] = sh[3 + k][0];
        }
}

// ---- brick masks -----------------------------------------------------------------------------------
__global__ void clear_mask_kernel(uint8_t* m, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) m[i] = 0;
}

// Marks every brick that a marker stencil of sphere p can touch: for each axis, marker
// coordinates span [c-R, c+R]; they are wrapped like the markers (fmod over SystemSize) and
// widened by the stencil reach (-2..+3 nodes); periodic axes wrap, wall axes clip.
__global__ void mark_kernel(const double* __restrict__ pos, Geom g, double radius, uint8_t* mask) {
    const int p = blockIdx.x;
    __shared__ int lists[3][64];
    __shared__ int counts[3];
    __shared__ uint8_t flags[3][1024];
    const int nd[3] = {g.nx, g.ny, g.nz};
    const int bn[3] = {g.bnx, g.bny, g.bnz};
    for (int d = 0; d < 3; d++)
        for (int i = threadIdx.x; i < bn[d] && i < 1024; i += blockDim.x) flags[d][i] = 0;
    __syncthreads();
    if (threadIdx.x < 3) {
        const int d = threadIdx.x;
        const double c = pos[3 * p + d];
        const double L = (double)g.sys[d];
        const bool finite = (c > -1e7 && c < 1e7);  // false for NaN: such a sphere marks nothing
        const int lo = finite ? (int)floor(c - radius) - 1 : 0, hi = finite ? (int)floor(c + radius) + 1 : -1;
        for (int v = lo; v <= hi; v++) {
            const double w = fmod((double)v + 0.5 + L, L);  // wrap like the markers
            const int wn = (int)floor(w);
            for (int k = -2; k <= 3; k++) {
                int node = wn + k;
                if (g.per[d]) node = ((node % nd[d]) + nd[d]) % nd[d];
                else if (node < 0 || node >= nd[d]) continue;
                const int b = node / BRICK;
                if (b < 1024) flags[d][b] = 1;
            }
        }
        int cnt = 0;
        for (int b = 0; b < bn[d] && b < 1024; b++)
            if (flags[d][b] && cnt < 64) lists[d][cnt++] = b;
        counts[d] = cnt;
    }
    __syncthreads();
    const int tot = counts[0] * counts[1] * counts[2];
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int kx = i % counts[0];
        const int ky = (i / counts[0]) % counts[1];
        const int kz = i / (counts[0] * counts[1]);
        mask[lists[0][kx] + g.bnx * (lists[1][ky] + g.bny * lists[2][kz])] = 1;
    }
}

int launch_clear_mask(Sim* s, uint8_t* m, hipStream_t st) {
    if (!m) return 0;
    hipLaunchKernelGGL(clear_mask_kernel, dim3((s->nbricks + 255) / 256), dim3(256), 0, st, m, s->nbricks);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_mark_bricks(Sim* s, uint8_t* m, hipStream_t st) {
    if (s->np == 0) return 0;
    // marker offsets are scaled unit vectors (|r0|^2 within 1% of 1, D3Q19-OpenCL.c:597-601)
    const double rmax = 0.5 * (double)s->fp.ParticleDiam * 1.01 + 1e-3;
    hipLaunchKernelGGL(mark_kernel, dim3(s->np), dim3(64), 0, st, s->pos, s->g, rmax, m);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_ibm(Sim* s, hipStream_t st) {
    if (s->np == 0) return 0;
    IbmArgs a;
    a.g = s->g;
    a.np = s->np;
    a.ppp = s->ppp;
    a.mode = s->opt.ibm_kernel;
    a.dA = (double)s->fp.PointArea;
    a.pos = s->pos;
    a.vel = s->vel;
    a.omega = s->omega;
    a.pts = s->pts;
    a.u = s->u;
    a.facc = s->facc;
    a.mforce = s->mforce;
    const int total = s->np * s->ppp;
    const int th = 256, blocks = (total + th - 1) / th;
    const bool dbl = s->storage == D3Q19_STORE_FP64;
    if (a.mode == D3Q19_IBM_PESKIN4) {
        if (dbl) hipLaunchKernelGGL((ibm_marker_kernel<double, D3Q19_IBM_PESKIN4>), dim3(blocks), dim3(th), 0, st, a);
        else hipLaunchKernelGGL((ibm_marker_kernel<float, D3Q19_IBM_PESKIN4>), dim3(blocks), dim3(th), 0, st, a);
    } else {
        if (dbl) hipLaunchKernelGGL((ibm_marker_kernel<double, D3Q19_IBM_LINEAR_SMOOTHED>), dim3(blocks), dim3(th), 0, st, a);
        else hipLaunchKernelGGL((ibm_marker_kernel<float, D3Q19_IBM_LINEAR_SMOOTHED>), dim3(blocks), dim3(th), 0, st, a);
    }
    hipLaunchKernelGGL(ibm_reduce_kernel, dim3(s->np), dim3(256), 0, st, s->mforce, s->ppp, s->ffluid, s->tfluid);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace d3q19
