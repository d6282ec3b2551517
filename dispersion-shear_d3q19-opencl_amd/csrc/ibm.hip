// ibm.hip -- immersed-boundary kernels: marker interpolation, penalty force, deterministic
// spreading and the per-sphere force/torque sums (particle_fluid_forces_linear_stencil,
// sum_particle_fluid_forces and reset_particle_fluid_forces, GPU_program.cl:81-313).
//
// Spreading is order independent: every contribution is rounded to a 2^-40 fixed-point integer
// and summed with 64-bit integer additions, whose result does not depend on the order of arrival,
// so the force field is bitwise reproducible (the reference's slot counter + float
// read-modify-write, GPU_program.cl:217-224, is not).  A block takes a group of neighbouring
// markers of one sphere and sums their contributions in an LDS box first, so that each node of the
// group's stencils takes one global atomic per component.  Peskin-4 markers run four to a
// wavefront (a 16-lane row each); the reference's trilinear kernel one per wavefront.  The
// reference's 27-point smoothing of the spread force at the next collide (GPU_program.cl:356-377,
// buffer shell contributing zero and no periodic wrap) is folded into the spread: each marker
// spreads directly with the convolution of its trilinear weights and the (1/4,1/2,1/4)^3 kernel,
// restricted to interior target nodes.
#include <hip/hip_ext.h>
#include "internal.h"

namespace d3q19 {

// Peskin's 4-point kernel; one square root per call (the lanes of a wavefront take either branch)
__device__ __forceinline__ double peskin4(double r) {
    r = fabs(r);
    const bool inner = r < 1.0;
    const double s = sqrt(inner ? 1.0 + 4.0 * r - 4.0 * r * r : -7.0 + 12.0 * r - 4.0 * r * r);
    const double w = (inner ? 3.0 - 2.0 * r + s : 5.0 - 2.0 * r - s) * 0.125;
    return r < 2.0 ? w : 0.0;
}

// fmod(x + L, L) (GPU_program.cl:141-144), exactly: for y = x + L in [0, 2L) fmod(y, L) is y or y - L,
// and y - L is exact there (Sterbenz); other values take the library fmod
__device__ __forceinline__ double wrap_coord(double x, double L) {
    const double y = x + L;
    if (y >= 0.0 && y < 2.0 * L) return y >= L ? y - L : y;
    return fmod(y, L);
}

// n mod nd in [0, nd); one conditional add or subtract covers the offsets a stencil reaches
__device__ __forceinline__ int wrap_index(int n, int nd) {
    n += n < 0 ? nd : 0;
    n -= n >= nd ? nd : 0;
    if ((unsigned)n >= (unsigned)nd) n = ((n % nd) + nd) % nd;
    return n;
}

// round to the nearest integer, ties to even (as __double2ll_rn): below 2^51 in magnitude, adding
// 1.5 * 2^52 leaves the integer in the low mantissa bits
__device__ __forceinline__ long long round_to_ll(double x) {
#pragma clang fp contract(off)
    if (!(fabs(x) < 2251799813685248.0)) return __double2ll_rn(x);
    return __double_as_longlong(x + 6755399441055744.0) - 0x4338000000000000ll;
}

// contribution -w F[c] of one stencil node, rounded to fixed point as a whole
__device__ __forceinline__ long long spread_value(double w, double Fc) { return round_to_ll((-w * Fc) * FIX_SCALE); }

// ... added with 64-bit integer atomics (order independent)
// (a node that receives a nonzero contribution gets its force tag; every writer stores the same
// value, so concurrent stores agree)
__device__ __forceinline__ void spread_add(D3Q19_GLOBAL long long* facc, D3Q19_GLOBAL uint8_t* ftag, long long nint, int node, double w,
                                           const double* F) {
    bool any = false;
    for (int c = 0; c < 3; c++) {
        const long long iv = spread_value(w, F[c]);
        if (iv != 0) {
            __hip_atomic_fetch_add((D3Q19_GLOBAL unsigned long long*)(facc + c * nint + node), (unsigned long long)iv,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            any = true;
        }
    }
    if (any) ftag[node] = 1;
}

// v of the lane that the DPP control CTRL names, within a row of 16 lanes
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double read_lane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sums in a fixed order, the same bits in every lane.  row_sum: over each row of 16 lanes, by pairs
// (lane ^ 1, lane ^ 2, then the mirrored half-row and row: each pair adds the same two values, so
// both lanes of a pair hold the same sum).  wave_sum: the four row sums as (r0 + r1) + (r2 + r3).
__device__ __forceinline__ double row_sum(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
    v = row_sum(v);
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// One 64-lane wavefront per marker; lane k = (kx, ky, kz) in [0,4)^3 owns one node of the
// marker's 4x4x4 support.  Interpolation is a wave reduction of the lanes' weighted velocities;
// spreading is three integer atomics per lane.
// Marker m of sphere p (mforce slot p*ppp + m), executed by one whole wavefront; each lane whose
// stencil node receives force calls sink(sn, weight, F) with the node's coordinates sn[3].
template <class R, int MODE, class Sink>
__device__ __forceinline__ void ibm_marker(const IbmArgs& a, const D3Q19_GLOBAL double* pos, const D3Q19_GLOBAL double* vel, int p,
                                           int m, int lane, Sink&& sink) {
    const int gid = p * a.ppp + m;
    const Geom& g = a.g;
    const double r0[3] = {(double)a.pts[4 * m], (double)a.pts[4 * m + 1], (double)a.pts[4 * m + 2]};
    const int nd[3] = {g.nx, g.ny, g.nz};
    double rpp[3];
    bool finite = true;
    for (int d = 0; d < 3; d++) {
        const double L = (double)g.sys[d];
        rpp[d] = wrap_coord(pos[3 * p + d] + r0[d], L);  // GPU_program.cl:141-144
        if (!(rpp[d] >= 0.0 && rpp[d] < L)) finite = false;  // NaN or out of range
    }
    if (!finite) {
        // a sphere whose state is no longer finite contributes nothing (the watchdog reports it)
        if (lane < 6) a.mforce[6 * (long long)gid + lane] = 0.0 / 0.0;
        return;
    }
    const int kk[3] = {lane & 3, (lane >> 2) & 3, lane >> 4};
    int sn[3];       // spread node of this lane, per axis
    double sw[3];    // spread weight factor per axis
    bool sok = true;
    double iw = 0.0; // interpolation weight of this lane (product over axes)
    int inode = 0;
    bool iok;
    if (MODE == D3Q19_IBM_LINEAR_SMOOTHED) {
        // interpolation: trilinear corners (lanes with kx,ky,kz in {0,1}), GPU_program.cl:150-196
        // spreading: trilinear pair convolved with the (1/4,1/2,1/4) smoothing, interior targets
        iok = (kk[0] < 2 && kk[1] < 2 && kk[2] < 2);
        double wi = 1.0;
        int ic[3];
        for (int d = 0; d < 3; d++) {
            const double fl = floor(rpp[d]);
            const int ia = (int)fl;
            const int ib = (ia == nd[d] - 1) ? 0 : ia + 1;  // GPU_program.cl:160-162
            const double A = 1.0 - (rpp[d] - fl), B = 1.0 - A;
            const int k = kk[d];
            ic[d] = (k & 1) ? ib : ia;
            wi *= (k & 1) ? B : A;
            int node;
            double w;
            if (ib == ia + 1) {
                node = ia - 1 + k;
                w = k == 0 ? 0.25 * A : (k == 1 ? 0.5 * A + 0.25 * B : (k == 2 ? 0.25 * A + 0.5 * B : 0.25 * B));
            } else {
                node = k == 0 ? ia - 1 : (k == 1 ? ia : (k == 2 ? ib : ib + 1));
                w = k == 0 ? 0.25 * A : (k == 1 ? 0.5 * A : (k == 2 ? 0.5 * B : 0.25 * B));
            }
            sn[d] = node;
            sw[d] = w;
            if (node < 0 || node >= nd[d]) sok = false;
        }
        iw = iok ? wi : 0.0;
        inode = ic[0] + g.nx * (ic[1] + g.ny * ic[2]);
    } else {
        double w3 = 1.0;
        for (int d = 0; d < 3; d++) {
            const int base = (int)floor(rpp[d]);
            int node = base - 1 + kk[d];
            const double w = peskin4(rpp[d] - (double)node);
            if (g.per[d]) node = wrap_index(node, nd[d]);
            else if (node < 0 || node >= nd[d]) sok = false;
            sn[d] = node;
            sw[d] = w;
            w3 *= w;
        }
        iok = sok;
        iw = sok ? w3 : 0.0;
        inode = sok ? sn[0] + g.nx * (sn[1] + g.ny * sn[2]) : 0;
    }
    const D3Q19_GLOBAL R* u = (const D3Q19_GLOBAL R*)a.u;
    double c0 = 0, c1 = 0, c2 = 0;
    if (iok) {
        const D3Q19_GLOBAL R* un = u + 3 * (long long)inode;  // components interleaved per node
        c0 = iw * (double)un[0];
        c1 = iw * (double)un[1];
        c2 = iw * (double)un[2];
    }
    const double up[3] = {wave_sum(c0), wave_sum(c1), wave_sum(c2)};
    const D3Q19_GLOBAL double* om = a.omega + 3 * p;
    const double wr[3] = {om[1] * r0[2] - om[2] * r0[1], om[2] * r0[0] - om[0] * r0[2], om[0] * r0[1] - om[1] * r0[0]};
    double F[3];
    for (int k = 0; k < 3; k++) F[k] = (up[k] - (vel[3 * p + k] + wr[k])) * a.dA;
    if (sok) sink(sn, sw[0] * sw[1] * sw[2], F);
    if (lane == 0) {
        D3Q19_GLOBAL double* mf = a.mforce + 6 * (long long)gid;
        mf[0] = F[0];
        mf[1] = F[1];
        mf[2] = F[2];
        mf[3] = r0[1] * F[2] - r0[2] * F[1];
        mf[4] = r0[2] * F[0] - r0[0] * F[2];
        mf[5] = r0[0] * F[1] - r0[1] * F[0];
    }
}

// Peskin-4 with four markers per wavefront.  Row r of the wavefront (lanes 16r .. 16r + 15, the
// groups that DPP operations work within) runs marker m of sphere p; lane j of the row owns the
// stencil row (ky, kz) = (j & 3, j >> 2), the four nodes kx = 0..3 along x.  The marker's twelve
// one-dimensional weights are evaluated once, by lanes 0..11 of the row (axis j >> 2, offset
// j & 3), and read by the others through lane shuffles; the interpolation is a sum over the lane's
// four nodes, then over the row.  Against one marker per wavefront this takes 40% of the vector
// instructions per marker.  Rows without a marker (has false) run along with zero weights
// and contribute nothing; every lane of the wavefront reaches the shuffles and row sums.
template <class R, class Sink>
__device__ __forceinline__ void ibm_rows(const IbmArgs& a, const D3Q19_GLOBAL double* pos, const D3Q19_GLOBAL double* vel, int p, int m,
                                         bool has, int lane, Sink&& sink) {
    const Geom& g = a.g;
    const int nd[3] = {g.nx, g.ny, g.nz};
    const int j = lane & 15, rowbase = lane & 48;
    const int ky = j & 3, kz = j >> 2;
    const int gid = p * a.ppp + m;
    double r0[3] = {0.0, 0.0, 0.0}, rpp[3] = {0.5, 0.5, 0.5};
    bool live = has;
    if (has) {
        for (int d = 0; d < 3; d++) {
            r0[d] = (double)a.pts[4 * m + d];
            const double L = (double)g.sys[d];
            rpp[d] = wrap_coord(pos[3 * p + d] + r0[d], L);  // GPU_program.cl:141-144
            if (!(rpp[d] >= 0.0 && rpp[d] < L)) live = false;  // NaN or out of range
        }
        // a sphere whose state is no longer finite contributes nothing (the watchdog reports it)
        if (!live && j < 6) a.mforce[6 * (long long)gid + j] = 0.0 / 0.0;
    }
    int base[3];
    for (int d = 0; d < 3; d++) base[d] = live ? (int)floor(rpp[d]) : 0;
    // this lane's one-dimensional weight (lanes 12..15 of a row evaluate an unused one)
    const int ax = (j >> 2) < 3 ? (j >> 2) : 0;
    const double rsel = ax == 0 ? rpp[0] : (ax == 1 ? rpp[1] : rpp[2]);
    const int bsel = ax == 0 ? base[0] : (ax == 1 ? base[1] : base[2]);
    const double wj = peskin4(rsel - (double)(bsel - 1 + (j & 3)));
    double wx[4];
    for (int k = 0; k < 4; k++) wx[k] = __shfl(wj, rowbase + k, 64);
    const double wy = __shfl(wj, rowbase + 4 + ky, 64), wz = __shfl(wj, rowbase + 8 + kz, 64);
    // stencil nodes: (base - 1 + k) per axis, wrapped on periodic axes, dropped beyond a wall
    int yn = base[1] - 1 + ky, zn = base[2] - 1 + kz;
    bool okyz = live;
    if (g.per[1]) yn = wrap_index(yn, nd[1]);
    else if (yn < 0 || yn >= nd[1]) okyz = false;
    if (g.per[2]) zn = wrap_index(zn, nd[2]);
    else if (zn < 0 || zn >= nd[2]) okyz = false;
    int xn[4];
    bool ok[4];
    double w[4];
    for (int k = 0; k < 4; k++) {
        int x = base[0] - 1 + k;
        bool okx = okyz;
        if (g.per[0]) x = wrap_index(x, nd[0]);
        else if (x < 0 || x >= nd[0]) okx = false;
        xn[k] = x;
        ok[k] = okx;
        w[k] = (wx[k] * wy) * wz;
    }
    // the lane's four nodes summed as (p0 + p1) + (p2 + p3), then the row over ky pairs and kz pairs:
    // the association of one marker per wavefront (lane = kx + 4 ky + 16 kz, wave_sum), so both
    // layouts give the same bits
    const D3Q19_GLOBAL R* u = (const D3Q19_GLOBAL R*)a.u;
    double pc[3][4];
    const long long rowoff = (long long)g.nx * (yn + (long long)g.ny * zn);
    for (int k = 0; k < 4; k++) {
        pc[0][k] = pc[1][k] = pc[2][k] = 0.0;
        if (ok[k]) {
            const D3Q19_GLOBAL R* un = u + 3 * (rowoff + xn[k]);  // components interleaved per node
            pc[0][k] = w[k] * (double)un[0];
            pc[1][k] = w[k] * (double)un[1];
            pc[2][k] = w[k] * (double)un[2];
        }
    }
    double up[3];
    for (int c = 0; c < 3; c++) up[c] = row_sum((pc[c][0] + pc[c][1]) + (pc[c][2] + pc[c][3]));
    if (!live) return;
    const D3Q19_GLOBAL double* om = a.omega + 3 * p;
    const double wr[3] = {om[1] * r0[2] - om[2] * r0[1], om[2] * r0[0] - om[0] * r0[2], om[0] * r0[1] - om[1] * r0[0]};
    double F[3];
    for (int k = 0; k < 3; k++) F[k] = (up[k] - (vel[3 * p + k] + wr[k])) * a.dA;
    for (int k = 0; k < 4; k++)
        if (ok[k]) {
            const int sn[3] = {xn[k], yn, zn};
            sink(sn, w[k], F);
        }
    if (j == 0) {
        D3Q19_GLOBAL double* mf = a.mforce + 6 * (long long)gid;
        mf[0] = F[0];
        mf[1] = F[1];
        mf[2] = F[2];
        mf[3] = r0[1] * F[2] - r0[2] * F[1];
        mf[4] = r0[2] * F[0] - r0[0] * F[2];
        mf[5] = r0[0] * F[1] - r0[1] * F[0];
    }
}

// Sphere p belongs to this pass: its stencils reach at most up to a z plane in [zlo, zhi).  The top
// plane of a marker's stencil is floor(z_marker) + 2 <= floor(z_sphere + zreach) + 2 (planes beyond a
// wall are dropped); a stencil that wraps around a periodic z axis and a non-finite position belong
// to the last slab, whose pass runs after every slab has collided.
__device__ __forceinline__ bool pass_owns(const IbmArgs& a, const D3Q19_GLOBAL double* pos, int p) {
    if (a.zlo <= 0 && a.zhi >= a.g.nz) return true;
    const double zc = pos[3 * p + 2];
    int top = a.g.nz - 1;
    if (zc > -1e7 && zc < 1e7) {
        const int hi = (int)floor(zc + a.zreach) + 2, lo = (int)floor(zc - a.zreach) - 1;
        if (!a.g.per[2]) top = hi < a.g.nz - 1 ? hi : a.g.nz - 1;
        else if (hi < a.g.nz && lo >= 0) top = hi;
    }
    return top >= a.zlo && top < a.zhi;
}

// ROWS: four Peskin-4 markers per wavefront (ibm_rows); otherwise one marker per wavefront
// (ibm_marker, also the reference's trilinear kernel)
template <int MODE, bool ROWS> constexpr bool rows_layout() { return MODE == D3Q19_IBM_PESKIN4 && ROWS; }

template <class R, int MODE, bool ROWS>
__global__ void __launch_bounds__(64 * IBM_WAVES) ibm_marker_kernel(IbmArgs a) {
    const int lane = threadIdx.x & 63;
    if constexpr (rows_layout<MODE, ROWS>()) {
        const int wave = blockIdx.x * IBM_WAVES + (int)(threadIdx.x >> 6);
        const int total = a.np * a.ppp;
        if (4 * wave >= total) return;  // uniform per wavefront
        if (wave == 0 && lane == 0 && a.commit) {
            const uint32_t e = a.epochs[0] + 1u;  // as below
            a.epochs[0] = e;
            a.epochs[1] = e;
        }
        const int gid = 4 * wave + (lane >> 4);  // uniform per row
        bool has = gid < total;
        const int p = has ? gid % a.np : 0, k = gid / a.np;
        has = has && pass_owns(a, a.pos, p);
        const int m = (int)(((long long)k * a.mstride) % a.ppp);
        const long long nint = a.g.nint;
        ibm_rows<R>(a, a.pos, a.vel, p, m, has, lane, [&](const int* sn, double w, const double* F) {
            spread_add(a.facc, a.ftag, nint, sn[0] + a.g.nx * (sn[1] + a.g.ny * sn[2]), w, F);
        });
        return;
    }
    const int gid = blockIdx.x * IBM_WAVES + (int)(threadIdx.x >> 6);
    if (gid >= a.np * a.ppp) return;  // uniform per wavefront
    if (gid == 0 && lane == 0 && a.commit) {
        // the step's epoch (read as epochs[0] + 1 by the kernels before this one) is committed,
        // and its spread force is now the pending one
        const uint32_t e = a.epochs[0] + 1u;
        a.epochs[0] = e;
        a.epochs[1] = e;
    }
    // wavefronts that run together take markers of different spheres, and a sphere's markers are
    // visited with a stride of about 0.618 ppp, so concurrent atomics rarely meet on one node
    const int p = gid % a.np, k = gid / a.np;
    if (!pass_owns(a, a.pos, p)) return;  // uniform per wavefront
    const int m = (int)(((long long)k * a.mstride) % a.ppp);
    const long long nint = a.g.nint;
    ibm_marker<R, MODE>(a, a.pos, a.vel, p, m, lane, [&](const int* sn, double w, const double* F) {
        spread_add(a.facc, a.ftag, nint, sn[0] + a.g.nx * (sn[1] + a.g.ny * sn[2]), w, F);
    });
}

// ---- spreading staged in LDS -----------------------------------------------------------------------
// A block takes a group of markers of one sphere that lie next to each other on its surface
// (groups of a.gm markers fixed on the host): each row (Peskin-4) or wavefront (trilinear) runs its marker as
// above, but adds its fixed-point contributions into an LDS box covering the group's stencils
// (64-bit integer LDS atomics); the block then adds the box's nonzero nodes to the spread force with
// one global atomic per node and component.  Stencils of neighbouring markers overlap, so the
// global atomics drop by the overlap factor; integer sums keep the result bitwise that of per-lane
// atomics.  Nodes are placed in the box by their offset from the sphere's centre node, taken into
// (-n/2, n/2] on periodic axes (used only where every periodic extent exceeds twice the sphere's
// stencil reach).  The box lives in dynamic LDS of `cap` nodes per component (the launch's
// largest group box); a group whose box exceeds the simulation's capacity spreads directly.
__device__ __forceinline__ int rel_offset(int n, int c, int nd, int per) {
    int o = n - c;
    if (per) {
        if (2 * o > nd) o -= nd;
        else if (2 * o <= -nd) o += nd;
    }
    return o;
}

// One group of markers of sphere p staged in an LDS box: zero the box, spread(sink) adds the group's
// contributions, then the box's nonzero nodes go to the spread force with one global atomic per
// node and component, and their force tags are set.
template <class Spread>
__device__ __forceinline__ void ibm_tile_box(const IbmArgs& a, const D3Q19_GLOBAL double* pos, int p, int grp, long long* tile, int cap,
                                             Spread spread) {
    const Geom& g = a.g;
    const int nd[3] = {g.nx, g.ny, g.nz};
    // the group's box, relative to the sphere's centre node, was computed from the offsets on the
    // host; the centre node is floor of the position (positions lie in [0, SystemSize))
    const D3Q19_GLOBAL int* gb = a.gbox + 8 * grp;
    const int lo[3] = {gb[0], gb[1], gb[2]};
    const int ex = gb[3], ey = gb[4], ez = gb[5];
    const int ctr[3] = {(int)floor(pos[3 * p]), (int)floor(pos[3 * p + 1]), (int)floor(pos[3 * p + 2])};
    const bool use_tile = a.tile_ok && gb[6] <= a.tile_nodes && gb[6] <= cap;  // uniform over the block
    const int vol = use_tile ? ex * ey * ez : 0;
    long long* const t0 = tile;
    long long* const t1 = tile + cap;
    long long* const t2 = tile + 2 * cap;
    for (int i = threadIdx.x; i < vol; i += blockDim.x) {
        t0[i] = 0;
        t1[i] = 0;
        t2[i] = 0;
    }
    __syncthreads();
    auto sink = [&](const int* sn, double w, const double* F) {
        int t = -1;
        if (use_tile) {
            const int tx = rel_offset(sn[0], ctr[0], nd[0], g.per[0]) - lo[0];
            const int ty = rel_offset(sn[1], ctr[1], nd[1], g.per[1]) - lo[1];
            const int tz = rel_offset(sn[2], ctr[2], nd[2], g.per[2]) - lo[2];
            if ((unsigned)tx < (unsigned)ex && (unsigned)ty < (unsigned)ey && (unsigned)tz < (unsigned)ez) t = tx + ex * (ty + ey * tz);
        }
        if (t >= 0) {
            long long* const tc[3] = {t0, t1, t2};
            for (int c = 0; c < 3; c++) {
                const long long iv = spread_value(w, F[c]);
                if (iv != 0) atomicAdd((unsigned long long*)&tc[c][t], (unsigned long long)iv);
            }
        } else {
            spread_add(a.facc, a.ftag, g.nint, sn[0] + g.nx * (sn[1] + g.ny * sn[2]), w, F);
        }
    };
    spread(sink);
    __syncthreads();
    // box position of node i by multiply and shift: for i < 2048 and extents up to 1024,
    // (i * ceil(2^21 / e)) >> 21 is i / e exactly (the excess stays below i / 2^21 < 1 / e), and
    // the product stays below 2^32
    static_assert(TILE_NODES <= 2048 && TILE_NODES_BATCH <= 2048, "the box index split below is exact for boxes up to 2048 nodes");
    const unsigned mx = vol ? ((1u << 21) + ex - 1) / ex : 0, my = vol ? ((1u << 21) + ey - 1) / ey : 0;
    for (int i = threadIdx.x; i < vol; i += blockDim.x) {
        const long long v0 = t0[i], v1 = t1[i], v2 = t2[i];
        if ((v0 | v1 | v2) == 0) continue;
        const int ixy = (int)(((unsigned)i * mx) >> 21), iz = (int)(((unsigned)ixy * my) >> 21);
        const int t3[3] = {i - ixy * ex, ixy - iz * ey, iz};
        int n3[3];
        for (int d = 0; d < 3; d++) {
            int n = ctr[d] + lo[d] + t3[d];
            if (g.per[d]) n = wrap_index(n, nd[d]);
            n3[d] = n;
        }
        const long long node = n3[0] + (long long)g.nx * (n3[1] + (long long)g.ny * n3[2]);
        const long long vs[3] = {v0, v1, v2};
        for (int c = 0; c < 3; c++)
            if (vs[c] != 0)
                __hip_atomic_fetch_add((D3Q19_GLOBAL unsigned long long*)(a.facc + c * (long long)g.nint + node), (unsigned long long)vs[c],
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a.ftag[node] = 1;
    }
}

// a single simulation's group: one marker per row (ROWS) or wavefront, marker m of this slot
template <class R, int MODE, bool ROWS>
__device__ __forceinline__ void ibm_tile_group(const IbmArgs& a, const D3Q19_GLOBAL double* pos, const D3Q19_GLOBAL double* vel, int p,
                                               int grp, int m, bool has, int lane, long long* tile, int cap) {
    ibm_tile_box(a, pos, p, grp, tile, cap, [&](auto& sink) __attribute__((always_inline)) {
        if constexpr (rows_layout<MODE, ROWS>()) ibm_rows<R>(a, pos, vel, p, m, has, lane, sink);  // every lane: row sums
        else if (has) ibm_marker<R, MODE>(a, pos, vel, p, m, lane, sink);
    });
}

// threads per marker of a group's block: a row or a wavefront; the most threads of such a block
// (groups of up to TILE_MARKERS_MAX markers in the row layout, else TILE_MARKERS)
template <int MODE, bool ROWS> constexpr int marker_threads() { return rows_layout<MODE, ROWS>() ? 16 : 64; }
template <int MODE, bool ROWS> constexpr int group_threads() {
    return marker_threads<MODE, ROWS>() * (rows_layout<MODE, ROWS>() ? TILE_MARKERS_MAX : TILE_MARKERS);
}

// one block per (sphere, group of a.gm markers); blocks of consecutive index take different spheres
template <class R, int MODE, bool ROWS>
__global__ void __launch_bounds__((group_threads<MODE, ROWS>())) ibm_tile_kernel(IbmArgs a) {
    extern __shared__ long long tile_mem[];  // 3 x a.tile_nodes
    const int lane = threadIdx.x & 63, w = (int)threadIdx.x / marker_threads<MODE, ROWS>();  // marker slot
    const int b = blockIdx.x;
    if (b == 0 && threadIdx.x == 0 && a.commit) {
        const uint32_t e = a.epochs[0] + 1u;  // commits the step's epoch, as ibm_marker_kernel does
        a.epochs[0] = e;
        a.epochs[1] = e;
    }
    const int p = b % a.np, grp = b / a.np;
    if (!pass_owns(a, a.pos, p)) return;  // uniform over the block
    const int j = grp * a.gm + w;
    const bool has = j < a.ppp;
    ibm_tile_group<R, MODE, ROWS>(a, a.pos, a.vel, p, grp, has ? a.perm[j] : 0, has, lane, tile_mem, a.tile_nodes);
}

// Batched, spreading staged in LDS: a persistent grid of one-group blocks walks the group list of
// all cells, (sphere << 12) | first position of the group in the sphere's group order.  A cell's
// groups take its own group_markers (ia.gm: up to TILE_MARKERS_MAX where its boxes fit
// TILE_NODES_BATCH, else TILE_MARKERS).  A block of 512 threads takes a group in passes: Peskin-4
// markers 32 at a time, one per 16-lane row (every row of the block runs each pass, rows past the
// group's end contribute nothing); the trilinear kernel's 8 at a time, one per wavefront.  The
// arithmetic per marker is a single simulation's, and the integer sums make the field independent
// of the grouping: a cell's bits are those of the cell alone.  The batch's epoch counter is
// committed by the first thread.  The tables come in as restrict arguments (nothing here writes
// them): the descriptor fields are scalar loads that the force stores do not invalidate.
constexpr int BATCH_IBM_THREADS = 512;
template <class R, int MODE>
__global__ void __launch_bounds__(BATCH_IBM_THREADS) ibm_batch_tile_kernel(BatchTables t, const CellDesc* __restrict__ cells,
                                                                          const int4* __restrict__ sphs, const int* __restrict__ msph,
                                                                          int sp) {
    constexpr bool ROWS = rows_layout<MODE, true>();
    constexpr int SLOTS = BATCH_IBM_THREADS / marker_threads<MODE, true>();  // markers per pass
    extern __shared__ long long tile_mem[];  // 3 x TILE_NODES_BATCH
    const int lane = threadIdx.x & 63, slot = (int)threadIdx.x / marker_threads<MODE, true>();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t e = t.epochs[0] + 1u;
        t.epochs[0] = e;
        t.epochs[1] = e;
    }
    const int ngrp = t.cnt->nmark;
    for (int k = blockIdx.x; k < ngrp; k += gridDim.x) {  // uniform over the block
        const int v = __builtin_amdgcn_readfirstlane(msph[k]);
        const int4 d = sphs[v >> 12];  // cell, local sphere, first global marker, markers per sphere
        const int cell = __builtin_amdgcn_readfirstlane(d.x), sph = __builtin_amdgcn_readfirstlane(d.y);
        const int ppp = __builtin_amdgcn_readfirstlane(d.w);
        const CellDesc& cd = cells[cell];
        const IbmArgs& ia = cd.i;
        const int par = __builtin_amdgcn_readfirstlane(cd.par);
        const int gm = __builtin_amdgcn_readfirstlane(ia.gm);
        const int j0 = v & 4095, cnt = min(gm, ppp - j0);
        // x(t+1), v(t+1): written by this step's turn
        const D3Q19_GLOBAL double* pos = cd.posb[par ^ sp ^ 1];
        const D3Q19_GLOBAL double* vel = cd.velb[par ^ sp ^ 1];
        ibm_tile_box(ia, pos, sph, j0 / gm, tile_mem, TILE_NODES_BATCH, [&](auto& sink) __attribute__((always_inline)) {
            for (int i0 = 0; i0 < cnt; i0 += SLOTS) {  // uniform over the block
                const int i = i0 + slot;
                const bool has = i < cnt;
                if constexpr (ROWS) ibm_rows<R>(ia, pos, vel, sph, has ? ia.perm[j0 + i] : 0, has, lane, sink);  // every lane: row sums
                else if (has) ibm_marker<R, MODE>(ia, pos, vel, sph, ia.perm[j0 + i], lane, sink);
            }
        });
        __syncthreads();  // the next group reuses the tile
    }
}

int marker_stride(int ppp) {
    if (ppp <= 2) return 1;
    int st = (int)(0.618034 * ppp + 0.5);
    auto gcd = [](int x, int y) {
        while (y) {
            const int t = x % y;
            x = y;
            y = t;
        }
        return x;
    };
    while (gcd(st, ppp) != 1) st++;
    return st;
}

IbmArgs make_ibm_args(Sim* s) {
    IbmArgs a;
    a.mstride = marker_stride(s->ppp > 0 ? s->ppp : 1);
    a.tile_ok = s->tile_ok ? 1 : 0;
    a.tile_nodes = s->tile_nodes;
    a.gm = s->group_markers;
    a.perm = gp(s->perm);
    a.gbox = gp(s->gbox);
    a.g = s->g;
    a.np = s->np;
    a.ppp = s->ppp;
    a.mode = s->opt.ibm_kernel;
    a.dA = (double)s->fp.PointArea;
    a.pos = gp(s->pos);
    a.vel = gp(s->vel);
    a.omega = gp(s->omega);
    a.pts = gp(s->pts);
    a.u = gp(s->u);
    a.facc = gp(s->facc);
    a.ftag = gp(s->ftag);
    a.mforce = gp(s->mforce);
    a.epochs = gp(s->epochs);
    a.zreach = s->mark_rout;
    a.zlo = 0;
    a.zhi = 1 << 30;
    a.commit = 1;
    return a;
}

// a kernel launch, timed by events at the kernel's own start and end when ev0 is set
template <class K, class... Args>
static void launch_timed(K kernel, dim3 grid, dim3 block, size_t lds, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1, Args... args) {
    if (ev0 && ev1) hipExtLaunchKernelGGL(kernel, grid, block, lds, st, ev0, ev1, 0, args...);
    else hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
}

template <class R, int MODE, bool ROWS>
static void launch_ibm_as(Sim* s, const IbmArgs& a, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    if (s->tile_ok) {
        const int gm = s->group_markers;
        const int blocks = s->np * ((s->ppp + gm - 1) / gm);
        const size_t lds = 3 * sizeof(long long) * (size_t)s->tile_nodes;
        launch_timed(ibm_tile_kernel<R, MODE, ROWS>, dim3(blocks), dim3(marker_threads<MODE, ROWS>() * gm), lds, st, ev0, ev1, a);
    } else {
        const int total = s->np * s->ppp;
        const int per_block = IBM_WAVES * (rows_layout<MODE, ROWS>() ? 4 : 1);
        launch_timed(ibm_marker_kernel<R, MODE, ROWS>, dim3((total + per_block - 1) / per_block), dim3(64 * IBM_WAVES), 0, st, ev0, ev1, a);
    }
}

int launch_ibm(Sim* s, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1, int zlo, int zhi, int commit) {
    if (s->np == 0) return 0;
    IbmArgs a = make_ibm_args(s);
    a.zlo = zlo;
    a.zhi = zhi;
    a.commit = commit;
    const bool dbl = s->storage == D3Q19_STORE_FP64;
    if (a.mode == D3Q19_IBM_PESKIN4) {
        if (s->ibm_rows) {
            if (dbl) launch_ibm_as<double, D3Q19_IBM_PESKIN4, true>(s, a, st, ev0, ev1);
            else launch_ibm_as<float, D3Q19_IBM_PESKIN4, true>(s, a, st, ev0, ev1);
        } else {
            if (dbl) launch_ibm_as<double, D3Q19_IBM_PESKIN4, false>(s, a, st, ev0, ev1);
            else launch_ibm_as<float, D3Q19_IBM_PESKIN4, false>(s, a, st, ev0, ev1);
        }
    } else {
        if (dbl) launch_ibm_as<double, D3Q19_IBM_LINEAR_SMOOTHED, false>(s, a, st, ev0, ev1);
        else launch_ibm_as<float, D3Q19_IBM_LINEAR_SMOOTHED, false>(s, a, st, ev0, ev1);
    }
    HIP_OK(hipGetLastError());
    return 0;
}

template <class R, int MODE>
static void launch_ibm_batch_as(const BatchLaunch& L, int sp, hipStream_t st) {
    const size_t lds = 3 * sizeof(long long) * (size_t)TILE_NODES_BATCH;
    hipLaunchKernelGGL((ibm_batch_tile_kernel<R, MODE>), dim3(L.grid_ibm), dim3(BATCH_IBM_THREADS), lds, st, L.t, L.t.cells, L.t.sph,
                       L.t.msph, sp);
}

// Batched cells run one marker per wavefront whatever the marker count: with groups of 8 markers a
// block of the row layout was only 2 wavefronts, and the small-cell server measured 63.8 against
// 59.5-59.9 solves/s without it (the 64-cell ensemble and the 24-cell C5 batch unchanged).
int launch_ibm_batch(const BatchLaunch& L, int sp, hipStream_t st) {
    const bool dbl = L.storage == D3Q19_STORE_FP64;
    if (L.ibm_mode == D3Q19_IBM_PESKIN4) {
        if (dbl) launch_ibm_batch_as<double, D3Q19_IBM_PESKIN4>(L, sp, st);
        else launch_ibm_batch_as<float, D3Q19_IBM_PESKIN4>(L, sp, st);
    } else {
        if (dbl) launch_ibm_batch_as<double, D3Q19_IBM_LINEAR_SMOOTHED>(L, sp, st);
        else launch_ibm_batch_as<float, D3Q19_IBM_LINEAR_SMOOTHED>(L, sp, st);
    }
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace d3q19
