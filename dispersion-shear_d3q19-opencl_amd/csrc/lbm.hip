// lbm.hip -- fused collide+stream for gfx950 and the lattice-field utility kernels.
//
// One thread per interior node.  Pull form: the thread gathers the 19 post-collision
// populations of its upstream neighbours (periodic axes wrap; on the wet wall layers the
// unknown populations are rebuilt from the knowns, GPU_program.cl:737-865), adds the pending
// immersed-boundary force, computes rho and u (GPU_program.cl:343-394), relaxes (MRT with the
// reference's rates, GPU_program.cl:396-623, or BGK; optional per-node tau from the strain rate,
// GPU_program.cl:457-564, 1041-1128) and writes its own 19 populations.  The reference's separate
// periodic copy and wall kernels (GPU_program.cl:626-865) are thereby part of the same pass.
//
// Algorithmic HBM traffic per lattice update: 19 loads + 19 stores of the storage type
// (152 B fp32, 76 B fp16/bf16, 304 B fp64); +8 B when tau is per node; +24 B read (+24 B
// written) only at nodes of bricks holding pending spread force; +12 B u stores only at nodes
// the next immersed-boundary pass reads (or everywhere when requested).
#include <hip/hip_ext.h>
#include "sphere_dev.h"

// No implicit contraction of multiplies and adds into FMAs in this file: the per-node arithmetic
// is inlined into several kernels (single simulation, batched cells, timing trials), and the same
// source must round the same way in all of them, so that a cell's results are bitwise those of
// the same cell alone.  The FMAs that pay are written out (fma_).
#pragma clang fp contract(off)

namespace d3q19 {

// ---- population storage formats --------------------------------------------------------------
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}

// Storage formats.  st rounds to nearest; st_sr takes B random bits r (the low bits of r; B = SR_BITS,
// the bits the 16-bit format drops from a float32) and rounds stochastically (unbiased: the result's
// expectation over uniform r is the stored value), which only the 16-bit formats need.  NT: the store
// is non-temporal.  The collide's population stores are, on lattices whose two population buffers
// exceed the 256 MB Infinity Cache (the kernels' NT argument, chosen by CollideArgs::nt_stores; a
// runtime branch between the two forms was merged by the compiler): the next read of a population is a
// whole sweep later, and the stores otherwise fill the L2 with dirty lines that the IBM pass, launched
// right after the collide, has to write back first.  On the 256x256x128 workload the fused step kernel
// went from 348 to 334 us (fp32) and the graph-replayed step by 5%; on a cache-resident lattice (128^3
// in fp16) they lost the cache residency: 10% slower.
template <bool NT, class T, class PT> __device__ __forceinline__ void st_pop(PT p, unsigned off, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, plane_at(p, off));
    else *plane_at(p, off) = v;
}
// The gather's population loads are non-temporal too on lattices well beyond the cache (nt_level):
// each stored value is read by the one node it streams to, and the stream otherwise evicts the u,
// force and tau lines that the IBM pass and the next step read (256x256x128 Casson workload, fp32:
// step kernel 254 -> 227 us and the IBM pass after it 58 -> 51 us; fp16: IBM pass 57 -> 53 us,
// kernel unchanged).  Below NT_LOAD_BYTES they cost the residency that ordinary loads keep (142^3
// nodes, 415 MB: 104 against 87 us per step; 184^3, 903 MB: 240 against 223 us), and on rows that
// are not a whole number of wavefronts they cost at every size (200^3, 1.2 GB: 257 against 231 us),
// so a single simulation takes them only on rows of a multiple of 64 nodes.
template <bool NT, class PT> __device__ __forceinline__ auto ld_pop(PT p, unsigned off) {
    if constexpr (NT) return __builtin_nontemporal_load(plane_at(p, off));
    else return *plane_at(p, off);
}
template <int ST> struct Pop;
template <> struct Pop<D3Q19_STORE_FP32> {
    using T = float;
    using R = float;
    static constexpr bool SR = false;
    static constexpr int SR_BITS = 0;
    template <int QI, bool NT = false, class PT> __device__ static R ld(PT p, unsigned off) { return ld_pop<NT>(p, off); }
    template <int QI, class PT> __device__ static double ldd(PT p, unsigned off) { return (double)*plane_at(p, off); }
    template <int QI, bool NT = false, class PT> __device__ static void st(PT p, unsigned off, R v) { st_pop<NT>(p, off, v); }
    template <int QI, bool NT = false, class PT> __device__ static void st_sr(PT p, unsigned off, R v, uint32_t) { st_pop<NT>(p, off, v); }
};
template <> struct Pop<D3Q19_STORE_FP64> {
    using T = double;
    using R = double;
    static constexpr bool SR = false;
    static constexpr int SR_BITS = 0;
    template <int QI, bool NT = false, class PT> __device__ static R ld(PT p, unsigned off) { return ld_pop<NT>(p, off); }
    template <int QI, class PT> __device__ static double ldd(PT p, unsigned off) { return *plane_at(p, off); }
    template <int QI, bool NT = false, class PT> __device__ static void st(PT p, unsigned off, R v) { st_pop<NT>(p, off, v); }
    template <int QI, bool NT = false, class PT> __device__ static void st_sr(PT p, unsigned off, R v, uint32_t) { st_pop<NT>(p, off, v); }
};
// binary16 of the relative deviation (f_q - w_q)/w_q scaled by 2^14: deviations down to ~1e-8
// stay normal numbers and the headroom reaches |dev| = 4.  (A power-of-two scale per population,
// which would make the shift and scale one exact FMA, measured much less accurate over long runs:
// relative viscosity 1.7e-3 from the fp64 build after 5000 steps of the 128^3 power-law workload,
// against 1.6e-5 with this scale.)
template <> struct Pop<D3Q19_STORE_FP16_SHIFT> {
    using T = _Float16;
    using R = float;
    template <int QI, bool NT = false, class PT> __device__ static R ld(PT p, unsigned off) {
        constexpr float w = (float)wq(QI);
        constexpr float sc = (float)(wq(QI) / (double)FP16_SCALE);
        return fmaf((float)ld_pop<NT>(p, off), sc, w);
    }
    // decoded in double around the float32 w_q the encoding is relative to
    template <int QI, class PT> __device__ static double ldd(PT p, unsigned off) {
        return (double)(float)wq(QI) + (double)*plane_at(p, off) * (wq(QI) / (double)FP16_SCALE);
    }
    static constexpr bool SR = true;
    static constexpr int SR_BITS = 13;
    template <int QI, bool NT = false, class PT> __device__ static void st(PT p, unsigned off, R v) {
        constexpr float w = (float)wq(QI);
        constexpr float isc = (float)((double)FP16_SCALE / wq(QI));
        st_pop<NT>(p, off, (_Float16)((v - w) * isc));
    }
    // the 13 magnitude bits that binary16 drops, plus 13 random bits, carry into the kept ones with
    // probability equal to the dropped fraction; the truncated value then converts exactly (normal
    // binary16 range; |f - w|/w below 2^-28 also rounds to nearest, which is immaterial).  The
    // collide measured 2% faster this way than with v_cvt_sr_f16_f32.
    template <int QI, bool NT = false, class PT> __device__ static void st_sr(PT p, unsigned off, R v, uint32_t r) {
        constexpr float w = (float)wq(QI);
        constexpr float isc = (float)((double)FP16_SCALE / wq(QI));
        const uint32_t b = (__float_as_uint((v - w) * isc) + r) & 0xFFFFE000u;
        st_pop<NT>(p, off, (_Float16)__uint_as_float(b));
    }
};
// bfloat16 of f_q - w_q
template <> struct Pop<D3Q19_STORE_BF16_SHIFT> {
    using T = uint16_t;
    using R = float;
    template <int QI, bool NT = false, class PT> __device__ static R ld(PT p, unsigned off) {
        constexpr float w = (float)wq(QI);
        return w + bf16_to_f32(ld_pop<NT>(p, off));
    }
    template <int QI, class PT> __device__ static double ldd(PT p, unsigned off) {
        return (double)(float)wq(QI) + (double)bf16_to_f32(*plane_at(p, off));
    }
    static constexpr bool SR = true;
    static constexpr int SR_BITS = 16;
    template <int QI, bool NT = false, class PT> __device__ static void st(PT p, unsigned off, R v) {
        constexpr float w = (float)wq(QI);
        st_pop<NT>(p, off, f32_to_bf16(v - w));
    }
    // the 16 discarded magnitude bits plus 16 random bits carry into the kept ones with
    // probability equal to the discarded fraction (finite inputs)
    template <int QI, bool NT = false, class PT> __device__ static void st_sr(PT p, unsigned off, R v, uint32_t r) {
        constexpr float w = (float)wq(QI);
        st_pop<NT>(p, off, (uint16_t)((__float_as_uint(v - w) + r) >> 16));
    }
};

// 32-bit finaliser of MurmurHash3 (Appleby): a bijection whose output bits each depend on every
// input bit
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// Random bits of population q for stochastic rounding at one node, B of them in the low bits: the
// top B bits of one of seven products of the node's hash h with fixed odd constants, read through
// three rotations (11 bits apart) so that the bits that decide a rounding differ between the
// populations.  Each rotated read is one funnel shift (v_alignbit_b32) and a mask: written as shifts
// of the product, the compiler turned every rotation into another quarter-rate multiply (19 per node
// instead of 7).
template <int q, int B> __device__ __forceinline__ uint32_t sr_bits(uint32_t h) {
    constexpr uint32_t K[7] = {0x9E3779B1u, 0x85EBCA77u, 0xC2B2AE3Du, 0x27D4EB2Fu, 0x165667B1u, 0xD3A2646Du, 0xFD7046C5u};
    const uint32_t x = h * K[q / 3];
    // (rotate left by 11 (q mod 3)) >> (32 - B)  ==  (rotate right by (32 - B - 11 (q mod 3)) mod 32) & (2^B - 1)
    constexpr unsigned sh = (unsigned)((32 - B - 11 * (q % 3) + 64) % 32);
    return __builtin_amdgcn_alignbit(x, x, sh) & ((1u << B) - 1u);
}

// Arithmetic helpers: the float build uses the hardware reciprocal, square root, exp2 and log2
// (about 1 ulp); the double build (the fp64 oracle build) keeps IEEE operations.
template <class R> __device__ __forceinline__ R rcp_(R x) { return (R)1 / x; }
template <> __device__ __forceinline__ float rcp_<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <class R> __device__ __forceinline__ R sqrt_(R x) { return sqrt(x); }
template <> __device__ __forceinline__ float sqrt_<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <class R> __device__ __forceinline__ R pow_(R x, R y) { return pow(x, y); }
template <> __device__ __forceinline__ float pow_<float>(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));  // x > 0 here (x >= 1e-8)
}
template <class R> __device__ __forceinline__ R exp_(R x) { return exp(x); }
template <> __device__ __forceinline__ float exp_<float>(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

// ---- moving-wall unknowns (GPU_program.cl:737-865) -----------------------------------------------
// Axis A, side S (0: lower wall, inward +A; 1: upper wall, inward -A).  Knowns: populations with
// zero or outward A-component.  Unknown q = its opposite + rho*(c_q . U)/3 (axis) or /6
// (diagonal); U = wall velocity with the A-component replaced by the inward normal velocity u_n,
// obtained from the knowns with rho = 1 (calc_rho = 0, sim_main.c:216) or, with calc_rho = 1,
// rho from the knowns and the prescribed u_n.  Once after an upload the raw uploaded unknowns
// are used instead (layout [side][k][plane], k = rank of q among the unknowns).
// Unknown values are selected, not branched on, so every store to f[] has a fixed index.
template <int A, int S, class R, class RWP>
__device__ __forceinline__ void wall_layer(R* f, const float* uw, int calc_rho, int use_raw, RWP raw, int pidx, int psz) {
    constexpr int inward = S == 0 ? 1 : -1;
    constexpr int a1 = A == 0 ? 1 : 0;
    constexpr int a2 = A == 2 ? 1 : 2;
    R sum0 = 0, sumout = 0;
    Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        if constexpr (cc(q, A) == 0) sum0 += f[q];
        if constexpr (cc(q, A) == -inward) sumout += f[q];
    });
    R un, rho;
    if (calc_rho) {
        un = (R)(S == 0 ? uw[A] : -uw[A]);
        rho = (sum0 + (R)2 * sumout) / ((R)1 - un);
    } else {
        rho = (R)1;
        un = (R)1 - (sum0 + (R)2 * sumout);
    }
    const R ua1 = (R)uw[a1], ua2 = (R)uw[a2];
    int k = 0;
    Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        if constexpr (cc(q, A) == inward) {
            constexpr int o = opposite(q);
            constexpr bool axial = (cc(q, a1) == 0 && cc(q, a2) == 0);
            R cu = un;
            if constexpr (cc(q, a1) != 0) cu = cu + signed_<cc(q, a1)>(ua1);
            if constexpr (cc(q, a2) != 0) cu = cu + signed_<cc(q, a2)>(ua2);
            R v = fma_(rho * cu, axial ? (R)(1.0 / 3.0) : (R)(1.0 / 6.0), f[o]);
            if (use_raw) v = (R)raw[((size_t)(S * 5 + k)) * psz + pidx];
            f[q] = v;
            k++;
        }
    });
}

__host__ __device__ inline int plane_size(const Geom& g, int A) {
    return A == 0 ? g.ny * g.nz : (A == 1 ? g.nx * g.nz : g.nx * g.ny);
}

// Gather the pre-collision populations of node (x,y,z): pull from the post-collision field with
// periodic wrap, then the wall unknowns (rebuilt, or the raw uploaded values once after an upload),
// then the optional tangential-wall passes (sim_main.c:388-404).
// RO = double decodes the stored populations in double precision (observables).
// AM: how the addresses are formed (pop_index is the layout, either form): ADDR_ZPLANE, z uniform
// over the wavefront (the single-simulation collide tiles): the source plane goes into each
// population's scalar base and the lane offset spans one plane; ADDR_CELL32: a 32-bit lane offset
// over the whole lattice (batch members); ADDR_WIDE: any z per lane, the plane term in 64 bits.
template <int ST, int AM = ADDR_WIDE, bool NT = false, class RO = typename Pop<ST>::R, class PIn = const typename Pop<ST>::T*>
__device__ __forceinline__ void gather_pre(PIn fin, const Geom& g, const FluidParams& fl,
                                           const void* bcraw, int use_raw, int x, int y, int z, RO* f) {
    using P = Pop<ST>;
    using R = RO;
    using RW = typename P::R;  // type of the raw wall-unknown buffer
    // Neighbour offsets with periodic wrap.  A source beyond a wall is replaced by the node
    // itself: those populations are wall unknowns, rebuilt below, so every load is in bounds and
    // no load is predicated.  An x shift is +-1 element (periodic: across the row); a row's offset
    // within its plane, and so a y shift, is the same for every population.
    const int oxm = (x > 0) ? -1 : (g.per[0] ? g.nx - 1 : 0);
    const int oxp = (x < g.nx - 1) ? 1 : (g.per[0] ? -(g.nx - 1) : 0);
    const int rmask = (1 << g.lr) - 1;
    auto row_at = [&](int yy) __attribute__((always_inline)) { return (yy >> g.lr) * g.cs + (yy & rmask) * g.nx; };
    const int e0 = row_at(y);
    const int oym = row_at((y > 0) ? y - 1 : (g.per[1] ? g.ny - 1 : y)) - e0;
    const int oyp = row_at((y < g.ny - 1) ? y + 1 : (g.per[1] ? 0 : y)) - e0;
    const auto f0 = (const D3Q19_GLOBAL typename P::T*)fin;
    auto load = [&](auto qc, const D3Q19_GLOBAL typename P::T* base, unsigned off) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        if constexpr (std::is_same<RO, typename P::R>::value) f[q] = P::template ld<q, NT>(base, off);
        else f[q] = P::template ldd<q>(base, off);
    };
    if constexpr (AM == ADDR_CELL32) {
        const int ps = (int)g.ps;
        const int ozm = (z > 0) ? -ps : (g.per[2] ? (g.nz - 1) * ps : 0);
        const int ozp = (z < g.nz - 1) ? ps : (g.per[2] ? -(g.nz - 1) * ps : 0);
        const int n = ps * z + e0 + x;
        Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            // pull: source node = n - c_q
            const int ox = cx(q) == 0 ? 0 : (cx(q) > 0 ? oxm : oxp);
            const int oy = cy(q) == 0 ? 0 : (cy(q) > 0 ? oym : oyp);
            const int oz = cz(q) == 0 ? 0 : (cz(q) > 0 ? ozm : ozp);
            load(qc, f0 + q * g.qs, (unsigned)(n + ox + oy + oz));
        });
    } else {
        const int zz = AM == ADDR_ZPLANE ? __builtin_amdgcn_readfirstlane(z) : z;
        const int zm = (zz > 0) ? zz - 1 : (g.per[2] ? g.nz - 1 : zz);
        const int zp = (zz < g.nz - 1) ? zz + 1 : (g.per[2] ? 0 : zz);
        const int n = e0 + x;
        Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            const int ox = cx(q) == 0 ? 0 : (cx(q) > 0 ? oxm : oxp);
            const int oy = cy(q) == 0 ? 0 : (cy(q) > 0 ? oym : oyp);
            const int zs = cz(q) == 0 ? zz : (cz(q) > 0 ? zm : zp);
            load(qc, f0 + q * g.qs + zs * g.ps, (unsigned)(n + ox + oy));
        });
    }
    const auto rw = to_global<RW>(bcraw);
    const int cr = fl.calc_rho;
    if (g.wall == 0) {
        if (x == 0) wall_layer<0, 0, R>(f, fl.uw[0], cr, use_raw, rw, y + g.ny * z, g.ny * g.nz);
        if (x == g.nx - 1) wall_layer<0, 1, R>(f, fl.uw[1], cr, use_raw, rw, y + g.ny * z, g.ny * g.nz);
    }
    if (g.wall == 1) {
        if (y == 0) wall_layer<1, 0, R>(f, fl.uw[0], cr, use_raw, rw, x + g.nx * z, g.nx * g.nz);
        if (y == g.ny - 1) wall_layer<1, 1, R>(f, fl.uw[1], cr, use_raw, rw, x + g.nx * z, g.nx * g.nz);
    }
    if (g.wall == 2) {
        if (z == 0) wall_layer<2, 0, R>(f, fl.uw[0], cr, use_raw, rw, x + g.nx * y, g.nx * g.ny);
        if (z == g.nz - 1) wall_layer<2, 1, R>(f, fl.uw[1], cr, use_raw, rw, x + g.nx * y, g.nx * g.ny);
    }
    // tangential passes in axis order after the wall pass (sim_main.c:388-404); skipped while the
    // raw uploaded unknowns are in use, since the periodic gather then returns the uploaded values
    if (fl.tan_mask && !use_raw) {
        if ((fl.tan_mask & 1) && x == 0) wall_layer<0, 0, R>(f, fl.uw[0], 0, 0, rw, 0, 1);
        if ((fl.tan_mask & 1) && x == g.nx - 1) wall_layer<0, 1, R>(f, fl.uw[1], 0, 0, rw, 0, 1);
        if ((fl.tan_mask & 2) && y == 0) wall_layer<1, 0, R>(f, fl.uw[0], 0, 0, rw, 0, 1);
        if ((fl.tan_mask & 2) && y == g.ny - 1) wall_layer<1, 1, R>(f, fl.uw[1], 0, 0, rw, 0, 1);
        if ((fl.tan_mask & 4) && z == 0) wall_layer<2, 0, R>(f, fl.uw[0], 0, 0, rw, 0, 1);
        if ((fl.tan_mask & 4) && z == g.nz - 1) wall_layer<2, 1, R>(f, fl.uw[1], 0, 0, rw, 0, 1);
    }
}

// ---- relaxation time law (compute_tau, GPU_program.cl:1041-1128) ----------------------------------
// Clamp: tau_min (reference 0.505), optional tau_max.  Casson with ViscosityParams[2] = m > 0 is
// Papanastasiou-regularised: nu = (sqrt(eta) + sqrt(tauY/g)(1 - exp(-sqrt(m g))))^2.
template <class R>
__device__ __forceinline__ R tau_law(const FluidParams& fl, R srt) {
    R tau = (R)fl.tau0;
    const int model = fl.model;
    if (model != D3Q19_VISC_POWER_LAW && model != D3Q19_VISC_CASSON && model != D3Q19_VISC_HB) return tau;
    srt = srt > (R)1e-8 ? srt : (R)1e-8;
    const R isrt = rcp_(srt);
    R nu;
    if (model == D3Q19_VISC_POWER_LAW) {
        const R k = (R)fl.vp[0], n = (R)fl.vp[1];
        if (n == (R)0.5) nu = k * rcp_(sqrt_(srt));
        else if (n == (R)1.0) nu = k;
        else if (n == (R)2.0) nu = k * srt;
        else nu = k * pow_(srt, n - (R)1);
    } else if (model == D3Q19_VISC_CASSON) {
        const R ty = (R)fl.vp[0], eta = (R)fl.vp[1], m = (R)fl.vp[2];
        R a = sqrt_(ty * isrt);
        if (m > (R)0) a = fma_(-a, exp_(-sqrt_(m * srt)), a);
        const R b = a + sqrt_(eta);
        nu = b * b;
    } else {
        const R ty = (R)fl.vp[0], k = (R)fl.vp[1], n = (R)fl.vp[2];
        if (n == (R)0.5) nu = fma_(ty, isrt, k * rcp_(sqrt_(srt)));
        else if (n == (R)1.0) nu = fma_(ty, isrt, k);
        else if (n == (R)2.0) nu = fma_(ty, isrt, k * srt);
        else nu = fma_(ty, isrt, k * pow_(srt, n - (R)1));
    }
    tau = fma_((R)3, nu, (R)0.5);
    if (tau < (R)fl.tau_min) tau = (R)fl.tau_min;
    if (fl.tau_max > 0.f && tau > (R)fl.tau_max) tau = (R)fl.tau_max;
    return tau;
}

// MRT relaxation of v: out = M^-1 S M v with M^-1 = M^T diag(1/|row|^2).
// Rates of GPU_program.cl:453 with 1/tau on moments 8, 9, 10, 14, 15 (GPU_program.cl:567-568).
template <class R>
__device__ __forceinline__ void mrt_relax(const R* v, R itau, R* out) {
    R a[Q];
    Unroll<0, Q>::run([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value;
        R m = (R)0;
        Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            constexpr int c = mrow(k, q);
            if constexpr (c == 1) m += v[q];
            else if constexpr (c == -1) m -= v[q];
            else if constexpr (c != 0) m = fma_((R)c, v[q], m);
        });
        constexpr double rate = (k == 0) ? 0.0 : (k == 1) ? 1.19 : (k == 2 || k == 12 || k == 13) ? 1.40
                              : (k == 4 || k == 6 || k == 11) ? 1.20 : (k >= 16) ? 1.98 : 1.0;
        constexpr bool tk = (k == 8 || k == 9 || k == 10 || k == 14 || k == 15);
        constexpr double inv_norm = 1.0 / (double)mnorm(k);
        if constexpr (k == 0) a[k] = (R)0;
        else if constexpr (tk) a[k] = m * itau * (R)inv_norm;
        else a[k] = m * (R)(rate * inv_norm);
    });
    Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        R o = (R)0;
        Unroll<0, Q>::run([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            constexpr int c = mrow(k, q);
            if constexpr (c == 1) o += a[k];
            else if constexpr (c == -1) o -= a[k];
            else if constexpr (c != 0) o = fma_((R)c, a[k], o);
        });
        out[q] = o;
    });
}

// strain-rate invariant from the collision change ss (GPU_program.cl:520-557)
template <class R>
__device__ __forceinline__ R strain_rate(const R* ss, R ux, R uy, R uz, R gx, R gy, R gz, R inv_rho) {
    R s[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    Unroll<1, Q>::run([&](auto qc) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        Unroll<0, 3>::run([&](auto ac) __attribute__((always_inline)) {
            constexpr int a = decltype(ac)::value;
            Unroll<a, 3>::run([&](auto bc) __attribute__((always_inline)) {
                constexpr int b = decltype(bc)::value;
                constexpr int w = cc(q, a) * cc(q, b);
                if constexpr (w == 1) s[a][b] += ss[q];
                else if constexpr (w == -1) s[a][b] -= ss[q];
            });
        });
    });
    const R u[3] = {ux, uy, uz}, g[3] = {gx, gy, gz};
    R sum = 0;
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
            const R v = s[a][b] - (u[a] * g[b] + u[b] * g[a]);
            sum += (a == b) ? v * v : (R)2 * v * v;
        }
    return sqrt_(sum) * (R)2.1213203435596424 * inv_rho;
}

// One lattice update at interior node (x, y, z).  VT: 0 constant tau, 1 per-node tau, 2 per-node tau
// when a.vt is set (batched cells of mixed rheology).  probe = 1 (autotuning trials) leaves the
// pending spread force, the tau field and u untouched so a trial launch does not change the state
// a later step reads (its populations go to the lattice the next step overwrites).
template <int ST, int COLL, int VT, int NT, bool SBASE_STORES = true>
__device__ __forceinline__ void collide_node(const CollideArgs& a, const void* fin, void* fout, int write_u_all, int x, int y,
                                             int z, int probe) {
    using P = Pop<ST>;
    using R = typename P::R;
    using T = typename P::T;
    const Geom& g = a.g;
    const int n = x + g.nx * (y + g.ny * z);

    // The node's force tag (set by the spreading wherever it added force) and its brick's mark (u
    // is stored where the next IBM pass reads it) are loaded before the populations, so that their
    // latency overlaps the population loads.  Both loads are unconditional (a flag that is off
    // points them at the first word of fin, always readable, and the value is then ignored): a
    // load made on one branch only makes the compiler wait for it where the branches join.  The
    // force itself is read only at tagged nodes: on the 256x256x128 workload about 0.7 M nodes
    // instead of the 3.2 M of the marked bricks.
    const int brick = (x >> 2) + g.bnx * ((y >> 2) + g.bny * (z >> 2));
    uint32_t ep = 0;
    const auto epochs = to_global<uint32_t>(a.epochs);
    if (a.use_umask) ep = epochs[0] + 1u;  // this step's epoch
    const bool force_on = a.has_force && !probe;
    const D3Q19_GLOBAL uint8_t* ftp = force_on ? to_global<uint8_t>(a.ftag) + n : to_global<uint8_t>(fin);
    const D3Q19_GLOBAL uint32_t* ump = a.use_umask ? to_global<uint32_t>(mask_of(a.mask0, a.mask1, a.mask2, ep)) + brick
                                                   : to_global<uint32_t>(fin);
    const uint32_t ftag = *ftp, umark = *ump;

    // The previous relaxation time (float) is copied into LDS by the memory unit while the
    // populations are in flight (global_load_lds: no register is held for it) and read from there
    // where the strain rate needs it.  Loaded there from memory, it cost each wavefront one more
    // memory round trip after the populations had arrived.  The wait for it is the one the first
    // use of the populations needs anyway.  Blocks have at most 256 threads; a lane's slot is
    // written and read by that lane's own wavefront only.
    const bool local_tau = VT == 1 || (VT == 2 && a.vt);
    const auto taup = to_global<R>(a.tau);
    constexpr bool stage_tau = sizeof(R) == 4;
    __shared__ float tau_stage[256];
    if constexpr (stage_tau) {
        if (local_tau)
            __builtin_amdgcn_global_load_lds((const void*)(taup + n), (__attribute__((address_space(3))) void*)&tau_stage[threadIdx.x & ~63u], 4,
                                             0, 0);
    }

    R f[Q];
    gather_pre<ST, SBASE_STORES ? ADDR_ZPLANE : ADDR_CELL32, NT == 2>(to_global<T>(fin), g, a.fl, a.bcraw, a.use_bcraw, x, y, z, f);
    if constexpr (stage_tau) {
        if (local_tau) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }

    R gx = (R)a.fl.gconst[0], gy = (R)a.fl.gconst[1], gz = (R)a.fl.gconst[2];
    if (force_on && ftag != 0u) {
        const auto fa = to_global<long long>(a.facc) + n;
        const long long vx = fa[0], vy = fa[g.nint], vz = fa[2 * (long long)g.nint];
        if ((vx | vy | vz) != 0) {
            gx += (R)((double)vx * FIX_INV);
            gy += (R)((double)vy * FIX_INV);
            gz += (R)((double)vz * FIX_INV);
            fa[0] = 0;
            fa[g.nint] = 0;
            fa[2 * (long long)g.nint] = 0;
        }
        to_global<uint8_t>(a.ftag)[n] = 0;
    }

    // density and momentum summed in the reference's order (GPU_program.cl:343-350, 387-394); the
    // strain rate takes the pair sums f_q + f_qb
    R rho = (R)0;
    Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) { rho += f[decltype(qc)::value]; });
    const R jx = f[1] - f[2] + f[7] + f[8] + f[9] + f[10] - f[11] - f[12] - f[13] - f[14];
    const R jy = f[3] - f[4] + f[7] - f[8] + f[11] - f[12] + f[15] + f[16] - f[17] - f[18];
    const R jz = f[5] - f[6] + f[9] - f[10] + f[13] - f[14] + f[15] - f[16] + f[17] - f[18];
    const R inv_rho = rcp_(rho);
    const R ux = fma_((R)0.5, gx, jx) * inv_rho;
    const R uy = fma_((R)0.5, gy, jy) * inv_rho;
    const R uz = fma_((R)0.5, gz, jz) * inv_rho;
    if (!probe && (write_u_all || (a.use_umask && umark == ep))) {
        // interleaved (x, y, z) per node: one write stream instead of three (the three separate
        // component arrays cost 7.5% of the fp32 collide on the 256x256x128 workload, whose marked
        // bricks hold a third of the nodes)
        const auto u = to_global<R>(a.u) + 3 * (long long)n;
        u[0] = ux;
        u[1] = uy;
        u[2] = uz;
    }

    // equilibrium (GPU_program.cl:978-1007), non-equilibrium n = feq - f, Guo term (1010-1038)
    const R usq = ux * ux + uy * uy + uz * uz;
    const R udg = ux * gx + uy * gy + uz * gz;
    R tau = (R)a.fl.tau0;
    R fo[Q];
    if constexpr (COLL == D3Q19_COLL_BGK) {
        // BGK in pair form.  A direction q and its opposite qb have c_qb = -c_q, so
        //   f_eq(q), f_eq(qb) = S +- A,   S = w rho (1 - 1.5 u.u + 4.5 (c.u)^2),   A = 3 w rho (c.u)
        //   F(q), F(qb)       = G +- H,   G = w' (3 (c.u)(c.g) - u.g),             H = w' (c.g)
        // (w' = 1/6 on the axes, 1/12 on the diagonals; the Guo term of GPU_program.cl:1010-1038),
        // each evaluated once per pair; the update is f + (f_eq - f)/tau + (1 - 1/(2 tau)) F as in
        // GPU_program.cl:943-944, which keeps the small non-equilibrium part f_eq - f exact to a
        // rounding of f_eq.
        const R c0 = fma_((R)-1.5, usq, (R)1);
        const R S0 = rho * (R)(1.0 / 3.0) * c0;
        R S[9], A[9], G[9], H[9];
        Unroll<0, 9>::run([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            constexpr int q = pair_dir(k);
            const R cu = cdot<q>(ux, uy, uz);
            const R cg = cdot<q>(gx, gy, gz);
            const R wr = rho * (R)wq(q);
            constexpr double wg = q < 7 ? 1.0 / 6.0 : 1.0 / 12.0;
            S[k] = wr * fma_((R)4.5 * cu, cu, c0);
            A[k] = ((R)3 * wr) * cu;
            G[k] = (R)wg * fma_((R)3 * cu, cg, -udg);
            H[k] = (R)wg * cg;
        });
        if (local_tau) {
            // collision change of BGK: f_neq/tau + (1 - 1/(2 tau)) F.  Its second moment minus
            // (uF + Fu) is (Pi(f_neq) - (uF + Fu)/2)/tau because the Guo term's second moment is
            // uF + Fu; Pi(f_neq) takes only the pair sums f_neq(q) + f_neq(qb) = 2 S - f_q - f_qb.
            const R told = stage_tau ? (R)tau_stage[threadIdx.x] : taup[n];
            const R itold = rcp_(told);
            R pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
            Unroll<0, 9>::run([&](auto kc) __attribute__((always_inline)) {
                constexpr int k = decltype(kc)::value;
                constexpr int q = pair_dir(k), qb = opposite(q);
                const R sp = fma_((R)2, S[k], -(f[q] + f[qb]));
                if constexpr (cx(q) != 0) pxx += sp;
                if constexpr (cy(q) != 0) pyy += sp;
                if constexpr (cz(q) != 0) pzz += sp;
                if constexpr (cx(q) * cy(q) == 1) pxy += sp;
                if constexpr (cx(q) * cy(q) == -1) pxy -= sp;
                if constexpr (cx(q) * cz(q) == 1) pxz += sp;
                if constexpr (cx(q) * cz(q) == -1) pxz -= sp;
                if constexpr (cy(q) * cz(q) == 1) pyz += sp;
                if constexpr (cy(q) * cz(q) == -1) pyz -= sp;
            });
            const R sxx = fma_(-ux, gx, pxx), syy = fma_(-uy, gy, pyy), szz = fma_(-uz, gz, pzz);
            const R sxy = fma_((R)-0.5, fma_(ux, gy, uy * gx), pxy);
            const R sxz = fma_((R)-0.5, fma_(ux, gz, uz * gx), pxz);
            const R syz = fma_((R)-0.5, fma_(uy, gz, uz * gy), pyz);
            const R n2 = fma_(sxx, sxx, fma_(syy, syy, fma_(szz, szz, (R)2 * fma_(sxy, sxy, fma_(sxz, sxz, syz * syz)))));
            tau = tau_law<R>(a.fl, sqrt_(n2) * itold * (R)2.1213203435596424 * inv_rho);
            if (!probe) taup[n] = tau;
        }
        const R itau = rcp_(tau);
        const R pf = fma_((R)-0.5, itau, (R)1);
        fo[0] = fma_(pf, -udg, fma_(S0 - f[0], itau, f[0]));
        Unroll<0, 9>::run([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            constexpr int q = pair_dir(k), qb = opposite(q);
            fo[q] = fma_(pf, G[k] + H[k], fma_((S[k] + A[k]) - f[q], itau, f[q]));
            fo[qb] = fma_(pf, G[k] - H[k], fma_((S[k] - A[k]) - f[qb], itau, f[qb]));
        });
    } else {
        R ne[Q], fg[Q];
        const R c0 = fma_((R)-1.5, usq, (R)1);
        Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            const R cu = cdot<q>(ux, uy, uz);
            const R cg = cdot<q>(gx, gy, gz);
            R feq;
            if constexpr (q == 0) {
                feq = rho * (R)(1.0 / 3.0) * c0;
                fg[q] = -udg;
            } else {
                feq = rho * (R)wq(q) * fma_((R)4.5 * cu, cu, fma_((R)3, cu, c0));
                fg[q] = fma_((R)3 * cu, cg, cg - udg) * (R)(q < 7 ? 1.0 / 6.0 : 1.0 / 12.0);
            }
            ne[q] = feq - f[q];
        });
        R v[Q];
        Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            v[q] = fma_((R)-0.5, fg[q], ne[q]);
        });
        if (local_tau) {
            const R told = stage_tau ? (R)tau_stage[threadIdx.x] : taup[n];
            R ss[Q];
            mrt_relax<R>(v, rcp_(told), ss);
            Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) { ss[decltype(qc)::value] += fg[decltype(qc)::value]; });
            tau = tau_law<R>(a.fl, strain_rate<R>(ss, ux, uy, uz, gx, gy, gz, inv_rho));
            if (!probe) taup[n] = tau;
        }
        R o[Q];
        mrt_relax<R>(v, rcp_(tau), o);
        Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            fo[q] = f[q] + o[q] + fg[q];
        });
    }
    // SBASE_STORES (the single-simulation tiles, whose z is uniform over the block): the plane
    // goes into the scalar base (ADDR_ZPLANE), and the stores' lane offset is redefined in each
    // store block by an empty asm statement (no instruction is emitted); otherwise the compiler
    // computes its 64-bit form once, before the branch, and every store adds it to the base with
    // a 64-bit VALU add instead of taking the scalar-base form.  The single-simulation step gains
    // (fp16 128^3: 60.8 -> 58.1 us); the batched collide (ADDR_CELL32) lost 4.7% with it and keeps
    // the 64-bit adds.
    const int erow = (y >> g.lr) * g.cs + (y & ((1 << g.lr) - 1)) * g.nx + x;  // pop_index less the q and plane terms
    D3Q19_GLOBAL T* out;
    unsigned ns;
    if constexpr (SBASE_STORES) {
        out = to_global<T>(fout) + (long long)__builtin_amdgcn_readfirstlane(z) * g.ps;
        ns = (unsigned)erow;
    } else {
        out = to_global<T>(fout);
        ns = (unsigned)((int)g.ps * z + erow);
    }
    if (P::SR && a.fl.sr) {
        // random bits from the node index and the bit patterns of the incoming populations: the
        // same for a cell alone and in a batch, and new whenever an input changes
        // (three-input exclusive ors: gfx950's v_bitop3_b32 with truth table 0x96)
        uint32_t x = ((uint32_t)n * 0x9E3779B1u) ^ __float_as_uint((float)f[0]);
        Unroll<0, 9>::run([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            x = __builtin_amdgcn_bitop3_b32(x, __float_as_uint((float)f[2 * k + 1]), __float_as_uint((float)f[2 * k + 2]), 0x96);
        });
        const uint32_t h = fmix32(x);
        if constexpr (SBASE_STORES) asm volatile("" : "+v"(ns));
        Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            P::template st_sr<q, NT != 0>(out + q * g.qs, ns, fo[q], sr_bits<q, P::SR_BITS>(h));
        });
    } else {
        if constexpr (SBASE_STORES) asm volatile("" : "+v"(ns));
        Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            P::template st<q, NT != 0>(out + q * g.qs, ns, fo[q]);
        });
    }
}

// Tile L of a single simulation's collide -> this lane's node; false past the lattice.  A tile is a
// bx x by block of one plane (x fastest), or with a.packed bx consecutive nodes of the plane in node
// order (a wavefront may then span two rows: its accesses are two contiguous stretches per
// population, each row its own chunk).  The float quotient is exact to +-1 for planes below 2^24
// nodes and corrected.
__device__ __forceinline__ bool tile_node(const CollideArgs& a, int L, int& x, int& y, int& z) {
    const int per_plane = a.nbx * a.nby;
    const int zl = L / per_plane;  // plane within this launch (a z-slab from plane a.z0)
    z = a.z0 + zl;
    const int rem = L - zl * per_plane;
    if (a.packed) {
        const int i = rem * a.bx + (int)threadIdx.x;
        int yy = (int)((float)i * a.inv_nx);
        int xx = i - yy * a.g.nx;
        if (xx < 0) { yy--; xx += a.g.nx; } else if (xx >= a.g.nx) { yy++; xx -= a.g.nx; }
        x = xx;
        y = yy;
    } else {
        const int tyb = rem / a.nbx;
        const int txb = rem - tyb * a.nbx;
        x = txb * a.bx + (int)(threadIdx.x % a.bx);
        y = tyb * a.by + (int)(threadIdx.x / a.bx);
    }
    return x >= 0 && x < a.g.nx && y < a.g.ny && z < a.g.nz && z >= 0;
}

template <int ST, int COLL, int VT, int NT>
__global__ void __launch_bounds__(256) collide_kernel(CollideArgs a) {
    // Tiles are taken in dispatch order (x fastest, then y, then z).  An XCD-contiguous remap of
    // the tiles (each XCD owning a z-slab) measured 2-7% slower on MI355X for 8-34 M node lattices:
    // with plain order all XCDs stream neighbouring addresses and the shared Infinity Cache serves
    // the y/z halo rows.
    int x, y, z;
    if (!tile_node(a, blockIdx.x, x, y, z)) return;
    collide_node<ST, COLL, VT, NT>(a, a.fin, a.fout, a.write_u_all, x, y, z, 0);
}

// Fused step: blocks [0, nsph) run the spheres' turn (sphere_turn: the force sums and contacts of
// the last step when do_post, the dynamics into the other position buffer, the dilated marks of
// the epoch after this one); the other blocks are the collide tiles of collide_kernel.  Nothing a
// sphere block writes is read by a collide block of the same launch: the collide reads the marks of
// this epoch and the pending force with its tags (written by the last IBM pass), the turn writes the
// marks of the next epoch (another array) and the other position and velocity buffers.
template <int ST, int COLL, int VT, int NT>
__device__ __forceinline__ void step_body(const CollideArgs& a, const PartArgs& p, int nsph, int do_post) {
    if ((int)blockIdx.x < nsph) {
        sphere_turn(p, blockIdx.x, do_post, p.epochs[0] + 2u);
        return;
    }
    int x, y, z;
    if (!tile_node(a, (int)blockIdx.x - nsph, x, y, z)) return;
    collide_node<ST, COLL, VT, NT>(a, a.fin, a.fout, a.write_u_all, x, y, z, 0);
}

template <int ST, int COLL, int VT, int NT>
__global__ void __launch_bounds__(256) step_kernel(CollideArgs a, PartArgs p, int nsph, int do_post) {
    step_body<ST, COLL, VT, NT>(a, p, nsph, do_post);
}

// The same step with a register budget: 5 waves per SIMD for the BGK form in fp32 (85 VGPRs, no
// spill; the 256x256x128 Casson workload: 254.2-255.5 us against 257.9-258.9 compiled for 6 waves,
// 80 VGPRs, in alternating runs) ...
template <int ST, int COLL, int VT, int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) step_kernel_w5(CollideArgs a, PartArgs p, int nsph,
                                                                                            int do_post) {
    step_body<ST, COLL, VT, NT>(a, p, nsph, do_post);
}

// ... and 7 waves per SIMD (72 VGPRs, 20 bytes spilled) for the 16-bit storages (fp16 on the same
// workload: 154.1-155.9 against 159.8-162.2 us at 6 waves in alternating runs)
template <int ST, int COLL, int VT, int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) step_kernel_w7(CollideArgs a, PartArgs p, int nsph,
                                                                                            int do_post) {
    step_body<ST, COLL, VT, NT>(a, p, nsph, do_post);
}

// flat node index of a cell -> (x, y, z); exact for n < 2^24 (float estimate within +-1, corrected)
__device__ __forceinline__ void split_node(const CellDesc& cd, int n, int& x, int& y, int& z) {
    int zz = (int)((float)n * cd.inv_nxy);
    int r = n - zz * cd.nxy;
    if (r < 0) { zz--; r += cd.nxy; } else if (r >= cd.nxy) { zz++; r -= cd.nxy; }
    int yy = (int)((float)r * cd.inv_nx);
    int xx = r - yy * cd.nx;
    if (xx < 0) { yy--; xx += cd.nx; } else if (xx >= cd.nx) { yy++; xx -= cd.nx; }
    x = xx; y = yy; z = zz;
}

// slot index of the batched collide -> (x, y, z); x >= nx is a padding slot
__device__ __forceinline__ void split_slot(const CellDesc& cd, int n, int& x, int& y, int& z) {
    int zz = (int)((float)n * cd.inv_pny);
    int r = n - zz * cd.pny;
    if (r < 0) { zz--; r += cd.pny; } else if (r >= cd.pny) { zz++; r -= cd.pny; }
    int yy = (int)((float)r * cd.inv_pitch);
    int xx = r - yy * cd.pitch;
    if (xx < 0) { yy--; xx += cd.pitch; } else if (xx >= cd.pitch) { yy++; xx -= cd.pitch; }
    x = xx; y = yy; z = zz;
}

// Batched collide+stream: a persistent grid walks the tile list of all cells (a tile is blockDim
// consecutive slots of one cell: rows of cd.pitch slots, of which the first nx are the row's
// nodes).  Cells are shear cells (periodic x and y, walls normal to z);
// each cell's tau law, wall speeds, spheres and lattice pointers come from its descriptor.  sp is
// the step parity: the cell's state is read from f[par ^ sp] and written to the other lattice.
// VT selects the tile range: 0 the constant-tau cells' tiles [0, ntiles_ct), 1 the per-node-tau
// cells' tiles [ntiles_ct, ntiles).  Two specialised launches keep the constant-tau cells free of
// the per-node-tau code's registers (occupancy).
template <int ST, int COLL, int VT, int NT>
__device__ __forceinline__ void collide_batch_tiles(const BatchTables& t, const CellDesc* __restrict__ cells, int sp, int write_u_all,
                                                    int probe, int first, int stride) {
    const int lo = VT ? t.cnt->ntiles_ct : 0;
    const int hi = VT ? t.cnt->ntiles : t.cnt->ntiles_ct;
    const int span = t.cnt->span;
    for (int k = lo + first; k < hi; k += stride) {
        const uint2 tl = t.tiles[k];
        const CellDesc& cd = cells[tl.x];
        const void* fin = cd.f[cd.par ^ sp];
        void* fout = (void*)cd.f[cd.par ^ sp ^ 1];
        for (int j = 0; j < span; j++) {
            const int sl = (int)tl.y + j * (int)blockDim.x + (int)threadIdx.x;
            if (sl >= cd.nslots) break;
            int x, y, z;
            split_slot(cd, sl, x, y, z);
            if (x < cd.nx) collide_node<ST, COLL, VT, NT, false>(cd.c, fin, fout, write_u_all, x, y, z, probe);
        }
    }
}

template <int ST, int COLL, int VT, int NT>
__global__ void __launch_bounds__(256) collide_batch_kernel(BatchTables t, const CellDesc* __restrict__ cells, int sp, int write_u_all,
                                                            int probe) {
    collide_batch_tiles<ST, COLL, VT, NT>(t, cells, sp, write_u_all, probe, blockIdx.x, gridDim.x);
}

// Batched step: blocks [0, nturn) walk the sphere list of all cells and run each sphere's turn
// (as step_body's sphere blocks: x(t), v(t) from p[par ^ sp], x(t+1), v(t+1) into the other
// buffer, the marks of the epoch after this one); the other blocks walk the collide tiles.  As in
// step_body, nothing a sphere block writes is read by a collide block of the same launch.
template <int ST, int COLL, int VT, int NT>
__device__ __forceinline__ void step_batch_body(const BatchTables& t, const CellDesc* __restrict__ cells, const int4* __restrict__ sph,
                                                int sp, int write_u_all, int nturn, int do_post) {
    if ((int)blockIdx.x < nturn) {
        const int nsph = t.cnt->nsph;
        const uint32_t e = t.epochs[0] + 2u;
        for (int k = blockIdx.x; k < nsph; k += nturn) {
            const int4 d = sph[k];
            const CellDesc& cd = cells[d.x];
            sphere_turn(cd.p[cd.par ^ sp], d.y, do_post, e);
            __syncthreads();
        }
        return;
    }
    collide_batch_tiles<ST, COLL, VT, NT>(t, cells, sp, write_u_all, 0, (int)blockIdx.x - nturn, (int)gridDim.x - nturn);
}

template <int ST, int COLL, int VT, int NT>
__global__ void __launch_bounds__(256) step_batch_kernel(BatchTables t, const CellDesc* __restrict__ cells, const int4* __restrict__ sph,
                                                         int sp, int write_u_all, int nturn, int do_post) {
    step_batch_body<ST, COLL, VT, NT>(t, cells, sph, sp, write_u_all, nturn, do_post);
}

// the same with the register budget of 6 waves per SIMD (80 VGPRs; the sphere code spills 60
// bytes), for the batched BGK collide in fp16 / bf16 storage, and of 5 waves (96 VGPRs, 8 bytes spilled) in fp32:
// on the 24-cell C5 batch (fp32) 0.748 / 0.735 against 0.778 / 0.760 ms per step at 6, on the
// 64-cell fp16 ensemble 0.665 against 0.642 ms (so fp16 keeps 6)
template <int ST, int COLL, int VT, int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
step_batch_kernel_w6(BatchTables t, const CellDesc* __restrict__ cells, const int4* __restrict__ sph, int sp, int write_u_all, int nturn,
                     int do_post) {
    step_batch_body<ST, COLL, VT, NT>(t, cells, sph, sp, write_u_all, nturn, do_post);
}
template <int ST, int COLL, int VT, int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
step_batch_kernel_w5(BatchTables t, const CellDesc* __restrict__ cells, const int4* __restrict__ sph, int sp, int write_u_all, int nturn,
                     int do_post) {
    step_batch_body<ST, COLL, VT, NT>(t, cells, sph, sp, write_u_all, nturn, do_post);
}

// With events ev0/ev1, they are recorded at the kernel's own start and end (hipExtLaunchKernelGGL),
// so their interval is the kernel's duration without the dispatch gap before it.
template <int ST, int COLL, int VT, int NT>
static void launch_one_nt(const CollideArgs& a, int threads, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    if (ev0) hipExtLaunchKernelGGL((collide_kernel<ST, COLL, VT, NT>), dim3(a.nblk), dim3(threads), 0, st, ev0, ev1, 0, a);
    else hipLaunchKernelGGL((collide_kernel<ST, COLL, VT, NT>), dim3(a.nblk), dim3(threads), 0, st, a);
}
template <int ST, int COLL, int VT>
static void launch_one(const CollideArgs& a, int threads, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    if (a.nt_stores == 2) launch_one_nt<ST, COLL, VT, 2>(a, threads, st, ev0, ev1);
    else if (a.nt_stores == 1) launch_one_nt<ST, COLL, VT, 1>(a, threads, st, ev0, ev1);
    else launch_one_nt<ST, COLL, VT, 0>(a, threads, st, ev0, ev1);
}

template <int ST>
static void dispatch_coll(const CollideArgs& a, int coll, int vt, int threads, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    if (coll == D3Q19_COLL_BGK) {
        if (vt) launch_one<ST, D3Q19_COLL_BGK, 1>(a, threads, st, ev0, ev1);
        else launch_one<ST, D3Q19_COLL_BGK, 0>(a, threads, st, ev0, ev1);
    } else {
        if (vt) launch_one<ST, D3Q19_COLL_MRT, 1>(a, threads, st, ev0, ev1);
        else launch_one<ST, D3Q19_COLL_MRT, 0>(a, threads, st, ev0, ev1);
    }
}

CollideArgs make_collide_args(Sim* s, const void* fin, void* fout, int write_u_all, int use_maskU) {
    CollideArgs a;
    a.fin = fin;
    a.fout = fout;
    a.u = s->u;
    a.tau = s->tau;
    a.facc = s->facc;
    a.ftag = s->ftag;
    a.mask0 = s->mask[0];
    a.mask1 = s->mask[1];
    a.mask2 = s->mask[2];
    a.epochs = s->epochs;
    a.bcraw = s->bcraw;
    a.g = s->g;
    a.fl = s->fl;
    a.write_u_all = write_u_all;
    // the brick masks and the force accumulator exist only when there are spheres
    a.use_umask = (use_maskU && s->mask[0] != nullptr) ? 1 : 0;
    a.use_bcraw = s->bcraw_valid;
    a.has_force = (s->np > 0 && s->facc != nullptr && s->ftag != nullptr) ? 1 : 0;
    a.vt = s->opt.variable_viscosity ? 1 : 0;
    a.packed = s->packed_tiles;
    a.inv_nx = 1.0f / (float)s->g.nx;
    if (a.packed) {
        a.bx = 256;
        a.by = 1;
        a.nbx = (int)(((long long)s->g.nx * s->g.ny + 255) / 256);
        a.nby = 1;
    } else {
        a.bx = s->block[0];
        a.by = s->block[1];
        a.nbx = (s->g.nx + a.bx - 1) / a.bx;
        a.nby = (s->g.ny + a.by - 1) / a.by;
    }
    a.nblk = a.nbx * a.nby * s->g.nz;
    a.z0 = 0;
    a.nt_stores = s->nt_stores;
    return a;
}

int launch_collide(Sim* s, const void* fin, void* fout, int write_u_all, int use_maskU, hipStream_t st, hipEvent_t ev0,
                   hipEvent_t ev1) {
    const CollideArgs a = make_collide_args(s, fin, fout, write_u_all, use_maskU);
    const int threads = a.bx * a.by;
    const int c = s->opt.collision, vt = s->opt.variable_viscosity;
    switch (s->storage) {
    case D3Q19_STORE_FP16_SHIFT: dispatch_coll<D3Q19_STORE_FP16_SHIFT>(a, c, vt, threads, st, ev0, ev1); break;
    case D3Q19_STORE_BF16_SHIFT: dispatch_coll<D3Q19_STORE_BF16_SHIFT>(a, c, vt, threads, st, ev0, ev1); break;
    case D3Q19_STORE_FP64: dispatch_coll<D3Q19_STORE_FP64>(a, c, vt, threads, st, ev0, ev1); break;
    default: dispatch_coll<D3Q19_STORE_FP32>(a, c, vt, threads, st, ev0, ev1); break;
    }
    HIP_OK(hipGetLastError());
    s->bcraw_valid = 0;
    return 0;
}

// A kernel launch, with start/end events recorded at the kernel's own start and end when ev0 is set
template <class K, class... Args>
static void launch_maybe_timed(K kernel, dim3 grid, dim3 block, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1, Args... args) {
    if (ev0 && ev1) hipExtLaunchKernelGGL(kernel, grid, block, 0, st, ev0, ev1, 0, args...);
    else hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
}

template <int ST, int COLL, int VT, int NT>
static void launch_fused_nt(const CollideArgs& a, const PartArgs& p, int nsph, int do_post, int threads, hipStream_t st,
                            hipEvent_t ev0, hipEvent_t ev1) {
    const dim3 grid(a.nblk + nsph);
    if constexpr (COLL == D3Q19_COLL_BGK && (ST == D3Q19_STORE_FP16_SHIFT || ST == D3Q19_STORE_BF16_SHIFT))
        launch_maybe_timed(step_kernel_w7<ST, COLL, VT, NT>, grid, dim3(threads), st, ev0, ev1, a, p, nsph, do_post);
    else if constexpr (COLL == D3Q19_COLL_BGK && ST == D3Q19_STORE_FP32)
        launch_maybe_timed(step_kernel_w5<ST, COLL, VT, NT>, grid, dim3(threads), st, ev0, ev1, a, p, nsph, do_post);
    else
        launch_maybe_timed(step_kernel<ST, COLL, VT, NT>, grid, dim3(threads), st, ev0, ev1, a, p, nsph, do_post);
}
template <int ST, int COLL, int VT>
static void launch_fused_one(const CollideArgs& a, const PartArgs& p, int nsph, int do_post, int threads, hipStream_t st,
                             hipEvent_t ev0, hipEvent_t ev1) {
    if (a.nt_stores == 2) launch_fused_nt<ST, COLL, VT, 2>(a, p, nsph, do_post, threads, st, ev0, ev1);
    else if (a.nt_stores == 1) launch_fused_nt<ST, COLL, VT, 1>(a, p, nsph, do_post, threads, st, ev0, ev1);
    else launch_fused_nt<ST, COLL, VT, 0>(a, p, nsph, do_post, threads, st, ev0, ev1);
}

template <int ST>
static void fused_st(const CollideArgs& a, const PartArgs& p, int nsph, int do_post, int coll, int vt, int threads, hipStream_t st,
                     hipEvent_t ev0, hipEvent_t ev1) {
    if (coll == D3Q19_COLL_BGK) {
        if (vt) launch_fused_one<ST, D3Q19_COLL_BGK, 1>(a, p, nsph, do_post, threads, st, ev0, ev1);
        else launch_fused_one<ST, D3Q19_COLL_BGK, 0>(a, p, nsph, do_post, threads, st, ev0, ev1);
    } else {
        if (vt) launch_fused_one<ST, D3Q19_COLL_MRT, 1>(a, p, nsph, do_post, threads, st, ev0, ev1);
        else launch_fused_one<ST, D3Q19_COLL_MRT, 0>(a, p, nsph, do_post, threads, st, ev0, ev1);
    }
}

int launch_step_fused(Sim* s, const void* fin, void* fout, int write_u_all, int do_post, hipStream_t st, hipEvent_t ev0,
                      hipEvent_t ev1, int nslab, hipEvent_t* slab_ev) {
    const CollideArgs all = make_collide_args(s, fin, fout, write_u_all, 1);
    PartArgs p = make_part_args(s);
    p.pos_out = gp(s->pos_alt);
    p.vel_out = gp(s->vel_alt);
    const int threads = all.bx * all.by;
    const int coll = s->opt.collision, vt = s->opt.variable_viscosity;
    if (nslab < 1) nslab = 1;
    // one launch: its own start and end events; slabs: stream events around the launches
    if (nslab > 1 && ev0) HIP_OK(hipEventRecord(ev0, st));
    for (int k = 0; k < nslab; k++) {
        CollideArgs a = all;
        if (nslab > 1) {
            a.z0 = s->zb[k];
            a.nblk = a.nbx * a.nby * (s->zb[k + 1] - s->zb[k]);
        }
        const int nsph = k == 0 ? s->np : 0;  // the spheres' turn goes with the first launch
        hipEvent_t e0 = nslab == 1 ? ev0 : nullptr, e1 = nslab == 1 ? ev1 : nullptr;
        switch (s->storage) {
        case D3Q19_STORE_FP16_SHIFT: fused_st<D3Q19_STORE_FP16_SHIFT>(a, p, nsph, do_post, coll, vt, threads, st, e0, e1); break;
        case D3Q19_STORE_BF16_SHIFT: fused_st<D3Q19_STORE_BF16_SHIFT>(a, p, nsph, do_post, coll, vt, threads, st, e0, e1); break;
        case D3Q19_STORE_FP64: fused_st<D3Q19_STORE_FP64>(a, p, nsph, do_post, coll, vt, threads, st, e0, e1); break;
        default: fused_st<D3Q19_STORE_FP32>(a, p, nsph, do_post, coll, vt, threads, st, e0, e1); break;
        }
        HIP_OK(hipGetLastError());
        if (slab_ev) HIP_OK(hipEventRecord(slab_ev[k], st));
    }
    if (nslab > 1 && ev1) HIP_OK(hipEventRecord(ev1, st));
    s->bcraw_valid = 0;
    return 0;
}

template <int ST, int COLL, int NT>
static void collide_batch_nt(const BatchLaunch& L, int sp, int write_u_all, int probe, hipStream_t st) {
    if (L.grid_collide > 0)
        hipLaunchKernelGGL((collide_batch_kernel<ST, COLL, 0, NT>), dim3(L.grid_collide), dim3(L.tile), 0, st, L.t, L.t.cells, sp,
                           write_u_all, probe);
    if (L.grid_collide_vt > 0)
        hipLaunchKernelGGL((collide_batch_kernel<ST, COLL, 1, NT>), dim3(L.grid_collide_vt), dim3(L.tile), 0, st, L.t, L.t.cells, sp,
                           write_u_all, probe);
}

template <int ST, int COLL>
static void collide_batch_coll(const BatchLaunch& L, int sp, int write_u_all, int probe, hipStream_t st) {
    if (L.nt_stores == 2) collide_batch_nt<ST, COLL, 2>(L, sp, write_u_all, probe, st);
    else if (L.nt_stores == 1) collide_batch_nt<ST, COLL, 1>(L, sp, write_u_all, probe, st);
    else collide_batch_nt<ST, COLL, 0>(L, sp, write_u_all, probe, st);
}

template <int ST>
static void collide_batch_st(const BatchLaunch& L, int sp, int write_u_all, int probe, hipStream_t st) {
    if (L.collision == D3Q19_COLL_BGK) collide_batch_coll<ST, D3Q19_COLL_BGK>(L, sp, write_u_all, probe, st);
    else collide_batch_coll<ST, D3Q19_COLL_MRT>(L, sp, write_u_all, probe, st);
}

// the step's collide launches with the sphere blocks in the first of them
template <int ST, int COLL, int VT, int NT>
static void step_batch_one(const BatchLaunch& L, int grid, int nturn, int sp, int write_u_all, int do_post, hipStream_t st) {
    const dim3 g(grid + nturn), blk(L.tile);
    if constexpr (COLL == D3Q19_COLL_BGK && ST == D3Q19_STORE_FP32)
        hipLaunchKernelGGL((step_batch_kernel_w5<ST, COLL, VT, NT>), g, blk, 0, st, L.t, L.t.cells, L.t.sph, sp, write_u_all, nturn, do_post);
    else if constexpr (COLL == D3Q19_COLL_BGK && ST != D3Q19_STORE_FP64)
        hipLaunchKernelGGL((step_batch_kernel_w6<ST, COLL, VT, NT>), g, blk, 0, st, L.t, L.t.cells, L.t.sph, sp, write_u_all, nturn, do_post);
    else
        hipLaunchKernelGGL((step_batch_kernel<ST, COLL, VT, NT>), g, blk, 0, st, L.t, L.t.cells, L.t.sph, sp, write_u_all, nturn, do_post);
}

template <int ST, int COLL, int NT>
static void step_batch_nt(const BatchLaunch& L, int sp, int write_u_all, int do_post, hipStream_t st) {
    const int nturn = L.grid_sphere;
    if (L.grid_collide > 0 || L.grid_collide_vt == 0) step_batch_one<ST, COLL, 0, NT>(L, L.grid_collide, nturn, sp, write_u_all, do_post, st);
    if (L.grid_collide_vt > 0)
        step_batch_one<ST, COLL, 1, NT>(L, L.grid_collide_vt, L.grid_collide > 0 ? 0 : nturn, sp, write_u_all, do_post, st);
}

template <int ST>
static void step_batch_st(const BatchLaunch& L, int sp, int write_u_all, int do_post, hipStream_t st) {
    const bool bgk = L.collision == D3Q19_COLL_BGK;
    if (L.nt_stores == 2) {
        if (bgk) step_batch_nt<ST, D3Q19_COLL_BGK, 2>(L, sp, write_u_all, do_post, st);
        else step_batch_nt<ST, D3Q19_COLL_MRT, 2>(L, sp, write_u_all, do_post, st);
    } else if (L.nt_stores == 1) {
        if (bgk) step_batch_nt<ST, D3Q19_COLL_BGK, 1>(L, sp, write_u_all, do_post, st);
        else step_batch_nt<ST, D3Q19_COLL_MRT, 1>(L, sp, write_u_all, do_post, st);
    } else {
        if (bgk) step_batch_nt<ST, D3Q19_COLL_BGK, 0>(L, sp, write_u_all, do_post, st);
        else step_batch_nt<ST, D3Q19_COLL_MRT, 0>(L, sp, write_u_all, do_post, st);
    }
}

int launch_step_batch(const BatchLaunch& L, int sp, int write_u_all, int do_post, hipStream_t st) {
    switch (L.storage) {
    case D3Q19_STORE_FP16_SHIFT: step_batch_st<D3Q19_STORE_FP16_SHIFT>(L, sp, write_u_all, do_post, st); break;
    case D3Q19_STORE_BF16_SHIFT: step_batch_st<D3Q19_STORE_BF16_SHIFT>(L, sp, write_u_all, do_post, st); break;
    case D3Q19_STORE_FP64: step_batch_st<D3Q19_STORE_FP64>(L, sp, write_u_all, do_post, st); break;
    default: step_batch_st<D3Q19_STORE_FP32>(L, sp, write_u_all, do_post, st); break;
    }
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_collide_batch(const BatchLaunch& L, int sp, int write_u_all, int probe, hipStream_t st) {
    switch (L.storage) {
    case D3Q19_STORE_FP16_SHIFT: collide_batch_st<D3Q19_STORE_FP16_SHIFT>(L, sp, write_u_all, probe, st); break;
    case D3Q19_STORE_BF16_SHIFT: collide_batch_st<D3Q19_STORE_BF16_SHIFT>(L, sp, write_u_all, probe, st); break;
    case D3Q19_STORE_FP64: collide_batch_st<D3Q19_STORE_FP64>(L, sp, write_u_all, probe, st); break;
    default: collide_batch_st<D3Q19_STORE_FP32>(L, sp, write_u_all, probe, st); break;
    }
    HIP_OK(hipGetLastError());
    return 0;
}

// ---- import / export of the reference-layout population field ------------------------------------
// import: the pre-collision field f_pre (reference layout, shell included) becomes the stored
// post-collision field so that the next gather reproduces it: f_post[q](r) = f_pre[q](r + c_q)
// (periodic wrap).  Wall unknowns of f_pre cannot be represented that way; they are kept in the
// raw side buffer and used by the next gather instead of the rebuilt values.
template <int ST, class IN>
__global__ void import_kernel(const IN* __restrict__ ref, Geom g, int Nx, int Ny, long long nn,
                              typename Pop<ST>::T* __restrict__ dst, typename Pop<ST>::R* raw) {
    using P = Pop<ST>;
    using R = typename P::R;
    const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (idx >= g.nint) return;
    const int x = (int)(idx % g.nx);
    const int y = (int)((idx / g.nx) % g.ny);
    const int z = (int)(idx / ((long long)g.nx * g.ny));
    Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        int s3[3] = {x + cx(q), y + cy(q), z + cz(q)};
        const int nd[3] = {g.nx, g.ny, g.nz};
        bool ok = true;
        for (int d = 0; d < 3; d++) {
            if (s3[d] < 0 || s3[d] >= nd[d]) {
                if (g.per[d]) s3[d] = (s3[d] + nd[d]) % nd[d];
                else ok = false;
            }
        }
        if (!ok) { s3[0] = x; s3[1] = y; s3[2] = z; }
        const long long ri = (s3[0] + 1) + (long long)Nx * ((s3[1] + 1) + (long long)Ny * (s3[2] + 1));
        P::template st<q>(to_global<typename P::T>(dst) + pop_index(g, q, x, y, z), 0u, (R)ref[q * nn + ri]);
    });
    if (g.wall >= 0 && raw) {
        const int A = g.wall;
        const int c = A == 0 ? x : (A == 1 ? y : z);
        const int n = A == 0 ? g.nx : (A == 1 ? g.ny : g.nz);
        const long long ri = (x + 1) + (long long)Nx * ((y + 1) + (long long)Ny * (z + 1));
        const int psz = plane_size(g, A);
        const int pidx = A == 0 ? y + g.ny * z : (A == 1 ? x + g.nx * z : x + g.nx * y);
        for (int S = 0; S < 2; S++) {
            if (!((S == 0 && c == 0) || (S == 1 && c == n - 1))) continue;
            const int inward = S == 0 ? 1 : -1;
            int k = 0;
            for (int q = 1; q < Q; q++) {
                if (cc(q, A) != inward) continue;
                raw[((size_t)(S * 5 + k)) * psz + pidx] = (R)ref[q * nn + ri];
                k++;
            }
        }
    }
}

template <int ST, class OUT>
__global__ void export_kernel(const typename Pop<ST>::T* __restrict__ fin, Geom g, FluidParams fl, const void* raw, int use_raw,
                              int Nx, int Ny, long long nn, OUT* __restrict__ ref) {
    using R = typename Pop<ST>::R;
    const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (idx >= g.nint) return;
    const int x = (int)(idx % g.nx);
    const int y = (int)((idx / g.nx) % g.ny);
    const int z = (int)(idx / ((long long)g.nx * g.ny));
    R f[Q];
    gather_pre<ST>(fin, g, fl, raw, use_raw, x, y, z, f);
    const long long ri = (x + 1) + (long long)Nx * ((y + 1) + (long long)Ny * (z + 1));
    Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) { ref[decltype(qc)::value * nn + ri] = (OUT)f[decltype(qc)::value]; });
}

template <int ST>
static void import_st(Sim* s, const void* ref, int is_double, void* dst) {
    using R = typename Pop<ST>::R;
    const int th = 256;
    const int blocks = (int)((s->g.nint + th - 1) / th);
    R* raw = (R*)s->bcraw;
    if (is_double)
        hipLaunchKernelGGL((import_kernel<ST, double>), dim3(blocks), dim3(th), 0, s->stream, (const double*)ref, s->g, s->Nx, s->Ny, s->nn,
                           (typename Pop<ST>::T*)dst, raw);
    else
        hipLaunchKernelGGL((import_kernel<ST, float>), dim3(blocks), dim3(th), 0, s->stream, (const float*)ref, s->g, s->Nx, s->Ny, s->nn,
                           (typename Pop<ST>::T*)dst, raw);
}

int launch_import_f(Sim* s, const void* ref, int is_double, void* dst) {
    switch (s->storage) {
    case D3Q19_STORE_FP16_SHIFT: import_st<D3Q19_STORE_FP16_SHIFT>(s, ref, is_double, dst); break;
    case D3Q19_STORE_BF16_SHIFT: import_st<D3Q19_STORE_BF16_SHIFT>(s, ref, is_double, dst); break;
    case D3Q19_STORE_FP64: import_st<D3Q19_STORE_FP64>(s, ref, is_double, dst); break;
    default: import_st<D3Q19_STORE_FP32>(s, ref, is_double, dst); break;
    }
    HIP_OK(hipGetLastError());
    s->bcraw_valid = (s->g.wall >= 0) ? 1 : 0;
    return 0;
}

template <int ST>
static void export_st(Sim* s, const void* fsrc, void* ref, int as_double) {
    const int th = 256;
    const int blocks = (int)((s->g.nint + th - 1) / th);
    if (as_double)
        hipLaunchKernelGGL((export_kernel<ST, double>), dim3(blocks), dim3(th), 0, s->stream, (const typename Pop<ST>::T*)fsrc, s->g, s->fl,
                           s->bcraw, s->bcraw_valid, s->Nx, s->Ny, s->nn, (double*)ref);
    else
        hipLaunchKernelGGL((export_kernel<ST, float>), dim3(blocks), dim3(th), 0, s->stream, (const typename Pop<ST>::T*)fsrc, s->g, s->fl,
                           s->bcraw, s->bcraw_valid, s->Nx, s->Ny, s->nn, (float*)ref);
}

int launch_export_f(Sim* s, const void* fsrc, void* ref, int as_double) {
    switch (s->storage) {
    case D3Q19_STORE_FP16_SHIFT: export_st<D3Q19_STORE_FP16_SHIFT>(s, fsrc, ref, as_double); break;
    case D3Q19_STORE_BF16_SHIFT: export_st<D3Q19_STORE_BF16_SHIFT>(s, fsrc, ref, as_double); break;
    case D3Q19_STORE_FP64: export_st<D3Q19_STORE_FP64>(s, fsrc, ref, as_double); break;
    default: export_st<D3Q19_STORE_FP32>(s, fsrc, ref, as_double); break;
    }
    HIP_OK(hipGetLastError());
    return 0;
}

// ---- scalar/vector fields between reference layout (float) and the interior layout ---------------
template <class R>
__global__ void field_in_kernel(const float* __restrict__ ref, Geom g, int Nx, int Ny, long long nn, int ncomp, R* __restrict__ dst) {
    const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (idx >= g.nint) return;
    const int x = (int)(idx % g.nx), y = (int)((idx / g.nx) % g.ny), z = (int)(idx / ((long long)g.nx * g.ny));
    const long long ri = (x + 1) + (long long)Nx * ((y + 1) + (long long)Ny * (z + 1));
    for (int c = 0; c < ncomp; c++) dst[idx * ncomp + c] = (R)ref[c * nn + ri];  // device: components interleaved
}
template <class R>
__global__ void field_out_kernel(const R* __restrict__ src, Geom g, int Nx, int Ny, long long nn, int ncomp, float* __restrict__ ref) {
    const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (idx >= g.nint) return;
    const int x = (int)(idx % g.nx), y = (int)((idx / g.nx) % g.ny), z = (int)(idx / ((long long)g.nx * g.ny));
    const long long ri = (x + 1) + (long long)Nx * ((y + 1) + (long long)Ny * (z + 1));
    for (int c = 0; c < ncomp; c++) ref[c * nn + ri] = (float)src[idx * ncomp + c];
}
__global__ void force_out_kernel(const long long* __restrict__ facc, Geom g, int Nx, int Ny, long long nn, float* __restrict__ ref) {
    const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (idx >= g.nint) return;
    const int x = (int)(idx % g.nx), y = (int)((idx / g.nx) % g.ny), z = (int)(idx / ((long long)g.nx * g.ny));
    const long long ri = (x + 1) + (long long)Nx * ((y + 1) + (long long)Ny * (z + 1));
    for (int c = 0; c < 3; c++) ref[c * nn + ri] = (float)((double)facc[c * (long long)g.nint + idx] * FIX_INV);
}

int launch_import_field(Sim* s, const float* ref, int ncomp, void* dst) {
    const int th = 256, blocks = (int)((s->g.nint + th - 1) / th);
    if (s->storage == D3Q19_STORE_FP64)
        hipLaunchKernelGGL(field_in_kernel<double>, dim3(blocks), dim3(th), 0, s->stream, ref, s->g, s->Nx, s->Ny, s->nn, ncomp, (double*)dst);
    else
        hipLaunchKernelGGL(field_in_kernel<float>, dim3(blocks), dim3(th), 0, s->stream, ref, s->g, s->Nx, s->Ny, s->nn, ncomp, (float*)dst);
    HIP_OK(hipGetLastError());
    return 0;
}
int launch_export_field(Sim* s, const void* src, int ncomp, float* ref) {
    const int th = 256, blocks = (int)((s->g.nint + th - 1) / th);
    if (s->storage == D3Q19_STORE_FP64)
        hipLaunchKernelGGL(field_out_kernel<double>, dim3(blocks), dim3(th), 0, s->stream, (const double*)src, s->g, s->Nx, s->Ny, s->nn, ncomp, ref);
    else
        hipLaunchKernelGGL(field_out_kernel<float>, dim3(blocks), dim3(th), 0, s->stream, (const float*)src, s->g, s->Nx, s->Ny, s->nn, ncomp, ref);
    HIP_OK(hipGetLastError());
    return 0;
}
int launch_export_force(Sim* s, float* ref) {
    const int th = 256, blocks = (int)((s->g.nint + th - 1) / th);
    hipLaunchKernelGGL(force_out_kernel, dim3(blocks), dim3(th), 0, s->stream, s->facc, s->g, s->Nx, s->Ny, s->nn, ref);
    HIP_OK(hipGetLastError());
    return 0;
}

// ---- initial field ---------------------------------------------------------------------------------
template <int ST>
__global__ void fill_eq_kernel(typename Pop<ST>::T* f, Geom g) {
    using P = Pop<ST>;
    using R = typename P::R;
    const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (idx >= g.nint) return;
    const int x = (int)(idx % g.nx);
    const int y = (int)((idx / g.nx) % g.ny);
    const int z = (int)(idx / ((long long)g.nx * g.ny));
    Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        P::template st<q>(to_global<typename P::T>(f) + pop_index(g, q, x, y, z), 0u, (R)wq(q));
    });
}
int launch_fill_equilibrium(Sim* s, void* f) {
    const int th = 256, blocks = (int)((s->g.nint + th - 1) / th);
    switch (s->storage) {
    case D3Q19_STORE_FP16_SHIFT: hipLaunchKernelGGL(fill_eq_kernel<D3Q19_STORE_FP16_SHIFT>, dim3(blocks), dim3(th), 0, s->stream, (_Float16*)f, s->g); break;
    case D3Q19_STORE_BF16_SHIFT: hipLaunchKernelGGL(fill_eq_kernel<D3Q19_STORE_BF16_SHIFT>, dim3(blocks), dim3(th), 0, s->stream, (uint16_t*)f, s->g); break;
    case D3Q19_STORE_FP64: hipLaunchKernelGGL(fill_eq_kernel<D3Q19_STORE_FP64>, dim3(blocks), dim3(th), 0, s->stream, (double*)f, s->g); break;
    default: hipLaunchKernelGGL(fill_eq_kernel<D3Q19_STORE_FP32>, dim3(blocks), dim3(th), 0, s->stream, (float*)f, s->g); break;
    }
    HIP_OK(hipGetLastError());
    return 0;
}

template <class R>
__global__ void fill_value_kernel(R* p, long long n, R v) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
int launch_fill_tau(Sim* s) {
    const int th = 256, blocks = (int)((s->g.nint + th - 1) / th);
    if (s->storage == D3Q19_STORE_FP64)
        hipLaunchKernelGGL(fill_value_kernel<double>, dim3(blocks), dim3(th), 0, s->stream, (double*)s->tau, (long long)s->g.nint,
                           (double)s->fp.NewtonianTau);
    else
        hipLaunchKernelGGL(fill_value_kernel<float>, dim3(blocks), dim3(th), 0, s->stream, (float*)s->tau, (long long)s->g.nint,
                           s->fp.NewtonianTau);
    HIP_OK(hipGetLastError());
    return 0;
}

// ---- observables ------------------------------------------------------------------------------------
// plane means of u over each interior z plane, FP64, one block per plane, fixed reduction order.
template <class R>
__global__ void plane_mean_kernel(const R* __restrict__ u, Geom g, double* out, int Nz) {
    const int z = blockIdx.x;
    __shared__ double sh[3][256];
    double acc[3] = {0, 0, 0};
    const int plane = g.nx * g.ny;
    for (int i = threadIdx.x; i < plane; i += blockDim.x) {
        const long long n = (long long)z * plane + i;
        for (int c = 0; c < 3; c++) acc[c] += (double)u[3 * n + c];
    }
    for (int c = 0; c < 3; c++) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int c = 0; c < 3; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++) out[c * Nz + z + 1] = sh[c][0] / (double)plane;
}
int launch_plane_mean(Sim* s, double* dout) {
    if (s->storage == D3Q19_STORE_FP64)
        hipLaunchKernelGGL(plane_mean_kernel<double>, dim3(s->g.nz), dim3(256), 0, s->stream, (const double*)s->u, s->g, dout, s->Nz);
    else
        hipLaunchKernelGGL(plane_mean_kernel<float>, dim3(s->g.nz), dim3(256), 0, s->stream, (const float*)s->u, s->g, dout, s->Nz);
    HIP_OK(hipGetLastError());
    return 0;
}

// total mass and momentum of the pre-collision field; block partials then one fixed-order pass
template <int ST>
__global__ void moments_kernel(const typename Pop<ST>::T* __restrict__ fin, Geom g, FluidParams fl, const void* raw, int use_raw,
                               double* partial, int* bad) {
    using R = typename Pop<ST>::R;
    __shared__ double sh[4][256];
    double acc[4] = {0, 0, 0, 0};
    int badl = 0;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < g.nint; idx += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(idx % g.nx), y = (int)((idx / g.nx) % g.ny), z = (int)(idx / ((long long)g.nx * g.ny));
        R f[Q];
        gather_pre<ST>(fin, g, fl, raw, use_raw, x, y, z, f);
        double r = 0, jx = 0, jy = 0, jz = 0;
        Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            r += (double)f[q];
            jx += cx(q) * (double)f[q];
            jy += cy(q) * (double)f[q];
            jz += cz(q) * (double)f[q];
        });
        if (!(r > 0.0) || !isfinite(r) || !isfinite(jx) || !isfinite(jy) || !isfinite(jz)) badl = 1;
        acc[0] += r; acc[1] += jx; acc[2] += jy; acc[3] += jz;
    }
    if (badl && bad) atomicOr(bad, 1);
    for (int c = 0; c < 4; c++) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int c = 0; c < 4; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 4; c++) partial[4 * blockIdx.x + c] = sh[c][0];
}
__global__ void sum_partials_kernel(const double* partial, int nb, double* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double a[4] = {0, 0, 0, 0};
        for (int b = 0; b < nb; b++)
            for (int c = 0; c < 4; c++) a[c] += partial[4 * b + c];
        for (int c = 0; c < 4; c++) out[c] = a[c];
    }
}

template <int ST>
static void moments_st(Sim* s, const void* fsrc, double* partial, int nb, int* bad) {
    hipLaunchKernelGGL((moments_kernel<ST>), dim3(nb), dim3(256), 0, s->stream, (const typename Pop<ST>::T*)fsrc, s->g, s->fl, s->bcraw,
                       s->bcraw_valid, partial, bad);
}
static int moments_any(Sim* s, const void* fsrc, double* dout, int* bad) {
    const int nb = 512;
    double* partial = s->red;  // 4*512 doubles + 4
    switch (s->storage) {
    case D3Q19_STORE_FP16_SHIFT: moments_st<D3Q19_STORE_FP16_SHIFT>(s, fsrc, partial, nb, bad); break;
    case D3Q19_STORE_BF16_SHIFT: moments_st<D3Q19_STORE_BF16_SHIFT>(s, fsrc, partial, nb, bad); break;
    case D3Q19_STORE_FP64: moments_st<D3Q19_STORE_FP64>(s, fsrc, partial, nb, bad); break;
    default: moments_st<D3Q19_STORE_FP32>(s, fsrc, partial, nb, bad); break;
    }
    if (dout) hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, s->stream, partial, nb, dout);
    HIP_OK(hipGetLastError());
    return 0;
}
int launch_moments(Sim* s, const void* fsrc, double* dout) { return moments_any(s, fsrc, dout, nullptr); }
int launch_check_finite(Sim* s, const void* fsrc, int* flag) { return moments_any(s, fsrc, nullptr, flag); }

// ---- shear-cell observables (FP64) ---------------------------------------------------------------
// Viscous shear stress sigma_FA at node (x, y, z) of the pre-collision field, evaluated in double
// precision from the stored populations:
//   sigma_FA = -(1 - 1/(2 tau)) (Pi_FA - rho u_F u_A + (u_F g_A + u_A g_F)/2),
//   Pi_FA = sum_q c_qF c_qA f_q,  u = (sum_q c_q f_q + g/2)/rho,  g = the constant body force,
// with tau the relaxation time of the F-A stress moment at the node (the per-node tau, NewtonianTau,
// or 1/1.4 for the MRT x-y moment, whose rate the reference fixes, GPU_program.cl:453,567-568).
template <int ST, class PIn>
__device__ double node_stress(const CellDesc& cd, PIn fin, int x, int y, int z) {
    using R = typename Pop<ST>::R;
    const CollideArgs& c = cd.c;
    double f[Q];
    gather_pre<ST, ADDR_WIDE, false, double>(fin, c.g, c.fl, c.bcraw, c.use_bcraw, x, y, z, f);
    const int A = cd.o.axisA, F = cd.o.axisF;
    double rho = 0, jx = 0, jy = 0, jz = 0, pi = 0;
    Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        rho += f[q];
        if constexpr (cx(q) != 0) jx += (double)cx(q) * f[q];
        if constexpr (cy(q) != 0) jy += (double)cy(q) * f[q];
        if constexpr (cz(q) != 0) jz += (double)cz(q) * f[q];
        const int w = cc(q, F) * cc(q, A);
        if (w != 0) pi += (double)w * f[q];
    });
    const double g3[3] = {(double)c.fl.gconst[0], (double)c.fl.gconst[1], (double)c.fl.gconst[2]};
    const double u3[3] = {(jx + 0.5 * g3[0]) / rho, (jy + 0.5 * g3[1]) / rho, (jz + 0.5 * g3[2]) / rho};
    const double uF = F == 0 ? u3[0] : (F == 1 ? u3[1] : u3[2]);
    const double uA = A == 0 ? u3[0] : (A == 1 ? u3[1] : u3[2]);
    const double gF = F == 0 ? g3[0] : (F == 1 ? g3[1] : g3[2]);
    const double gA = A == 0 ? g3[0] : (A == 1 ? g3[1] : g3[2]);
    double tau = (double)c.fl.tau0;
    if (cd.o.rate_fixed > 0.f) tau = 1.0 / (double)cd.o.rate_fixed;
    else if (c.vt) tau = (double)to_global<R>(c.tau)[x + c.g.nx * (y + c.g.ny * z)];
    return -(1.0 - 0.5 / tau) * (pi - rho * uF * uA + 0.5 * (uF * gA + uA * gF));
}

// Block b: cell b/3; role 0/1 sums sigma_FA over the lower/upper plane next to the walls (plane
// index 1 and n_A - 2 along the wall axis), role 2 sums |fluid force| over the cell's spheres.
// Fixed-order block reductions: deterministic.
template <int ST>
__global__ void __launch_bounds__(256) observables_kernel(const CellDesc* __restrict__ cells, int sp, double* __restrict__ out) {
    const int cell = blockIdx.x / 3, role = blockIdx.x % 3;
    const CellDesc& cd = cells[cell];
    double v0 = 0.0, v1 = 0.0;
    if (role < 2) {
        const int A = cd.o.axisA;
        if (A >= 0) {
            const Geom& g = cd.c.g;
            const int nA = A == 0 ? g.nx : (A == 1 ? g.ny : g.nz);
            const int n1 = A == 0 ? g.ny : g.nx;
            const int n2 = A == 2 ? g.ny : g.nz;
            const int pidx = role == 0 ? 1 : nA - 2;
            const auto fin = to_global<typename Pop<ST>::T>(cd.f[cd.par ^ sp]);
            for (int i = threadIdx.x; i < n1 * n2; i += blockDim.x) {
                const int u1 = i % n1, u2 = i / n1;
                const int x = A == 0 ? pidx : u1;
                const int y = A == 1 ? pidx : (A == 0 ? u1 : u2);
                const int z = A == 2 ? pidx : u2;
                v0 += node_stress<ST>(cd, fin, x, y, z);
                v1 += 1.0;
            }
        }
    } else {
        const PartArgs& p = cd.p[0];
        for (int k = threadIdx.x; k < p.np; k += blockDim.x) {
            const auto F = p.ffluid + 3 * k;
            v0 += sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
            v1 += 1.0;
        }
    }
    __shared__ double sh[2][256];
    sh[0][threadIdx.x] = v0;
    sh[1][threadIdx.x] = v1;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + st];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[OBS_STRIDE * cell + 2 * role] = sh[0][0];
        out[OBS_STRIDE * cell + 2 * role + 1] = sh[1][0];
        if (role == 2) {
            out[OBS_STRIDE * cell + 6] = 0.0;
            out[OBS_STRIDE * cell + 7] = 0.0;
        }
    }
}

// Watchdog of the cells of a batch, one launch for all of them (the per-cell checks took four queue
// operations per resident cell at every check, about 2 ms per check with 64 small cells resident).
// Block b scans part b % bpc of cell b / bpc: the density and momentum of every node's
// pre-collision populations (as the single simulation's check), non-finite or non-positive density
// raising the cell's first word.
template <int ST>
__global__ void __launch_bounds__(256) watchdog_kernel(const CellDesc* __restrict__ cells, int sp, int bpc, int* __restrict__ out) {
    using R = typename Pop<ST>::R;
    const int cell = blockIdx.x / bpc, part = blockIdx.x % bpc;
    const CellDesc& cd = cells[cell];
    if (cd.nint <= 0) return;
    const auto fin = to_global<typename Pop<ST>::T>(cd.f[cd.par ^ sp]);
    const CollideArgs& c = cd.c;
    int bad = 0;
    for (int n = part * (int)blockDim.x + (int)threadIdx.x; n < cd.nint; n += bpc * (int)blockDim.x) {
        int x, y, z;
        split_node(cd, n, x, y, z);
        R f[Q];
        gather_pre<ST>(fin, c.g, c.fl, c.bcraw, c.use_bcraw, x, y, z, f);
        double r = 0, jx = 0, jy = 0, jz = 0;
        Unroll<0, Q>::run([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            r += (double)f[q];
            jx += cx(q) * (double)f[q];
            jy += cy(q) * (double)f[q];
            jz += cz(q) * (double)f[q];
        });
        if (!(r > 0.0) || !isfinite(r) || !isfinite(jx) || !isfinite(jy) || !isfinite(jz)) bad = 1;
    }
    if (bad) atomicOr(&out[2 * cell], 1);
    if (part == 0 && threadIdx.x == 0 && cd.p[0].np > 0 && cd.p[0].flag) out[2 * cell + 1] = *cd.p[0].flag;
}

int launch_watchdog(const CellDesc* cells, int ncells, int storage, int sp, int bpc, int* out, hipStream_t st) {
    if (ncells <= 0) return 0;
    const dim3 grid(ncells * bpc), block(256);
    switch (storage) {
    case D3Q19_STORE_FP16_SHIFT: hipLaunchKernelGGL(watchdog_kernel<D3Q19_STORE_FP16_SHIFT>, grid, block, 0, st, cells, sp, bpc, out); break;
    case D3Q19_STORE_BF16_SHIFT: hipLaunchKernelGGL(watchdog_kernel<D3Q19_STORE_BF16_SHIFT>, grid, block, 0, st, cells, sp, bpc, out); break;
    case D3Q19_STORE_FP64: hipLaunchKernelGGL(watchdog_kernel<D3Q19_STORE_FP64>, grid, block, 0, st, cells, sp, bpc, out); break;
    default: hipLaunchKernelGGL(watchdog_kernel<D3Q19_STORE_FP32>, grid, block, 0, st, cells, sp, bpc, out); break;
    }
    HIP_OK(hipGetLastError());
    return 0;
}

int launch_observables(const CellDesc* cells, int ncells, int storage, int sp, double* out, hipStream_t st) {
    if (ncells <= 0) return 0;
    const dim3 grid(3 * ncells), block(256);
    switch (storage) {
    case D3Q19_STORE_FP16_SHIFT: hipLaunchKernelGGL(observables_kernel<D3Q19_STORE_FP16_SHIFT>, grid, block, 0, st, cells, sp, out); break;
    case D3Q19_STORE_BF16_SHIFT: hipLaunchKernelGGL(observables_kernel<D3Q19_STORE_BF16_SHIFT>, grid, block, 0, st, cells, sp, out); break;
    case D3Q19_STORE_FP64: hipLaunchKernelGGL(observables_kernel<D3Q19_STORE_FP64>, grid, block, 0, st, cells, sp, out); break;
    default: hipLaunchKernelGGL(observables_kernel<D3Q19_STORE_FP32>, grid, block, 0, st, cells, sp, out); break;
    }
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace d3q19
