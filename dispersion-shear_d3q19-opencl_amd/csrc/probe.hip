// probe.hip -- attainable HBM bandwidth for the collide's access pattern (a measurement utility;
// the reference has no counterpart).
//
// A streaming kernel with the memory traffic of the collide+stream kernel and none of its
// arithmetic: `streams` input planes are read and as many output planes written, each lane
// handling `W` consecutive nodes (one W-wide load and store per plane), optionally at the pull
// offsets of the D3Q19 velocity set (plane q of node n read at n - c_q, periodic over the flat
// index).  With streams = 19 and W = 1 this is the collide's pattern exactly; W = 2 / 4 give the
// same bytes with dwordx2 / dwordx4 lane accesses; streams = 1 is a plain copy.  width = -64 / -256
// keeps one node per lane but interleaves the 19 populations in chunks of 64 / 256 nodes.  `lds_bytes` of
// dynamic LDS per 256-thread block caps the blocks per CU, to see how the bandwidth depends on
// the waves in flight.
#include "internal.h"

namespace d3q19 {

template <int W> struct VecOf;
template <> struct VecOf<1> { typedef float T; };
typedef float f32x2u __attribute__((ext_vector_type(2), aligned(4)));
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
template <> struct VecOf<2> { typedef f32x2u T; };
template <> struct VecOf<4> { typedef f32x4u T; };

// Element of population q at node n: planes (LAYOUT 0, q * nint + n) or chunks of C nodes holding
// all populations of the chunk (LAYOUT C > 0, ((n / C) * S + q) * C + n % C): one wavefront's loads
// of all populations then fall in one contiguous block instead of S planes far apart.
template <int LAYOUT, int S>
__device__ __forceinline__ unsigned pop_elem(int q, unsigned n, unsigned nint) {
    if constexpr (LAYOUT == 0) return (unsigned)q * nint + n;
    else return ((n / LAYOUT) * S + (unsigned)q) * LAYOUT + n % LAYOUT;
}

template <int W, int S, bool PULL, int LAYOUT = 0, bool NTM = false>
__global__ void __launch_bounds__(256) stream_probe_kernel(const float* __restrict__ in, float* __restrict__ out, unsigned nint,
                                                           int nx, int nxy, float a) {
    extern __shared__ float lds_pad[];
    typedef typename VecOf<W>::T V;
    const unsigned n = (blockIdx.x * 256u + threadIdx.x) * (unsigned)W;
    if (n + W > nint) return;
    if (a == 12345.0f) lds_pad[threadIdx.x] = 0.f;  // keeps the dynamic LDS allocation (never true)
    const auto gin = to_global<float>(in);
    const auto gout = to_global<float>(out);
    V v[S];
    Unroll<0, S>::run([&](auto qc) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        unsigned src = n;
        if constexpr (PULL && S == Q) {
            const long long off = (long long)cx(q) + (long long)nx * cy(q) + (long long)nxy * cz(q);
            long long s = (long long)n - off;
            if (s < 0) s += nint;
            if (s >= (long long)nint) s -= nint;
            src = (s + W <= (long long)nint) ? (unsigned)s : n;
        }
        if constexpr (LAYOUT == 0) v[q] = *(const D3Q19_GLOBAL V*)plane_at(gin + (long long)q * nint, src);
        else if constexpr (NTM) v[q] = __builtin_nontemporal_load(&gin[pop_elem<LAYOUT, S>(q, src, nint)]);
        else v[q] = gin[pop_elem<LAYOUT, S>(q, src, nint)];
    });
    Unroll<0, S>::run([&](auto qc) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        if constexpr (LAYOUT == 0) *(D3Q19_GLOBAL V*)plane_at(gout + (long long)q * nint, n) = v[q] * a;
        else if constexpr (NTM) __builtin_nontemporal_store(v[q] * a, &gout[pop_elem<LAYOUT, S>(q, n, nint)]);
        else gout[pop_elem<LAYOUT, S>(q, n, nint)] = v[q] * a;
    });
}

template <int W, int S, bool PULL, int LAYOUT = 0, bool NTM = false>
static void probe_launch(const float* in, float* out, unsigned nint, int nx, int nxy, int lds, hipStream_t st) {
    const unsigned threads = nint / W;
    const dim3 grid((threads + 255) / 256);
    hipLaunchKernelGGL((stream_probe_kernel<W, S, PULL, LAYOUT, NTM>), grid, dim3(256), lds, st, in, out, nint, nx, nxy, 0.999f);
}

typedef void (*probe_fn)(const float*, float*, unsigned, int, int, int, hipStream_t);

// width 1 / 2 / 4: nodes per lane, planes; width -64 / -256: one node per lane, chunk layout of
// 64 / 256 nodes, non-temporal accesses when nt is set
static probe_fn pick(int streams, int width, int pull, bool nt) {
    if (width == -64 && nt) return pull ? probe_launch<1, Q, true, 64, true> : probe_launch<1, Q, false, 64, true>;
    if (width == -256 && nt) return pull ? probe_launch<1, Q, true, 256, true> : probe_launch<1, Q, false, 256, true>;
    if (width == -64) return pull ? probe_launch<1, Q, true, 64> : probe_launch<1, Q, false, 64>;
    if (width == -256) return pull ? probe_launch<1, Q, true, 256> : probe_launch<1, Q, false, 256>;
    if (streams == 1) return width == 4 ? probe_launch<4, 1, false> : (width == 2 ? probe_launch<2, 1, false> : probe_launch<1, 1, false>);
    if (pull) return width == 4 ? probe_launch<4, Q, true> : (width == 2 ? probe_launch<2, Q, true> : probe_launch<1, Q, true>);
    return width == 4 ? probe_launch<4, Q, false> : (width == 2 ? probe_launch<2, Q, false> : probe_launch<1, Q, false>);
}

}  // namespace d3q19

using namespace d3q19;

extern "C" int d3q19_stream_probe(int device, int nx, int ny, int nz, int streams, int width, int pull, int lds_bytes, int launches,
                                  double* gbs, double* ms) {
    const bool chunked = width == -64 || width == -256;
    if (pull < 0 || pull > 3 || ((pull & 2) && !chunked)) {
        set_error("d3q19_stream_probe: pull is 0..3 (bit 1, non-temporal accesses, with the chunk layouts only)");
        return D3Q19_E_INVALID;
    }
    const bool nt = (pull & 2) != 0;
    pull &= 1;
    if (chunked && streams != Q) {
        set_error("d3q19_stream_probe: the chunk layouts hold 19 populations");
        return D3Q19_E_INVALID;
    }
    if (!gbs || nx <= 0 || ny <= 0 || nz <= 0 || (streams != 1 && streams != Q) || (width != 1 && width != 2 && width != 4 && !chunked) ||
        launches < 1 || launches > 4096 || lds_bytes < 0 || lds_bytes > 65536) {
        set_error("d3q19_stream_probe: bad argument");
        return D3Q19_E_INVALID;
    }
    const long long nint = (long long)nx * ny * nz;
    const int wdiv = chunked ? -width : width;
    if (nint % wdiv != 0 || nint * 4 * (chunked ? Q : 1) >= (1ll << 32)) {
        set_error("d3q19_stream_probe: the node count must be a multiple of width (chunk) and the lattice below 4 GiB");
        return D3Q19_E_INVALID;
    }
    HIP_OK(hipSetDevice(device));
    const size_t bytes = (size_t)streams * nint * sizeof(float);
    float *in = nullptr, *out = nullptr;
    if (hipMalloc(&in, bytes) != hipSuccess) {
        set_error("d3q19_stream_probe: device allocation failed");
        return D3Q19_E_NOMEM;
    }
    if (hipMalloc(&out, bytes) != hipSuccess) {
        (void)hipFree(in);
        set_error("d3q19_stream_probe: device allocation failed");
        return D3Q19_E_NOMEM;
    }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = D3Q19_OK;
    do {
        if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            set_error("d3q19_stream_probe: stream/event creation failed");
            rc = D3Q19_E_DEVICE;
            break;
        }
        if (hipMemsetAsync(in, 0, bytes, st) != hipSuccess) { rc = D3Q19_E_DEVICE; break; }
        const probe_fn fn = pick(streams, width, pull, nt);
        for (int i = 0; i < 2; i++) fn(in, out, (unsigned)nint, nx, nx * ny, lds_bytes, st);
        (void)hipEventRecord(e0, st);
        for (int i = 0; i < launches; i++) fn(in, out, (unsigned)nint, nx, nx * ny, lds_bytes, st);
        (void)hipEventRecord(e1, st);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            set_error("d3q19_stream_probe: launch failed");
            rc = D3Q19_E_DEVICE;
            break;
        }
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        const double per = (double)t / launches;
        if (ms) *ms = per;
        *gbs = 2.0 * (double)bytes / (per * 1e-3) / 1e9;
    } while (0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(in);
    (void)hipFree(out);
    return rc;
}
