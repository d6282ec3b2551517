// internal.h -- device layout, simulation object and kernel entry points of libd3q19.
//
// Device layout (interior nodes only; the reference's one-node buffer shell is not stored):
//   node n = x + nx*(y + ny*z), 0 <= x < nx = LatticeSize[0]-2, etc.  Lattice index i_1D of the
//   reference = (x+1) + Nx*((y+1) + Ny*(z+1)).
//   populations  two lattices A/B (pull ping-pong).  Chunk-interleaved (pop_index): chunks of R
//                rows (R * nx a multiple of 64 nodes where ny allows, R = 1 for nx = 256), each
//                holding its 19 populations as 19 consecutive runs of R * nx values, so a
//                wavefront's loads and stores touch one block of 19 runs instead of 38 planes a
//                lattice apart.  Batch members whose runs are not whole wavefronts: planar,
//                f[q*nint + n].  The state between steps is the post-collision field;
//                a step gathers f[q](n - c_q) (periodic wrap, moving-wall unknowns rebuilt from
//                the knowns), collides and writes at n.
//   u, tau       R[3*nint], R[nint]           (R = float, or double for the fp64 build)
//   force        int64[3*nint] fixed point (2^-40 units), spread with integer atomics so the
//                sum is independent of order; consumed and zeroed by the next collide.
//   force tags   uint8[nint]: 1 where the spreading added a nonzero contribution; the next collide
//                reads the force only where its node's tag is set, and clears both.
//   masks        uint32 epoch stamp per 4x4x4 brick, three arrays by epoch mod 3: a brick is
//                marked for the step whose epoch it holds (no per-step clearing).  The step with
//                epoch e = epochs[0] + 1 writes u in the bricks marked e, and its spreading kernel
//                commits epochs[0] = epochs[1] = e.  In the fused step the spheres' turn,
//                running beside the collide, marks epoch e + 1 (array (e+1) mod 3, which no other
//                block of the launch reads) for its new positions dilated by one lattice unit.
//   particles    double SoA; positions and velocities double-buffered (the fused turn reads
//                x(t), v(t) of every partner while writing x(t+1), v(t+1)).
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>
#include <map>
#include <string>
#include "d3q19_abi.h"

namespace d3q19 {

constexpr int Q = 19;
constexpr double FIX_SCALE = 1099511627776.0;       // 2^40
constexpr double FIX_INV = 1.0 / 1099511627776.0;
constexpr int BRICK = 4;                              // brick edge (nodes)
constexpr int IBM_WAVES = 4;                          // markers (wavefronts) per IBM block
constexpr int TILE_MARKERS = 8;                       // markers per block of the LDS-staged spreading
constexpr long long INFINITY_CACHE_BYTES = 256ll << 20;  // MI355X's last-level (memory-side) cache
constexpr long long NT_LOAD_BYTES = 1024ll << 20;        // population buffers from which non-temporal loads pay
// Non-temporal population accesses of the collide for population buffers of `bytes` (both
// lattices): none while they fit the Infinity Cache; stores beyond it; stores and loads from
// NT_LOAD_BYTES (nearer the cache size the loads give up residency the ordinary ones keep; a
// single simulation takes the loads only on rows of a multiple of 64 nodes).
inline int nt_level(long long bytes) { return bytes > NT_LOAD_BYTES ? 2 : (bytes > INFINITY_CACHE_BYTES ? 1 : 0); }
constexpr long long IBM_ROWS_MIN_MARKERS = 16384;        // markers from which four per wavefront pay
constexpr int IBM_LARGE_GROUP_MIN_BLOCKS = 512;          // blocks needed to take groups larger than TILE_MARKERS
constexpr int TILE_MARKERS_MAX = 64;                  // ... of a single simulation's Peskin-4 groups whose boxes fit
constexpr int TILE_NODES = 2048;                      // LDS box of that spreading (nodes), single simulations
constexpr int TILE_NODES_BATCH = 2048;                // ... cells of a batch (groups of up to TILE_MARKERS_MAX markers)
constexpr float FP16_SCALE = 16384.0f;                // 2^14, relative-deviation scale of fp16 storage
constexpr int MARK_BRICKS = 64;                       // bricks per axis a sphere's marking can list

// ---- velocity set (reference order, D3Q19-OpenCL.c:9-13) ---------------------------------
__host__ __device__ constexpr int cx(int q) {
    return (q == 1 || q == 7 || q == 8 || q == 9 || q == 10) ? 1 : ((q == 2 || q == 11 || q == 12 || q == 13 || q == 14) ? -1 : 0);
}
__host__ __device__ constexpr int cy(int q) {
    return (q == 3 || q == 7 || q == 11 || q == 15 || q == 16) ? 1 : ((q == 4 || q == 8 || q == 12 || q == 17 || q == 18) ? -1 : 0);
}
__host__ __device__ constexpr int cz(int q) {
    return (q == 5 || q == 9 || q == 13 || q == 15 || q == 17) ? 1 : ((q == 6 || q == 10 || q == 14 || q == 16 || q == 18) ? -1 : 0);
}
__host__ __device__ constexpr int cc(int q, int a) { return a == 0 ? cx(q) : (a == 1 ? cy(q) : cz(q)); }
__host__ __device__ constexpr double wq(int q) { return q == 0 ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0); }

// opposite direction computed from the velocity components
__host__ __device__ constexpr int opposite(int q) {
    int r = 0;
    for (int p = 0; p < Q; p++)
        if (cx(p) == -cx(q) && cy(p) == -cy(q) && cz(p) == -cz(q)) r = p;
    return r;
}

// the 9 direction pairs (q, opposite(q)) of the velocity set, one member each: 1, 3, 5, 7, 8, 9, 10, 15, 16
__host__ __device__ constexpr int pair_dir(int k) {
    return k == 0 ? 1 : k == 1 ? 3 : k == 2 ? 5 : k == 3 ? 7 : k == 4 ? 8 : k == 5 ? 9 : k == 6 ? 10 : k == 7 ? 15 : 16;
}

// fused multiply-add written out (the collide compiles with implicit contraction off, so the
// rounding of every node's arithmetic is fixed by its source, whatever kernel it is inlined into)
template <class R> __device__ __forceinline__ R fma_(R a, R b, R c);
template <> __device__ __forceinline__ float fma_<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double fma_<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

// c_q . v for a velocity-set direction q, with only its nonzero components (a product by a zero
// component cannot be dropped by the compiler under IEEE semantics)
template <int s, class R> __device__ __forceinline__ R signed_(R v) {
    if constexpr (s > 0) return v;
    else return -v;
}
template <int q, class R> __device__ __forceinline__ R cdot(R x, R y, R z) {
    constexpr int X = cx(q), Y = cy(q), Z = cz(q);
    if constexpr (X != 0 && Y != 0) return signed_<X>(x) + signed_<Y>(y);
    else if constexpr (X != 0 && Z != 0) return signed_<X>(x) + signed_<Z>(z);
    else if constexpr (Y != 0 && Z != 0) return signed_<Y>(y) + signed_<Z>(z);
    else if constexpr (X != 0) return signed_<X>(x);
    else if constexpr (Y != 0) return signed_<Y>(y);
    else if constexpr (Z != 0) return signed_<Z>(z);
    else return (R)0;
}

// compile-time loop
template <int I, int N> struct Unroll {
    template <class F> __device__ __forceinline__ static void run(F&& f) {
        f(std::integral_constant<int, I>{});
        Unroll<I + 1, N>::run(f);
    }
};
template <int N> struct Unroll<N, N> {
    template <class F> __device__ __forceinline__ static void run(F&&) {}
};

// ---- d'Humieres moment basis, row order of GPU_program.cl:410-449 -----------------------------
__host__ __device__ constexpr int mrow(int k, int q) {
    int x = cx(q), y = cy(q), z = cz(q);
    int c2 = x * x + y * y + z * z;
    switch (k) {
    case 0: return 1;
    case 1: return 19 * c2 - 30;
    case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
    case 3: return x;
    case 4: return (5 * c2 - 9) * x;
    case 5: return y;
    case 6: return (5 * c2 - 9) * y;
    case 7: return z;
    case 8: return (5 * c2 - 9) * z;
    case 9: return 3 * x * x - c2;
    case 10: return (3 * c2 - 5) * (3 * x * x - c2);
    case 11: return y * y - z * z;
    case 12: return (3 * c2 - 5) * (y * y - z * z);
    case 13: return x * y;
    case 14: return y * z;
    case 15: return x * z;
    case 16: return (y * y - z * z) * x;
    case 17: return (z * z - x * x) * y;
    default: return (x * x - y * y) * z;
    }
}
__host__ __device__ constexpr int mnorm(int k) {
    int s = 0;
    for (int q = 0; q < Q; q++) s += mrow(k, q) * mrow(k, q);
    return s;
}

// ---- lattice geometry passed to kernels --------------------------------------------------------
struct Geom {
    int nx, ny, nz;       // interior extents
    int nint;             // nx*ny*nz
    // population layout (pop_index): log2 of the rows per chunk, and the element strides from one
    // population to the next at a node, from one z-plane to the next and from one chunk of rows
    // to the next within a plane
    int lr;
    int cs;
    long long qs, ps;
    int per[3];           // 1 periodic, 0 wall
    int wall;             // wall axis or -1
    int bnx, bny, bnz;    // brick grid
    int sys[3];           // SystemSize
};

// Element index of population q at node (x, y, z).  Chunked: each z-plane is cut into chunks of
// R = 2^lr consecutive rows (R divides ny); a chunk holds its Q populations as Q consecutive runs
// of R * nx values (row y % R of the chunk at offset (y % R) * nx in each run): qs = R * nx,
// cs = Q * R * nx, ps = Q * nx * ny.  R is the smallest that makes a run a multiple of 64 nodes
// (one wavefront) where ny allows.  The collide tiles of a single simulation are row segments, so
// a wavefront's access is one stretch of a run and a chunk's 19 runs lie together.  Planar (plane
// q holds population q of every node): qs = nint, ps = nx * ny, lr = PLANAR_LR (every row in
// chunk 0), cs = 0; batched tiles are runs of consecutive nodes of any row length, which a planar
// field keeps contiguous, so batch members whose chunk runs are not whole wavefronts stay planar.
constexpr int PLANAR_LR = 30;
__host__ __device__ inline long long pop_index(const Geom& g, int q, int x, int y, int z) {
    return (long long)q * g.qs + (long long)z * g.ps + (long long)(y >> g.lr) * g.cs + (long long)(y & ((1 << g.lr) - 1)) * g.nx + x;
}
// how a kernel forms population addresses
enum : int {
    ADDR_ZPLANE = 0,  // z uniform over the wavefront: the plane in each population's scalar base
    ADDR_CELL32 = 1,  // a 32-bit lane offset over the whole lattice (batch members, Q * nint values below 4 GB)
    ADDR_WIDE = 2,    // any z per lane: the plane term in 64 bits (cold kernels)
};

struct FluidParams {
    float gconst[3];
    float uw[2][3];       // [lower, upper] wall velocity
    float tau0;           // NewtonianTau
    int model;            // ViscosityModel
    float vp[4];          // ViscosityParams
    float tau_min, tau_max;
    int calc_rho;
    int tan_mask;         // bit a: tangential velocity BC on axis a
    int sr;               // stochastic rounding of 16-bit population stores
};

// Pointer fields that device code dereferences in place are declared in the global address space
// for the device compilation (same size and layout as the host's plain pointers).
#ifdef __HIP_DEVICE_COMPILE__
#define D3Q19_GLOBAL __attribute__((address_space(1)))
#else
#define D3Q19_GLOBAL
#endif
// host-side stores into such fields (the cast is an identity in the host compilation)
template <class T> inline D3Q19_GLOBAL T* gp(T* p) { return (D3Q19_GLOBAL T*)p; }
template <class T> inline const D3Q19_GLOBAL T* gp(const T* p) { return (const D3Q19_GLOBAL T*)p; }
// Device code reaching memory through a generic pointer that the compiler cannot trace to a kernel
// argument (for example one read from the batched kernels' descriptor table) gets FLAT
// instructions, whose waits also cover LDS traffic; the batched collide measured up to 30% slower
// that way.  to_global states the address space so that loads, stores and atomics are global_*.
template <class T> __host__ __device__ __forceinline__ const D3Q19_GLOBAL T* to_global(const void* p) {
    return (const D3Q19_GLOBAL T*)p;
}
template <class T> __host__ __device__ __forceinline__ D3Q19_GLOBAL T* to_global(void* p) { return (D3Q19_GLOBAL T*)p; }
// Element i of a population plane that starts at p (the same address for the whole wavefront).
// The byte offset is formed in 32 bits, so the compiler addresses the element as a scalar base plus
// one 32-bit register (the saddr form of global loads and stores) rather than a 64-bit sum per lane:
// 2-3 VALU instructions fewer per access.  Planes stay below 4 GiB (checked when a lattice is
// accepted).
template <class T> __device__ __forceinline__ const D3Q19_GLOBAL T* plane_at(const D3Q19_GLOBAL T* p, unsigned i) {
    return (const D3Q19_GLOBAL T*)((const D3Q19_GLOBAL char*)p + (unsigned long long)(i * (unsigned)sizeof(T)));
}
template <class T> __device__ __forceinline__ D3Q19_GLOBAL T* plane_at(D3Q19_GLOBAL T* p, unsigned i) {
    return (D3Q19_GLOBAL T*)((D3Q19_GLOBAL char*)p + (unsigned long long)(i * (unsigned)sizeof(T)));
}

// ---- per-cell kernel arguments ------------------------------------------------------------------
// The same structs drive the single-simulation kernels (passed by value) and the batched kernels
// (read from a device table of CellDesc, one per cell).
struct CollideArgs {
    const void* fin;
    void* fout;
    void* u;
    void* tau;
    long long* facc;
    uint8_t* ftag;         // per node: 1 where the last IBM pass spread force (cleared by the collide that consumes it)
    const uint32_t* mask0;
    const uint32_t* mask1;
    const uint32_t* mask2;
    const uint32_t* epochs;
    const void* bcraw;
    Geom g;
    FluidParams fl;
    int write_u_all;
    int use_umask;
    int use_bcraw;
    int has_force;
    int vt;                // per-node tau (runtime form of the kernels' VT template argument)
    int bx, by;            // block tile
    int nbx, nby;          // tiles per plane
    int packed;            // 1: a tile is bx consecutive nodes of its plane in node order (rows of any length)
    float inv_nx;          // 1 / nx, for the node-order split of packed tiles
    int nblk;              // tiles of this launch
    int z0;                // first z plane of this launch's tiles (a z-slab of the lattice)
    int nt_stores;         // non-temporal population accesses (nt_level): 1 stores, 2 stores and loads
};

struct IbmArgs {
    Geom g;
    int np, ppp, mode;
    int mstride;  // marker visiting order: wavefront k of sphere p takes marker (k * mstride) mod ppp
    int tile_ok;  // spreading may be staged in LDS (every periodic extent exceeds twice the stencil reach)
    int tile_nodes;  // LDS box capacity of this simulation's launches (the largest group box that fits)
    int gm;          // markers per group (TILE_MARKERS up to TILE_MARKERS_MAX)
    const D3Q19_GLOBAL int* perm;  // ppp marker indices in Morton order of their offsets
    const D3Q19_GLOBAL int* gbox;  // per group of gm markers: LDS box relative to the centre node (lo[3], extent[3], volume, 0)
    double dA;
    const D3Q19_GLOBAL double* pos;
    const D3Q19_GLOBAL double* vel;
    const D3Q19_GLOBAL double* omega;
    const D3Q19_GLOBAL float* pts;
    const D3Q19_GLOBAL void* u;
    D3Q19_GLOBAL long long* facc;
    D3Q19_GLOBAL uint8_t* ftag;  // set to 1 at every node that receives a nonzero contribution
    D3Q19_GLOBAL double* mforce;
    D3Q19_GLOBAL uint32_t* epochs;
    // slab-owned passes: this launch runs the spheres whose stencils reach at most up to a z plane
    // in [zlo, zhi) (the top of a stencil is floor(z + zreach) + 2; a stencil that leaves the lattice
    // in z belongs to the last slab); commit: this launch commits the step's epoch
    double zreach;
    int zlo, zhi;
    int commit;
};

struct PartArgs {
    int np, ppp;
    Geom g;
    double sys[3];
    double gconst[3];
    double mass, inertia;
    double diam;
    double kcol;
    double radius;  // brick-marking radius (cube extent)
    double rin, rout;  // bounds of the marker offsets' lengths: bricks outside the shell are not marked
    int force_model;
    int visc_model;
    double tau0;
    double zbuf;
    D3Q19_GLOBAL double* pos;       // x(t), v(t): read
    D3Q19_GLOBAL double* vel;
    D3Q19_GLOBAL double* pos_out;   // dynamics writes here (== pos/vel except in the fused turn)
    D3Q19_GLOBAL double* vel_out;
    D3Q19_GLOBAL double* quat;
    D3Q19_GLOBAL double* omega;
    D3Q19_GLOBAL double* ffluid;
    D3Q19_GLOBAL double* tfluid;
    D3Q19_GLOBAL double* fcont;
    D3Q19_GLOBAL double* tcont;
    const D3Q19_GLOBAL double* mforce;
    D3Q19_GLOBAL uint32_t* mask0;
    D3Q19_GLOBAL uint32_t* mask1;
    D3Q19_GLOBAL uint32_t* mask2;
    const D3Q19_GLOBAL uint32_t* epochs;
    D3Q19_GLOBAL int* flag;         // set when a sphere moves a lattice unit or more in one step
    // contact pairs through the reference's zone lists (D3Q19_CONTACT_ZONES): zone of each sphere
    // at the last rebuild, two generations (rebuild count mod 2) of np entries; the contacts of
    // step t = epochs[0] + tofs read generation (t - 1) / rebuild
    int zone_mode;
    int rebuild;
    int tofs;
    int nz[3];
    float zw[3];
    D3Q19_GLOBAL int* zone;
};



// Observable definition of a cell: wall-normal axis A, flow axis F (largest wall-velocity
// difference among the other axes), MRT rate of the F-A stress moment.
struct ObsArgs {
    int axisA, axisF;
    int coll;            // D3Q19_COLL_*
    float rate_fixed;    // > 0: the F-A stress moment relaxes at this fixed rate (MRT x-y moment)
};

// One cell of a batch (or the single simulation): everything its kernels need.  f[par ^ sp] holds
// the post-collision state before a step of parity sp.
struct CellDesc {
    CollideArgs c;
    IbmArgs i;
    // the sphere arguments for either buffer of x(t), v(t): p[k] reads posb[k] / velb[k] and its
    // dynamics writes posb[k ^ 1] / velb[k ^ 1].  Kept whole in the table (read through a restrict
    // kernel argument, so with scalar loads) rather than patched into a kernel's own copy.
    PartArgs p[2];
    ObsArgs o;
    const void* f[2];
    D3Q19_GLOBAL double* posb[2];  // sphere positions / velocities: [par ^ sp] is x(t) before a step of parity sp
    D3Q19_GLOBAL double* velb[2];
    int par;
    int nint;
    int nx, nxy;
    float inv_nx, inv_nxy;  // flat node index -> (x, y, z) by float reciprocal + exact correction
    // the batched collide's slots: rows of `pitch` slots (nx, or nx padded to a multiple of 64 so
    // that each wavefront stays within one row), slot index -> (x, y, z) as for nodes
    int pitch, pny, nslots;
    float inv_pitch, inv_pny;
};

// p[1] from p[0] and both buffer pairs (host, after posb / velb are set)
inline void set_cell_spheres(CellDesc& d) {
    d.p[1] = d.p[0];
    for (int k = 0; k < 2; k++) {
        d.p[k].pos = d.posb[k];
        d.p[k].vel = d.velb[k];
        d.p[k].pos_out = d.posb[k ^ 1];
        d.p[k].vel_out = d.velb[k ^ 1];
    }
}

// Work-list sizes of a batch, read by the persistent kernels from device memory so that a captured
// graph stays valid while cells join and leave.
struct BatchCounts {
    int ntiles;     // collide work-list entries (span x blockDim consecutive nodes of one cell)
    int ntiles_ct;  // the first ntiles_ct entries belong to cells of constant tau
    int span;       // tiles per entry
    int nsph;     // spheres over all cells
    int nmark;    // entries of the group list (each cell's groups over all its spheres)
    int ncells;
};

struct BatchTables {
    const CellDesc* cells;
    const uint2* tiles;      // collide work list: (cell, first node); an entry covers span tiles
    const int4* sph;         // (cell, local sphere, first global marker, markers per sphere)
    const int* msph;         // group list: (global sphere << 12) | first position of the group in the sphere's group order
    const BatchCounts* cnt;
    uint32_t* epochs;        // the batch's epoch counter (shared by all its cells)
};

// ---- device memory arena: one allocation, first-fit sub-allocation (256-byte aligned) ------------
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    size_t used = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    std::map<size_t, size_t> live_;  // offset -> length
    void* alloc(size_t bytes);
    bool release(void* p);
};

// ---- simulation object ------------------------------------------------------------------------
struct Sim {
    int device = 0;
    hipStream_t stream = nullptr;
    int_param_struct ip;
    flp_param_struct fp;
    d3q19_options opt;
    Geom g;
    FluidParams fl;
    int Nx, Ny, Nz;
    long long nn;         // reference N_C
    int storage = 0;
    size_t pop_bytes = 0; // bytes per population value
    void* fA = nullptr;
    void* fB = nullptr;
    void* u = nullptr;    // R[3*nint]
    void* tau = nullptr;  // R[nint]
    long long* facc = nullptr;
    uint8_t* ftag = nullptr;
    uint32_t* mask[3] = {nullptr, nullptr, nullptr};  // brick epoch stamps, by epoch mod 3
    uint32_t* epochs = nullptr;              // [current epoch, epoch of the pending spread force]
    int nbricks = 0;
    // raw wall unknowns of an uploaded pre-collision field: [2 sides][5][plane]
    void* bcraw = nullptr;
    int bcraw_valid = 0;
    int bc_plane = 0;
    // particles
    int np = 0, ppp = 0, nfa = 1;
    double* pos = nullptr;     // 3*np, current
    double* vel = nullptr;     // 3*np, current
    double* pos_alt = nullptr; // the other buffer of the fused turn
    double* vel_alt = nullptr;
    int* d_flag = nullptr;     // displacement watchdog of the dilated marks
    bool pending_post = false; // the last step's force sums and contacts are not done yet
    bool marks_next = false;   // marks of the next epoch exist (made from the current positions)
    int* zone = nullptr;       // 2*np zone of each sphere (contact zones), by rebuild generation
    int* perm = nullptr;       // ppp marker indices in Morton order of their offsets (LDS-staged spreading)
    int* gbox = nullptr;       // the LDS box of each group of group_markers markers (8 ints per group; sized for groups of TILE_MARKERS)
    bool tile_ok = false;      // every periodic extent exceeds twice the spheres' stencil reach
    int tile_nodes = 0;        // the largest group box (nodes) within the LDS cap: the LDS the IBM launches reserve
    int group_markers = TILE_MARKERS;  // markers per group of the LDS-staged spreading
    bool ibm_rows = false;     // Peskin-4 markers run four to a wavefront (many markers); else one per wavefront
    int nt_stores = 0;         // non-temporal population accesses of the collide (nt_level)
    double mark_rin = 0.0, mark_rout = 0.0;  // shortest / longest marker offset, with a margin
    long long h_epoch = 0;     // host count of committed epochs (one per step)
    int tofs = 0;              // step number t of the contacts = committed epoch + tofs
    double* quat = nullptr;    // 4*np
    double* omega = nullptr;   // 3*np
    double* ffluid = nullptr;  // 3*np
    double* tfluid = nullptr;  // 3*np
    double* fcont = nullptr;   // 3*np contact force
    double* tcont = nullptr;   // 3*np contact torque slot
    float* pts = nullptr;      // 4*ppp marker offsets
    double* mforce = nullptr;  // 6*np*ppp per-marker force, torque
    double* red = nullptr;     // reduction scratch
    // step graph
    hipGraphExec_t graph2 = nullptr;   // two steps (ping-pong closes)
    const void* graph_f = nullptr;     // state the graph was captured from
    const void* graph_pos = nullptr;
    int graph_tofs = 0;
    int parity = 0;            // 0: current post-collision field in fA
    // collide in z-slabs overlapped with the IBM pass (single simulations with spheres)
    static constexpr int MAX_SLABS = 16;
    int slabs = 1;
    int zb[MAX_SLABS + 1] = {};        // slab k holds planes [zb[k], zb[k+1])
    hipStream_t aux = nullptr;         // the IBM pass's stream
    hipEvent_t sev[MAX_SLABS + 1] = {};  // slab k collided; aux done
    int block[3] = {64, 4, 1};
    int packed_tiles = 0;              // collide tiles of 256 consecutive nodes of a plane (CollideArgs::packed)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // batch membership: device memory comes from the batch's arena, the batch's stream and epoch
    // counter are used, and the batch owns neither
    Arena* arena = nullptr;
    bool shares_stream = false;
    int batch_pitch = 0;  // batch members: slots per row of the batched collide (0: nx)
    bool shares_epochs = false;
    // single-cell descriptor for the observables kernel (host copy kept for the async upload)
    CellDesc* d_desc = nullptr;
    double* d_obs = nullptr;
    CellDesc h_desc;
};

// allocation + initial state (shared by d3q19_create and batch members) and release
int sim_init(Sim* s, int device, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt,
             const d3q19_float4* sphere_points, Arena* arena, hipStream_t shared_stream, uint32_t* shared_epochs);
void sim_release(Sim* s);
int sim_validate(const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt);
// watchdog: non-finite density on the current lattice, sphere displacement flag (see abi.hip)
int sim_watchdog_enqueue(Sim* s, int* h_nonfinite, int* h_moved);
int sim_watchdog(Sim* s, int* nonfinite, int* moved);
size_t sim_state_bytes(const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt);
// observables of one cell from the 8 raw sums of the observables kernel
void derive_observables(const Sim* s, const double* raw, d3q19_observables* out);

// ---- argument builders and launchers (defined in the .hip files) ---------------------------------
CollideArgs make_collide_args(Sim* s, const void* fin, void* fout, int write_u_all, int use_maskU);
IbmArgs make_ibm_args(Sim* s);
int marker_stride(int ppp);  // coprime with ppp, about 0.618 ppp
PartArgs make_part_args(Sim* s);
CellDesc make_cell_desc(Sim* s);
// ev0/ev1 (optional): recorded at the kernel's own start and end, for timing
int launch_collide(Sim* s, const void* fin, void* fout, int write_u_all, int use_maskU, hipStream_t st, hipEvent_t ev0 = nullptr,
                   hipEvent_t ev1 = nullptr);
int launch_import_f(Sim* s, const void* ref_f, int is_double, void* fdst);
int launch_export_f(Sim* s, const void* fsrc, void* ref_out, int as_double);
int launch_import_field(Sim* s, const float* ref, int ncomp, void* dst);
int launch_export_field(Sim* s, const void* src, int ncomp, float* ref_out);
int launch_export_force(Sim* s, float* ref_out);
int launch_plane_mean(Sim* s, double* dev_out);
int launch_moments(Sim* s, const void* fsrc, double* dev_out);
int launch_check_finite(Sim* s, const void* fsrc, int* dev_flag);
// zlo / zhi / commit: the slab-owned pass of the overlapped step (see IbmArgs); the defaults run
// every sphere and commit the epoch
int launch_ibm(Sim* s, hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, int zlo = 0, int zhi = 1 << 30,
               int commit = 1);
int launch_sphere_pre(Sim* s, int do_dyn, int do_mark, hipStream_t st);       // dynamics, brick marks
int launch_sphere_post(Sim* s, int do_reduce, int do_contacts, hipStream_t st); // force sums, contacts
int launch_zone_init(Sim* s, hipStream_t st);  // zones of the current positions, both generations
int launch_sphere_prime(Sim* s, hipStream_t st);                                // dilated marks, next epoch
// fused step: the spheres' turn (post of the last step if pending, dynamics into the other
// position buffer, dilated marks of the epoch after this one) in np extra blocks of the collide
// nslab > 1: the collide tiles go out as s->slabs launches of the planes [zb[k], zb[k+1]) (the
// spheres' turn with the first), each followed by a record of slab_ev[k] on st; ev0 / ev1 time the
// first launch's start and the last one's end
int launch_step_fused(Sim* s, const void* fin, void* fout, int write_u_all, int do_post, hipStream_t st, hipEvent_t ev0 = nullptr,
                      hipEvent_t ev1 = nullptr, int nslab = 1, hipEvent_t* slab_ev = nullptr);
int launch_fill_equilibrium(Sim* s, void* f);
int launch_fill_tau(Sim* s);

// observables: per cell 8 doubles (FP64): [sum of the F-A shear stress over the lower wall-adjacent
// plane, node count, same for the upper plane, sum over spheres of |fluid force|, spheres, 0, 0]
constexpr int OBS_STRIDE = 8;
int launch_observables(const CellDesc* cells, int ncells, int storage, int sp, double* out, hipStream_t st);
// watchdog of every cell slot in one launch: out[2c] |= 1 when cell c's current lattice holds a
// non-finite or non-positive density, out[2c+1] = its sphere displacement flag (out zeroed first);
// empty slots (nint 0) are skipped; bpc blocks scan each cell
int launch_watchdog(const CellDesc* cells, int ncells, int storage, int sp, int bpc, int* out, hipStream_t st);

// batched (persistent) kernels over the work lists of BatchTables; grid sizes are fixed per batch
struct BatchLaunch {
    BatchTables t;
    int storage, collision, ibm_mode;
    int tile;              // threads (= nodes) per collide tile
    int span;              // tiles per collide work-list entry
    int grid_collide;      // blocks of the collide kernel over the constant-tau cells' tiles
    int grid_collide_vt;   // blocks of the collide kernel over the per-node-tau cells' tiles
    int grid_sphere;       // persistent blocks of the sphere kernels
    int grid_ibm;          // persistent blocks (one marker group each per iteration) of the IBM kernel
    int nt_stores;         // non-temporal population accesses (nt_level of the members' lattices together)
};
int launch_collide_batch(const BatchLaunch& L, int sp, int write_u_all, int probe, hipStream_t st);
// a batched step's collide with the spheres' turn in the leading blocks of its first launch
int launch_step_batch(const BatchLaunch& L, int sp, int write_u_all, int do_post, hipStream_t st);
int launch_ibm_batch(const BatchLaunch& L, int sp, hipStream_t st);
int launch_sphere_prime_batch(const BatchLaunch& L, int sp, hipStream_t st);
int launch_sphere_post_batch(const BatchLaunch& L, int sp, hipStream_t st);

// error reporting
void set_error(const std::string& msg);

}  // namespace d3q19

// the opaque handle of the C-ABI is the simulation object itself
struct d3q19_sim : public d3q19::Sim {};

#define HIP_OK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::d3q19::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));         \
            return D3Q19_E_DEVICE;                                                         \
        }                                                                                  \
    } while (0)
