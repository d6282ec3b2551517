// batch.hip -- batched ensembles of independent shear cells, the serving scheduler (continuous
// batching with admission by device-memory footprint) and the plain sequential solve.
//
// A batch owns one HIP stream, one device arena from which every member cell's state is carved
// as a single block, one epoch counter shared by the members' brick marks, and the work lists of
// the persistent kernels: a table of CellDesc (one slot per cell id), the collide tiles (cell,
// first node), the spheres (cell, local index, first marker, markers per sphere) and a
// marker -> sphere map.  The lists live in device memory and the persistent kernels read their
// lengths from there, so the captured two-step graph stays valid while cells join and leave; it
// is re-captured only when a list outgrows its allocation or the launch shape changes.
//
// Step parity: every member's current post-collision field is f[par ^ phase], where par is fixed
// when the cell joins and phase is the batch's step parity.  Graph pairs leave phase unchanged; a
// single direct step flips it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <string>
#include <vector>
#include "internal.h"

using namespace d3q19;

namespace d3q19 {

static int bfail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

struct Member {
    d3q19_sim* sim = nullptr;
    Arena sub;            // the member's own block of the batch arena
    void* block = nullptr;
    int par = 0;
    double* pb[2] = {nullptr, nullptr};  // position / velocity buffers as they were at join
    double* vb[2] = {nullptr, nullptr};
};

// Point a member's single-simulation view (current lattice and position buffers) at its state.
static void sync_view(Member& m, int phase) {
    Sim* s = m.sim;
    const int i = m.par ^ phase;
    s->parity = i;
    if (s->np > 0) {
        s->pos = m.pb[i];
        s->pos_alt = m.pb[i ^ 1];
        s->vel = m.vb[i];
        s->vel_alt = m.vb[i ^ 1];
    }
}

struct Batch {
    int device = 0;
    int ncu = 256;
    hipStream_t stream = nullptr;
    d3q19_batch_options bo;
    Arena arena;
    uint32_t* epochs = nullptr;
    int max_cells = 1024;
    std::vector<Member> cells;  // index = cell id
    int phase = 0;
    long long h_epoch = 0;      // batch steps run (one epoch each)
    // device work lists
    CellDesc* d_cells = nullptr;
    uint2* d_tiles = nullptr;
    size_t cap_tiles = 0;
    int4* d_sph = nullptr;
    size_t cap_sph = 0;
    int* d_msph = nullptr;
    size_t cap_mark = 0;
    BatchCounts* d_cnt = nullptr;
    double* d_obs = nullptr;
    int* d_wd = nullptr;  // watchdog words, 2 per cell slot
    // host mirrors, kept alive until the next rebuild (which first synchronises the stream)
    std::vector<CellDesc> h_cells;
    std::vector<uint2> h_tiles;
    std::vector<int4> h_sph;
    std::vector<int> h_msph;
    BatchCounts h_cnt{};
    bool dirty = true;
    d3q19_tuning tuning{};
    // launch grids in effect: a tuning field of 0 means "one work item per block / wavefront",
    // rounded up to a power of two so that the captured graph survives most membership changes
    int eff_collide = 1, eff_collide_vt = 0, eff_sphere = 1, eff_ibm = 1;
    int eff_nt = 0;        // non-temporal population stores (decided from the members' lattice bytes)
    hipGraphExec_t graph2 = nullptr;
    bool graph_valid = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

static BatchLaunch launch_of(const Batch* b) {
    BatchLaunch L;
    L.t.cells = b->d_cells;
    L.t.tiles = b->d_tiles;
    L.t.sph = b->d_sph;
    L.t.msph = b->d_msph;
    L.t.cnt = b->d_cnt;
    L.t.epochs = b->epochs;
    L.storage = b->bo.storage;
    L.collision = b->bo.collision;
    L.ibm_mode = b->bo.ibm_kernel;
    L.tile = b->tuning.tile;
    L.span = b->tuning.span > 0 ? b->tuning.span : 1;
    L.grid_collide = b->eff_collide;
    L.grid_collide_vt = b->eff_collide_vt;
    L.grid_sphere = b->eff_sphere;
    L.grid_ibm = b->eff_ibm;
    L.nt_stores = b->eff_nt;
    return L;
}

static int active_slots(const Batch* b) {
    int hi = 0;
    for (int i = 0; i < (int)b->cells.size(); i++)
        if (b->cells[i].sim) hi = i + 1;
    return hi;
}

template <class T>
static int grow(T** p, size_t* cap, size_t need, bool* moved) {
    if (need <= *cap && *p) return 0;
    size_t n = std::max<size_t>(need + need / 2, 1024);
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        *cap = 0;
        return bfail(D3Q19_E_NOMEM, std::string("hipMalloc (batch work list): ") + hipGetErrorString(e));
    }
    *cap = n;
    *moved = true;
    return 0;
}

// one block per work item, rounded up to a multiple of 1024 blocks (at least 256)
static int grid_for(long long n) {
    if (n <= 256) return 256;
    return (int)(((n + 1023) / 1024) * 1024);
}

// Rebuild the work lists from the member set and upload them (stream ordered).
static int upload_tables(Batch* b) {
    HIP_OK(hipStreamSynchronize(b->stream));  // host mirrors may still be read by earlier copies
    const int nslots = active_slots(b);
    const int tile = b->tuning.tile * (b->tuning.span > 0 ? b->tuning.span : 1);  // nodes per work-list entry
    b->h_tiles.clear();
    b->h_sph.clear();
    b->h_msph.clear();
    // the members' population buffers together beyond the Infinity Cache: non-temporal population
    // stores, as for a single simulation of that size
    long long pop_total = 0;
    for (int c = 0; c < b->max_cells; c++)
        if (const Sim* s = b->cells[c].sim) pop_total += 2LL * Q * s->pop_bytes * s->g.nint;
    const int nt = nt_level(pop_total);
    for (int c = 0; c < b->max_cells; c++) {
        Member& m = b->cells[c];
        if (!m.sim) {
            memset(&b->h_cells[c], 0, sizeof(CellDesc));
            continue;
        }
        Sim* s = m.sim;
        sync_view(m, b->phase);
        CellDesc d = make_cell_desc(s);
        d.c.nt_stores = nt;
        d.par = m.par;
        d.f[0] = s->fA;
        d.f[1] = s->fB;
        for (int k = 0; k < 2; k++) {
            d.posb[k] = gp(m.pb[k]);
            d.velb[k] = gp(m.vb[k]);
        }
        set_cell_spheres(d);
        b->h_cells[c] = d;
    }
    // tiles: the constant-tau cells first, then the per-node-tau cells
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) b->h_cnt.ntiles_ct = (int)b->h_tiles.size();
        for (int c = 0; c < b->max_cells; c++) {
            const Sim* s = b->cells[c].sim;
            if (!s || (s->opt.variable_viscosity ? 1 : 0) != pass) continue;
            const long long nslots = (long long)(s->batch_pitch > 0 ? s->batch_pitch : s->g.nx) * s->g.ny * s->g.nz;
            for (long long base = 0; base < nslots; base += tile) b->h_tiles.push_back(make_uint2((unsigned)c, (unsigned)base));
        }
    }
    int max_ppp = 0;
    long long nmark = 0;
    for (int c = 0; c < b->max_cells; c++) {
        const Sim* s = b->cells[c].sim;
        if (!s) continue;
        for (int p = 0; p < s->np; p++) {
            b->h_sph.push_back(make_int4(c, p, (int)nmark, s->ppp));
            nmark += s->ppp;
        }
        max_ppp = std::max(max_ppp, s->ppp);
    }
    if (b->h_sph.size() >= (1u << 19) || nmark >= (1LL << 31))
        return bfail(D3Q19_E_CAPACITY, "batch holds more than 2^19 spheres or 2^31 markers");
    // LDS-staged IBM: group list, (sphere << 12) | first position of the group in the sphere's
    // group order (groups of the cell's group_markers); round k takes group k of every sphere with
    // more than k groups, so that blocks that run together spread into different spheres' nodes
    {
        b->h_msph.reserve((size_t)(nmark / TILE_MARKERS + b->h_sph.size()));
        for (int k = 0; k * TILE_MARKERS < max_ppp; k++)
            for (size_t i = 0; i < b->h_sph.size(); i++) {
                const int gm = b->cells[b->h_sph[i].x].sim->group_markers;
                if (k * gm < b->h_sph[i].w) b->h_msph.push_back(((int)i << 12) | (k * gm));
            }
    }
    (void)nslots;
    b->h_cnt.ntiles = (int)b->h_tiles.size();
    b->h_cnt.nsph = (int)b->h_sph.size();
    b->h_cnt.nmark = (int)b->h_msph.size();
    b->h_cnt.ncells = active_slots(b);
    b->h_cnt.span = b->tuning.span > 0 ? b->tuning.span : 1;
    const int n_ct = b->h_cnt.ntiles_ct, n_vt = b->h_cnt.ntiles - b->h_cnt.ntiles_ct;
    const int ec = n_ct == 0 ? 0 : (b->tuning.grid_collide > 0 ? b->tuning.grid_collide : grid_for(n_ct));
    const int ev = n_vt == 0 ? 0 : (b->tuning.grid_collide > 0 ? b->tuning.grid_collide : grid_for(n_vt));
    const int es = b->tuning.grid_sphere > 0 ? b->tuning.grid_sphere : grid_for(b->h_cnt.nsph);
    const int ei = b->tuning.grid_ibm > 0 ? b->tuning.grid_ibm : grid_for(b->h_cnt.nmark);  // one group per block
    if (ec != b->eff_collide || ev != b->eff_collide_vt || es != b->eff_sphere || ei != b->eff_ibm || nt != b->eff_nt) {
        b->eff_collide = ec;
        b->eff_collide_vt = ev;
        b->eff_sphere = es;
        b->eff_ibm = ei;
        b->eff_nt = nt;
        b->graph_valid = false;
    }
    bool moved = false;
    int rc;
    if ((rc = grow(&b->d_tiles, &b->cap_tiles, b->h_tiles.size(), &moved))) return rc;
    if ((rc = grow(&b->d_sph, &b->cap_sph, b->h_sph.size(), &moved))) return rc;
    if ((rc = grow(&b->d_msph, &b->cap_mark, b->h_msph.size(), &moved))) return rc;
    if (moved) b->graph_valid = false;
    HIP_OK(hipMemcpyAsync(b->d_cells, b->h_cells.data(), (size_t)b->max_cells * sizeof(CellDesc), hipMemcpyHostToDevice, b->stream));
    if (!b->h_tiles.empty())
        HIP_OK(hipMemcpyAsync(b->d_tiles, b->h_tiles.data(), b->h_tiles.size() * sizeof(uint2), hipMemcpyHostToDevice, b->stream));
    if (!b->h_sph.empty())
        HIP_OK(hipMemcpyAsync(b->d_sph, b->h_sph.data(), b->h_sph.size() * sizeof(int4), hipMemcpyHostToDevice, b->stream));
    if (!b->h_msph.empty())
        HIP_OK(hipMemcpyAsync(b->d_msph, b->h_msph.data(), b->h_msph.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipMemcpyAsync(b->d_cnt, &b->h_cnt, sizeof(BatchCounts), hipMemcpyHostToDevice, b->stream));
    b->dirty = false;
    return 0;
}

// One batched step of every cell with step parity sp: one launch whose leading blocks run the
// spheres' turn while the others collide (independent, as in the single simulation's fused step:
// the turn writes the next epoch's marks and the other position and velocity buffers, the collide
// reads this epoch's marks and the pending force), then the IBM pass.  The turn as a launch of its
// own on a second stream beside the collide (two graph branches) cost the fork and join: on one
// 127x127x128 cell with 63 spheres 115.0 us per step against 101.2 us fused; on the 64-cell fp16
// ensemble 0.660 against 0.649 ms.
static int enqueue_batch_step(Batch* b, int sp, int write_u_all, int do_post, hipStream_t st) {
    const BatchLaunch L = launch_of(b);
    int rc;
    if ((rc = launch_step_batch(L, sp, write_u_all, do_post, st))) return rc;
    return launch_ibm_batch(L, sp, st);
}

static int capture_graph(Batch* b) {
    if (b->graph2) {
        (void)hipGraphExecDestroy(b->graph2);
        b->graph2 = nullptr;
    }
    hipGraph_t g;
    HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_batch_step(b, 0, 0, 1, b->stream);
    if (!rc) rc = enqueue_batch_step(b, 1, 0, 1, b->stream);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return bfail(D3Q19_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph2, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(D3Q19_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    b->graph_valid = true;
    return 0;
}

static int direct_step(Batch* b, int write_u_all, int do_post) {
    int rc = enqueue_batch_step(b, b->phase, write_u_all, do_post, b->stream);
    if (rc) return rc;
    b->phase ^= 1;
    for (Member& m : b->cells)
        if (m.sim) sync_view(m, b->phase);
    return 0;
}

// n steps of every cell: dilated marks of the next epoch from the current positions, a first step
// with no pending force sums, graph-replayed pairs at phase 0, the remaining steps (the last one
// carries flags), then the last step's force sums and contacts.
static int batch_run(Batch* b, int n, int flags) {
    if (n <= 0) return 0;
    int rc;
    if (b->dirty && (rc = upload_tables(b))) return rc;
    if (b->h_cnt.ncells == 0) return 0;
    b->h_epoch += n;
    const int wu = (flags & D3Q19_STEP_WRITE_U) ? 1 : 0;
    const BatchLaunch L0 = launch_of(b);
    if ((rc = launch_sphere_prime_batch(L0, b->phase, b->stream))) return rc;
    if ((rc = direct_step(b, n == 1 ? wu : 0, 0))) return rc;
    int left = n - 1;
    if (left > 0 && b->phase == 1) {  // realign with the graph's parity
        if ((rc = direct_step(b, left == 1 ? wu : 0, 1))) return rc;
        left--;
    }
    const int tail = (wu && left > 0) ? 1 : 0;
    const int pairs = (left - tail) / 2;
    if (pairs > 0) {
        if (!b->graph_valid && (rc = capture_graph(b))) return rc;
        for (int i = 0; i < pairs; i++) HIP_OK(hipGraphLaunch(b->graph2, b->stream));
        left -= 2 * pairs;
    }
    while (left > 0) {
        if ((rc = direct_step(b, left == 1 ? wu : 0, 1))) return rc;
        left--;
    }
    return launch_sphere_post_batch(launch_of(b), b->phase, b->stream);
}

static void default_tuning(Batch* b) {
    b->tuning.tile = 256;
    b->tuning.span = 1;
    b->tuning.grid_collide = 0;  // one work-list entry per block
    b->tuning.grid_sphere = 0;   // one sphere per block
    b->tuning.grid_ibm = 0;      // one marker group per block
    b->tuning.collide_ms = 0.0;
}

static int valid_tuning(const d3q19_tuning* t) {
    if (t->tile != 64 && t->tile != 128 && t->tile != 256) return 0;
    if (t->span != 0 && t->span != 1 && t->span != 2 && t->span != 4 && t->span != 8) return 0;
    if (t->grid_collide < 0 || t->grid_sphere < 0 || t->grid_ibm < 0) return 0;
    if (t->grid_collide > (1 << 20) || t->grid_sphere > (1 << 20) || t->grid_ibm > (1 << 20)) return 0;
    return 1;
}

// mean duration of `reps` trial collide launches (state unchanged) under the current tuning
static int time_probe(Batch* b, int reps, double* ms) {
    const BatchLaunch L = launch_of(b);
    int rc;
    for (int i = 0; i < 2; i++)
        if ((rc = launch_collide_batch(L, b->phase, 0, 1, b->stream))) return rc;
    HIP_OK(hipEventRecord(b->ev0, b->stream));
    for (int i = 0; i < reps; i++)
        if ((rc = launch_collide_batch(L, b->phase, 0, 1, b->stream))) return rc;
    HIP_OK(hipEventRecord(b->ev1, b->stream));
    HIP_OK(hipEventSynchronize(b->ev1));
    float t = 0;
    HIP_OK(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = (double)t / reps;
    return 0;
}

static int autotune(Batch* b) {
    int rc;
    if (b->dirty && (rc = upload_tables(b))) return rc;
    if (b->h_cnt.ncells == 0) return bfail(D3Q19_E_STATE, "autotune needs at least one cell");
    const d3q19_tuning keep = b->tuning;
    d3q19_tuning best = keep;
    double best_ms = 1e30;
    // candidates: tile of 64/128/256 nodes x a persistent grid of k*256 threads per CU (k blocks
    // of 256 threads per CU, or the same thread count in smaller tiles), or one tile per block (0)
    const int tiles[3] = {64, 128, 256};
    const int spans[4] = {1, 2, 4, 8};
    const int ks[4] = {0, 4, 8, 16};
    for (int ti = 0; ti < 3; ti++)
        for (int si = 0; si < 4; si++) {
            b->tuning.tile = tiles[ti];
            b->tuning.span = spans[si];
            b->tuning.grid_collide = 0;
            b->dirty = true;
            if ((rc = upload_tables(b))) return rc;
            for (int pi = 0; pi < 4; pi++) {
                const int grid = ks[pi] * b->ncu * (256 / tiles[ti]);
                b->tuning.grid_collide = grid;
                const int n_ct = b->h_cnt.ntiles_ct, n_vt = b->h_cnt.ntiles - n_ct;
                b->eff_collide = n_ct == 0 ? 0 : (grid > 0 ? grid : grid_for(n_ct));
                b->eff_collide_vt = n_vt == 0 ? 0 : (grid > 0 ? grid : grid_for(n_vt));
                double ms = 0;
                if ((rc = time_probe(b, 6, &ms))) return rc;
                if (ms < best_ms) {
                    best_ms = ms;
                    best = b->tuning;
                }
            }
        }
    best.collide_ms = best_ms;
    b->tuning = best;
    b->dirty = true;
    b->graph_valid = false;
    return upload_tables(b);
}

static int batch_observables(Batch* b, const int* ids, int n, d3q19_observables* out) {
    int rc;
    if (b->dirty && (rc = upload_tables(b))) return rc;
    const int nslots = active_slots(b);
    if (nslots == 0) return 0;
    if ((rc = launch_observables(b->d_cells, nslots, b->bo.storage, b->phase, b->d_obs, b->stream))) return rc;
    std::vector<double> raw((size_t)OBS_STRIDE * nslots);
    HIP_OK(hipMemcpyAsync(raw.data(), b->d_obs, raw.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    for (int i = 0; i < n; i++) {
        const int c = ids[i];
        if (c < 0 || c >= nslots || !b->cells[c].sim) return bfail(D3Q19_E_INVALID, "no such cell");
        derive_observables(b->cells[c].sim, raw.data() + (size_t)OBS_STRIDE * c, &out[i]);
    }
    return 0;
}

// A server check: the observables and the watchdog words of cells ids[0..n), two launches, two
// copies and one synchronisation.
static int batch_check(Batch* b, const int* ids, int n, d3q19_observables* out, int* nonfinite, int* moved) {
    int rc;
    if (b->dirty && (rc = upload_tables(b))) return rc;
    const int nslots = active_slots(b);
    if (nslots == 0) return 0;
    long long maxn = 1;
    for (int c = 0; c < nslots; c++)
        if (b->cells[c].sim) maxn = std::max(maxn, (long long)b->cells[c].sim->g.nint);
    const int bpc = (int)std::min<long long>(256, std::max<long long>(1, (maxn + 16383) / 16384));
    if ((rc = launch_observables(b->d_cells, nslots, b->bo.storage, b->phase, b->d_obs, b->stream))) return rc;
    HIP_OK(hipMemsetAsync(b->d_wd, 0, (size_t)2 * nslots * sizeof(int), b->stream));
    if ((rc = launch_watchdog(b->d_cells, nslots, b->bo.storage, b->phase, bpc, b->d_wd, b->stream))) return rc;
    std::vector<double> raw((size_t)OBS_STRIDE * nslots);
    std::vector<int> w((size_t)2 * nslots);
    HIP_OK(hipMemcpyAsync(raw.data(), b->d_obs, raw.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipMemcpyAsync(w.data(), b->d_wd, w.size() * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    for (int i = 0; i < n; i++) {
        const int c = ids[i];
        if (c < 0 || c >= nslots || !b->cells[c].sim) return bfail(D3Q19_E_INVALID, "no such cell");
        derive_observables(b->cells[c].sim, raw.data() + (size_t)OBS_STRIDE * c, &out[i]);
        nonfinite[i] = w[2 * (size_t)c];
        moved[i] = w[2 * (size_t)c + 1];
    }
    return 0;
}

static int batch_add(Batch* b, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt,
                     const d3q19_float4* pts, const d3q19_float4* kin, int* id) {
    if (!ip || !fp || !id) return bfail(D3Q19_E_INVALID, "null argument");
    d3q19_options o;
    if (opt) o = *opt;
    else memset(&o, 0, sizeof(o));
    o.storage = b->bo.storage;
    o.collision = b->bo.collision;
    o.ibm_kernel = b->bo.ibm_kernel;
    o.block[0] = o.block[1] = o.block[2] = 0;
    int rc = sim_validate(ip, fp, &o);
    if (rc) return rc;
    if (ip->BoundaryConds[0] != 0 || ip->BoundaryConds[1] != 0 || ip->BoundaryConds[2] != 1)
        return bfail(D3Q19_E_UNSUPPORTED, "batched cells are shear cells: periodic x and y, moving walls normal to z");
    for (int d = 0; d < 3; d++)
        if (o.tangential_bc[d] == 1) return bfail(D3Q19_E_UNSUPPORTED, "batched cells take no tangential wall passes");
    const long long nint = (long long)(ip->LatticeSize[0] - 2) * (ip->LatticeSize[1] - 2) * (ip->LatticeSize[2] - 2);
    if (nint >= (1LL << 24)) return bfail(D3Q19_E_UNSUPPORTED, "a batched cell holds fewer than 2^24 nodes");
    if (ip->NumParticles > 0 && !pts) return bfail(D3Q19_E_INVALID, "sphere_points required when NumParticles > 0");
    int slot = -1;
    for (int i = 0; i < b->max_cells; i++)
        if (!b->cells[i].sim) {
            slot = i;
            break;
        }
    if (slot < 0) return bfail(D3Q19_E_CAPACITY, "batch holds its maximum number of cells");
    const size_t bytes = sim_state_bytes(ip, fp, &o);
    HIP_OK(hipSetDevice(b->device));
    HIP_OK(hipStreamSynchronize(b->stream));
    void* block = b->arena.alloc(bytes);
    if (!block) return bfail(D3Q19_E_CAPACITY, "batch arena cannot hold the cell");
    Member& m = b->cells[slot];
    m.block = block;
    m.sub = Arena();
    m.sub.base = (char*)block;
    m.sub.size = bytes;
    m.sub.free_[0] = bytes;
    d3q19_sim* s = new d3q19_sim();
    rc = sim_init(s, b->device, ip, fp, &o, pts, &m.sub, b->stream, b->epochs);
    // a member's own step numbers start at 1 with the batch's next step
    s->h_epoch = b->h_epoch;
    s->tofs = (int)-b->h_epoch;
    if (!rc && kin && s->np > 0) rc = d3q19_write_particles(s, kin, nullptr, nullptr);
    if (rc) {
        sim_release(s);
        delete s;
        b->arena.release(block);
        m = Member();
        return rc;
    }
    m.sim = s;
    m.par = b->phase;  // f[par ^ phase] = fA and posb[par ^ phase] = the initial positions
    m.pb[0] = s->pos;
    m.pb[1] = s->pos_alt;
    m.vb[0] = s->vel;
    m.vb[1] = s->vel_alt;
    s->parity = 0;
    b->dirty = true;
    *id = slot;
    return 0;
}

static int batch_remove(Batch* b, int id) {
    if (id < 0 || id >= b->max_cells || !b->cells[id].sim) return bfail(D3Q19_E_INVALID, "no such cell");
    HIP_OK(hipSetDevice(b->device));
    HIP_OK(hipStreamSynchronize(b->stream));
    Member& m = b->cells[id];
    sim_release(m.sim);
    delete m.sim;
    b->arena.release(m.block);
    m = Member();
    b->dirty = true;
    return 0;
}

static void batch_free(Batch* b) {
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (Member& m : b->cells)
        if (m.sim) {
            sim_release(m.sim);
            delete m.sim;
        }
    b->cells.clear();
    void* ptrs[] = {b->arena.base, b->epochs, b->d_cells, b->d_tiles, b->d_sph, b->d_msph, b->d_cnt, b->d_obs, b->d_wd};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b->graph2) (void)hipGraphExecDestroy(b->graph2);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

static int batch_create(int device, const d3q19_batch_options* bo, Batch** out) {
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return bfail(D3Q19_E_DEVICE, "no such HIP device");
    Batch* b = new Batch();
    if (bo) b->bo = *bo;
    else memset(&b->bo, 0, sizeof(b->bo));
    if (b->bo.storage < 0 || b->bo.storage > 3 || (b->bo.collision != D3Q19_COLL_MRT && b->bo.collision != D3Q19_COLL_BGK) ||
        (b->bo.ibm_kernel != D3Q19_IBM_LINEAR_SMOOTHED && b->bo.ibm_kernel != D3Q19_IBM_PESKIN4)) {
        delete b;
        return bfail(D3Q19_E_INVALID, "bad batch options");
    }
    b->device = device;
    b->max_cells = b->bo.max_cells > 0 ? b->bo.max_cells : 1024;
    if (b->max_cells > 65536) {
        delete b;
        return bfail(D3Q19_E_INVALID, "max_cells above 65536");
    }
    hipError_t e = hipSetDevice(device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e == hipSuccess) b->ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    size_t freeb = 0, totalb = 0;
    if (e == hipSuccess) e = hipMemGetInfo(&freeb, &totalb);
    size_t ab = b->bo.arena_bytes ? (size_t)b->bo.arena_bytes : freeb / 2;
    ab &= ~(size_t)255;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void**)&b->arena.base, ab);
    if (e == hipSuccess) {
        b->arena.size = ab;
        b->arena.free_[0] = ab;
    }
    if (e == hipSuccess) e = hipMalloc((void**)&b->epochs, 2 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(b->epochs, 0, 2 * sizeof(uint32_t), b->stream);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_cells, (size_t)b->max_cells * sizeof(CellDesc));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_cnt, sizeof(BatchCounts));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_obs, (size_t)OBS_STRIDE * b->max_cells * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_wd, (size_t)2 * b->max_cells * sizeof(int));
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        const int code = (e == hipErrorOutOfMemory) ? D3Q19_E_NOMEM : D3Q19_E_DEVICE;
        const std::string msg = std::string("batch creation: ") + hipGetErrorString(e);
        batch_free(b);
        delete b;
        return bfail(code, msg);
    }
    b->cells.resize(b->max_cells);
    b->h_cells.resize(b->max_cells);
    default_tuning(b);
    b->dirty = true;
    *out = b;
    return 0;
}

// ---- serving -----------------------------------------------------------------------------------------
static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    int_param_struct ip;
    flp_param_struct fp;
    d3q19_options opt;
    std::vector<d3q19_float4> pts, kin;
    d3q19_solve_params sp;
    double t_submit = 0, t_admit = 0;
    int cell = -1;
    int steps = 0;
    double last = 0.0;
    bool have_last = false;
};

struct Server {
    Batch* b = nullptr;
    std::deque<Request*> queue;
    std::vector<Request*> resident;
    std::deque<d3q19_result> outbox;  // finished, not yet handed to the caller
    int64_t next_id = 1;
    int interval = 0;  // steps per scheduling round = the requests' check interval
};

// convergence rule shared by the server and the sequential solve; returns 1 when finished.
// nonfinite / moved are the cell's watchdog words at this check: a non-finite or non-positive
// density anywhere in its lattice, or a sphere that moved a lattice unit or more in one step (its
// dilated brick marks no longer covered its stencils, so the run is not trusted).  Either ends the
// request with an error status (D3Q19_E_STATE for the displacement, D3Q19_E_NONFINITE otherwise).
static int check_request(Request* r, const d3q19_observables& o, int nonfinite, int moved, d3q19_result* res) {
    const double sw = o.wall_stress;
    int status = -99;
    if (moved) status = D3Q19_E_STATE;
    else if (nonfinite || !std::isfinite(sw)) status = D3Q19_E_NONFINITE;
    else if (r->have_last && r->steps >= r->sp.min_steps && fabs(sw - r->last) <= r->sp.tolerance * fabs(sw)) status = 0;
    else if (r->steps >= r->sp.max_steps) status = 1;
    r->last = sw;
    r->have_last = true;
    if (status == -99) return 0;
    memset(res, 0, sizeof(*res));
    res->id = r->id;
    res->status = status;
    res->steps = r->steps;
    res->obs = o;
    res->t_submit = r->t_submit;
    res->t_admit = r->t_admit;
    res->t_finish = now_s();
    return 1;
}

static int valid_solve_params(const d3q19_solve_params* sp) {
    return sp && sp->check_interval > 0 && (sp->check_interval % 2) == 0 && sp->max_steps > 0 && sp->tolerance >= 0.0 && sp->min_steps >= 0;
}

static void drain(Server* v, d3q19_result* done, int cap, int* n_done) {
    while (!v->outbox.empty() && *n_done < cap) {
        done[(*n_done)++] = v->outbox.front();
        v->outbox.pop_front();
    }
}

static int server_progress(Server* v, d3q19_result* done, int cap, int* n_done) {
    *n_done = 0;
    Batch* b = v->b;
    int rc;
    // admission in arrival order while the arena holds the next request
    while (!v->queue.empty()) {
        Request* r = v->queue.front();
        int id = -1;
        rc = batch_add(b, &r->ip, &r->fp, &r->opt, r->pts.empty() ? nullptr : r->pts.data(), r->kin.empty() ? nullptr : r->kin.data(), &id);
        if (rc == D3Q19_E_CAPACITY) {
            if (v->resident.empty()) return bfail(D3Q19_E_CAPACITY, "request larger than the whole batch arena");
            break;
        }
        if (rc) return rc;
        r->cell = id;
        r->t_admit = now_s();
        v->queue.pop_front();
        v->resident.push_back(r);
    }
    if (v->resident.empty()) {
        drain(v, done, cap, n_done);
        return 0;
    }
    if ((rc = batch_run(b, v->interval, 0))) return rc;
    std::vector<int> ids;
    for (Request* r : v->resident) ids.push_back(r->cell);
    // observables and watchdog of every resident cell on its current lattice (one launch each for
    // all cells, one synchronisation)
    std::vector<d3q19_observables> obs(ids.size());
    std::vector<int> bad(ids.size(), 0), moved(ids.size(), 0);
    if ((rc = batch_check(b, ids.data(), (int)ids.size(), obs.data(), bad.data(), moved.data()))) return rc;
    std::vector<Request*> keep;
    for (size_t i = 0; i < v->resident.size(); i++) {
        Request* r = v->resident[i];
        r->steps += v->interval;
        d3q19_result res;
        if (check_request(r, obs[i], bad[i], moved[i], &res)) {
            v->outbox.push_back(res);
            rc = batch_remove(b, r->cell);
            delete r;
            if (rc) return rc;
        } else {
            keep.push_back(r);
        }
    }
    v->resident.swap(keep);
    drain(v, done, cap, n_done);
    return 0;
}

}  // namespace d3q19

struct d3q19_batch : public d3q19::Batch {};
struct d3q19_server : public d3q19::Server {};

#define CHECK_BATCH(b)                                                         \
    do {                                                                       \
        if (!(b)) return d3q19::bfail(D3Q19_E_INVALID, "null batch");          \
        HIP_OK(hipSetDevice((b)->device));                                     \
    } while (0)

extern "C" {

int d3q19_batch_create(int device, const d3q19_batch_options* bo, d3q19_batch** out) {
    if (!out) return bfail(D3Q19_E_INVALID, "null output");
    *out = nullptr;
    Batch* b = nullptr;
    int rc = batch_create(device, bo, &b);
    if (rc) return rc;
    *out = static_cast<d3q19_batch*>(b);
    return 0;
}

int d3q19_batch_destroy(d3q19_batch* b) {
    if (!b) return 0;
    (void)hipSetDevice(b->device);
    batch_free(b);
    delete static_cast<Batch*>(b);
    return 0;
}

int d3q19_batch_add(d3q19_batch* b, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt,
                    const d3q19_float4* pts, const d3q19_float4* kin, int* id) {
    CHECK_BATCH(b);
    return batch_add(b, ip, fp, opt, pts, kin, id);
}

int d3q19_batch_remove(d3q19_batch* b, int id) {
    CHECK_BATCH(b);
    return batch_remove(b, id);
}

int d3q19_batch_cell_bytes(const d3q19_batch* b, const int_param_struct* ip, size_t* bytes) {
    if (!b || !ip || !bytes) return bfail(D3Q19_E_INVALID, "null argument");
    d3q19_options o;
    memset(&o, 0, sizeof(o));
    o.storage = b->bo.storage;
    flp_param_struct fp;
    memset(&fp, 0, sizeof(fp));
    *bytes = sim_state_bytes(ip, &fp, &o);
    return 0;
}

int d3q19_batch_memory(const d3q19_batch* b, size_t* arena_bytes, size_t* used_bytes, int* cells) {
    if (!b) return bfail(D3Q19_E_INVALID, "null batch");
    if (arena_bytes) *arena_bytes = b->arena.size;
    if (used_bytes) *used_bytes = b->arena.used;
    if (cells) {
        int n = 0;
        for (const Member& m : b->cells)
            if (m.sim) n++;
        *cells = n;
    }
    return 0;
}

int d3q19_batch_run(d3q19_batch* b, int n, int flags) {
    CHECK_BATCH(b);
    return batch_run(b, n, flags);
}

int d3q19_batch_sync(d3q19_batch* b) {
    CHECK_BATCH(b);
    HIP_OK(hipStreamSynchronize(b->stream));
    return 0;
}

int d3q19_batch_observables(d3q19_batch* b, const int* ids, int n, d3q19_observables* out) {
    CHECK_BATCH(b);
    if (n < 0 || (n > 0 && (!ids || !out))) return bfail(D3Q19_E_INVALID, "bad observables request");
    return batch_observables(b, ids, n, out);
}

static d3q19_sim* member(d3q19_batch* b, int id) {
    if (!b || id < 0 || id >= b->max_cells || !b->cells[id].sim) return nullptr;
    Member& m = b->cells[id];
    sync_view(m, b->phase);
    return m.sim;
}

int d3q19_batch_read_u(d3q19_batch* b, int id, float* u) {
    CHECK_BATCH(b);
    d3q19_sim* s = member(b, id);
    if (!s) return bfail(D3Q19_E_INVALID, "no such cell");
    return d3q19_read_u(s, u);
}

int d3q19_batch_read_f(d3q19_batch* b, int id, float* f) {
    CHECK_BATCH(b);
    d3q19_sim* s = member(b, id);
    if (!s) return bfail(D3Q19_E_INVALID, "no such cell");
    return d3q19_read_f(s, f);
}

int d3q19_batch_read_particles_f64(d3q19_batch* b, int id, double* pos, double* vel, double* av, double* ff, double* tf) {
    CHECK_BATCH(b);
    d3q19_sim* s = member(b, id);
    if (!s) return bfail(D3Q19_E_INVALID, "no such cell");
    return d3q19_read_particles_f64(s, pos, vel, av, ff, tf);
}

int d3q19_batch_autotune(d3q19_batch* b, d3q19_tuning* out) {
    CHECK_BATCH(b);
    int rc = autotune(b);
    if (rc) return rc;
    if (out) *out = b->tuning;
    return 0;
}

int d3q19_batch_get_tuning(const d3q19_batch* b, d3q19_tuning* out) {
    if (!b || !out) return bfail(D3Q19_E_INVALID, "null argument");
    *out = b->tuning;
    return 0;
}

int d3q19_batch_set_tuning(d3q19_batch* b, const d3q19_tuning* t) {
    CHECK_BATCH(b);
    if (!t || !valid_tuning(t)) return bfail(D3Q19_E_INVALID, "bad tuning");
    HIP_OK(hipStreamSynchronize(b->stream));
    b->tuning = *t;
    b->dirty = true;
    b->graph_valid = false;
    return 0;
}

int d3q19_batch_time_steps(d3q19_batch* b, int n, double* total_ms, double* collide_ms) {
    CHECK_BATCH(b);
    if (n <= 0 || !total_ms) return bfail(D3Q19_E_INVALID, "bad timing request");
    int rc;
    if (b->dirty && (rc = upload_tables(b))) return rc;
    if (n >= 2 && !b->graph_valid && (rc = capture_graph(b))) return rc;
    HIP_OK(hipStreamSynchronize(b->stream));
    HIP_OK(hipEventRecord(b->ev0, b->stream));
    if ((rc = batch_run(b, n, 0))) return rc;
    HIP_OK(hipEventRecord(b->ev1, b->stream));
    HIP_OK(hipEventSynchronize(b->ev1));
    float t = 0;
    HIP_OK(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *total_ms = (double)t;
    if (collide_ms && (rc = time_probe(b, 16, collide_ms))) return rc;
    return 0;
}

int d3q19_server_create(int device, const d3q19_batch_options* bo, d3q19_server** out) {
    if (!out) return bfail(D3Q19_E_INVALID, "null output");
    *out = nullptr;
    Batch* b = nullptr;
    int rc = batch_create(device, bo, &b);
    if (rc) return rc;
    d3q19_server* v = new d3q19_server();
    v->b = b;
    *out = v;
    return 0;
}

int d3q19_server_destroy(d3q19_server* v) {
    if (!v) return 0;
    for (Request* r : v->queue) delete r;
    for (Request* r : v->resident) delete r;
    if (v->b) d3q19_batch_destroy(static_cast<d3q19_batch*>(v->b));
    delete v;
    return 0;
}

int d3q19_server_submit(d3q19_server* v, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt,
                        const d3q19_float4* pts, const d3q19_float4* kin, const d3q19_solve_params* sp, int64_t* id) {
    if (!v || !ip || !fp || !id) return bfail(D3Q19_E_INVALID, "null argument");
    if (!valid_solve_params(sp)) return bfail(D3Q19_E_INVALID, "bad solve parameters (check_interval must be even and > 0)");
    if (v->interval == 0) v->interval = sp->check_interval;
    if (sp->check_interval != v->interval) return bfail(D3Q19_E_INVALID, "all requests of a server share one check_interval");
    if (ip->NumParticles > 0 && !pts) return bfail(D3Q19_E_INVALID, "sphere_points required when NumParticles > 0");
    Request* r = new Request();
    r->id = v->next_id++;
    r->ip = *ip;
    r->fp = *fp;
    if (opt) r->opt = *opt;
    else memset(&r->opt, 0, sizeof(r->opt));
    r->sp = *sp;
    if (ip->NumParticles > 0) {
        r->pts.assign(pts, pts + ip->PointsPerParticle);
        if (kin) r->kin.assign(kin, kin + 4 * (size_t)ip->NumParticles);
    }
    r->t_submit = now_s();
    v->queue.push_back(r);
    *id = r->id;
    return 0;
}

int d3q19_server_progress(d3q19_server* v, d3q19_result* done, int cap, int* n_done) {
    if (!v || !n_done || (cap > 0 && !done)) return bfail(D3Q19_E_INVALID, "null argument");
    HIP_OK(hipSetDevice(v->b->device));
    return server_progress(v, done, cap, n_done);
}

int d3q19_server_status(const d3q19_server* v, int* queued, int* resident, size_t* used) {
    if (!v) return bfail(D3Q19_E_INVALID, "null server");
    if (queued) *queued = (int)v->queue.size();
    if (resident) *resident = (int)(v->resident.size() + v->outbox.size());
    if (used) *used = v->b->arena.used;
    return 0;
}

int d3q19_server_get_batch(d3q19_server* v, d3q19_batch** b) {
    if (!v || !b) return bfail(D3Q19_E_INVALID, "null argument");
    *b = static_cast<d3q19_batch*>(v->b);
    return 0;
}

int d3q19_solve(int device, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt,
                const d3q19_float4* pts, const d3q19_float4* kin, const d3q19_solve_params* sp, d3q19_result* out) {
    if (!ip || !fp || !out) return bfail(D3Q19_E_INVALID, "null argument");
    if (!valid_solve_params(sp)) return bfail(D3Q19_E_INVALID, "bad solve parameters (check_interval must be even and > 0)");
    Request r;
    r.id = 0;
    r.sp = *sp;
    r.t_submit = r.t_admit = now_s();
    d3q19_sim* s = nullptr;
    int rc = d3q19_create(device, ip, fp, opt, pts, &s);
    if (rc) return rc;
    if (kin && ip->NumParticles > 0) rc = d3q19_write_particles(s, kin, nullptr, nullptr);
    int t = 1;
    while (!rc) {
        if ((rc = d3q19_run(s, t, sp->check_interval, 0))) break;
        t += sp->check_interval;
        r.steps += sp->check_interval;
        d3q19_observables o;
        if ((rc = d3q19_observables_get(s, &o))) break;
        int bad = 0, moved = 0;
        if ((rc = sim_watchdog(s, &bad, &moved))) break;
        if (check_request(&r, o, bad, moved, out)) break;
    }
    d3q19_destroy(s);
    return rc;
}

double d3q19_clock(void) { return now_s(); }

}  // extern "C"
