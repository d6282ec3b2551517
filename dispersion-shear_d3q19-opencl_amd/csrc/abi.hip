// abi.hip -- the C-ABI of include/d3q19_abi.h over HIP.
//
// Replaces the reference's OpenCL plumbing: device discovery (analyse_platform,
// D3Q19-OpenCL.c:847-971), buffer creation and upload (sim_main.c:115-206), kernel creation and
// argument binding (create_LB_kernels D3Q19-OpenCL.c:465-556, sim_main.c:241-306), the per-step
// enqueue sequence (sim_main.c:315-471) and error_check (D3Q19-OpenCL.c:1169-1246; here errors
// are returned, never exit()).
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <cmath>
#include <utility>
#include <string>
#include <mutex>
#include <vector>
#include "internal.h"

namespace d3q19 {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

static int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

static size_t pop_size(int storage) {
    switch (storage) {
    case D3Q19_STORE_FP16_SHIFT:
    case D3Q19_STORE_BF16_SHIFT: return 2;
    case D3Q19_STORE_FP64: return 8;
    default: return 4;
    }
}
static size_t real_size(int storage) { return storage == D3Q19_STORE_FP64 ? 8 : 4; }

static int validate(const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt) {
    for (int d = 0; d < 3; d++) {
        if (ip->LatticeSize[d] < 5) return fail(D3Q19_E_INVALID, "LatticeSize must be >= 5 on every axis");
        if (ip->LatticeSize[d] - 2 > 4096) return fail(D3Q19_E_UNSUPPORTED, "interior extent above 4096 on an axis");
        if (ip->BufferSize[d] != 1) return fail(D3Q19_E_INVALID, "BufferSize must be 1 (one-node shell)");
        if (ip->BoundaryConds[d] != 0 && ip->BoundaryConds[d] != 1) return fail(D3Q19_E_INVALID, "BoundaryConds must be 0 or 1");
        const int want = ip->BoundaryConds[d] == 0 ? ip->LatticeSize[d] - 2 : ip->LatticeSize[d] - 3;
        if (ip->SystemSize[d] != want) return fail(D3Q19_E_INVALID, "SystemSize inconsistent with LatticeSize/BoundaryConds (D3Q19-OpenCL.c:119-128)");
    }
    if (ip->BoundaryConds[0] + ip->BoundaryConds[1] + ip->BoundaryConds[2] > 1)
        return fail(D3Q19_E_INVALID, "more than one pair of velocity-boundary faces");
    const long long nint = (long long)(ip->LatticeSize[0] - 2) * (ip->LatticeSize[1] - 2) * (ip->LatticeSize[2] - 2);
    if (nint >= (1LL << 31)) return fail(D3Q19_E_UNSUPPORTED, "more than 2^31 interior nodes");
    // per-node arrays and planar population planes are addressed by 32-bit byte offsets (no
    // options: fp32 storage); chunked populations by a 32-bit lane offset within one z-plane (Q x
    // nx x ny values, below 4 GiB for any extents up to 4096)
    if (nint * (long long)pop_size(opt ? opt->storage : 0) >= (1LL << 32))
        return fail(D3Q19_E_UNSUPPORTED, "a population plane above 4 GiB (interior nodes x bytes per value)");
    if (!(fp->NewtonianTau > 0.5f)) return fail(D3Q19_E_INVALID, "NewtonianTau must exceed 0.5");
    if (ip->NumParticles < 0) return fail(D3Q19_E_INVALID, "NumParticles < 0");
    if (ip->NumParticles > 0) {
        if (ip->PointsPerParticle <= 0 || ip->PointsPerParticle > 4096) return fail(D3Q19_E_INVALID, "PointsPerParticle out of range");
        if (!(fp->ParticleMass > 0.f) || !(fp->ParticleMomInertia > 0.f)) return fail(D3Q19_E_INVALID, "particle mass / inertia must be positive");
        if ((long long)ip->NumParticles * ip->PointsPerParticle >= (1LL << 31) / 6) return fail(D3Q19_E_UNSUPPORTED, "too many markers");
    }
    if (opt) {
        if (opt->collision != D3Q19_COLL_MRT && opt->collision != D3Q19_COLL_BGK) return fail(D3Q19_E_INVALID, "unknown collision operator");
        if (opt->ibm_kernel != D3Q19_IBM_LINEAR_SMOOTHED && opt->ibm_kernel != D3Q19_IBM_PESKIN4) return fail(D3Q19_E_INVALID, "unknown IBM kernel");
        if (opt->storage < 0 || opt->storage > 3) return fail(D3Q19_E_INVALID, "unknown population storage");
        if (opt->block[0] < 0 || opt->block[1] < 0) return fail(D3Q19_E_INVALID, "negative block size");
        if (opt->block[0] > 0 && (opt->block[0] * (opt->block[1] > 0 ? opt->block[1] : 1) > 256 ||
                                   (opt->block[0] * (opt->block[1] > 0 ? opt->block[1] : 1)) % 64 != 0))
            return fail(D3Q19_E_INVALID, "block must hold a multiple of 64 and at most 256 threads");
        if (opt->rounding != D3Q19_ROUND_STOCHASTIC && opt->rounding != D3Q19_ROUND_NEAREST)
            return fail(D3Q19_E_INVALID, "unknown rounding mode");
        if (opt->overlap_slabs < 0 || opt->overlap_slabs > Sim::MAX_SLABS)
            return fail(D3Q19_E_INVALID, "overlap_slabs must be 0 (automatic) to 16");
        if (opt->ibm_layout < D3Q19_IBM_LAYOUT_AUTO || opt->ibm_layout > D3Q19_IBM_LAYOUT_ROWS_LARGE)
            return fail(D3Q19_E_INVALID, "ibm_layout must be 0 (automatic) to 3");
        if (opt->contact_detection != D3Q19_CONTACT_ZONES && opt->contact_detection != D3Q19_CONTACT_ALL_PAIRS)
            return fail(D3Q19_E_INVALID, "unknown contact detection");
        if (opt->contact_detection == D3Q19_CONTACT_ZONES && ip->NumParticles > 0) {
            long long nzt = 1;
            for (int d = 0; d < 3; d++) {
                if (ip->NumZones[d] < 1 || !(fp->ZoneWidth[d] > 0.f))
                    return fail(D3Q19_E_INVALID, "contact zones need NumZones >= 1 and ZoneWidth > 0 (D3Q19-OpenCL.c:366-375)");
                nzt *= ip->NumZones[d];
            }
            if (nzt >= (1LL << 30)) return fail(D3Q19_E_UNSUPPORTED, "more than 2^30 contact zones");
        }
    }
    return 0;
}

// population layouts (pop_index): chunks of 2^lr rows, or planes
static void set_chunked(Geom& g, int lr) {
    g.lr = lr;
    g.qs = (long long)g.nx << lr;
    g.cs = (int)(Q * g.qs);
    g.ps = (long long)Q * g.nx * g.ny;
}
static void set_planar(Geom& g) {
    g.lr = PLANAR_LR;
    g.qs = g.nint;
    g.cs = 0;
    g.ps = (long long)g.nx * g.ny;
}

static void geom_from(Sim* s) {
    Geom& g = s->g;
    g.nx = s->Nx - 2;
    g.ny = s->Ny - 2;
    g.nz = s->Nz - 2;
    g.nint = g.nx * g.ny * g.nz;
    // rows per population chunk (pop_index) of a batch member: the smallest power of two R dividing
    // ny with R * nx a multiple of 64 nodes; failing that, the largest R dividing ny with R * nx below
    // 128 (sim_init sets one row per chunk for a single simulation)
    int lr = -1;
    for (int k = 0; k <= 6 && g.ny % (1 << k) == 0; k++)
        if ((g.nx << k) % 64 == 0) {
            lr = k;
            break;
        }
    if (lr < 0) {
        lr = 0;
        for (int k = 1; k <= 6 && g.ny % (1 << k) == 0; k++)
            if ((g.nx << k) < 128) lr = k;
    }
    set_chunked(g, lr);
    g.wall = -1;
    for (int d = 0; d < 3; d++) {
        g.per[d] = s->ip.BoundaryConds[d] == 0 ? 1 : 0;
        if (!g.per[d]) g.wall = d;
        g.sys[d] = s->ip.SystemSize[d];
    }
    g.bnx = (g.nx + BRICK - 1) / BRICK;
    g.bny = (g.ny + BRICK - 1) / BRICK;
    g.bnz = (g.nz + BRICK - 1) / BRICK;
    FluidParams& f = s->fl;
    for (int d = 0; d < 3; d++) {
        f.gconst[d] = s->fp.ConstBodyForce[d];
        f.uw[0][d] = s->fp.VelLower[d];
        f.uw[1][d] = s->fp.VelUpper[d];
    }
    f.tau0 = s->fp.NewtonianTau;
    f.model = s->ip.ViscosityModel;
    for (int k = 0; k < 4; k++) f.vp[k] = s->fp.ViscosityParams[k];
    f.tau_min = s->opt.tau_min > 0.f ? s->opt.tau_min : 0.505f;
    f.tau_max = s->opt.tau_max > 0.f ? s->opt.tau_max : 0.f;
    f.calc_rho = s->opt.calc_rho;
    f.tan_mask = 0;
    f.sr = (s->storage == D3Q19_STORE_FP16_SHIFT || s->storage == D3Q19_STORE_BF16_SHIFT) &&
           s->opt.rounding == D3Q19_ROUND_STOCHASTIC;
    if (g.wall >= 0)  // tangential passes only run when a velocity boundary exists (sim_main.c:383-405)
        for (int d = 0; d < 3; d++)
            if (s->opt.tangential_bc[d] == 1) f.tan_mask |= (1 << d);
}

// Default tile: one block along x (a row segment of up to 256 nodes, or rows of nx rounded up to 64
// stacked to 256 threads).  Row tiles measured 3-8% faster than 64x4 on MI355X.  The sphere sums of
// the fused step do not depend on the block size, so the tile changes no result bit.
static void default_block(Sim* s) {
    int bx = s->opt.block[0], by = s->opt.block[1];
    if (bx <= 0) {
        // one x-row segment of up to 256 nodes per block; rows of 128 and more alone (measured on
        // the 128^3 workload: 128x1 tiles 2% faster than 128x2, fp32 and fp16), shorter rows stacked
        const int nx = s->g.nx;
        bx = nx >= 256 ? 256 : ((nx + 63) / 64) * 64;
        if (bx > 256) bx = 256;
        by = bx >= 128 ? 1 : 256 / bx;
    }
    if (by <= 0) by = 256 / bx;
    s->block[0] = bx;
    s->block[1] = by;
    s->block[2] = 1;
    // Rows of 257-511 nodes (not a multiple of 64) take two row tiles, the second mostly idle (a
    // 288-node row: 32 of 256 lanes).  The default tile there is 256 consecutive nodes of the plane
    // in node order, across rows.  Measured step kernel, fp32: 288^2x100 263 -> 218 us, 300^2x90
    // 361 -> 289, 360^2x64 317 -> 258, 400^2x50 248 -> 213, 480^2x36 232 -> 219.  Rows of at most
    // 256 nodes lost with it (200^3 231 -> 266 us, 142^3 89 -> 101), rows of 520-700 nodes 2%.
    // Planes stay below 2^24 nodes, where the float split of the node index is exact to +-1.
    const int nx = s->g.nx;
    s->packed_tiles = (s->opt.block[0] <= 0 && nx > 256 && nx < 512 && nx % 64 != 0 &&
                       (long long)nx * s->g.ny < (1LL << 24)) ? 1 : 0;
}

static void* cur_f(Sim* s) { return s->parity == 0 ? s->fA : s->fB; }
static void* nxt_f(Sim* s) { return s->parity == 0 ? s->fB : s->fA; }

// Watchdog of a simulation or batch member on its current lattice: a density pass over the
// populations (non-finite or non-positive density anywhere) and the spheres' displacement flag
// (set when a sphere moves a lattice unit or more in one step, which its dilated brick marks do not
// cover).  The pass writes its flag into a spare word of the simulation's reduction scratch, so
// nothing is allocated.  The enqueue form copies both words to host memory asynchronously; they are
// valid after the stream is synchronised.
int sim_watchdog_enqueue(Sim* s, int* h_nonfinite, int* h_moved) {
    int* flag = (int*)(s->red + 4 * 512 + 4);
    HIP_OK(hipMemsetAsync(flag, 0, sizeof(int), s->stream));
    const int rc = launch_check_finite(s, cur_f(s), flag);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(h_nonfinite, flag, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    *h_moved = 0;
    if (s->np > 0 && s->d_flag) HIP_OK(hipMemcpyAsync(h_moved, s->d_flag, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    return 0;
}

int sim_watchdog(Sim* s, int* nonfinite, int* moved) {
    int h[2] = {0, 0};
    const int rc = sim_watchdog_enqueue(s, &h[0], &h[1]);
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(s->stream));
    *nonfinite = h[0];
    *moved = h[1];
    return 0;
}

void sim_release(Sim* s) {
    void* ptrs[] = {s->fA, s->fB, s->u, s->tau, s->facc, s->ftag, s->mask[0], s->mask[1], s->mask[2], s->bcraw, s->pos, s->vel,
                    s->pos_alt, s->vel_alt, s->quat, s->omega, s->ffluid, s->tfluid, s->fcont, s->tcont, s->pts, s->mforce,
                    s->red, s->d_desc, s->d_obs, s->d_flag, s->zone, s->perm, s->gbox};
    for (void* p : ptrs)
        if (p) {
            if (s->arena) s->arena->release(p);
            else (void)hipFree(p);
        }
    if (s->epochs && !s->shares_epochs) (void)hipFree(s->epochs);
    if (s->graph2) (void)hipGraphExecDestroy(s->graph2);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    for (hipEvent_t& e : s->sev)
        if (e) (void)hipEventDestroy(e);
    if (s->aux) (void)hipStreamDestroy(s->aux);
    if (s->stream && !s->shares_stream) (void)hipStreamDestroy(s->stream);
}
static void free_all(Sim* s) { sim_release(s); }

// ---- arena ---------------------------------------------------------------------------------------------
void* Arena::alloc(size_t bytes) {
    const size_t need = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < need) continue;
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > need) free_[off + need] = len - need;
        live_[off] = need;
        used += need;
        return base + off;
    }
    return nullptr;
}

bool Arena::release(void* p) {
    const size_t off = (size_t)((char*)p - base);
    auto it = live_.find(off);
    if (it == live_.end()) return false;
    size_t len = it->second;
    live_.erase(it);
    used -= len;
    size_t start = off;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && nx->first == off + len) {  // merge with the following hole
        len += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {  // merge with the preceding hole
            start = pv->first;
            len += pv->second;
            free_.erase(pv);
        }
    }
    free_[start] = len;
    return true;
}

template <class T>
static int dalloc(T** p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(D3Q19_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return 0;
}

// device memory of a simulation: from its arena (batch members) or hipMalloc
template <class T>
static int salloc(Sim* s, T** p, size_t bytes) {
    if (!s->arena) return dalloc(p, bytes);
    if (bytes == 0) bytes = 8;
    *p = (T*)s->arena->alloc(bytes);
    if (!*p) return fail(D3Q19_E_CAPACITY, "batch arena exhausted");
    return 0;
}

CellDesc make_cell_desc(Sim* s) {
    CellDesc d;
    memset(&d, 0, sizeof(d));
    d.c = make_collide_args(s, nullptr, nullptr, 0, 1);
    d.i = make_ibm_args(s);
    d.p[0] = make_part_args(s);
    d.f[0] = s->fA;
    d.f[1] = s->fB;
    d.posb[0] = gp(s->pos);
    d.posb[1] = gp(s->pos_alt);
    d.velb[0] = gp(s->vel);
    d.velb[1] = gp(s->vel_alt);
    set_cell_spheres(d);
    d.par = s->parity;
    d.nint = s->g.nint;
    d.nx = s->g.nx;
    d.nxy = s->g.nx * s->g.ny;
    d.inv_nx = 1.0f / (float)d.nx;
    d.inv_nxy = 1.0f / (float)d.nxy;
    d.pitch = s->batch_pitch > 0 ? s->batch_pitch : d.nx;
    d.pny = d.pitch * s->g.ny;
    d.nslots = d.pny * s->g.nz;
    d.inv_pitch = 1.0f / (float)d.pitch;
    d.inv_pny = 1.0f / (float)d.pny;
    ObsArgs& o = d.o;
    o.axisA = s->g.wall;
    o.coll = s->opt.collision;
    o.rate_fixed = 0.f;
    o.axisF = -1;
    if (o.axisA >= 0) {
        double best = -1.0;
        for (int a = 0; a < 3; a++) {
            if (a == o.axisA) continue;
            const double du = fabs((double)s->fp.VelUpper[a] - (double)s->fp.VelLower[a]);
            if (du > best) {
                best = du;
                o.axisF = a;
            }
        }
        // the reference's MRT relaxes the x-y stress moment at the fixed rate 1.4 (GPU_program.cl:453)
        if (o.coll == D3Q19_COLL_MRT && o.axisA + o.axisF == 1) o.rate_fixed = 1.4f;
    }
    return d;
}

// ---- step building blocks --------------------------------------------------------------------------
static int do_collide_standalone(Sim* s, int flags) {
    (void)flags;
    int rc = launch_collide(s, cur_f(s), nxt_f(s), 1, 0, s->stream);
    if (rc) return rc;
    s->parity ^= 1;
    return 0;
}

// One step in sim_main.c:315-471 order, fused: the spheres' turn -- the previous step's force sums
// and contacts (do_post) and this step's particle_dynamics, which the reference runs on the CPU
// queue beside the collide -- runs in extra blocks of the collide launch, then the IBM pass.  The
// collide writes u in the bricks marked for this epoch, made by the previous fused step (or by the
// prime kernel) from the positions before this step's dynamics, dilated by one lattice unit.
// ev (optional): four events, recorded at the start and end of the step kernel and of the IBM pass
// (with slabs: the first slab's start and the last one's end, for each of the two).
// With s->slabs = K > 1 the collide goes out as K z-slab launches on st, and the IBM pass as K
// launches on the aux stream: pass k runs the spheres whose stencils end in slab k, after an event
// that slab k has collided, while st collides slabs k+1 ...  A pass therefore reads u only where this
// step's collide has written it and spreads only into nodes whose collide has consumed the previous
// force; the last pass commits the epoch, after every slab.  st then waits for the aux stream.
static int enqueue_fused(Sim* s, int flags, int do_post, hipStream_t st, const hipEvent_t* ev = nullptr) {
    int rc;
    const int K = s->slabs;
    if ((rc = launch_step_fused(s, cur_f(s), nxt_f(s), (flags & D3Q19_STEP_WRITE_U) ? 1 : 0, do_post, st, ev ? ev[0] : nullptr,
                                ev ? ev[1] : nullptr, K, K > 1 ? s->sev : nullptr)))
        return rc;
    s->parity ^= 1;
    if (s->np > 0) {
        std::swap(s->pos, s->pos_alt);
        std::swap(s->vel, s->vel_alt);
    }
    if (K <= 1) return launch_ibm(s, st, ev ? ev[2] : nullptr, ev ? ev[3] : nullptr);
    for (int k = 0; k < K; k++) {
        HIP_OK(hipStreamWaitEvent(s->aux, s->sev[k], 0));
        if (ev && k == 0) HIP_OK(hipEventRecord(ev[2], s->aux));
        if ((rc = launch_ibm(s, s->aux, nullptr, nullptr, s->zb[k], s->zb[k + 1], k == K - 1))) return rc;
    }
    if (ev) HIP_OK(hipEventRecord(ev[3], s->aux));
    HIP_OK(hipEventRecord(s->sev[K], s->aux));
    HIP_OK(hipStreamWaitEvent(st, s->sev[K], 0));
    return 0;
}

static int build_graph(Sim* s) {
    if (s->graph2) {
        (void)hipGraphExecDestroy(s->graph2);
        s->graph2 = nullptr;
    }
    hipGraph_t graph;
    HIP_OK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_fused(s, 0, 1, s->stream);
    if (!rc) rc = enqueue_fused(s, 0, 1, s->stream);
    hipError_t e = hipStreamEndCapture(s->stream, &graph);
    if (rc) return rc;
    if (e != hipSuccess) return fail(D3Q19_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&s->graph2, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(D3Q19_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    // capture enqueued nothing; the two flips of the lattice and of the position buffers cancel
    s->graph_f = cur_f(s);
    s->graph_pos = s->pos;
    s->graph_tofs = s->tofs;
    return 0;
}

// n steps: the prime marks if needed, a first fused step without a pending post, graph-replayed
// pairs, the remaining steps (the last one carries flags), then the last step's force sums and
// contacts, so that the state after the call is the reference's state after n steps.  With ev
// (4 events per step) every step is launched directly, its kernels timed by the events.
static int run_steps(Sim* s, int t_first, int n, int flags, const hipEvent_t* ev = nullptr) {
    if (n <= 0) return 0;
    int rc;
    // the contacts of step t read the zones of generation (t - 1) / rebuild_freq
    s->tofs = (int)((long long)t_first - (s->h_epoch + 1));
    s->h_epoch += n;
    if (s->np > 0 && !s->marks_next) {
        if ((rc = launch_sphere_prime(s, s->stream))) return rc;
        s->marks_next = true;
    }
    if ((rc = enqueue_fused(s, n == 1 ? flags : 0, 0, s->stream, ev))) return rc;
    int done = 1;
    const int tail = (flags && n > 1) ? 1 : 0;
    const int pairs = ev ? 0 : (n - 1 - tail) / 2;
    if (pairs > 0) {
        if (!s->graph2 || s->graph_f != cur_f(s) || s->graph_pos != (const void*)s->pos || s->graph_tofs != s->tofs) {
            if ((rc = build_graph(s))) return rc;
        }
        for (int i = 0; i < pairs; i++) HIP_OK(hipGraphLaunch(s->graph2, s->stream));
        done += 2 * pairs;
    }
    while (done < n) {
        if ((rc = enqueue_fused(s, done == n - 1 ? flags : 0, 1, s->stream, ev ? ev + 4 * done : nullptr))) return rc;
        done++;
    }
    return launch_sphere_post(s, 1, 1, s->stream);
}

}  // namespace d3q19

using namespace d3q19;

namespace d3q19 {
// Fields of a simulation derived from its parameter structs (no device work).
static void sim_setup(Sim* s, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt) {
    s->ip = *ip;
    s->fp = *fp;
    if (opt) s->opt = *opt;
    else memset(&s->opt, 0, sizeof(s->opt));
    s->Nx = ip->LatticeSize[0];
    s->Ny = ip->LatticeSize[1];
    s->Nz = ip->LatticeSize[2];
    s->nn = (long long)s->Nx * s->Ny * s->Nz;
    s->storage = s->opt.storage;
    s->pop_bytes = pop_size(s->storage);
    geom_from(s);
    default_block(s);
    s->np = ip->NumParticles;
    s->ppp = s->np > 0 ? ip->PointsPerParticle : 0;
    s->nfa = ip->WorkGroupsPerParticle > 0 ? ip->WorkGroupsPerParticle : 1;
    const int wallA = s->g.wall;
    s->bc_plane = wallA >= 0 ? (wallA == 0 ? s->g.ny * s->g.nz : (wallA == 1 ? s->g.nx * s->g.nz : s->g.nx * s->g.ny)) : 1;
    s->nbricks = s->g.bnx * s->g.bny * s->g.bnz;
}

struct AllocItem {
    void** ptr;
    size_t bytes;
};
template <class T>
static AllocItem item(T** p, size_t bytes) {
    return AllocItem{(void**)p, bytes};
}

// Every device array of a simulation, in allocation order.
static void alloc_list(Sim* s, std::vector<AllocItem>& v) {
    const size_t nint = (size_t)s->g.nint;
    const size_t rs = real_size(s->storage);
    v.push_back(item(&s->fA, (size_t)Q * nint * s->pop_bytes));
    v.push_back(item(&s->fB, (size_t)Q * nint * s->pop_bytes));
    v.push_back(item(&s->u, 3 * nint * rs));
    v.push_back(item(&s->tau, nint * rs));
    v.push_back(item(&s->bcraw, (size_t)10 * s->bc_plane * rs));
    v.push_back(item(&s->red, (size_t)(4 * 512 + 8) * sizeof(double)));
    v.push_back(item(&s->d_desc, sizeof(CellDesc)));
    v.push_back(item(&s->d_obs, (size_t)OBS_STRIDE * sizeof(double)));
    v.push_back(item(&s->d_flag, sizeof(int)));
    if (s->np > 0) {
        const size_t npd = (size_t)s->np;
        v.push_back(item(&s->facc, 3 * nint * sizeof(long long)));
        v.push_back(item(&s->ftag, nint));
        for (int k = 0; k < 3; k++) v.push_back(item(&s->mask[k], (size_t)s->nbricks * sizeof(uint32_t)));
        v.push_back(item(&s->pos, 3 * npd * 8));
        v.push_back(item(&s->vel, 3 * npd * 8));
        v.push_back(item(&s->pos_alt, 3 * npd * 8));
        v.push_back(item(&s->vel_alt, 3 * npd * 8));
        v.push_back(item(&s->quat, 4 * npd * 8));
        v.push_back(item(&s->omega, 3 * npd * 8));
        v.push_back(item(&s->ffluid, 3 * npd * 8));
        v.push_back(item(&s->tfluid, 3 * npd * 8));
        v.push_back(item(&s->fcont, 3 * npd * 8));
        v.push_back(item(&s->tcont, 3 * npd * 8));
        v.push_back(item(&s->pts, 4 * (size_t)s->ppp * sizeof(float)));
        v.push_back(item(&s->mforce, 6 * npd * (size_t)s->ppp * 8));
        v.push_back(item(&s->zone, 2 * npd * sizeof(int)));
        v.push_back(item(&s->perm, (size_t)s->ppp * sizeof(int)));
        v.push_back(item(&s->gbox, 8 * (size_t)((s->ppp + TILE_MARKERS - 1) / TILE_MARKERS) * sizeof(int)));
    }
}

// Device bytes of one simulation's state, each array rounded to the arena's 256-byte granule.
size_t sim_state_bytes(const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt) {
    Sim t;
    sim_setup(&t, ip, fp, opt);
    std::vector<AllocItem> items;
    alloc_list(&t, items);
    size_t b = 0;
    for (const AllocItem& it : items) b += ((it.bytes ? it.bytes : 8) + 255) & ~(size_t)255;
    return b;
}

int sim_validate(const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt) { return validate(ip, fp, opt); }

// The brick marking (mark_sphere) lists at most MARK_BRICKS bricks per axis.  Along an axis a
// sphere's markers have floor coordinates over at most floor(2R) + 2 values, whose stencils (padded
// by up to 3 nodes for a step's motion) cover E = floor(2R) + 9 consecutive nodes, so at most
// ceil(E / 4) + 1 bricks; R is the marking radius (the longest marker offset, at least 0.505 d).
// A larger sphere is refused rather than marked incompletely.
static int check_marking_span(const int_param_struct* ip, const flp_param_struct* fp, const d3q19_float4* pts) {
    double rmax = 0.505 * (double)fp->ParticleDiam + 1e-3;
    for (int m = 0; m < ip->PointsPerParticle; m++)
        rmax = std::max(rmax, std::sqrt((double)pts[m].x * pts[m].x + (double)pts[m].y * pts[m].y + (double)pts[m].z * pts[m].z));
    const double E = std::floor(2.0 * rmax) + 9.0;
    if (!(rmax < 1e6) || (long long)std::ceil(E / 4.0) + 1 > MARK_BRICKS)
        return fail(D3Q19_E_UNSUPPORTED, "sphere too large for the brick marking (at most " + std::to_string(MARK_BRICKS) +
                                             " bricks of 4 nodes per axis: diameter up to about 240 lattice units)");
    return 0;
}

// Allocation and initial state of a simulation (d3q19_create) or of a batch member, whose memory
// comes from the batch arena and which uses the batch's stream and epoch counter.
int sim_init(Sim* s, int device, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt,
             const d3q19_float4* sphere_points, Arena* arena, hipStream_t shared_stream, uint32_t* shared_epochs) {
    int rc = validate(ip, fp, opt);
    if (rc) return rc;
    if (ip->NumParticles > 0 && !sphere_points) return fail(D3Q19_E_INVALID, "sphere_points required when NumParticles > 0");
    if (ip->NumParticles > 0 && (rc = check_marking_span(ip, fp, sphere_points))) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(D3Q19_E_DEVICE, "no such HIP device");
    HIP_OK(hipSetDevice(device));
    s->device = device;
    s->arena = arena;
    sim_setup(s, ip, fp, opt);
    // Batch members: the batched collide walks runs of consecutive slots.  A member keeps the
    // chunks where each run is a whole number of wavefronts; else where its rows, padded to a
    // multiple of 64 slots, stay at least 70% full (the padding slots idle, so that every
    // wavefront stays within one row's run); else it takes planar populations.  Chunks also need
    // the populations within the 32-bit lane offsets of that kernel.
    // Single simulations keep one row per chunk: their collide tiles cover whole rows, and runs of
    // several rows measured slower (184^3 nodes: 189 against 225 us per step) with no shape faster.
    if (!shared_stream) set_chunked(s->g, 0);
    if (shared_stream) {
        const int nx = s->g.nx, pad = ((nx + 63) / 64) * 64;
        const bool fits = (long long)Q * s->g.nint * (long long)pop_size(s->storage) < (1LL << 32) &&
                          (long long)pad * s->g.ny * s->g.nz < (1LL << 31);
        if (fits && ((nx << s->g.lr) % 64) == 0) s->batch_pitch = nx;
        else if (fits && nx * 10 >= pad * 7) s->batch_pitch = pad;
        else {
            set_planar(s->g);
            s->batch_pitch = nx;
        }
    }
    const size_t nint = (size_t)s->g.nint;
    const size_t rs = real_size(s->storage);
    s->nbricks = s->g.bnx * s->g.bny * s->g.bnz;
    // non-temporal population stores beyond the Infinity Cache, loads too beyond NT_LOAD_BYTES where
    // every row's run is a whole number of wavefronts (a batch decides for its cells).  On rows of
    // other lengths non-temporal loads measured slower at every size (200^3 nodes, 1.2 GB: 257
    // against 231 us per step; 256x256x128, rows of 256: 211 against 221 us)
    s->nt_stores = shared_stream ? 0 : nt_level((long long)(2 * Q * pop_size(s->storage) * nint));
    if (s->nt_stores == 2 && s->g.nx % 64 != 0) s->nt_stores = 1;
    hipError_t e = hipSuccess;
    if (shared_stream) {
        s->stream = shared_stream;
        s->shares_stream = true;
    } else {
        e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) return fail(D3Q19_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    {
        std::vector<AllocItem> items;
        alloc_list(s, items);
        for (const AllocItem& it : items)
            if ((rc = salloc(s, it.ptr, it.bytes))) {
                free_all(s);
                return rc;
            }
    }
    if (s->np > 0) {
        if (shared_epochs) {
            s->epochs = shared_epochs;
            s->shares_epochs = true;
        } else if ((rc = dalloc(&s->epochs, 2 * sizeof(uint32_t)))) {
            free_all(s);
            return rc;
        }
    }
    // A batch member's whole state is one block of the batch arena: one fill zeroes it here, before
    // anything is written into it (a server admitting small cells otherwise spent some 20 fill
    // launches per cell); a single simulation zeroes the arrays that start at zero one by one below.
    const bool block_zeroed = s->arena != nullptr;
    if (block_zeroed) {
        const hipError_t ez = hipMemsetAsync(s->arena->base, 0, s->arena->used, s->stream);
        if (ez != hipSuccess) {
            free_all(s);
            return fail(D3Q19_E_DEVICE, std::string("hipMemsetAsync: ") + hipGetErrorString(ez));
        }
    }
    // initial state: f = f_eq(1, 0), u = 0, tau = NewtonianTau (D3Q19-OpenCL.c:158-187)
    rc = launch_fill_equilibrium(s, s->fA);
    if (!rc) rc = launch_fill_equilibrium(s, s->fB);
    if (rc) {
        free_all(s);
        return rc;
    }
    // marker groups of the LDS-staged spreading (gm consecutive entries of perm), made compact so
    // that their stencils overlap and their LDS boxes stay small: seeds are taken from one pole to
    // the other (by z, then azimuth), and each takes its gm - 1 nearest unassigned markers by the
    // largest per-axis distance (ties by index, so the grouping is deterministic).  Against chunks
    // of a Morton order this shrinks the largest box on the 192-marker, d = 12 set from 1890 to 832
    // nodes (groups of 8).  The group size is chosen below.  The grouping decides only which block
    // adds a contribution: integer sums make the spread force independent of it.  Staging is
    // enabled where every periodic extent exceeds twice the stencil reach (|offset| + 3 nodes) plus
    // a margin.
    std::vector<int> perm, gbox;
    if (s->np > 0) {
        double rmax = 0.0, rmin = 1e300;
        for (int m = 0; m < s->ppp; m++) {
            const d3q19_float4 q = sphere_points[m];
            const double r = std::sqrt((double)q.x * q.x + (double)q.y * q.y + (double)q.z * q.z);
            rmax = std::max(rmax, r);
            rmin = std::min(rmin, r);
        }
        s->mark_rin = std::max(0.0, rmin * (1.0 - 1e-6) - 1e-6);
        s->mark_rout = rmax * (1.0 + 1e-6) + 1e-6;
        auto pt = [&](int m, int d) { return (double)(d == 0 ? sphere_points[m].x : d == 1 ? sphere_points[m].y : sphere_points[m].z); };
        // groups of gm markers: perm and 8 ints per group (box lo[3], extent[3], volume, 0); returns
        // the largest box volume
        auto build_groups = [&](int gm, std::vector<int>& pm, std::vector<int>& gb) {
            const int n = s->ppp;
            std::vector<int> order((size_t)n);
            for (int m = 0; m < n; m++) order[(size_t)m] = m;
            std::stable_sort(order.begin(), order.end(), [&](int i, int j) {
                if (pt(i, 2) != pt(j, 2)) return pt(i, 2) > pt(j, 2);
                return std::atan2(pt(i, 1), pt(i, 0)) < std::atan2(pt(j, 1), pt(j, 0));
            });
            std::vector<char> used((size_t)n, 0);
            std::vector<std::pair<double, int>> cand;
            pm.clear();
            pm.reserve((size_t)n);
            for (int sd : order) {
                if (used[(size_t)sd]) continue;
                cand.clear();
                for (int j = 0; j < n; j++)
                    if (!used[(size_t)j]) {
                        double dist = 0.0;
                        for (int d = 0; d < 3; d++) dist = std::max(dist, std::fabs(pt(j, d) - pt(sd, d)));
                        cand.push_back({dist, j});
                    }
                const size_t take = std::min(cand.size(), (size_t)gm);
                std::partial_sort(cand.begin(), cand.begin() + take, cand.end());
                for (size_t k = 0; k < take; k++) {
                    pm.push_back(cand[k].second);
                    used[(size_t)cand[k].second] = 1;
                }
            }
            // LDS box of each group, relative to the sphere's centre node: a marker at offset r has
            // its stencil nodes at floor(c + r) - floor(c) - 1 ... + 2, and floor(c + r) - floor(c)
            // is floor(r) or floor(r) + 1 whatever the centre c; a lane outside the box (rounding at
            // a periodic wrap) spreads directly
            const int ngrp = (n + gm - 1) / gm;
            gb.assign(8 * (size_t)ngrp, 0);
            long long vmax = 0;
            for (int k = 0; k < ngrp; k++) {
                int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
                for (int j = k * gm; j < std::min(n, (k + 1) * gm); j++) {
                    const d3q19_float4 q = sphere_points[pm[(size_t)j]];
                    const float c[3] = {q.x, q.y, q.z};
                    for (int d = 0; d < 3; d++) {
                        const int f = (int)std::floor(c[d]);
                        lo[d] = std::min(lo[d], f - 1);
                        hi[d] = std::max(hi[d], f + 3);
                    }
                }
                long long vol = 1;
                for (int d = 0; d < 3; d++) {
                    gb[8 * (size_t)k + d] = lo[d];
                    gb[8 * (size_t)k + 3 + d] = hi[d] - lo[d] + 1;
                    vol *= hi[d] - lo[d] + 1;
                }
                gb[8 * (size_t)k + 6] = (int)std::min(vol, (long long)(1 << 30));
                vmax = std::max(vmax, vol);
            }
            return vmax;
        };
        // Peskin-4 markers four to a wavefront when there are many (a quarter of the wavefronts,
        // each with 40% of the instructions per marker); with few markers one per wavefront keeps
        // more wavefronts in flight (the 128^3 workload's 6,144 markers: 12.4 against 14.2 us).
        // In the row layout a single simulation takes the largest groups, up to TILE_MARKERS_MAX,
        // whose boxes all fit in TILE_NODES while at least IBM_LARGE_GROUP_MIN_BLOCKS blocks remain:
        // fewer global atomics, since a group's box adds each node once.  The 256x256x128
        // workload's IBM pass (64,000 markers): groups of 8 74 us, 16 57 us, 32 52 us, 64 50 us.
        // Cells of a batch take the largest groups whose boxes fit TILE_NODES_BATCH, whatever their
        // marker count (the batched kernel runs a group in passes of its block).  Both layouts sum
        // a marker's stencil in the same order and the grouping decides only which block adds a
        // contribution, so no choice here changes a result bit.
        const int layout = s->opt.ibm_layout;
        const bool peskin = s->opt.ibm_kernel == D3Q19_IBM_PESKIN4;
        bool large;
        if (layout == D3Q19_IBM_LAYOUT_AUTO) {
            s->ibm_rows = peskin && (long long)s->np * s->ppp >= IBM_ROWS_MIN_MARKERS;
            large = s->ibm_rows;
        } else {
            s->ibm_rows = peskin && layout != D3Q19_IBM_LAYOUT_WAVE;
            large = s->ibm_rows && layout == D3Q19_IBM_LAYOUT_ROWS_LARGE;
        }
        const int cap = shared_stream ? TILE_NODES_BATCH : TILE_NODES;
        s->group_markers = TILE_MARKERS;
        bool chosen = false;
        // A batch member takes the largest groups whose boxes fit, with any layout and marker count:
        // the batched IBM kernel's wavefronts loop over a group's markers, and the batch as a whole
        // has blocks enough.
        for (int gm = TILE_MARKERS_MAX; shared_stream && gm > TILE_MARKERS && !chosen; gm /= 2)
            if (build_groups(gm, perm, gbox) <= cap) {
                s->group_markers = gm;
                chosen = true;
            }
        for (int gm = TILE_MARKERS_MAX; large && !shared_stream && gm > TILE_MARKERS && !chosen; gm /= 2) {
            const long long blocks = (long long)s->np * ((s->ppp + gm - 1) / gm);
            if (blocks < IBM_LARGE_GROUP_MIN_BLOCKS && layout == D3Q19_IBM_LAYOUT_AUTO) continue;
            if (build_groups(gm, perm, gbox) <= cap) {
                s->group_markers = gm;
                chosen = true;
            }
        }
        if (!chosen) (void)build_groups(TILE_MARKERS, perm, gbox);
        s->tile_nodes = 0;
        for (size_t k = 0; k < gbox.size(); k += 8)
            if (gbox[k + 6] <= cap) s->tile_nodes = std::max(s->tile_nodes, gbox[k + 6]);
        const int need = 2 * ((int)std::ceil(rmax) + 3) + 4;
        s->tile_ok = true;
        const int nd[3] = {s->g.nx, s->g.ny, s->g.nz};
        for (int d = 0; d < 3; d++)
            if (s->g.per[d] && nd[d] < need) s->tile_ok = false;
    }
    if (!block_zeroed) e = hipMemsetAsync(s->u, 0, 3 * nint * rs, s->stream);
    if (e == hipSuccess && !block_zeroed) e = hipMemsetAsync(s->d_flag, 0, sizeof(int), s->stream);
    if (e == hipSuccess && s->np > 0) {
        if (!block_zeroed) {
            e = hipMemsetAsync(s->facc, 0, 3 * nint * sizeof(long long), s->stream);
            if (e == hipSuccess) e = hipMemsetAsync(s->ftag, 0, nint, s->stream);
            for (int k = 0; k < 3; k++)
                if (e == hipSuccess) e = hipMemsetAsync(s->mask[k], 0, s->nbricks * sizeof(uint32_t), s->stream);
        }
        static const uint32_t ep0[2] = {0u, 0u};  // next step has epoch 1; no spread force pending
        if (e == hipSuccess && !s->shares_epochs) e = hipMemcpyAsync(s->epochs, ep0, sizeof(ep0), hipMemcpyHostToDevice, s->stream);
        double* zs[] = {s->pos, s->vel, s->pos_alt, s->vel_alt, s->omega, s->ffluid, s->tfluid, s->fcont, s->tcont};
        for (double* z : zs)
            if (e == hipSuccess && !block_zeroed) e = hipMemsetAsync(z, 0, 3 * (size_t)s->np * 8, s->stream);
        if (e == hipSuccess) {
            std::vector<double> q(4 * (size_t)s->np, 0.0);
            for (int p = 0; p < s->np; p++) q[4 * p + 3] = 1.0;
            e = hipMemcpyAsync(s->quat, q.data(), q.size() * 8, hipMemcpyHostToDevice, s->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(s->pts, sphere_points, 4 * (size_t)s->ppp * sizeof(float), hipMemcpyHostToDevice, s->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(s->perm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, s->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(s->gbox, gbox.data(), gbox.size() * sizeof(int), hipMemcpyHostToDevice, s->stream);
            if (e == hipSuccess && launch_zone_init(s, s->stream)) e = hipErrorLaunchFailure;
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        }
    }
    if (e == hipSuccess && launch_fill_tau(s)) e = hipErrorLaunchFailure;  // tau = NewtonianTau
    if (e == hipSuccess) e = hipEventCreate(&s->ev0);
    if (e == hipSuccess) e = hipEventCreate(&s->ev1);
    // Collide in z-slabs overlapped with the IBM pass (single simulations with spheres), on request
    // only: on MI355X the 256x256x128, 500-sphere fp32 step took 0.458 ms in one launch each and
    // 0.478 / 0.512 / 0.565 ms with 2 / 4 / 8 slabs (fp16: 0.329 against 0.354 / 0.392 / 0.450 ms).
    // Running beside the streaming collide, the latency-bound IBM blocks slow it down by more than
    // they save, and every slab adds a launch tail.  Slabs of at least 8 planes.
    if (e == hipSuccess && !s->shares_stream && s->np > 0) {
        int k = s->opt.overlap_slabs;
        if (k == 0) k = 1;
        k = std::min(k, std::max(1, s->g.nz / 8));
        if (k > 1) {
            e = hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking);
            for (int i = 0; i <= k && e == hipSuccess; i++) e = hipEventCreateWithFlags(&s->sev[i], hipEventDisableTiming);
            if (e == hipSuccess) {
                s->slabs = k;
                for (int i = 0; i <= k; i++) s->zb[i] = (int)(((long long)i * s->g.nz) / k);
            }
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        free_all(s);
        return fail(D3Q19_E_DEVICE, std::string("initialisation: ") + hipGetErrorString(e));
    }
    return 0;
}

// The fluid's kinematic viscosity at shear rate gdot: NewtonianTau, or the tau law with its clamps
// when tau is per node (the law of the collide kernel, evaluated in double).
static double fluid_nu(const Sim* s, double gdot) {
    const FluidParams& fl = s->fl;
    if (!s->opt.variable_viscosity) return ((double)fl.tau0 - 0.5) / 3.0;
    const int model = fl.model;
    double tau = (double)fl.tau0;
    if (model == D3Q19_VISC_POWER_LAW || model == D3Q19_VISC_CASSON || model == D3Q19_VISC_HB) {
        const double g = gdot > 1e-8 ? gdot : 1e-8;
        double nu;
        if (model == D3Q19_VISC_POWER_LAW) {
            nu = (double)fl.vp[0] * pow(g, (double)fl.vp[1] - 1.0);
        } else if (model == D3Q19_VISC_CASSON) {
            double a = sqrt((double)fl.vp[0] / g);
            if (fl.vp[2] > 0.f) a *= 1.0 - exp(-sqrt((double)fl.vp[2] * g));
            const double b = a + sqrt((double)fl.vp[1]);
            nu = b * b;
        } else {
            nu = (double)fl.vp[0] / g + (double)fl.vp[1] * pow(g, (double)fl.vp[2] - 1.0);
        }
        tau = 3.0 * nu + 0.5;
        if (tau < (double)fl.tau_min) tau = (double)fl.tau_min;
        if (fl.tau_max > 0.f && tau > (double)fl.tau_max) tau = (double)fl.tau_max;
    }
    return (tau - 0.5) / 3.0;
}

void derive_observables(const Sim* s, const double* r, d3q19_observables* o) {
    memset(o, 0, sizeof(*o));
    const CellDesc d = make_cell_desc(const_cast<Sim*>(s));
    const int A = d.o.axisA, F = d.o.axisF;
    o->wall_axis = A;
    o->flow_axis = F;
    if (r[5] > 0) o->mean_fluid_force = r[4] / r[5];
    if (A < 0 || F < 0) return;
    o->wall_stress_lower = r[1] > 0 ? r[0] / r[1] : 0.0;
    o->wall_stress_upper = r[3] > 0 ? r[2] / r[3] : 0.0;
    o->wall_stress = 0.5 * (o->wall_stress_lower + o->wall_stress_upper);
    const int nA = A == 0 ? s->g.nx : (A == 1 ? s->g.ny : s->g.nz);
    const double du = (double)s->fp.VelUpper[F] - (double)s->fp.VelLower[F];
    o->shear_rate = du / (double)(nA - 1);
    o->fluid_viscosity = fluid_nu(s, fabs(o->shear_rate));
    if (o->shear_rate != 0.0) o->apparent_viscosity = o->wall_stress / o->shear_rate;
    if (o->fluid_viscosity > 0.0) o->relative_viscosity = o->apparent_viscosity / o->fluid_viscosity;
    if (r[5] > 0) {
        const double U = 0.5 * fabs(du), a = 0.5 * (double)s->fp.ParticleDiam;
        if (U > 0 && a > 0) o->drag_coefficient = o->mean_fluid_force / (0.5 * U * U * 3.14159265358979323846 * a * a);
    }
}
}  // namespace d3q19

extern "C" {



const char* d3q19_strerror(int code) {
    switch (code) {
    case D3Q19_OK: return "ok";
    case D3Q19_E_INVALID: return "invalid argument";
    case D3Q19_E_DEVICE: return "HIP runtime error";
    case D3Q19_E_NOMEM: return "out of device memory";
    case D3Q19_E_UNSUPPORTED: return "unsupported configuration";
    case D3Q19_E_STATE: return "invalid state";
    case D3Q19_E_NONFINITE: return "non-finite or non-positive density";
    case D3Q19_E_CAPACITY: return "capacity exceeded";
    default: return "unknown error";
    }
}

const char* d3q19_last_error(void) { return g_last_error.c_str(); }

int d3q19_device_count(int* count) {
    if (!count) return fail(D3Q19_E_INVALID, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail(D3Q19_E_DEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *count = n;
    return 0;
}

int d3q19_device_name(int device, char* buf, size_t len) {
    if (!buf || len == 0) return fail(D3Q19_E_INVALID, "null buffer");
    hipDeviceProp_t p;
    HIP_OK(hipGetDeviceProperties(&p, device));
    snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return 0;
}

int d3q19_footprint(const int_param_struct* ip, const d3q19_options* opt, size_t* bytes) {
    if (!ip || !bytes) return fail(D3Q19_E_INVALID, "null argument");
    flp_param_struct fp;
    memset(&fp, 0, sizeof(fp));
    *bytes = sim_state_bytes(ip, &fp, opt) + (ip->NumParticles > 0 ? 256 : 0);  // + epoch counter
    return 0;
}

int d3q19_create(int device, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt,
                 const d3q19_float4* sphere_points, d3q19_sim** out) {
    if (!ip || !fp || !out) return fail(D3Q19_E_INVALID, "null argument");
    *out = nullptr;
    d3q19_sim* s = new d3q19_sim();
    const int rc = sim_init(s, device, ip, fp, opt, sphere_points, nullptr, nullptr, nullptr);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return 0;
}

int d3q19_destroy(d3q19_sim* s) {
    if (!s) return 0;
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return 0;
}

#define CHECK_SIM(s)                                                   \
    do {                                                               \
        if (!(s)) return fail(D3Q19_E_INVALID, "null simulation");     \
        HIP_OK(hipSetDevice((s)->device));                             \
    } while (0)

// staging copies: host buffer <-> temporary device buffer in reference layout
static int stage_in(d3q19_sim* s, const void* host, size_t bytes, void** dev) {
    int rc = dalloc((char**)dev, bytes);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(*dev, host, bytes, hipMemcpyHostToDevice, s->stream));
    return 0;
}

static int write_f_any(d3q19_sim* s, const void* f, int is_double) {
    CHECK_SIM(s);
    if (!f) return fail(D3Q19_E_INVALID, "null field");
    void* d = nullptr;
    const size_t bytes = (size_t)Q * s->nn * (is_double ? 8 : 4);
    int rc = stage_in(s, f, bytes, &d);
    if (!rc) rc = launch_import_f(s, d, is_double, cur_f(s));
    hipError_t e = hipStreamSynchronize(s->stream);
    if (d) (void)hipFree(d);
    if (rc) return rc;
    if (e != hipSuccess) return fail(D3Q19_E_DEVICE, hipGetErrorString(e));
    return 0;
}

static int read_f_any(d3q19_sim* s, void* f, int as_double) {
    CHECK_SIM(s);
    if (!f) return fail(D3Q19_E_INVALID, "null field");
    void* d = nullptr;
    const size_t bytes = (size_t)Q * s->nn * (as_double ? 8 : 4);
    int rc = dalloc((char**)&d, bytes);
    if (rc) return rc;
    hipError_t e = hipMemsetAsync(d, 0, bytes, s->stream);
    if (e == hipSuccess) rc = launch_export_f(s, cur_f(s), d, as_double);
    if (e == hipSuccess && !rc) e = hipMemcpyAsync(f, d, bytes, hipMemcpyDeviceToHost, s->stream);
    hipError_t e2 = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (rc) return rc;
    if (e != hipSuccess || e2 != hipSuccess) return fail(D3Q19_E_DEVICE, hipGetErrorString(e != hipSuccess ? e : e2));
    return 0;
}

int d3q19_write_f(d3q19_sim* s, const float* f) { return write_f_any(s, f, 0); }
int d3q19_read_f(d3q19_sim* s, float* f) { return read_f_any(s, f, 0); }
int d3q19_write_f64(d3q19_sim* s, const double* f) { return write_f_any(s, f, 1); }
int d3q19_read_f64(d3q19_sim* s, double* f) { return read_f_any(s, f, 1); }

static int write_field(d3q19_sim* s, const float* h, int ncomp, void* dst) {
    CHECK_SIM(s);
    if (!h) return fail(D3Q19_E_INVALID, "null field");
    void* d = nullptr;
    int rc = stage_in(s, h, (size_t)ncomp * s->nn * 4, &d);
    if (!rc) rc = launch_import_field(s, (const float*)d, ncomp, dst);
    hipError_t e = hipStreamSynchronize(s->stream);
    if (d) (void)hipFree(d);
    if (rc) return rc;
    if (e != hipSuccess) return fail(D3Q19_E_DEVICE, hipGetErrorString(e));
    return 0;
}

static int read_field(d3q19_sim* s, float* h, int ncomp, const void* src, int is_force) {
    CHECK_SIM(s);
    if (!h) return fail(D3Q19_E_INVALID, "null field");
    const size_t bytes = (size_t)ncomp * s->nn * 4;
    if (is_force && !s->facc) {
        memset(h, 0, bytes);
        return 0;
    }
    void* d = nullptr;
    int rc = dalloc((char**)&d, bytes);
    if (rc) return rc;
    hipError_t e = hipMemsetAsync(d, 0, bytes, s->stream);
    if (e == hipSuccess) rc = is_force ? launch_export_force(s, (float*)d) : launch_export_field(s, src, ncomp, (float*)d);
    if (e == hipSuccess && !rc) e = hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s->stream);
    hipError_t e2 = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (rc) return rc;
    if (e != hipSuccess || e2 != hipSuccess) return fail(D3Q19_E_DEVICE, hipGetErrorString(e != hipSuccess ? e : e2));
    return 0;
}

int d3q19_write_u(d3q19_sim* s, const float* u) { return write_field(s, u, 3, s ? s->u : nullptr); }
int d3q19_read_u(d3q19_sim* s, float* u) { return read_field(s, u, 3, s ? s->u : nullptr, 0); }
int d3q19_write_tau(d3q19_sim* s, const float* t) { return write_field(s, t, 1, s ? s->tau : nullptr); }
int d3q19_read_tau(d3q19_sim* s, float* t) { return read_field(s, t, 1, s ? s->tau : nullptr, 0); }
int d3q19_read_force(d3q19_sim* s, float* g) { return read_field(s, g, 3, nullptr, 1); }

int d3q19_write_particles(d3q19_sim* s, const d3q19_float4* pk, const d3q19_float4* pf, const d3q19_float4* pff) {
    CHECK_SIM(s);
    const int np = s->np;
    if (np == 0) return 0;
    s->marks_next = false;
    HIP_OK(hipStreamSynchronize(s->stream));
    std::vector<double> a(3 * (size_t)np), b(3 * (size_t)np), c(4 * (size_t)np), d(3 * (size_t)np);
    if (pk) {
        for (int p = 0; p < np; p++) {
            const d3q19_float4 x = pk[p], v = pk[p + np], q = pk[p + 2 * np], w = pk[p + 3 * np];
            a[3 * p] = x.x; a[3 * p + 1] = x.y; a[3 * p + 2] = x.z;
            b[3 * p] = v.x; b[3 * p + 1] = v.y; b[3 * p + 2] = v.z;
            c[4 * p] = q.x; c[4 * p + 1] = q.y; c[4 * p + 2] = q.z; c[4 * p + 3] = q.w;
            d[3 * p] = w.x; d[3 * p + 1] = w.y; d[3 * p + 2] = w.z;
        }
        HIP_OK(hipMemcpy(s->pos, a.data(), a.size() * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->vel, b.data(), b.size() * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->quat, c.data(), c.size() * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->omega, d.data(), d.size() * 8, hipMemcpyHostToDevice));
        // the zone lists start from the uploaded positions (initialize_particle_zones)
        const int rc = launch_zone_init(s, s->stream);
        if (rc) return rc;
    }
    if (pf) {
        for (int p = 0; p < np; p++) {
            a[3 * p] = pf[p].x; a[3 * p + 1] = pf[p].y; a[3 * p + 2] = pf[p].z;
            b[3 * p] = pf[p + np].x; b[3 * p + 1] = pf[p + np].y; b[3 * p + 2] = pf[p + np].z;
        }
        HIP_OK(hipMemcpy(s->fcont, a.data(), a.size() * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->tcont, b.data(), b.size() * 8, hipMemcpyHostToDevice));
    }
    if (pff) {
        // sum of the per-work-group partials (CPU_program.cl:218-223)
        const int nfa = s->nfa;
        for (int p = 0; p < np; p++) {
            double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
            for (int fa = 0; fa < nfa; fa++) {
                const d3q19_float4 f = pff[p * nfa + fa], t = pff[(p + np) * nfa + fa];
                F[0] += f.x; F[1] += f.y; F[2] += f.z;
                T[0] += t.x; T[1] += t.y; T[2] += t.z;
            }
            for (int k = 0; k < 3; k++) { a[3 * p + k] = F[k]; b[3 * p + k] = T[k]; }
        }
        HIP_OK(hipMemcpy(s->ffluid, a.data(), a.size() * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->tfluid, b.data(), b.size() * 8, hipMemcpyHostToDevice));
    }
    return 0;
}

int d3q19_read_particles(d3q19_sim* s, d3q19_float4* pk, d3q19_float4* pf, d3q19_float4* pff) {
    CHECK_SIM(s);
    const int np = s->np;
    if (np == 0) return 0;
    HIP_OK(hipStreamSynchronize(s->stream));
    std::vector<double> a(3 * (size_t)np), b(3 * (size_t)np), c(4 * (size_t)np), d(3 * (size_t)np);
    if (pk) {
        HIP_OK(hipMemcpy(a.data(), s->pos, a.size() * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(b.data(), s->vel, b.size() * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(c.data(), s->quat, c.size() * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(d.data(), s->omega, d.size() * 8, hipMemcpyDeviceToHost));
        for (int p = 0; p < np; p++) {
            pk[p] = {(float)a[3 * p], (float)a[3 * p + 1], (float)a[3 * p + 2], 0.f};
            pk[p + np] = {(float)b[3 * p], (float)b[3 * p + 1], (float)b[3 * p + 2], 0.f};
            pk[p + 2 * np] = {(float)c[4 * p], (float)c[4 * p + 1], (float)c[4 * p + 2], (float)c[4 * p + 3]};
            pk[p + 3 * np] = {(float)d[3 * p], (float)d[3 * p + 1], (float)d[3 * p + 2], 0.f};
        }
    }
    if (pf) {
        HIP_OK(hipMemcpy(a.data(), s->fcont, a.size() * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(b.data(), s->tcont, b.size() * 8, hipMemcpyDeviceToHost));
        for (int p = 0; p < np; p++) {
            pf[p] = {(float)a[3 * p], (float)a[3 * p + 1], (float)a[3 * p + 2], 0.f};
            pf[p + np] = {(float)b[3 * p], (float)b[3 * p + 1], (float)b[3 * p + 2], 0.f};
        }
    }
    if (pff) {
        HIP_OK(hipMemcpy(a.data(), s->ffluid, a.size() * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(b.data(), s->tfluid, b.size() * 8, hipMemcpyDeviceToHost));
        const int nfa = s->nfa;
        for (int p = 0; p < np; p++)
            for (int fa = 0; fa < nfa; fa++) {
                pff[p * nfa + fa] = fa == 0 ? d3q19_float4{(float)a[3 * p], (float)a[3 * p + 1], (float)a[3 * p + 2], 0.f} : d3q19_float4{0, 0, 0, 0};
                pff[(p + np) * nfa + fa] = fa == 0 ? d3q19_float4{(float)b[3 * p], (float)b[3 * p + 1], (float)b[3 * p + 2], 0.f} : d3q19_float4{0, 0, 0, 0};
            }
    }
    return 0;
}

int d3q19_read_particles_f64(d3q19_sim* s, double* pos, double* vel, double* av, double* ff, double* tf) {
    CHECK_SIM(s);
    if (s->np == 0) return 0;
    HIP_OK(hipStreamSynchronize(s->stream));
    const size_t b = 3 * (size_t)s->np * 8;
    if (pos) HIP_OK(hipMemcpy(pos, s->pos, b, hipMemcpyDeviceToHost));
    if (vel) HIP_OK(hipMemcpy(vel, s->vel, b, hipMemcpyDeviceToHost));
    if (av) HIP_OK(hipMemcpy(av, s->omega, b, hipMemcpyDeviceToHost));
    if (ff) HIP_OK(hipMemcpy(ff, s->ffluid, b, hipMemcpyDeviceToHost));
    if (tf) HIP_OK(hipMemcpy(tf, s->tfluid, b, hipMemcpyDeviceToHost));
    return 0;
}

int d3q19_collide_stream(d3q19_sim* s, int flags) {
    CHECK_SIM(s);
    return do_collide_standalone(s, flags);
}

int d3q19_particle_dynamics(d3q19_sim* s) {
    CHECK_SIM(s);
    s->marks_next = false;  // the positions move
    return launch_sphere_pre(s, 1, 0, s->stream);
}

int d3q19_particle_fluid_forces(d3q19_sim* s) {
    CHECK_SIM(s);
    if (s->np == 0) return 0;
    s->marks_next = false;  // this pass commits an epoch
    s->h_epoch++;
    int rc = launch_sphere_pre(s, 0, 1, s->stream);
    if (!rc) rc = launch_ibm(s, s->stream);
    if (!rc) rc = launch_sphere_post(s, 1, 0, s->stream);
    return rc;
}

int d3q19_particle_particle_forces(d3q19_sim* s) {
    CHECK_SIM(s);
    return launch_sphere_post(s, 0, 1, s->stream);
}

int d3q19_step(d3q19_sim* s, int t, int flags) {
    CHECK_SIM(s);
    return run_steps(s, t, 1, flags);
}

int d3q19_run(d3q19_sim* s, int t_first, int n, int flags) {
    CHECK_SIM(s);
    return run_steps(s, t_first, n, flags);
}

int d3q19_sync(d3q19_sim* s) {
    CHECK_SIM(s);
    HIP_OK(hipStreamSynchronize(s->stream));
    return 0;
}

int d3q19_plane_mean_u(d3q19_sim* s, double* means) {
    CHECK_SIM(s);
    if (!means) return fail(D3Q19_E_INVALID, "null output");
    double* d = nullptr;
    const size_t bytes = 3 * (size_t)s->Nz * 8;
    int rc = dalloc(&d, bytes);
    if (rc) return rc;
    hipError_t e = hipMemsetAsync(d, 0, bytes, s->stream);
    if (e == hipSuccess) rc = launch_plane_mean(s, d);
    if (e == hipSuccess && !rc) e = hipMemcpyAsync(means, d, bytes, hipMemcpyDeviceToHost, s->stream);
    hipError_t e2 = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (rc) return rc;
    if (e != hipSuccess || e2 != hipSuccess) return fail(D3Q19_E_DEVICE, "plane_mean_u copy failed");
    return 0;
}

// d3q19_plane_mean_u without allocation.  The plane means are staged in `scratch`, a file-scope
// device array of PLANE_MEAN_MAX_NZ planes x 3 components.  Host code must not pass the array's name
// to kernels or copies (in host code it is a stand-in, not a device address): its address on each
// device comes from hipGetSymbolAddress, taken once per device with that device current, at the
// entry point's first call there.  Every call on a device shares that device's array, so calls
// that overlap on one device (different simulations on different streams) race; the two shell
// planes it reports as zero are cleared by a small kernel on every call.
}  // extern "C"

namespace {
constexpr int PLANE_MEAN_MAX_NZ = 2048;
static __device__ __align__(256) double scratch[3 * PLANE_MEAN_MAX_NZ];
constexpr int MAX_DEVICES = 64;
double* g_scratch[MAX_DEVICES] = {};
std::once_flag g_scratch_once[MAX_DEVICES];
int g_scratch_rc[MAX_DEVICES] = {};

__global__ void zero_shell_planes_kernel(double* out, int Nz) {
    const int c = threadIdx.x;
    if (c < 3) {
        out[c * Nz] = 0.0;
        out[c * Nz + Nz - 1] = 0.0;
    }
}

// the device address of `scratch` on the current device (set up once per device)
int scratch_on(int device, double** ptr) {
    if (device < 0 || device >= MAX_DEVICES) return fail(D3Q19_E_UNSUPPORTED, "device index above the scratch table");
    std::call_once(g_scratch_once[device], [device]() {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_scratch_rc[device] = fail(D3Q19_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return;
        }
        g_scratch[device] = (double*)p;
    });
    if (g_scratch_rc[device]) return g_scratch_rc[device];
    *ptr = g_scratch[device];
    return 0;
}
}  // namespace

extern "C" {

int d3q19_plane_mean_u_noalloc(d3q19_sim* s, double* means) {
    CHECK_SIM(s);
    if (!means) return fail(D3Q19_E_INVALID, "null output");
    if (s->Nz > PLANE_MEAN_MAX_NZ)
        return fail(D3Q19_E_CAPACITY, "d3q19_plane_mean_u_noalloc: LatticeSize[2] above " + std::to_string(PLANE_MEAN_MAX_NZ) +
                                          " (its device scratch holds that many planes); use d3q19_plane_mean_u");
    double* d = nullptr;
    int rc = scratch_on(s->device, &d);
    if (rc) return rc;
    hipLaunchKernelGGL(zero_shell_planes_kernel, dim3(1), dim3(64), 0, s->stream, d, s->Nz);
    HIP_OK(hipGetLastError());
    if ((rc = launch_plane_mean(s, d))) return rc;
    HIP_OK(hipMemcpyAsync(means, d, 3 * (size_t)s->Nz * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return 0;
}

int d3q19_moments(d3q19_sim* s, double* out4) {
    CHECK_SIM(s);
    if (!out4) return fail(D3Q19_E_INVALID, "null output");
    double* d = s->red + 4 * 512;
    int rc = launch_moments(s, cur_f(s), d);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(out4, d, 4 * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return 0;
}

int d3q19_observables_get(d3q19_sim* s, d3q19_observables* out) {
    CHECK_SIM(s);
    if (!out) return fail(D3Q19_E_INVALID, "null output");
    s->h_desc = make_cell_desc(s);
    double raw[OBS_STRIDE];
    HIP_OK(hipMemcpyAsync(s->d_desc, &s->h_desc, sizeof(CellDesc), hipMemcpyHostToDevice, s->stream));
    int rc = launch_observables(s->d_desc, 1, s->storage, 0, s->d_obs, s->stream);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(raw, s->d_obs, sizeof(raw), hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    derive_observables(s, raw, out);
    return 0;
}

int d3q19_check_finite(d3q19_sim* s) {
    CHECK_SIM(s);
    int nonfinite = 0, moved = 0;
    const int rc = sim_watchdog(s, &nonfinite, &moved);
    if (rc) return rc;
    if (moved) return fail(D3Q19_E_STATE, "a sphere moved one lattice unit or more in one step");
    if (nonfinite) return fail(D3Q19_E_NONFINITE, "non-finite or non-positive density in the population field");
    return 0;
}

int d3q19_time_collide(d3q19_sim* s, int n, int flags, double* ms) {
    CHECK_SIM(s);
    if (n <= 0 || n > 4096 || !ms) return fail(D3Q19_E_INVALID, "bad timing request");
    // each launch's own start and end (no dispatch gaps), averaged
    std::vector<hipEvent_t> evs(2 * (size_t)n, nullptr);
    int rc = 0;
    for (auto& ev : evs)
        if (hipEventCreate(&ev) != hipSuccess) {
            rc = fail(D3Q19_E_DEVICE, "hipEventCreate failed");
            break;
        }
    for (int i = 0; i < n && !rc; i++) {
        rc = launch_collide(s, cur_f(s), nxt_f(s), (flags & D3Q19_STEP_WRITE_U) ? 1 : 0, 1, s->stream, evs[2 * i], evs[2 * i + 1]);
        if (!rc) s->parity ^= 1;
    }
    double acc = 0;
    if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(D3Q19_E_DEVICE, "hipStreamSynchronize failed");
    for (int i = 0; i < n && !rc; i++) {
        float dt = 0;
        if (hipEventElapsedTime(&dt, evs[2 * i], evs[2 * i + 1]) != hipSuccess) rc = fail(D3Q19_E_DEVICE, "hipEventElapsedTime failed");
        acc += dt;
    }
    for (auto& ev : evs)
        if (ev) (void)hipEventDestroy(ev);
    if (rc) return rc;
    *ms = acc / n;
    return 0;
}

int d3q19_time_steps(d3q19_sim* s, int t_first, int n, double* total_ms, double* kernel_ms) {
    CHECK_SIM(s);
    if (n <= 0 || n > 4096 || !total_ms) return fail(D3Q19_E_INVALID, "bad timing request");
    // 4 events per step (step kernel, IBM pass: each kernel's own start and end)
    std::vector<hipEvent_t> evs(4 * (size_t)n, nullptr);
    int rc = 0;
    for (auto& ev : evs)
        if (hipEventCreate(&ev) != hipSuccess) {
            rc = fail(D3Q19_E_DEVICE, "hipEventCreate failed");
            break;
        }
    float t = 0.f;
    if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(D3Q19_E_DEVICE, "hipStreamSynchronize failed");
    if (!rc && hipEventRecord(s->ev0, s->stream) != hipSuccess) rc = fail(D3Q19_E_DEVICE, "hipEventRecord failed");
    if (!rc) rc = run_steps(s, t_first, n, 0, evs.data());
    if (!rc && (hipEventRecord(s->ev1, s->stream) != hipSuccess || hipEventSynchronize(s->ev1) != hipSuccess ||
                hipEventElapsedTime(&t, s->ev0, s->ev1) != hipSuccess))
        rc = fail(D3Q19_E_DEVICE, "step timing failed");
    double acc[2] = {0.0, 0.0};
    for (int i = 0; i < n && !rc && kernel_ms; i++)
        for (int k = 0; k < 2; k++) {
            if (k == 1 && s->np == 0) continue;
            float dt = 0.f;
            if (hipEventElapsedTime(&dt, evs[4 * i + 2 * k], evs[4 * i + 2 * k + 1]) != hipSuccess) {
                rc = fail(D3Q19_E_DEVICE, "hipEventElapsedTime failed");
                break;
            }
            acc[k] += dt;
        }
    for (auto& ev : evs)
        if (ev) (void)hipEventDestroy(ev);
    if (rc) return rc;
    *total_ms = (double)t;
    if (kernel_ms) {
        kernel_ms[0] = acc[0] / n;
        kernel_ms[1] = acc[1] / n;
    }
    return 0;
}

}  // extern "C"
