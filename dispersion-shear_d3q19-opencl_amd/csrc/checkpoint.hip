// checkpoint.hip -- checkpoint / restart of a simulation's whole state (no counterpart in the
// reference, whose only output is u at the end, sim_main.c:476-481).
//
// The image is self-describing: a header with a magic word, a version, the parameter structs and
// options the simulation was created with, the step bookkeeping (lattice parity, epoch counters,
// step-number offset, raw-wall-unknowns flag, marks flag), then one section per device array, each
// (id, byte count, bytes).  Populations are kept in their storage format, so a restart continues
// bit for bit.  A load checks that the parameter structs and options equal those of the target
// simulation and that every section has the size the target expects.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "internal.h"

using namespace d3q19;

namespace {

constexpr char MAGIC[8] = {'D', '3', 'Q', '1', '9', 'C', 'K', '1'};
constexpr uint32_t VERSION = 6;  // 2: u stored with its components interleaved per node; 3: displacement watchdog;
                                 // 4: per-node force tags; 5: chunk-interleaved populations; 6: one row per
                                 // chunk for single simulations

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int_param_struct ip;
    flp_param_struct fp;
    d3q19_options opt;
    int32_t parity;
    int32_t bcraw_valid;
    int32_t marks_next;
    int32_t tofs;
    int64_t h_epoch;
    int32_t nsections;
    int32_t reserved[7];
};

struct Section {
    uint32_t id;
    uint32_t pad;
    uint64_t bytes;
};

enum : uint32_t {
    S_F = 1, S_U, S_TAU, S_FACC, S_MASK0, S_MASK1, S_MASK2, S_EPOCHS, S_POS, S_VEL, S_QUAT, S_OMEGA, S_FFLUID,
    S_TFLUID, S_FCONT, S_TCONT, S_MFORCE, S_ZONE, S_BCRAW, S_PTS, S_DFLAG, S_FTAG
};

struct Array {
    uint32_t id;
    void* ptr;
    size_t bytes;
    bool match;  // configuration, not state: a load checks that it equals the target's
};

int ck_fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

// the device arrays that make up the state, with their sizes
std::vector<Array> state_arrays(Sim* s) {
    const size_t nint = (size_t)s->g.nint;
    const size_t rs = s->storage == D3Q19_STORE_FP64 ? 8 : 4;
    std::vector<Array> v;
    v.push_back({S_F, s->parity == 0 ? s->fA : s->fB, (size_t)Q * nint * s->pop_bytes, false});
    v.push_back({S_U, s->u, 3 * nint * rs, false});
    v.push_back({S_TAU, s->tau, nint * rs, false});
    v.push_back({S_BCRAW, s->bcraw, (size_t)10 * s->bc_plane * rs, false});
    if (s->np > 0) {
        const size_t np = (size_t)s->np;
        v.push_back({S_FACC, s->facc, 3 * nint * sizeof(long long), false});
        v.push_back({S_FTAG, s->ftag, nint, false});
        v.push_back({S_MASK0, s->mask[0], (size_t)s->nbricks * sizeof(uint32_t), false});
        v.push_back({S_MASK1, s->mask[1], (size_t)s->nbricks * sizeof(uint32_t), false});
        v.push_back({S_MASK2, s->mask[2], (size_t)s->nbricks * sizeof(uint32_t), false});
        v.push_back({S_EPOCHS, s->epochs, 2 * sizeof(uint32_t), false});
        v.push_back({S_POS, s->pos, 3 * np * 8, false});
        v.push_back({S_VEL, s->vel, 3 * np * 8, false});
        v.push_back({S_QUAT, s->quat, 4 * np * 8, false});
        v.push_back({S_OMEGA, s->omega, 3 * np * 8, false});
        v.push_back({S_FFLUID, s->ffluid, 3 * np * 8, false});
        v.push_back({S_TFLUID, s->tfluid, 3 * np * 8, false});
        v.push_back({S_FCONT, s->fcont, 3 * np * 8, false});
        v.push_back({S_TCONT, s->tcont, 3 * np * 8, false});
        v.push_back({S_MFORCE, s->mforce, 6 * np * (size_t)s->ppp * 8, false});
        v.push_back({S_ZONE, s->zone, 2 * np * sizeof(int), false});
        // the watchdog travels with the state: a restored run reports a displacement that the run
        // it continues had raised
        v.push_back({S_DFLAG, s->d_flag, sizeof(int), false});
        v.push_back({S_PTS, s->pts, 4 * (size_t)s->ppp * sizeof(float), true});
    }
    return v;
}

size_t image_bytes(Sim* s) {
    size_t b = sizeof(Header);
    for (const Array& a : state_arrays(s)) b += sizeof(Section) + a.bytes;
    return b;
}

}  // namespace

extern "C" {

int d3q19_checkpoint_size(d3q19_sim* s, size_t* bytes) {
    if (!s || !bytes) return ck_fail(D3Q19_E_INVALID, "null argument");
    if (s->arena) return ck_fail(D3Q19_E_UNSUPPORTED, "a batch member is checkpointed through its batch");
    *bytes = image_bytes(s);
    return 0;
}

int d3q19_checkpoint_save(d3q19_sim* s, void* buf, size_t bytes) {
    if (!s || !buf) return ck_fail(D3Q19_E_INVALID, "null argument");
    if (s->arena) return ck_fail(D3Q19_E_UNSUPPORTED, "a batch member is checkpointed through its batch");
    if (bytes < image_bytes(s)) return ck_fail(D3Q19_E_INVALID, "checkpoint buffer too small");
    HIP_OK(hipSetDevice(s->device));
    HIP_OK(hipStreamSynchronize(s->stream));
    const std::vector<Array> arrays = state_arrays(s);
    Header h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, MAGIC, sizeof(MAGIC));
    h.version = VERSION;
    h.header_bytes = sizeof(Header);
    h.ip = s->ip;
    h.fp = s->fp;
    h.opt = s->opt;
    h.parity = s->parity;
    h.bcraw_valid = s->bcraw_valid;
    h.marks_next = s->marks_next ? 1 : 0;
    h.tofs = s->tofs;
    h.h_epoch = s->h_epoch;
    h.nsections = (int32_t)arrays.size();
    char* p = (char*)buf;
    memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    for (const Array& a : arrays) {
        Section sec{a.id, 0u, (uint64_t)a.bytes};
        memcpy(p, &sec, sizeof(sec));
        p += sizeof(sec);
        if (a.bytes) HIP_OK(hipMemcpy(p, a.ptr, a.bytes, hipMemcpyDeviceToHost));
        p += a.bytes;
    }
    return 0;
}

int d3q19_checkpoint_load(d3q19_sim* s, const void* buf, size_t bytes) {
    if (!s || !buf) return ck_fail(D3Q19_E_INVALID, "null argument");
    if (s->arena) return ck_fail(D3Q19_E_UNSUPPORTED, "a batch member is checkpointed through its batch");
    if (bytes < sizeof(Header)) return ck_fail(D3Q19_E_INVALID, "checkpoint image too short");
    Header h;
    memcpy(&h, buf, sizeof(h));
    if (memcmp(h.magic, MAGIC, sizeof(MAGIC)) != 0 || h.version != VERSION || h.header_bytes != sizeof(Header))
        return ck_fail(D3Q19_E_INVALID, "not a checkpoint image of this library version");
    // every parameter but the run length must agree
    int_param_struct hip_ = h.ip;
    hip_.MaxIterations = s->ip.MaxIterations;
    if (memcmp(&hip_, &s->ip, sizeof(hip_)) != 0 || memcmp(&h.fp, &s->fp, sizeof(h.fp)) != 0 ||
        memcmp(&h.opt, &s->opt, sizeof(h.opt)) != 0)
        return ck_fail(D3Q19_E_INVALID, "checkpoint was written by a simulation with other parameters");
    if (h.parity != 0 && h.parity != 1) return ck_fail(D3Q19_E_INVALID, "checkpoint header corrupt");
    HIP_OK(hipSetDevice(s->device));
    HIP_OK(hipStreamSynchronize(s->stream));
    // the populations go to the buffer that the image's parity names as current
    const int parity0 = s->parity;
    s->parity = h.parity;
    const std::vector<Array> arrays = state_arrays(s);
    if (h.nsections != (int32_t)arrays.size()) {
        s->parity = parity0;
        return ck_fail(D3Q19_E_INVALID, "checkpoint section count differs");
    }
    const char* p = (const char*)buf + sizeof(Header);
    const char* end = (const char*)buf + bytes;
    // validate every section before any device array changes
    {
        const char* q = p;
        for (const Array& a : arrays) {
            Section sec;
            if (q + sizeof(sec) > end) {
                s->parity = parity0;
                return ck_fail(D3Q19_E_INVALID, "checkpoint image truncated");
            }
            memcpy(&sec, q, sizeof(sec));
            q += sizeof(sec);
            if (sec.id != a.id || sec.bytes != (uint64_t)a.bytes || q + a.bytes > end) {
                s->parity = parity0;
                return ck_fail(D3Q19_E_INVALID, "checkpoint section does not match this simulation");
            }
            if (a.match && a.bytes) {
                std::vector<char> mine(a.bytes);
                HIP_OK(hipMemcpy(mine.data(), a.ptr, a.bytes, hipMemcpyDeviceToHost));
                if (memcmp(mine.data(), q, a.bytes) != 0) {
                    s->parity = parity0;
                    return ck_fail(D3Q19_E_INVALID, "checkpoint was written with other sphere marker offsets");
                }
            }
            q += a.bytes;
        }
    }
    for (const Array& a : arrays) {
        p += sizeof(Section);
        if (a.bytes && !a.match) HIP_OK(hipMemcpy(a.ptr, p, a.bytes, hipMemcpyHostToDevice));
        p += a.bytes;
    }
    s->bcraw_valid = h.bcraw_valid;
    s->marks_next = h.marks_next != 0;
    s->tofs = h.tofs;
    s->h_epoch = h.h_epoch;
    return 0;
}

int d3q19_checkpoint_write(d3q19_sim* s, const char* path) {
    if (!s || !path) return ck_fail(D3Q19_E_INVALID, "null argument");
    size_t n = 0;
    int rc = d3q19_checkpoint_size(s, &n);
    if (rc) return rc;
    std::vector<char> buf(n);
    if ((rc = d3q19_checkpoint_save(s, buf.data(), n))) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) return ck_fail(D3Q19_E_INVALID, std::string("cannot open for writing: ") + path);
    const size_t w = fwrite(buf.data(), 1, n, f);
    const int cl = fclose(f);
    if (w != n || cl != 0) return ck_fail(D3Q19_E_INVALID, std::string("write failed: ") + path);
    return 0;
}

int d3q19_checkpoint_read(d3q19_sim* s, const char* path) {
    if (!s || !path) return ck_fail(D3Q19_E_INVALID, "null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return ck_fail(D3Q19_E_INVALID, std::string("cannot open: ") + path);
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t r;
    while ((r = fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + r);
    fclose(f);
    return d3q19_checkpoint_load(s, buf.data(), buf.size());
}

}  // extern "C"
