// Host runtime and C ABI (include/pgeof_hip.h) of the gfx950 pgeof path.
//
// Responsibilities:
//   - argument validation with the reference's error conditions and messages
//     (include/pgeof.hpp:81, 165-168, 250; include/nn_search.hpp:37, 92-95)
//   - grid sizing for the neighbour search (bounding box read back, one
//     occupancy probe for kNN)
//   - workspace carving, host<->device staging for the host-memory API
//   - optional per-kernel HIP-event profiling
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pgeof_hip.h"
#include "pgeof_internal.hpp"

using namespace pgeof_amd;

namespace
{

thread_local std::string t_err;

int fail(int code, const std::string& msg)
{
    t_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* where)
{
    return fail(e == hipErrorOutOfMemory ? PGEOF_ENOMEM : PGEOF_ERUNTIME,
                std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
    do                                                  \
    {                                                   \
        hipError_t _e = (expr);                         \
        if (_e != hipSuccess) return hip_fail(_e, #expr); \
    } while (0)

// ------------------------------------------------------------- tuning ----
struct Tuning
{
    double knn_occupancy = 0.35;       // target points per cell = max(0.75, min(this * k, knn_cell_points_max))
    double knn_cell_points_max = 25.0;  // cap of that target (large k)
    double grid_cells_per_point = 8.0;  // cap on G / n
    double lds_budget = 65536.0;        // bytes of LDS per workgroup for heaps
    double radius_cell_factor = 0.5;    // cell edge = this * radius
    double grid_levels = 3.0;           // kNN grid pyramid depth (1..4)
    double near_first = 1.0;            // visit rows nearest-first (1) or in index order (0)
    double budget_cells = 2.0;          // budgeted kNN pass: ball half-width limit in cells
    double budget_shells = 1.0;         // budgeted kNN pass: cubic shells scanned before deferring
    double budget_levels = 1.0;         // pyramid levels that get a budgeted pass (the rest: unbudgeted only)
    double grid_level_ratio = 2.0;      // cell edge ratio between consecutive pyramid levels (2..8)
    double csr_order = 1.0;             // CSR features: visit large inputs' rows in spatial order
    double heap_tail = 1.0;             // search heaps: nodes past 20 in global memory (k > 21); 2: per wave slot always
    double staging = 0.0;               // LDS staging of cell rows: 0 never, 1 radius + budgeted kNN at k >= 32, 2 always
    double fuse_multiscale = 0.0;       // radius + multiscale: features in the search kernel (1) or a CSR kernel after it (0)
};
Tuning g_tune;
std::mutex g_tune_mu;

// name -> member, for pgeof_set_tuning / pgeof_get_tuning and the autotuner
struct KnobName
{
    const char* name;
    double Tuning::*field;
};
const KnobName kKnobs[] = {
    {"knn_occupancy", &Tuning::knn_occupancy},
    {"knn_cell_points_max", &Tuning::knn_cell_points_max},
    {"grid_cells_per_point", &Tuning::grid_cells_per_point},
    {"lds_budget", &Tuning::lds_budget},
    {"radius_cell_factor", &Tuning::radius_cell_factor},
    {"grid_levels", &Tuning::grid_levels},
    {"near_first", &Tuning::near_first},
    {"budget_cells", &Tuning::budget_cells},
    {"budget_shells", &Tuning::budget_shells},
    {"budget_levels", &Tuning::budget_levels},
    {"grid_level_ratio", &Tuning::grid_level_ratio},
    {"csr_order", &Tuning::csr_order},
    {"heap_tail", &Tuning::heap_tail},
    {"staging", &Tuning::staging},
    {"fuse_multiscale", &Tuning::fuse_multiscale},
};

double* knob(Tuning& t, const std::string& name)
{
    for (const KnobName& k : kKnobs)
        if (name == k.name) return &(t.*(k.field));
    return nullptr;
}

// Every entry point works on one snapshot of the knobs, taken when it starts,
// so that a concurrent pgeof_set_tuning cannot make, e.g., the grid sizing
// disagree with the workspace carved for it within one call.
thread_local const Tuning* t_snap = nullptr;

// Knob sets chosen by the autotuner (pgeof_autotune) per search family, k and
// size class (ceil(log2 n)): a call whose sizes match takes that set instead
// of the global knobs.
enum
{
    kFamKnn = 0,
    kFamRadius = 1
};
struct TuneKey
{
    int family;
    uint32_t k;
    int size_class;
    bool operator<(const TuneKey& o) const
    {
        if (family != o.family) return family < o.family;
        if (k != o.k) return k < o.k;
        return size_class < o.size_class;
    }
};
std::map<TuneKey, Tuning> g_tuned;  // guarded by g_tune_mu

int size_class(int64_t n)
{
    int c = 0;
    while (c < 62 && (int64_t(1) << c) < n) ++c;
    return c;
}

struct TuningSnapshot
{
    Tuning v;
    const Tuning* prev;
    TuningSnapshot() : TuningSnapshot(-1, 0, 0) {}
    // family >= 0: the autotuned set for (family, k, size class of n) if one exists
    TuningSnapshot(int family, int64_t n, uint32_t k) : prev(t_snap)
    {
        {
            std::lock_guard<std::mutex> l(g_tune_mu);
            v = g_tune;
            if (family >= 0 && !prev)
            {
                auto it = g_tuned.find(TuneKey{family, k, size_class(n)});
                if (it != g_tuned.end()) v = it->second;
            }
        }
        if (!prev) t_snap = &v;
    }
    ~TuningSnapshot()
    {
        if (!prev) t_snap = nullptr;
    }
    TuningSnapshot(const TuningSnapshot&) = delete;
    TuningSnapshot& operator=(const TuningSnapshot&) = delete;
};

Tuning tuning()
{
    if (t_snap) return *t_snap;
    std::lock_guard<std::mutex> l(g_tune_mu);
    return g_tune;
}

// ---------------------------------------------------------- profiling ----
std::atomic<int> g_prof{0};
struct ProfRec
{
    std::string name;
    hipEvent_t a, b;
};
std::mutex g_prof_mu;
std::vector<ProfRec> g_prof_pending;
std::map<std::string, std::pair<int64_t, double>> g_prof_acc;

// ------------------------------------------------------- device state ----
struct DevState
{
    int lds_max = 65536;
    int xccs = 0;  // 0: unknown
    bool init = false;
};
std::mutex g_dev_mu;
std::map<int, DevState> g_dev;

// attributes of the current device (defaults when there is none)
DevState dev_state()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return DevState{};
    std::lock_guard<std::mutex> l(g_dev_mu);
    DevState& d = g_dev[dev];
    if (!d.init)
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            d.lds_max = v;
        v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeNumberOfXccs, dev) == hipSuccess && v > 0) d.xccs = v;
        d.init = true;
    }
    return d;
}

int lds_limit() { return dev_state().lds_max; }

// whether hw_wave_slot() names every resident wave of the current device
// uniquely (its XCC id fits the bits it keeps)
bool hw_slots_ok()
{
    const int x = dev_state().xccs;
    return x > 0 && x <= (int)kHwMaxXcc;
}

// Scratch of device calls made without a workspace: device buffers kept by
// the library (hipMalloc, per device), each handed out to one call at a time
// and handed out again once an event recorded on that call's stream after
// its last launch has completed.  Stream-ordered allocation
// (hipMallocAsync / hipFreeAsync) is not used: the buffers stay mapped for
// the life of the process.
struct ScratchBuf
{
    int dev;
    void* p;
    size_t bytes;
    hipEvent_t done;  // recorded on the last user's stream when it returned the buffer
    bool busy;
    hipStream_t last;  // that stream (a call on the same stream may reuse the buffer at once)
};
std::mutex g_scratch_mu;
std::vector<std::unique_ptr<ScratchBuf>> g_scratch;  // stable addresses: leases hold pointers

// A buffer of at least `need` bytes on the current device, not in use by any
// call still running.  Free buffers too small for the request are released
// first (hipFree), so the cache holds about one buffer per concurrent call.
int scratch_acquire(size_t need, hipStream_t s, ScratchBuf** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    need = (need + 4095) & ~size_t(4095);
    std::lock_guard<std::mutex> l(g_scratch_mu);
    ScratchBuf* best = nullptr;
    for (auto& up : g_scratch)
    {
        ScratchBuf* b = up.get();
        if (b->dev != dev || b->busy || b->bytes < need) continue;
        if (b->last != s && hipEventQuery(b->done) != hipSuccess) continue;  // still in use on another stream
        if (!best || b->bytes < best->bytes) best = b;
    }
    if (!best)
    {
        for (size_t i = 0; i < g_scratch.size();)
        {
            ScratchBuf* b = g_scratch[i].get();
            if (b->dev == dev && !b->busy && b->bytes < need && hipEventQuery(b->done) == hipSuccess)
            {
                (void)hipFree(b->p);
                (void)hipEventDestroy(b->done);
                g_scratch.erase(g_scratch.begin() + (ptrdiff_t)i);
            }
            else
                ++i;
        }
        auto b = std::make_unique<ScratchBuf>(ScratchBuf{dev, nullptr, need, nullptr, false, s});
        HIP_TRY(hipMalloc(&b->p, need));
        hipError_t e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
        if (e != hipSuccess)
        {
            (void)hipFree(b->p);
            return hip_fail(e, "hipEventCreateWithFlags");
        }
        best = b.get();
        g_scratch.push_back(std::move(b));
    }
    best->busy = true;
    *out = best;
    return PGEOF_OK;
}

// the caller's launches on s are queued: the buffer is free once s gets here
void scratch_release(ScratchBuf* b, hipStream_t s)
{
    std::lock_guard<std::mutex> l(g_scratch_mu);
    (void)hipEventRecord(b->done, s);
    b->last = s;
    b->busy = false;
}

struct ScratchLease
{
    ScratchBuf* buf = nullptr;
    hipStream_t s = nullptr;
    void* p = nullptr;
    int take(size_t need, hipStream_t st)
    {
        const int rc = scratch_acquire(need, st, &buf);
        if (rc) return rc;
        s = st;
        p = buf->p;
        return PGEOF_OK;
    }
    ~ScratchLease()
    {
        if (buf) scratch_release(buf, s);
    }
};

// Per-thread host-API context: a stream, a grow-only device buffer and a
// small pinned host buffer for read-backs.
struct ThreadCtx
{
    int dev = -1;
    hipStream_t stream = nullptr;
    char* dbuf = nullptr;
    size_t dbytes = 0;
    void* pinned = nullptr;
    ~ThreadCtx()
    {
        // device resources are left to process teardown
    }
};
thread_local ThreadCtx t_ctx;

int ctx_prepare(hipStream_t* s)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (t_ctx.dev != dev)
    {
        // this thread moved to another device: release the old device's
        // stream and buffer on that device (the pinned host buffer is usable
        // from every device and is kept)
        if (t_ctx.dev >= 0 && (t_ctx.stream || t_ctx.dbuf) && hipSetDevice(t_ctx.dev) == hipSuccess)
        {
            if (t_ctx.stream) (void)hipStreamSynchronize(t_ctx.stream);
            if (t_ctx.dbuf) (void)hipFree(t_ctx.dbuf);
            if (t_ctx.stream) (void)hipStreamDestroy(t_ctx.stream);
            (void)hipSetDevice(dev);
        }
        t_ctx.dev = dev;
        t_ctx.stream = nullptr;
        t_ctx.dbuf = nullptr;
        t_ctx.dbytes = 0;
    }
    if (!t_ctx.stream) HIP_TRY(hipStreamCreateWithFlags(&t_ctx.stream, hipStreamNonBlocking));
    *s = t_ctx.stream;
    return PGEOF_OK;
}

int ctx_reserve(size_t bytes)
{
    if (bytes <= t_ctx.dbytes) return PGEOF_OK;
    if (t_ctx.dbuf)
    {
        (void)hipStreamSynchronize(t_ctx.stream);
        (void)hipFree(t_ctx.dbuf);
        t_ctx.dbuf = nullptr;
        t_ctx.dbytes = 0;
    }
    size_t want = bytes + bytes / 8 + (1 << 20);
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, want));
    t_ctx.dbuf = (char*)p;
    t_ctx.dbytes = want;
    return PGEOF_OK;
}

void* pinned_scratch()
{
    if (!t_ctx.pinned)
    {
        void* p = nullptr;
        if (hipHostMalloc(&p, 4096, hipHostMallocDefault) != hipSuccess) return nullptr;
        t_ctx.pinned = p;
    }
    return t_ctx.pinned;
}

inline size_t al(size_t b) { return (b + 255) & ~size_t(255); }

// ------------------------------------------------------------- grids ----
uint64_t grid_cap(uint64_t n)
{
    const Tuning t = tuning();
    double c = std::max(1.0, t.grid_cells_per_point) * (double)n;
    c = std::max(c, 65536.0);
    c = std::min(c, 2147483000.0);
    return (uint64_t)c;
}

uint64_t cells_for(const double ext[3], double h, int g[3])
{
    uint64_t G = 1;
    for (int a = 0; a < 3; ++a)
    {
        double c = std::floor(ext[a] / h) + 1.0;
        if (!(c >= 1.0)) c = 1.0;
        if (c > 1048576.0) c = 1048576.0;
        g[a] = (int)c;
        G *= (uint64_t)g[a];
    }
    return G;
}

// smallest h (up to 0.1%) with cells_for(h) <= cap
double h_for_cap(const double ext[3], uint64_t cap)
{
    const double emax = std::max(ext[0], std::max(ext[1], ext[2]));
    if (!(emax > 0.0)) return 1.0;
    int g[3];
    double hi = emax * 2.0, lo = emax * 1e-7;
    if (cells_for(ext, lo, g) <= cap) return lo;
    for (int it = 0; it < 80; ++it)
    {
        const double mid = std::sqrt(lo * hi);
        if (cells_for(ext, mid, g) <= cap)
            hi = mid;
        else
            lo = mid;
        if (hi / lo < 1.001) break;
    }
    return hi;
}

GridParams make_grid(const float bb[6], double h)
{
    GridParams gp;
    double ext[3];
    for (int a = 0; a < 3; ++a)
    {
        gp.lo[a] = (double)bb[a];
        ext[a] = std::max(0.0, (double)bb[3 + a] - (double)bb[a]);
    }
    gp.h = h;
    gp.inv_h = 1.0 / h;
    gp.G = cells_for(ext, h, gp.g);
    return gp;
}

// Layout of a search workspace.
struct SearchWs
{
    float* bbox_partial;
    float* bbox_out;
    unsigned long long* sumsq;
    uint32_t* count;
    uint32_t* cs_all;  // cell starts of every level, 2 * (cap + 1) + 64 entries
    uint32_t* scan_part;
    uint32_t* cell;
    uint32_t* rank;
    float4* pts[kMaxLevels];
    uint32_t* qcell;
    uint32_t* qrank;
    float4* qs;
    uint64_t* gheap_keys;
    uint32_t* dense;   // radius csr: n_query * cap
    uint32_t* counts;  // radius csr: n_query (+1 for the scan)
    uint32_t* csr_part;
    uint32_t* flags;   // kNN: deferred queries of the budgeted pass (n_query + 1)
    uint32_t* foff;    // their exclusive scan
    uint32_t* fpart;
    float4* qlist;     // compacted deferred queries
    float4* qlist2;    // the same, ping-pong partner for the next level
    float* dbound;     // per slot of a budgeted pass: bound of the k-th distance of a deferred query
    float* blist;      // bounds of the compacted deferred queries
    float* blist2;
    LevelTable* tab;   // the grid pyramid, planned on the device
};

constexpr uint32_t kBboxBlocks = 1024;

struct HeapPlan
{
    int block;
    size_t lds;
    bool global;
    uint32_t tail_pairs = 0;  // pairs of heap nodes per lane in global memory (nodes 21..k-1, whole groups of four)
};

HeapPlan heap_plan_bytes(uint32_t k, size_t entry_bytes)
{
    const Tuning t = tuning();
    const size_t per = (size_t)k * entry_bytes;
    const size_t budget = (size_t)std::min<double>(t.lds_budget, (double)lds_limit());
    for (int b : {256, 128, 64})
        if (per * (size_t)b <= budget) return {b, per * (size_t)b, false};
    return {256, 0, true};
}

// Search kernels: per wave 64 heaps plus a 128-point staging buffer (2 KiB)
// in LDS.  A heap keeps nodes 0..4 in registers and nodes 5..k-1 in groups of
// four slots: ceil((k - 5) / 4) groups of four 8-byte keys per lane.
// With the heap tail (k > 21), nodes 5..20 stay in LDS (128 bytes per lane)
// and nodes 21..k-1 go to global memory as lane-interleaved 16-byte pairs.
HeapPlan heap_plan(uint32_t k)
{
    const Tuning t = tuning();
    const bool tail = t.heap_tail > 0.5 && k > 21;
    const uint32_t groups = k > 5 ? (k - 2) / 4 : 0;
    const size_t per = (size_t)4 * (tail ? 4 : groups) * 8;
    const size_t stage = 128 * 16;
    const size_t budget = (size_t)std::min<double>(t.lds_budget, (double)lds_limit());
    for (int b : {256, 128, 64})
        if (per * (size_t)b + stage * (size_t)(b / 64) <= budget)
            return {b, per * (size_t)b + stage * (size_t)(b / 64), false, tail ? 2 * ((k - 21 + 3) / 4) : 0u};
    return {256, 0, true, 0u};
}

// Where the heap tails of a search over m queries live: 2 one per hardware
// wave slot, 1 one per lane of the launch, 0 none.  Per wave slot, only the
// slots of the resident waves are touched (a few MB, which stay in cache,
// instead of a tail per query, rewritten through to HBM); its storage,
// kHwWaveSlots * 64 lanes, is taken once the launch has at least half as
// many lanes, or always with heap_tail 2.
int tail_layout(const HeapPlan& hp, uint64_t m)
{
    if (!hp.tail_pairs) return 0;
    if (!hw_slots_ok()) return 1;
    const bool force = tuning().heap_tail > 1.5;
    return force || (m + 256) * 2 >= (uint64_t)kHwWaveSlots * 64 ? 2 : 1;
}

// global-memory heap storage of a search over m queries: the whole heap
// (fallback), or the tail of every hardware wave slot, or of every lane of
// the launch (grid rounded up to 256)
size_t heap_global_bytes(const HeapPlan& hp, uint64_t m, uint32_t k)
{
    if (hp.global) return (size_t)m * k * 8;
    switch (tail_layout(hp, m))
    {
        case 2: return (size_t)kHwWaveSlots * 64 * hp.tail_pairs * 16;
        case 1: return (size_t)(m + 256) * hp.tail_pairs * 16;
        default: return 0;
    }
}

size_t search_ws_bytes(uint64_t n, uint64_t m, uint32_t k, bool csr)
{
    const uint64_t G = grid_cap(n) + 1;
    size_t b = 0;
    b += al(kBboxBlocks * 6 * sizeof(float)) + al(64) + al(64);
    b += al(G * 4) + al((2 * G + 64) * 4) + al(scan_partials_needed(G + 64) * 4);
    b += 2 * al(n * 4) + kMaxLevels * al(n * 16);
    b += 2 * al(m * 4) + al(m * 16);
    b += 2 * al((m + 1) * 4) + al(scan_partials_needed(m + 1) * 4) + 2 * al(m * 16) + 3 * al(m * 4);
    b += al(sizeof(LevelTable));
    const HeapPlan hp = heap_plan(k);
    if (const size_t hb = heap_global_bytes(hp, m, k)) b += al(hb);
    if (csr) b += al(m * k * 4) + al((m + 1) * 4) + al(scan_partials_needed(m + 1) * 4);
    return b + 4096;
}

SearchWs carve(Arena& ar, uint64_t n, uint64_t m, uint32_t k, bool csr)
{
    SearchWs w{};
    const uint64_t G = grid_cap(n) + 1;
    w.bbox_partial = ar.take<float>(kBboxBlocks * 6);
    w.bbox_out = ar.take<float>(16);
    w.sumsq = ar.take<unsigned long long>(8);
    w.count = ar.take<uint32_t>(G);
    w.cs_all = ar.take<uint32_t>(2 * G + 64);
    w.scan_part = ar.take<uint32_t>(scan_partials_needed(G + 64));
    w.cell = ar.take<uint32_t>(n);
    w.rank = ar.take<uint32_t>(n);
    for (int l = 0; l < kMaxLevels; ++l) w.pts[l] = ar.take<float4>(n);
    w.qcell = ar.take<uint32_t>(m);
    w.qrank = ar.take<uint32_t>(m);
    w.qs = ar.take<float4>(m);
    w.flags = ar.take<uint32_t>(m + 1);
    w.foff = ar.take<uint32_t>(m + 1);
    w.fpart = ar.take<uint32_t>(scan_partials_needed(m + 1));
    w.qlist = ar.take<float4>(m);
    w.qlist2 = ar.take<float4>(m);
    w.dbound = ar.take<float>(m);
    w.blist = ar.take<float>(m);
    w.blist2 = ar.take<float>(m);
    w.tab = ar.take<LevelTable>(1);
    const HeapPlan hp = heap_plan(k);
    if (const size_t hb = heap_global_bytes(hp, m, k)) w.gheap_keys = ar.take<uint64_t>(hb / 8);
    if (csr)
    {
        w.dense = ar.take<uint32_t>(m * k);
        w.counts = ar.take<uint32_t>(m + 1);
        w.csr_part = ar.take<uint32_t>(scan_partials_needed(m + 1));
    }
    return w;
}

// Read `bytes` from device to host through pinned memory (synchronises s).
int readback(void* host, const void* dev, size_t bytes, hipStream_t s)
{
    void* pin = pinned_scratch();
    static_assert(sizeof(LevelTable) <= 4096, "level table read-back");
    if (!pin || bytes > 4096) return fail(PGEOF_ERUNTIME, "pinned staging buffer unavailable");
    HIP_TRY(hipMemcpyAsync(pin, dev, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::memcpy(host, pin, bytes);
    return PGEOF_OK;
}

// Cell-start table of level l inside cs_all (2 * (cap + 1) + 64 entries):
// level 0 gets cap + 1 entries, level l > 0 about (cap + 1) / 2^l (a coarser
// level that does not fit is not built), all 16-byte aligned.
void level_tables(uint32_t* cs_all, uint64_t cap, uint32_t* cs[kMaxLevels], uint64_t room[kMaxLevels])
{
    uint64_t off = 0;
    for (int l = 0; l < kMaxLevels; ++l)
    {
        const uint64_t r = l == 0 ? ((cap + 1 + 3) & ~uint64_t(3)) : ((((cap + 1) >> l) + 4 + 3) & ~uint64_t(3));
        cs[l] = cs_all + off;
        room[l] = r;
        off += r;
    }
}

// Sort `data` into a pyramid of uniform grids (cell edge h_0 * ratio^l) and,
// when query != data, sort the queries into level 0.  kNN sizes h_0 from the
// target occupancy (k used for sizing); radius searches use one level with
// cells of radius_cell_factor * radius.  The sizes are planned on the device
// (k_plan_probe, k_plan_levels): every launch is queued without a read-back,
// the kernels take their grid from the level table and the launches are
// sized for the largest grid the workspace allows.
int build_grid(const float* data, uint32_t n, const float* query, uint32_t m, uint32_t k, float radius, bool is_radius,
               SearchWs& w, SearchArgs& sa, const float4** qs_out, hipStream_t s)
{
    const Tuning t = tuning();
    const uint64_t cap = grid_cap(n);
    PlanArgs p{};
    p.n = n;
    p.k = k;
    p.radius = radius;
    p.is_radius = is_radius ? 1 : 0;
    p.cap = cap;
    p.max_levels = is_radius ? 1 : std::max(1, std::min(kMaxLevels, (int)t.grid_levels));
    p.knn_occupancy = t.knn_occupancy;
    p.knn_cell_points_max = t.knn_cell_points_max;
    p.radius_cell_factor = t.radius_cell_factor;
    p.level_ratio = t.grid_level_ratio;
    p.probe_stride = std::max<uint32_t>(1, std::min<uint32_t>(64, n >> 20));
    level_tables(w.cs_all, cap, p.cs, p.cs_room);
    for (int l = 0; l < kMaxLevels; ++l) p.pts[l] = w.pts[l];

    const uint32_t nbb = std::max<uint32_t>(1, std::min<uint32_t>(kBboxBlocks, (n + 255) / 256));
    HIP_TRY(launch_bbox(data, n, w.bbox_partial, nbb, w.tab->bb, s));
    HIP_TRY(launch_plan_probe(p, w.tab, s));
    if (!is_radius) HIP_TRY(launch_probe(data, n, p.probe_stride, w.tab, w.count, cap + 1, s));
    HIP_TRY(launch_plan_levels(p, w.tab, s));
    for (int l = 0; l < p.max_levels; ++l)
    {
        // counts straight into the level's table, scanned in place
        uint32_t* csl = p.cs[l];
        const uint64_t* G = &w.tab->G[l];
        HIP_TRY(launch_fill_u32_dyn(csl, p.cs_room[l], G, 1, 0u, s));
        if (l == 0)
            HIP_TRY(launch_hist(data, n, w.tab, 0, csl, w.cell, w.rank, s));
        else  // from the level-0 order: runs of equal cells, one atomic each
            HIP_TRY(launch_hist_sorted(w.pts[0], n, w.tab, l, csl, w.cell, w.rank, s));
        HIP_TRY(launch_scan_exclusive_dyn(csl, csl, p.cs_room[l], G, 1, w.scan_part, s));
        if (l == 0)
            HIP_TRY(launch_scatter(data, n, w.cell, w.rank, csl, w.pts[0], w.tab, 0, s));
        else
            HIP_TRY(launch_scatter_sorted(w.pts[0], n, w.cell, w.rank, csl, w.pts[l], w.tab, l, s));
    }
    if (query == data && m == n)
        *qs_out = w.pts[0];
    else
    {
        const uint64_t* G0 = &w.tab->G[0];
        HIP_TRY(launch_fill_u32_dyn(w.count, cap + 1, G0, 1, 0u, s));
        HIP_TRY(launch_hist(query, m, w.tab, 0, w.count, w.qcell, w.qrank, s));
        HIP_TRY(launch_scan_exclusive_dyn(w.count, w.count, cap + 1, G0, 1, w.scan_part, s));
        HIP_TRY(launch_scatter(query, m, w.qcell, w.qrank, w.count, w.qs, w.tab, 0, s));
        *qs_out = w.qs;
    }
    sa.tab = w.tab;
    sa.base_level = 0;
    return PGEOF_OK;
}

// Device search core shared by the host and device entry points.
int run_search(int mode, const float* data, uint32_t n, const float* query, uint32_t m, uint32_t k, float radius,
               void* out_idx, float* out_d2, float* out_feat, uint32_t* out_count, Arena& ar, hipStream_t s,
               SearchWs* ws_out = nullptr, const uint32_t* scales = nullptr, uint32_t n_scales = 0)
{
    const bool csr = mode == kModeRadiusCsr || mode == kModeRadiusMulti;
    SearchWs w = carve(ar, n, m, k, csr);
    if (ar.off > ar.size) return fail(PGEOF_ENOMEM, "workspace too small");
    SearchArgs a{};
    a.scales = scales;
    a.n_scales = n_scales;
    const float4* qs = nullptr;
    const bool is_radius = mode == kModeRadius || mode == kModeRadiusCsr || mode == kModeRadiusMulti;
    int rc = build_grid(data, n, query, m, k, radius, is_radius, w, a, &qs, s);
    if (rc) return rc;
    a.data = data;
    a.queries = qs;
    a.n_query = m;
    a.k = k;
    a.margin = 1e-6f;
    a.r2 = radius * radius;
    a.near_first = tuning().near_first > 0.5 ? 1 : 0;
    // staged rows pay off in the budgeted kNN pass only for large k (measured)
    const int staging = std::max(0, std::min(2, (int)tuning().staging));
    a.staging = staging == 1 && k >= 32 ? 2 : staging;
    a.budget_cells = std::max(1, std::min(64, (int)tuning().budget_cells));
    a.budget_shells = std::max(1, std::min(a.budget_cells, (int)tuning().budget_shells));
    a.out_idx = csr ? (void*)w.dense : out_idx;
    a.out_d2 = out_d2;
    a.out_feat = out_feat;
    a.out_count = csr ? w.counts : out_count;
    const HeapPlan hp = heap_plan(k);
    a.gheap_keys = w.gheap_keys;
    a.heap_tail = tail_layout(hp, m);
    a.tail_pairs = hp.tail_pairs;
    if (mode == kModeKnn || mode == kModeKnnFeat)
    {
        // Budgeted passes, finest level first: a query whose k-th neighbour
        // lies beyond +-2 cells of the level is flagged and, compacted and in
        // sorted order, searched again one level coarser (cells 2x larger).
        // What is left after the last budgeted level goes to the unbudgeted
        // pass, which picks a level per query.
        // A deferred query carries the k-th distance found so far (when its
        // heap was full) to the next pass, which scans only that ball.
        // The number of queries left after pass l is a device-side counter
        // (tab->count[l + 1]): the later passes are launched for up to m
        // queries and take their count from it.
        const float4* list = qs;
        const float* blist = nullptr;
        const uint64_t* n_dev = nullptr;  // queries in `list` (nullptr: all m)
        float4* next = w.qlist;
        float* bnext = w.blist;
        const int max_levels = std::max(1, std::min(kMaxLevels, (int)tuning().grid_levels));
        const int budget_levels = std::max(1, std::min(max_levels, (int)tuning().budget_levels));
        for (int l = 0; l < budget_levels; ++l)
        {
            SearchArgs b = a;
            b.base_level = l;
            b.queries = list;
            b.n_query = m;
            b.n_query_dev = n_dev;
            b.defer_flag = w.flags;
            b.query_bound = blist;
            b.defer_bound = w.dbound;
            HIP_TRY(launch_search(mode, b, hp.block, hp.lds, hp.global, true, s));
            if (n_dev)
                HIP_TRY(launch_put_u32(w.flags, n_dev, 0u, s));
            else
                HIP_TRY(launch_fill_u32(w.flags + m, 1, 0u, s));
            HIP_TRY(launch_scan_exclusive_dyn(w.flags, w.foff, (uint64_t)m + 1, n_dev, 1, w.fpart, s));
            uint64_t* n_next = &w.tab->count[l + 1];
            HIP_TRY(launch_compact_queries(list, m, n_dev, w.flags, w.foff, next, w.dbound, bnext, n_next, s));
            list = next;
            blist = bnext;
            n_dev = n_next;
            next = next == w.qlist ? w.qlist2 : w.qlist;
            bnext = bnext == w.blist ? w.blist2 : w.blist;
        }
        // slowest queries first, lanes of similar cost in each wave
        HIP_TRY(launch_order_by_bound(list, blist, m, n_dev, w.count, w.flags, w.foff, w.scan_part, next, bnext, s));
        SearchArgs b = a;
        b.queries = next;
        b.n_query = m;
        b.n_query_dev = n_dev;
        b.defer_flag = nullptr;
        b.query_bound = bnext;
        b.defer_bound = nullptr;
        HIP_TRY(launch_search(mode, b, hp.block, hp.lds, hp.global, false, s));
    }
    else
        HIP_TRY(launch_search(mode, a, hp.block, hp.lds, hp.global, false, s));
    if (ws_out) *ws_out = w;
    return PGEOF_OK;
}

bool ok_size(int64_t v) { return v >= 0 && v < (int64_t)0xFFFFFFFFLL; }

int validate_knn(int64_t n_data, int64_t n_query, uint64_t knn, bool radius)
{
    if (!ok_size(n_data) || !ok_size(n_query)) return fail(PGEOF_EINVAL, "point cloud too large for uint32 indices");
    if ((int64_t)knn > n_data)
        return fail(PGEOF_EINVAL, radius ? "max knn size is greater than the data point cloud size"
                                         : "knn size is greater than the data point cloud size");
    if ((uint64_t)n_query * knn >= (1ull << 40)) return fail(PGEOF_EINVAL, "output too large");
    return PGEOF_OK;
}

int validate_csr(int64_t n_xyz, int64_t n_nn, int64_t n_nn_ptr)
{
    if (n_nn_ptr < 1) return fail(PGEOF_EINVAL, "nn_ptr must hold at least one entry");
    if (!ok_size(n_xyz) || !ok_size(n_nn_ptr - 1) || n_nn < 0) return fail(PGEOF_EINVAL, "input too large");
    return PGEOF_OK;
}

bool check_scales(const uint32_t* s, int64_t n)
{
    uint32_t prev = 1;  // pgeof.hpp:123-132
    for (int64_t i = 0; i < n; ++i)
    {
        if (s[i] < prev) return false;
        prev = s[i];
    }
    return true;
}

int err_from_word(uint32_t w)
{
    if (w & kErrNnPtr) return fail(PGEOF_EINVAL, "nn_ptr is not non-decreasing or points past the end of nn");
    if (w & kErrNnIndex) return fail(PGEOF_EINVAL, "nn holds an index >= the number of xyz points");
    return PGEOF_OK;
}

// Device CSR driver: modes 0 plain, 1 multiscale, 2 optimal.  The row order
// of large inputs takes its scratch from `order_ws` (order_ws_bytes >=
// csr_order_bytes(rows)) when given, else stream-ordered from the pool.
int csr_device(int mode, const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn, const uint32_t* nn_ptr,
               int64_t n_nn_ptr, uint32_t k_min, uint32_t k_step, uint32_t k_min_search, const uint32_t* d_scales,
               uint32_t n_scales, float* out, uint32_t* err, hipStream_t s, void* order_ws = nullptr,
               size_t order_ws_bytes = 0)
{
    CsrArgs a{};
    a.xyz = xyz;
    a.n_xyz = (uint32_t)n_xyz;
    a.nn = nn;
    a.n_nn = (uint64_t)n_nn;
    a.nn_ptr = nn_ptr;
    a.n_points = (uint32_t)(n_nn_ptr - 1);
    a.k_min = k_min;
    a.k_step = k_step;
    a.k_min_search = k_min_search;
    a.scales = d_scales;
    a.n_scales = n_scales;
    a.out = out;
    a.err = err;
    a.order = nullptr;
    // Features of large inputs: visit the rows in the spatial order of their
    // points (a Morton-like key sorted in two counting passes, stream-ordered
    // scratch), so the coordinate gathers of a wave hit the same cache lines.
    // Measured, order included: optimal k, 1 M rows of 100: 2.7 -> 1.8 ms;
    // plain, 10 M rows of 20: 3.7-3.8 -> 2.3 ms; multiscale (scales
    // 10/20/50/100), 2 M radius rows of up to 64: 2.0 -> 1.89 ms.
    ScratchLease tmp;
    if (a.n_points >= 65536 && n_xyz > 0 && tuning().csr_order > 0.5)
    {
        void* ows = order_ws;
        if (ows)
        {
            if (order_ws_bytes < csr_order_bytes(a.n_points)) return fail(PGEOF_ENOMEM, "row-order scratch too small");
        }
        else
        {
            const int rc = tmp.take(csr_order_bytes(a.n_points), s);
            if (rc) return rc;
            ows = tmp.p;
        }
        uint32_t* ord[1] = {nullptr};
        hipError_t e = launch_csr_order(a, ows, ord, s);
        if (e != hipSuccess) return hip_fail(e, "launch_csr_order");
        a.order = ord[0];
    }
    hipError_t e = launch_csr_features(mode, a, s);
    if (e != hipSuccess) return hip_fail(e, "launch_csr_features");
    return PGEOF_OK;
}

// Host-memory CSR driver.
int csr_host(int mode, const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn, const uint32_t* nn_ptr,
             int64_t n_nn_ptr, uint32_t k_min, uint32_t k_step, uint32_t k_min_search, const uint32_t* scales,
             int64_t n_scales, float* out)
{
    const uint64_t n_points = (uint64_t)(n_nn_ptr - 1);
    const uint64_t n_out = n_points * (mode == 1 ? (uint64_t)n_scales * 11 : (mode == 2 ? 12 : 11));
    if (n_points == 0) return PGEOF_OK;
    hipStream_t s;
    int rc = ctx_prepare(&s);
    if (rc) return rc;
    const size_t b_xyz = al((size_t)n_xyz * 12 + 4), b_nn = al((size_t)n_nn * 4 + 4),
                 b_ptr = al((size_t)n_nn_ptr * 4), b_sc = al((size_t)n_scales * 4 + 4), b_out = al(n_out * 4),
                 b_err = al(4);
    rc = ctx_reserve(b_xyz + b_nn + b_ptr + b_sc + b_out + b_err);
    if (rc) return rc;
    Arena ar{t_ctx.dbuf, t_ctx.dbytes, 0};
    float* d_xyz = ar.take<float>((size_t)n_xyz * 3 + 1);
    uint32_t* d_nn = ar.take<uint32_t>((size_t)n_nn + 1);
    uint32_t* d_ptr = ar.take<uint32_t>((size_t)n_nn_ptr);
    uint32_t* d_sc = ar.take<uint32_t>((size_t)n_scales + 1);
    float* d_out = ar.take<float>(n_out);
    uint32_t* d_err = ar.take<uint32_t>(1);
    if (n_xyz) HIP_TRY(hipMemcpyAsync(d_xyz, xyz, (size_t)n_xyz * 12, hipMemcpyHostToDevice, s));
    if (n_nn) HIP_TRY(hipMemcpyAsync(d_nn, nn, (size_t)n_nn * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_ptr, nn_ptr, (size_t)n_nn_ptr * 4, hipMemcpyHostToDevice, s));
    if (n_scales) HIP_TRY(hipMemcpyAsync(d_sc, scales, (size_t)n_scales * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_fill_u32(d_err, 1, 0u, s));
    rc = csr_device(mode, d_xyz, n_xyz, d_nn, n_nn, d_ptr, n_nn_ptr, k_min, k_step, k_min_search, d_sc,
                    (uint32_t)n_scales, d_out, d_err, s);
    if (rc) return rc;
    uint32_t werr = 0;
    HIP_TRY(hipMemcpyAsync(out, d_out, n_out * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&werr, d_err, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return err_from_word(werr);
}

// Device-memory wrapper: workspace from the caller or from the library's
// scratch cache (returned when the entry point returns, reusable once the
// stream has passed its last launch).
using WsGuard = ScratchLease;

int acquire_ws(WsGuard& g, void* ws, size_t ws_bytes, size_t need, hipStream_t s, Arena& ar)
{
    if (ws)
    {
        if (ws_bytes < need) return fail(PGEOF_EINVAL, "workspace smaller than the size returned by *_workspace_size");
        ar = Arena{(char*)ws, ws_bytes, 0};
        return PGEOF_OK;
    }
    const int rc = g.take(need, s);
    if (rc) return rc;
    ar = Arena{(char*)g.p, need, 0};
    return PGEOF_OK;
}

// compute_features_selected on host arrays (include/pgeof.hpp:325-375).
template <typename T>
int selected_host(const T* xyz, int64_t n, T search_radius, uint32_t max_knn, const int32_t* selected,
                  int64_t n_selected, T* out)
{
    if (!ok_size(n) || n_selected < 0) return fail(PGEOF_EINVAL, "input too large");
    for (int64_t i = 0; i < n_selected; ++i)
        if (selected[i] < 0 || selected[i] > 13) return fail(PGEOF_EINVAL, "selected_features holds a value outside EFeatureID");
    if (n == 0 || n_selected == 0) return PGEOF_OK;
    const bool f64 = sizeof(T) == 8;
    const uint32_t cap = std::max<uint32_t>(max_knn, 2);
    hipStream_t s;
    int rc = ctx_prepare(&s);
    if (rc) return rc;
    const HeapPlan hp = heap_plan_bytes(cap, sizeof(T) + 4);
    const size_t b_x = al((size_t)n * 3 * sizeof(T)), b_xf = f64 ? al((size_t)n * 12) + al((size_t)n * 24) : 0,
                 b_o = al((size_t)n * (size_t)n_selected * sizeof(T)), b_f = al((size_t)n_selected * 4),
                 b_h = hp.global ? al((size_t)n * cap * sizeof(T)) + al((size_t)n * cap * 4) : 0;
    const size_t b_ws = search_ws_bytes((uint64_t)n, (uint64_t)n, 1, false);
    rc = ctx_reserve(b_x + b_xf + b_o + b_f + b_h + b_ws);
    if (rc) return rc;
    Arena ar{t_ctx.dbuf, t_ctx.dbytes, 0};
    T* d_x = ar.take<T>((size_t)n * 3);
    float* d_xf = f64 ? ar.take<float>((size_t)n * 3) : reinterpret_cast<float*>(d_x);
    double* d_xs = f64 ? ar.take<double>((size_t)n * 3) : nullptr;  // sorted copy of the doubles
    T* d_out = ar.take<T>((size_t)n * (size_t)n_selected);
    int32_t* d_f = ar.take<int32_t>((size_t)n_selected);
    void* d_hk = hp.global ? (void*)ar.take<T>((size_t)n * cap) : nullptr;
    uint32_t* d_hi = hp.global ? ar.take<uint32_t>((size_t)n * cap) : nullptr;
    HIP_TRY(hipMemcpyAsync(d_x, xyz, (size_t)n * 3 * sizeof(T), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_f, selected, (size_t)n_selected * 4, hipMemcpyHostToDevice, s));
    if (f64) HIP_TRY(launch_f64_to_f32(reinterpret_cast<const double*>(d_x), d_xf, (uint64_t)n * 3, s));
    Arena wa{t_ctx.dbuf + ar.off, t_ctx.dbytes - ar.off, 0};
    SearchWs w = carve(wa, (uint64_t)n, (uint64_t)n, 1, false);
    if (wa.off > wa.size) return fail(PGEOF_ENOMEM, "workspace too small");
    SearchArgs sa{};
    const float4* qs = nullptr;
    rc = build_grid(d_xf, (uint32_t)n, d_xf, (uint32_t)n, cap, (float)search_radius, true, w, sa, &qs, s);
    if (rc) return rc;
    // this host entry point synchronises anyway: the kernel takes level 0 by value
    LevelTable tab;
    rc = readback(&tab, w.tab, sizeof tab, s);
    if (rc) return rc;
    const Level& L0 = tab.lv[0];
    GridView gv;
    for (int ax = 0; ax < 3; ++ax)
    {
        gv.p.lo[ax] = L0.lo[ax];
        gv.p.g[ax] = L0.g[ax];
    }
    gv.p.h = L0.h;
    gv.p.inv_h = L0.inv_h;
    gv.p.G = (uint64_t)L0.g[0] * L0.g[1] * L0.g[2];
    gv.cell_start = L0.cs;
    gv.pts = L0.pts;
    SelectedArgs a{};
    a.grid = gv;
    a.xyz = d_x;
    a.n = (uint32_t)n;
    a.cap = cap;
    a.max_knn = max_knn;
    a.r2 = (double)(search_radius * search_radius);
    double maxabs = 0.0;
    for (int ax = 0; ax < 3; ++ax)
        maxabs = std::max(maxabs, std::max(std::fabs(gv.p.lo[ax]), std::fabs(gv.p.lo[ax] + gv.p.g[ax] * gv.p.h)));
    // the grid holds float32 copies: widen the cell range by their rounding
    const double margin = gv.p.h * 1e-6 + (f64 ? 2.4e-7 * maxabs : 0.0);
    a.reach = std::fabs((double)search_radius) + margin;
    a.margin = margin;
    a.fids = d_f;
    a.n_sel = (uint32_t)n_selected;
    a.out = d_out;
    a.gheap_keys = d_hk;
    a.gheap_idx = d_hi;
    if (f64)
    {
        HIP_TRY(launch_sorted_copy_f64(L0.pts, reinterpret_cast<const double*>(d_x), (uint32_t)n, d_xs, s));
        a.xs = d_xs;
    }
    HIP_TRY(launch_selected(f64, a, hp.block, hp.lds, hp.global, s));
    HIP_TRY(hipMemcpyAsync(out, d_out, (size_t)n * (size_t)n_selected * sizeof(T), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return PGEOF_OK;
}

}  // namespace

// ===================================================== profiling hooks ====
namespace pgeof_amd
{
void prof_begin(const char* name, hipStream_t s, void** tok)
{
    *tok = nullptr;
    if (!g_prof.load(std::memory_order_relaxed)) return;
    ProfRec* r = new ProfRec{name, nullptr, nullptr};
    if (hipEventCreate(&r->a) != hipSuccess || hipEventCreate(&r->b) != hipSuccess)
    {
        delete r;
        return;
    }
    (void)hipEventRecord(r->a, s);
    *tok = r;
}

void prof_end(void* tok, hipStream_t s)
{
    if (!tok) return;
    ProfRec* r = (ProfRec*)tok;
    (void)hipEventRecord(r->b, s);
    std::lock_guard<std::mutex> l(g_prof_mu);
    g_prof_pending.push_back(*r);
    delete r;
}
}  // namespace pgeof_amd

// ============================================================== C ABI =====
extern "C" {

const char* pgeof_last_error(void) { return t_err.c_str(); }

int pgeof_abi_version(void) { return 100; }

int pgeof_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pgeof_set_tuning(const char* key, double v)
{
    std::lock_guard<std::mutex> l(g_tune_mu);
    const std::string k = key ? key : "";
    if (!(v >= 0.0) || !std::isfinite(v)) return fail(PGEOF_EINVAL, "tuning value must be non-negative");
    double* f = knob(g_tune, k);
    if (!f) return fail(PGEOF_EINVAL, "unknown tuning key " + k);
    *f = v;
    return PGEOF_OK;
}

double pgeof_get_tuning(const char* key)
{
    std::lock_guard<std::mutex> l(g_tune_mu);
    const double* f = knob(g_tune, key ? key : "");
    return f ? *f : -1.0;
}

void pgeof_profile_enable(int on) { g_prof.store(on ? 1 : 0); }

void pgeof_profile_reset(void)
{
    std::lock_guard<std::mutex> l(g_prof_mu);
    for (auto& r : g_prof_pending)
    {
        (void)hipEventSynchronize(r.b);
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_pending.clear();
    g_prof_acc.clear();
}

int pgeof_profile_report(char* buf, size_t len)
{
    std::lock_guard<std::mutex> l(g_prof_mu);
    for (auto& r : g_prof_pending)
    {
        (void)hipEventSynchronize(r.b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        auto& acc = g_prof_acc[r.name];
        acc.first += 1;
        acc.second += ms;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_pending.clear();
    std::string js = "{";
    bool first = true;
    for (auto& kv : g_prof_acc)
    {
        char tmp[256];
        std::snprintf(tmp, sizeof tmp, "%s\"%s\": [%lld, %.6f]", first ? "" : ", ", kv.first.c_str(),
                      (long long)kv.second.first, kv.second.second);
        js += tmp;
        first = false;
    }
    js += "}";
    if (buf && len)
    {
        std::strncpy(buf, js.c_str(), len - 1);
        buf[len - 1] = 0;
    }
    return (int)js.size();
}

// ------------------------------------------------------ host searches ----
int pgeof_knn_search(const float* data, int64_t n_data, const float* query, int64_t n_query, uint32_t knn,
                     uint32_t* indices, float* sqr_dist)
{
    TuningSnapshot snapshot(kFamKnn, n_data, knn);
    int rc = validate_knn(n_data, n_query, knn, false);
    if (rc) return rc;
    if (n_query == 0 || knn == 0) return PGEOF_OK;
    hipStream_t s;
    rc = ctx_prepare(&s);
    if (rc) return rc;
    const bool self = (query == data && n_query == n_data);
    const size_t b_d = al((size_t)n_data * 12), b_q = self ? 0 : al((size_t)n_query * 12),
                 b_i = al((size_t)n_query * knn * 4), b_o = al((size_t)n_query * knn * 4);
    const size_t b_ws = search_ws_bytes((uint64_t)n_data, (uint64_t)n_query, knn, false);
    rc = ctx_reserve(b_d + b_q + b_i + b_o + b_ws);
    if (rc) return rc;
    Arena ar{t_ctx.dbuf, t_ctx.dbytes, 0};
    float* d_data = ar.take<float>((size_t)n_data * 3);
    float* d_query = self ? d_data : ar.take<float>((size_t)n_query * 3);
    uint32_t* d_idx = ar.take<uint32_t>((size_t)n_query * knn);
    float* d_d2 = ar.take<float>((size_t)n_query * knn);
    HIP_TRY(hipMemcpyAsync(d_data, data, (size_t)n_data * 12, hipMemcpyHostToDevice, s));
    if (!self) HIP_TRY(hipMemcpyAsync(d_query, query, (size_t)n_query * 12, hipMemcpyHostToDevice, s));
    Arena wa{t_ctx.dbuf + ar.off, t_ctx.dbytes - ar.off, 0};
    rc = run_search(kModeKnn, d_data, (uint32_t)n_data, d_query, (uint32_t)n_query, knn, 0.0f, d_idx, d_d2, nullptr,
                    nullptr, wa, s);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(indices, d_idx, (size_t)n_query * knn * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(sqr_dist, d_d2, (size_t)n_query * knn * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return PGEOF_OK;
}

int pgeof_radius_search(const float* data, int64_t n_data, const float* query, int64_t n_query,
                        float search_radius, uint32_t max_knn, int32_t* indices, float* sqr_dist)
{
    TuningSnapshot snapshot(kFamRadius, n_data, max_knn);
    int rc = validate_knn(n_data, n_query, max_knn, true);
    if (rc) return rc;
    if (n_query == 0 || max_knn == 0) return PGEOF_OK;
    hipStream_t s;
    rc = ctx_prepare(&s);
    if (rc) return rc;
    const bool self = (query == data && n_query == n_data);
    const size_t b_d = al((size_t)n_data * 12), b_q = self ? 0 : al((size_t)n_query * 12),
                 b_i = al((size_t)n_query * max_knn * 4), b_o = al((size_t)n_query * max_knn * 4);
    const size_t b_ws = search_ws_bytes((uint64_t)n_data, (uint64_t)n_query, max_knn, false);
    rc = ctx_reserve(b_d + b_q + b_i + b_o + b_ws);
    if (rc) return rc;
    Arena ar{t_ctx.dbuf, t_ctx.dbytes, 0};
    float* d_data = ar.take<float>((size_t)n_data * 3);
    float* d_query = self ? d_data : ar.take<float>((size_t)n_query * 3);
    int32_t* d_idx = ar.take<int32_t>((size_t)n_query * max_knn);
    float* d_d2 = ar.take<float>((size_t)n_query * max_knn);
    HIP_TRY(hipMemcpyAsync(d_data, data, (size_t)n_data * 12, hipMemcpyHostToDevice, s));
    if (!self) HIP_TRY(hipMemcpyAsync(d_query, query, (size_t)n_query * 12, hipMemcpyHostToDevice, s));
    Arena wa{t_ctx.dbuf + ar.off, t_ctx.dbytes - ar.off, 0};
    rc = run_search(kModeRadius, d_data, (uint32_t)n_data, d_query, (uint32_t)n_query, max_knn, search_radius, d_idx,
                    d_d2, nullptr, nullptr, wa, s);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(indices, d_idx, (size_t)n_query * max_knn * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(sqr_dist, d_d2, (size_t)n_query * max_knn * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return PGEOF_OK;
}

// ----------------------------------------------------- host features ----
int pgeof_compute_features(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                           const uint32_t* nn_ptr, int64_t n_nn_ptr, uint64_t k_min, float* features)
{
    TuningSnapshot snapshot;
    if (k_min < 1) return fail(PGEOF_EINVAL, "k_min should be > 1");
    int rc = validate_csr(n_xyz, n_nn, n_nn_ptr);
    if (rc) return rc;
    const uint32_t km = (uint32_t)std::min<uint64_t>(k_min, 0xFFFFFFFFull);
    return csr_host(0, xyz, n_xyz, nn, n_nn, nn_ptr, n_nn_ptr, km, 1, 1, nullptr, 0, features);
}

int pgeof_compute_features_multiscale(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                      const uint32_t* nn_ptr, int64_t n_nn_ptr, const uint32_t* k_scales,
                                      int64_t n_scales, float* features)
{
    TuningSnapshot snapshot;
    if (n_scales < 0 || !check_scales(k_scales, n_scales))
        return fail(PGEOF_EINVAL, "k_scales should be > 1 and sorted in ascending order");
    int rc = validate_csr(n_xyz, n_nn, n_nn_ptr);
    if (rc) return rc;
    if (n_scales == 0) return PGEOF_OK;
    return csr_host(1, xyz, n_xyz, nn, n_nn, nn_ptr, n_nn_ptr, 1, 1, 1, k_scales, n_scales, features);
}

int pgeof_compute_features_optimal(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                   const uint32_t* nn_ptr, int64_t n_nn_ptr, uint32_t k_min, uint32_t k_step,
                                   uint32_t k_min_search, float* features)
{
    TuningSnapshot snapshot;
    if (k_min < 1 && k_min_search < 1) return fail(PGEOF_EINVAL, "k_min and k_min_search should be > 1");
    if (k_step < 1) return fail(PGEOF_EINVAL, "k_step should be >= 1");
    int rc = validate_csr(n_xyz, n_nn, n_nn_ptr);
    if (rc) return rc;
    return csr_host(2, xyz, n_xyz, nn, n_nn, nn_ptr, n_nn_ptr, k_min, k_step, k_min_search, nullptr, 0, features);
}

int pgeof_compute_features_selected_f32(const float* xyz, int64_t n, float search_radius, uint32_t max_knn,
                                        const int32_t* selected, int64_t n_selected, float* features)
{
    TuningSnapshot snapshot;
    return selected_host<float>(xyz, n, search_radius, max_knn, selected, n_selected, features);
}

int pgeof_compute_features_selected_f64(const double* xyz, int64_t n, double search_radius, uint32_t max_knn,
                                        const int32_t* selected, int64_t n_selected, double* features)
{
    TuningSnapshot snapshot;
    return selected_host<double>(xyz, n, search_radius, max_knn, selected, n_selected, features);
}

// ---------------------------------------------------- device searches ----
size_t pgeof_knn_workspace_size(int64_t n_data, int64_t n_query, uint32_t knn)
{
    TuningSnapshot snapshot(kFamKnn, n_data, knn);
    return search_ws_bytes((uint64_t)std::max<int64_t>(n_data, 0), (uint64_t)std::max<int64_t>(n_query, 0), knn, false);
}

size_t pgeof_radius_workspace_size(int64_t n_data, int64_t n_query, uint32_t max_knn)
{
    TuningSnapshot snapshot(kFamRadius, n_data, max_knn);
    return search_ws_bytes((uint64_t)std::max<int64_t>(n_data, 0), (uint64_t)std::max<int64_t>(n_query, 0), max_knn,
                           false);
}

size_t pgeof_radius_csr_workspace_size(int64_t n, uint32_t max_knn)
{
    TuningSnapshot snapshot(kFamRadius, n, max_knn);
    return search_ws_bytes((uint64_t)std::max<int64_t>(n, 0), (uint64_t)std::max<int64_t>(n, 0), max_knn, true);
}

int pgeof_knn_search_device(const float* data, int64_t n_data, const float* query, int64_t n_query, uint32_t knn,
                            uint32_t* indices, float* sqr_dist, void* workspace, size_t workspace_bytes,
                            void* stream)
{
    TuningSnapshot snapshot(kFamKnn, n_data, knn);
    int rc = validate_knn(n_data, n_query, knn, false);
    if (rc) return rc;
    if (n_query == 0 || knn == 0) return PGEOF_OK;
    hipStream_t s = (hipStream_t)stream;
    WsGuard g;
    Arena ar;
    rc = acquire_ws(g, workspace, workspace_bytes, pgeof_knn_workspace_size(n_data, n_query, knn), s, ar);
    if (rc) return rc;
    return run_search(kModeKnn, data, (uint32_t)n_data, query, (uint32_t)n_query, knn, 0.0f, indices, sqr_dist,
                      nullptr, nullptr, ar, s);
}

int pgeof_radius_search_device(const float* data, int64_t n_data, const float* query, int64_t n_query,
                               float search_radius, uint32_t max_knn, int32_t* indices, float* sqr_dist,
                               void* workspace, size_t workspace_bytes, void* stream)
{
    TuningSnapshot snapshot(kFamRadius, n_data, max_knn);
    int rc = validate_knn(n_data, n_query, max_knn, true);
    if (rc) return rc;
    if (n_query == 0 || max_knn == 0) return PGEOF_OK;
    hipStream_t s = (hipStream_t)stream;
    WsGuard g;
    Arena ar;
    rc = acquire_ws(g, workspace, workspace_bytes, pgeof_radius_workspace_size(n_data, n_query, max_knn), s, ar);
    if (rc) return rc;
    return run_search(kModeRadius, data, (uint32_t)n_data, query, (uint32_t)n_query, max_knn, search_radius, indices,
                      sqr_dist, nullptr, nullptr, ar, s);
}

int pgeof_knn_features_device(const float* xyz, int64_t n, uint32_t knn, uint32_t* nn, uint32_t* nn_ptr,
                              float* sqr_dist, float* features, void* workspace, size_t workspace_bytes,
                              void* stream)
{
    TuningSnapshot snapshot(kFamKnn, n, knn);
    int rc = validate_knn(n, n, knn, false);
    if (rc) return rc;
    if (knn < 1) return fail(PGEOF_EINVAL, "knn should be >= 1");
    if ((uint64_t)n * knn > 0xFFFFFFFFull) return fail(PGEOF_EINVAL, "n * knn exceeds the uint32 CSR range");
    if (n == 0) return PGEOF_OK;
    hipStream_t s = (hipStream_t)stream;
    WsGuard g;
    Arena ar;
    rc = acquire_ws(g, workspace, workspace_bytes, pgeof_knn_workspace_size(n, n, knn), s, ar);
    if (rc) return rc;
    rc = run_search(kModeKnnFeat, xyz, (uint32_t)n, xyz, (uint32_t)n, knn, 0.0f, nn, sqr_dist, features, nullptr, ar,
                    s);
    if (rc) return rc;
    HIP_TRY(launch_arange_ptr(nn_ptr, (uint32_t)n, knn, s));
    return PGEOF_OK;
}

int pgeof_radius_csr_device(const float* xyz, int64_t n, float search_radius, uint32_t max_knn, uint32_t* nn,
                            uint32_t* nn_ptr, int64_t* nnz_out, void* workspace, size_t workspace_bytes,
                            void* stream)
{
    TuningSnapshot snapshot(kFamRadius, n, max_knn);
    int rc = validate_knn(n, n, max_knn, true);
    if (rc) return rc;
    if ((uint64_t)n * max_knn > 0xFFFFFFFFull) return fail(PGEOF_EINVAL, "n * max_knn exceeds the uint32 CSR range");
    hipStream_t s = (hipStream_t)stream;
    if (n == 0 || max_knn == 0)
    {
        HIP_TRY(launch_fill_u32(nn_ptr, (uint64_t)n + 1, 0u, s));
        if (nnz_out) *nnz_out = 0;
        return PGEOF_OK;
    }
    WsGuard g;
    Arena ar;
    rc = acquire_ws(g, workspace, workspace_bytes, pgeof_radius_csr_workspace_size(n, max_knn), s, ar);
    if (rc) return rc;
    SearchWs w;
    rc = run_search(kModeRadiusCsr, xyz, (uint32_t)n, xyz, (uint32_t)n, max_knn, search_radius, nullptr, nullptr,
                    nullptr, nullptr, ar, s, &w);
    if (rc) return rc;
    HIP_TRY(launch_fill_u32(w.counts + n, 1, 0u, s));
    HIP_TRY(launch_scan_exclusive(w.counts, nn_ptr, (uint64_t)n + 1, w.csr_part, s));
    HIP_TRY(launch_compact_rows(w.dense, (uint32_t)n, max_knn, w.counts, nn_ptr, nn, s));
    if (nnz_out)
    {
        uint32_t nnz = 0;
        rc = readback(&nnz, nn_ptr + n, 4, s);
        if (rc) return rc;
        *nnz_out = nnz;
    }
    return PGEOF_OK;
}

size_t pgeof_radius_multiscale_workspace_size(int64_t n, uint32_t max_knn, int64_t n_scales)
{
    TuningSnapshot snapshot(kFamRadius, n, max_knn);
    return al((size_t)std::max<int64_t>(n_scales, 0) * 4 + 4) + pgeof_radius_csr_workspace_size(n, max_knn);
}

int pgeof_radius_multiscale_device(const float* xyz, int64_t n, float search_radius, uint32_t max_knn,
                                   const uint32_t* k_scales, int64_t n_scales, uint32_t* nn, uint32_t* nn_ptr,
                                   float* features, void* workspace, size_t workspace_bytes, void* stream)
{
    TuningSnapshot snapshot(kFamRadius, n, max_knn);
    if (n_scales < 0 || (n_scales > 0 && !k_scales) || !check_scales(k_scales, n_scales))
        return fail(PGEOF_EINVAL, "k_scales should be > 1 and sorted in ascending order");
    int rc = validate_knn(n, n, max_knn, true);
    if (rc) return rc;
    if ((uint64_t)n * max_knn > 0xFFFFFFFFull) return fail(PGEOF_EINVAL, "n * max_knn exceeds the uint32 CSR range");
    if ((uint64_t)n * (uint64_t)n_scales * 11 >= (1ull << 40)) return fail(PGEOF_EINVAL, "output too large");
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) return PGEOF_OK;
    if (max_knn == 0)
    {
        // no hits: empty rows, every scale zero (k_scales entries are >= 1)
        HIP_TRY(launch_fill_u32(nn_ptr, (uint64_t)n + 1, 0u, s));
        HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(features), (uint64_t)n * n_scales * 11, 0u, s));
        return PGEOF_OK;
    }
    WsGuard g;
    Arena ar;
    rc = acquire_ws(g, workspace, workspace_bytes, pgeof_radius_multiscale_workspace_size(n, max_knn, n_scales), s,
                    ar);
    if (rc) return rc;
    uint32_t* d_sc = ar.take<uint32_t>((size_t)n_scales + 1);  // the scales, then the error word of the CSR kernel
    HIP_TRY(launch_put_values(d_sc, k_scales, (uint64_t)n_scales, s));
    const bool fused = tuning().fuse_multiscale > 0.5;
    SearchWs w;
    rc = run_search(fused ? kModeRadiusMulti : kModeRadiusCsr, xyz, (uint32_t)n, xyz, (uint32_t)n, max_knn,
                    search_radius, nullptr, nullptr, features, nullptr, ar, s, &w, d_sc, (uint32_t)n_scales);
    if (rc) return rc;
    HIP_TRY(launch_fill_u32(w.counts + n, 1, 0u, s));
    HIP_TRY(launch_scan_exclusive(w.counts, nn_ptr, (uint64_t)n + 1, w.csr_part, s));
    HIP_TRY(launch_compact_rows(w.dense, (uint32_t)n, max_knn, w.counts, nn_ptr, nn, s));
    if (!fused && n_scales)
    {
        // the multiscale CSR kernel over the compacted rows (its error word
        // stays 0: the rows hold search hits only)
        uint32_t* d_err = d_sc + n_scales;
        HIP_TRY(launch_fill_u32(d_err, 1, 0u, s));
        CsrArgs a{};
        a.xyz = xyz;
        a.n_xyz = (uint32_t)n;
        a.nn = nn;
        a.n_nn = (uint64_t)n * max_knn;  // a bound of nnz: the kernel checks rows against it
        a.nn_ptr = nn_ptr;
        a.n_points = (uint32_t)n;
        a.k_min = a.k_step = a.k_min_search = 1;
        a.scales = d_sc;
        a.n_scales = (uint32_t)n_scales;
        a.out = features;
        a.err = d_err;
        // rows in the search's spatial order (the sorted points of the self
        // search), kept in the query-cell array a self search does not use
        HIP_TRY(launch_order_from_sorted(w.pts[0], (uint32_t)n, w.qcell, s));
        a.order = w.qcell;
        HIP_TRY(launch_csr_features(1, a, s));
    }
    return PGEOF_OK;
}

int pgeof_knn_optimal_device(const float* xyz, int64_t n, uint32_t knn, uint32_t k_min, uint32_t k_step,
                             uint32_t k_min_search, uint32_t* nn, uint32_t* nn_ptr, float* features, void* workspace,
                             size_t workspace_bytes, void* stream)
{
    TuningSnapshot snapshot(kFamKnn, n, knn);
    int rc = validate_knn(n, n, knn, false);
    if (rc) return rc;
    if (knn < 1) return fail(PGEOF_EINVAL, "knn should be >= 1");
    if (k_min < 1 && k_min_search < 1) return fail(PGEOF_EINVAL, "k_min and k_min_search should be > 1");
    if (k_step < 1) return fail(PGEOF_EINVAL, "k_step should be >= 1");
    if ((uint64_t)n * knn > 0xFFFFFFFFull) return fail(PGEOF_EINVAL, "n * knn exceeds the uint32 CSR range");
    if (n == 0) return PGEOF_OK;
    hipStream_t s = (hipStream_t)stream;
    WsGuard g;
    Arena ar;
    rc = acquire_ws(g, workspace, workspace_bytes, pgeof_knn_workspace_size(n, n, knn) + al(4), s, ar);
    if (rc) return rc;
    uint32_t* d_err = ar.take<uint32_t>(1);
    HIP_TRY(launch_fill_u32(d_err, 1, 0u, s));
    SearchWs w;
    rc = run_search(kModeKnn, xyz, (uint32_t)n, xyz, (uint32_t)n, knn, 0.0f, nn, nullptr, nullptr, nullptr, ar, s,
                    &w);
    if (rc) return rc;
    HIP_TRY(launch_arange_ptr(nn_ptr, (uint32_t)n, knn, s));
    // rows in the search's spatial order (the sorted points of a self
    // search), kept in the query-cell array a self search does not use
    HIP_TRY(launch_order_from_sorted(w.pts[0], (uint32_t)n, w.qcell, s));
    CsrArgs a{};
    a.xyz = xyz;
    a.n_xyz = (uint32_t)n;
    a.nn = nn;
    a.n_nn = (uint64_t)n * knn;
    a.nn_ptr = nn_ptr;
    a.n_points = (uint32_t)n;
    a.k_min = k_min;
    a.k_step = k_step;
    a.k_min_search = k_min_search;
    a.out = features;
    a.err = d_err;
    a.order = w.qcell;
    HIP_TRY(launch_csr_features(2, a, s));
    return PGEOF_OK;
}

// ---------------------------------------------------- device features ----
static int csr_device_checked(int mode, const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                              const uint32_t* nn_ptr, int64_t n_nn_ptr, uint32_t k_min, uint32_t k_step,
                              uint32_t k_min_search, const uint32_t* scales_host, int64_t n_scales, float* out,
                              uint32_t* err, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (n_nn_ptr - 1 == 0) return PGEOF_OK;
    uint32_t* d_err = err;
    ScratchLease tmp;
    int rc = tmp.take(al((size_t)n_scales * 4 + 4) + al(4), s);
    if (rc) return rc;
    uint32_t* d_sc = (uint32_t*)tmp.p;
    // the scales travel as kernel arguments (no copy from host memory)
    if (n_scales) HIP_TRY(launch_put_values(d_sc, scales_host, (uint64_t)n_scales, s));
    if (!d_err)
    {
        d_err = (uint32_t*)((char*)tmp.p + al((size_t)n_scales * 4 + 4));
        HIP_TRY(launch_fill_u32(d_err, 1, 0u, s));
    }
    rc = csr_device(mode, xyz, n_xyz, nn, n_nn, nn_ptr, n_nn_ptr, k_min, k_step, k_min_search, d_sc,
                    (uint32_t)n_scales, out, d_err, s);
    if (rc) return rc;
    if (!err)
    {
        uint32_t w = 0;
        rc = readback(&w, d_err, 4, s);
        if (rc) return rc;
        return err_from_word(w);
    }
    return PGEOF_OK;
}

int pgeof_compute_features_device(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                  const uint32_t* nn_ptr, int64_t n_nn_ptr, uint64_t k_min, float* features,
                                  uint32_t* err, void* stream)
{
    TuningSnapshot snapshot;
    if (k_min < 1) return fail(PGEOF_EINVAL, "k_min should be > 1");
    int rc = validate_csr(n_xyz, n_nn, n_nn_ptr);
    if (rc) return rc;
    const uint32_t km = (uint32_t)std::min<uint64_t>(k_min, 0xFFFFFFFFull);
    return csr_device_checked(0, xyz, n_xyz, nn, n_nn, nn_ptr, n_nn_ptr, km, 1, 1, nullptr, 0, features, err,
                              stream);
}

int pgeof_compute_features_device_noalloc(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                          const uint32_t* nn_ptr, int64_t n_nn_ptr, uint64_t k_min, float* features,
                                          uint32_t* err, void* stream)
{
    TuningSnapshot snapshot;
    if (k_min < 1) return fail(PGEOF_EINVAL, "k_min should be > 1");
    int rc = validate_csr(n_xyz, n_nn, n_nn_ptr);
    if (rc) return rc;
    if (!err) return fail(PGEOF_EINVAL, "err must be a device word: the no-allocation entry point never synchronises");
    if ((uint64_t)(n_nn_ptr - 1) > PGEOF_NOALLOC_MAX_ROWS)
        return fail(PGEOF_EINVAL, "more rows than the no-allocation entry point supports (PGEOF_NOALLOC_MAX_ROWS = " +
                                      std::to_string(PGEOF_NOALLOC_MAX_ROWS) + ")");
    if (n_nn_ptr - 1 == 0) return PGEOF_OK;
    // the library's static scratch on this device (row-order buffers); it
    // holds no state between calls, everything in it is rewritten per call
    void* ws = nullptr;
    size_t ws_bytes = 0;
    hipError_t e = noalloc_scratch(&ws, &ws_bytes);
    if (e != hipSuccess) return hip_fail(e, "no-allocation scratch");
    const uint32_t km = (uint32_t)std::min<uint64_t>(k_min, 0xFFFFFFFFull);
    return csr_device(0, xyz, n_xyz, nn, n_nn, nn_ptr, n_nn_ptr, km, 1, 1, nullptr, 0, features, err,
                      (hipStream_t)stream, ws, ws_bytes);
}

int pgeof_compute_features_multiscale_device(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                             const uint32_t* nn_ptr, int64_t n_nn_ptr,
                                             const uint32_t* k_scales_host, int64_t n_scales, float* features,
                                             uint32_t* err, void* stream)
{
    TuningSnapshot snapshot;
    if (n_scales < 0 || !check_scales(k_scales_host, n_scales))
        return fail(PGEOF_EINVAL, "k_scales should be > 1 and sorted in ascending order");
    int rc = validate_csr(n_xyz, n_nn, n_nn_ptr);
    if (rc) return rc;
    if (n_scales == 0) return PGEOF_OK;
    return csr_device_checked(1, xyz, n_xyz, nn, n_nn, nn_ptr, n_nn_ptr, 1, 1, 1, k_scales_host, n_scales, features,
                              err, stream);
}

int pgeof_compute_features_optimal_device(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                          const uint32_t* nn_ptr, int64_t n_nn_ptr, uint32_t k_min,
                                          uint32_t k_step, uint32_t k_min_search, float* features, uint32_t* err,
                                          void* stream)
{
    TuningSnapshot snapshot;
    if (k_min < 1 && k_min_search < 1) return fail(PGEOF_EINVAL, "k_min and k_min_search should be > 1");
    if (k_step < 1) return fail(PGEOF_EINVAL, "k_step should be >= 1");
    int rc = validate_csr(n_xyz, n_nn, n_nn_ptr);
    if (rc) return rc;
    return csr_device_checked(2, xyz, n_xyz, nn, n_nn, nn_ptr, n_nn_ptr, k_min, k_step, k_min_search, nullptr, 0,
                              features, err, stream);
}

// ----------------------------------------------------------- autotuner ----
int pgeof_autotune(const char* workload, const float* xyz, int64_t n, uint32_t k, float radius,
                   const uint32_t* k_scales, int64_t n_scales, int reps, void* stream, char* report,
                   size_t report_len)
{
    const std::string wl = workload ? workload : "";
    int family;
    if (wl == "knn_features" || wl == "knn_optimal")
        family = kFamKnn;
    else if (wl == "radius_multiscale")
        family = kFamRadius;
    else
        return fail(PGEOF_EINVAL, "unknown autotune workload " + wl);
    if (n < 1 || k < 1) return fail(PGEOF_EINVAL, "autotune needs n >= 1 and k >= 1");
    reps = std::max(1, std::min(reps, 100));
    hipStream_t s = (hipStream_t)stream;

    // Candidates: the current knobs first, then a sweep of the knobs that
    // move the speed of this workload: grid resolution (cells per point,
    // points per cell or cell edge / radius), block size (through the LDS
    // budget of the heaps), row staging (wave-shared LDS tiles or per-lane
    // reads) and, for radius + multiscale, where the features are computed.
    Tuning base;
    {
        std::lock_guard<std::mutex> l(g_tune_mu);
        base = g_tune;
    }
    std::vector<Tuning> cand{base};
    if (family == kFamKnn)
    {
        for (double cpp : {4.0, 8.0, 16.0})
            for (double occ : {0.25, 0.35, 0.5})
                for (double lds : {65536.0, 32768.0})
                    for (double st : {0.0, 2.0})
                    {
                        Tuning t = base;
                        t.grid_cells_per_point = cpp;
                        t.knn_occupancy = occ;
                        t.lds_budget = lds;
                        t.staging = st;
                        cand.push_back(t);
                    }
    }
    else
    {
        for (double rcf : {0.35, 0.5, 0.7, 1.0})
            for (double fu : {0.0, 1.0})
                for (double st : {0.0, 1.0})
                {
                    Tuning t = base;
                    t.radius_cell_factor = rcf;
                    t.fuse_multiscale = fu;
                    t.staging = st;
                    cand.push_back(t);
                }
    }

    // buffers: outputs, and the largest workspace over the candidates
    const uint64_t nn_cap = (uint64_t)n * k;
    const uint64_t n_feat = wl == "knn_features" ? 11 : (wl == "knn_optimal" ? 12 : (uint64_t)n_scales * 11);
    size_t ws_bytes = 0;
    for (const Tuning& t : cand)
    {
        t_snap = &t;
        size_t b = family == kFamKnn ? pgeof_knn_workspace_size(n, n, k) + 256
                                     : pgeof_radius_multiscale_workspace_size(n, k, n_scales);
        t_snap = nullptr;
        ws_bytes = std::max(ws_bytes, b);
    }
    void *d_nn = nullptr, *d_ptr = nullptr, *d_f = nullptr, *d_ws = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        t_snap = nullptr;
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (void* p : {d_nn, d_ptr, d_f, d_ws})
            if (p) (void)hipFree(p);
    };
    hipError_t he = hipMalloc(&d_nn, nn_cap * 4 + 4);
    if (he == hipSuccess) he = hipMalloc(&d_ptr, ((size_t)n + 1) * 4);
    if (he == hipSuccess) he = hipMalloc(&d_f, (size_t)n * std::max<uint64_t>(n_feat, 1) * 4);
    if (he == hipSuccess) he = hipMalloc(&d_ws, ws_bytes);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he != hipSuccess)
    {
        cleanup();
        return hip_fail(he, "autotune buffers");
    }
    auto run = [&]() -> int {
        if (wl == "knn_features")
            return pgeof_knn_features_device(xyz, n, k, (uint32_t*)d_nn, (uint32_t*)d_ptr, nullptr, (float*)d_f, d_ws,
                                             ws_bytes, s);
        if (wl == "knn_optimal")
            return pgeof_knn_optimal_device(xyz, n, k, 10, 1, 10, (uint32_t*)d_nn, (uint32_t*)d_ptr, (float*)d_f, d_ws,
                                            ws_bytes, s);
        return pgeof_radius_multiscale_device(xyz, n, radius, k, k_scales, n_scales, (uint32_t*)d_nn,
                                              (uint32_t*)d_ptr, (float*)d_f, d_ws, ws_bytes, s);
    };
    // one timed run of candidate c (its knobs for the duration of the call)
    auto time_once = [&](size_t c, float* v) -> int {
        t_snap = &cand[c];
        hipError_t e = hipEventRecord(e0, s);
        int rc = e == hipSuccess ? run() : 0;
        if (!rc && e == hipSuccess) e = hipEventRecord(e1, s);
        if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
        if (!rc && e == hipSuccess) e = hipEventElapsedTime(v, e0, e1);
        t_snap = nullptr;
        if (!rc && e != hipSuccess) rc = hip_fail(e, "autotune timing");
        return rc;
    };
    auto median = [](std::vector<float>& v) -> double {
        std::sort(v.begin(), v.end());
        return v[v.size() / 2];
    };
    // Stage 1: every candidate, a warm-up and `reps` timed runs (median).
    std::vector<double> ms(cand.size(), 0.0);
    auto stage1 = [&](size_t c) -> int {
        t_snap = &cand[c];
        int rc = run();  // warm-up
        t_snap = nullptr;
        std::vector<float> tt;
        for (int r = 0; r < reps && !rc; ++r)
        {
            float v = 0.0f;
            rc = time_once(c, &v);
            tt.push_back(v);
        }
        if (!rc) ms[c] = median(tt);
        return rc;
    };
    for (size_t c = 0; c < cand.size(); ++c)
        if (const int rc = stage1(c))
        {
            cleanup();
            return rc;
        }
    // kNN: the pyramid's depth, tried on the fastest set only (the workspace
    // does not depend on it).  One level wins on clouds without sparse
    // regions for the coarse levels to serve (10 k uniform points: 0.197 vs
    // 0.226 ms) and loses on clustered ones (10 k LiDAR-like: 1.20 vs 0.64).
    if (family == kFamKnn)
    {
        const size_t b0 = (size_t)(std::min_element(ms.begin(), ms.end()) - ms.begin());
        Tuning v = cand[b0];
        v.grid_levels = v.grid_levels > 1.5 ? 1.0 : 3.0;
        cand.push_back(v);
        ms.push_back(0.0);
        if (const int rc = stage1(cand.size() - 1))
        {
            cleanup();
            return rc;
        }
    }
    // Stage 2: the four fastest again, 2 * reps + 1 runs each, taken in
    // turns so that clock drift affects them alike; their medians replace
    // the first ones and the lowest time wins (close candidates are
    // otherwise picked by the noise of a few runs).
    std::vector<size_t> order(cand.size());
    for (size_t c = 0; c < order.size(); ++c) order[c] = c;
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return ms[x] < ms[y]; });
    const size_t n_fin = std::min<size_t>(4, order.size());
    std::vector<std::vector<float>> fin(n_fin);
    for (int r = 0; r < 2 * reps + 1; ++r)
        for (size_t f = 0; f < n_fin; ++f)
        {
            float v = 0.0f;
            const int rc = time_once(order[f], &v);
            if (rc)
            {
                cleanup();
                return rc;
            }
            fin[f].push_back(v);
        }
    for (size_t f = 0; f < n_fin; ++f) ms[order[f]] = median(fin[f]);
    size_t best = 0;
    for (size_t c = 1; c < ms.size(); ++c)
        if (ms[c] < ms[best]) best = c;
    cleanup();
    const int sc = size_class(n);
    {
        std::lock_guard<std::mutex> l(g_tune_mu);
        g_tuned[TuneKey{family, k, sc}] = cand[best];
    }
    if (report && report_len)
    {
        auto knobs_json = [](const Tuning& t) {
            std::string js = "{";
            for (size_t i = 0; i < sizeof(kKnobs) / sizeof(kKnobs[0]); ++i)
            {
                char tmp[96];
                std::snprintf(tmp, sizeof tmp, "%s\"%s\": %g", i ? ", " : "", kKnobs[i].name, t.*(kKnobs[i].field));
                js += tmp;
            }
            return js + "}";
        };
        std::string js = "{\"workload\": \"" + wl + "\", \"n\": " + std::to_string(n) + ", \"k\": " +
                         std::to_string(k) + ", \"size_class\": " + std::to_string(sc) + ", \"ms\": [";
        for (size_t c = 0; c < ms.size(); ++c)
        {
            char tmp[32];
            std::snprintf(tmp, sizeof tmp, "%s%.4f", c ? ", " : "", ms[c]);
            js += tmp;
        }
        char tmp[96];
        std::snprintf(tmp, sizeof tmp, "], \"chosen\": %zu, \"default_ms\": %.4f, \"best_ms\": %.4f", best, ms[0],
                      ms[best]);
        js += tmp;
        js += ", \"knobs\": " + knobs_json(cand[best]) + ", \"candidates\": [";
        for (size_t c = 0; c < cand.size(); ++c) js += (c ? ", " : "") + knobs_json(cand[c]);
        js += "]}";
        std::strncpy(report, js.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return PGEOF_OK;
}

void pgeof_autotune_clear(void)
{
    std::lock_guard<std::mutex> l(g_tune_mu);
    g_tuned.clear();
}

}  // extern "C"
