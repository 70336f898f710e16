// Host-side interfaces between the C-ABI layer and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace pgeof_amd
{

// error bits reported by kernels through a device word
enum : uint32_t
{
    kErrNnIndex = 1u,  // a neighbour index >= number of xyz rows
    kErrNnPtr = 2u,    // nn_ptr not non-decreasing or beyond len(nn)
};

// Uniform grid over a point set.  Cells are linearised row-major with x
// fastest, so one (y, z) row of cells is one contiguous range of the sorted
// point array.
struct GridParams
{
    double lo[3];
    double h, inv_h;
    int g[3];
    uint64_t G;  // g[0] * g[1] * g[2]
};

struct GridView
{
    GridParams p;
    const uint32_t* cell_start;  // G + 1 entries
    const float4* pts;           // sorted points: x, y, z, bit pattern of the original index
};

// One level of the multi-level grid used by the neighbour search.
struct Level
{
    double lo[3];
    double h, inv_h;
    int g[3];
    const uint32_t* cs;  // cell starts, G + 1 entries
    const float4* pts;   // points sorted by cell of this level
};

constexpr int kMaxLevels = 4;

// Hardware wave slots a kernel can name with hw_wave_slot(): 3 bits of XCC,
// 2 of shader engine, 1 of shader array, 4 of CU, 2 of SIMD and 4 of wave
// slot.  Waves resident at the same time never share one.
constexpr uint32_t kHwWaveSlots = 1u << 16;
constexpr uint32_t kHwMaxXcc = 8;

// The grid pyramid of one search, planned and written by kernels (bounding
// box, sizing probe, levels), so that no value travels to the host between
// the launches of a search.  Level l has cell edge h_0 * ratio^l; a level
// that was not built has G[l] = 0.
struct LevelTable
{
    Level lv[kMaxLevels];
    uint64_t G[kMaxLevels];  // cells of level l (0: absent)
    int n_levels;
    Level probe;             // kNN sizing probe grid (cs, pts unused)
    uint64_t probe_G;
    float bb[8];             // bounding box of the finite data coordinates: lo xyz, hi xyz
    unsigned long long sumsq;  // probe: sum of squared cell counts
    double hcap, T, h1;      // planning intermediates (grid-size floor, target occupancy, first guess)
    uint64_t count[kMaxLevels + 2];  // kNN: queries left after each budgeted pass (device-side work counters)
};

// Everything the planning kernels need to size the grids of a search.
struct PlanArgs
{
    uint32_t n;  // data points
    uint32_t k;  // kNN: k (sizes the cells); radius: unused
    float radius;
    int is_radius;
    uint64_t cap;  // largest G of level 0 (grid_cells_per_point * n)
    int max_levels;
    double knn_occupancy, knn_cell_points_max, radius_cell_factor, level_ratio;
    uint32_t probe_stride;
    uint32_t* cs[kMaxLevels];       // cell-start table of level l (fixed offsets in the workspace)
    uint64_t cs_room[kMaxLevels];   // entries available there
    float4* pts[kMaxLevels];        // sorted points of level l
};

// Arguments of the neighbour-search kernels.
struct SearchArgs
{
    const LevelTable* tab;  // the pyramid (device memory)
    int base_level;         // budgeted passes and radius searches: the level they scan
    const float* data;      // the caller's data points, (n, 3) float32 (fused features gather from it)
    const float4* queries;  // sorted queries (x, y, z, bits of original query index)
    uint32_t n_query;       // queries, or their upper bound when n_query_dev is set
    const uint64_t* n_query_dev;  // device-side query count (deferred passes), or nullptr
    uint32_t k;    // kNN: k; radius: max_knn
    float r2;      // radius: search_radius * search_radius (float)
    float margin;  // absolute distance slack for cell pruning, relative to h (times h_l)
    void* out_idx;        // (n_query, k) u32 (kNN) or i32 (radius), rows = original query index
    float* out_d2;        // (n_query, k) or nullptr
    float* out_feat;      // (n_query, 11) or nullptr (fused features)
    uint32_t* out_count;  // per original query: number of hits (radius) or nullptr
    uint64_t* gheap_keys;  // global-memory heap fallback: k * n_query keys; or, with heap_tail, heap nodes past 20
    // heap nodes past 20 in gheap_keys, as 16-byte pairs: 0 none; 1 one slot
    // per lane of the launch, pair p of lane t at (p * lanes + t) * 2; 2 one
    // slot per hardware wave slot (kHwWaveSlots), pair p of lane l of wave
    // slot w at ((w * tail_pairs + p) * 64 + l) * 2
    int heap_tail;
    uint32_t tail_pairs;   // heap_tail 2: pairs per lane
    uint32_t* defer_flag;  // budgeted pass: 1 where the query is left to the next pass
    const float* query_bound;  // per query: upper bound of its k-th squared distance, or +inf (nullable)
    float* defer_bound;        // budgeted pass: that bound for the deferred queries
    int near_first;        // visit cell rows nearest-first
    int staging;           // wave-shared LDS staging of cell rows: 0 never, 1 radius search, 2 also the budgeted kNN pass
    int budget_cells;      // budgeted pass: largest ball half-width, in cells around the query's cell
    int budget_shells;     // budgeted pass: cubic shells scanned before a query without k candidates is deferred
    const uint32_t* scales;  // kModeRadiusMulti: the k_scales (device), n_scales of them
    uint32_t n_scales;
};

// Simple linear device arena carved out of one allocation.  Every region
// starts on a 256-byte boundary of the address space, not only of the
// offset: regions hold 64-bit atomics and 16-byte vectors, and the base of an
// arena (a caller's workspace, or the rest of a buffer after outputs of odd
// length) may be aligned to 4 bytes only.  A misaligned base costs at most
// 255 bytes, inside the 4096 bytes of slack every workspace size adds.
struct Arena
{
    char* base = nullptr;
    size_t size = 0;
    size_t off = 0;
    template <typename T>
    T* take(size_t count)
    {
        const uintptr_t b = reinterpret_cast<uintptr_t>(base);
        const size_t a = (size_t)(((b + off + 255) & ~uintptr_t(255)) - b);
        off = a + count * sizeof(T);
        return reinterpret_cast<T*>(base + a);
    }
};

// ---- launchers (pgeof_search.hip) ----
// Sizes given as (n_max, n_dev): the kernels work on min(*n_dev + add, n_max)
// entries when n_dev is set (a count only the device knows), else on n_max.
hipError_t launch_bbox(const float* xyz, uint32_t n, float* partial, uint32_t n_partial, float* out6,
                       hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
hipError_t launch_fill_u32_dyn(uint32_t* p, uint64_t n_max, const uint64_t* n_dev, uint64_t add, uint32_t v,
                               hipStream_t s);
size_t scan_partials_needed(uint64_t n);
hipError_t launch_scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* partials,
                                 hipStream_t s);
hipError_t launch_scan_exclusive_dyn(const uint32_t* in, uint32_t* out, uint64_t n_max, const uint64_t* n_dev,
                                     uint64_t add, uint32_t* partials, hipStream_t s);
// Grid planning on the device: pgeof_plan_probe after the bounding box (into
// tab->bb), pgeof_plan_levels after the probe histogram's sum of squares.
hipError_t launch_plan_probe(const PlanArgs& p, LevelTable* tab, hipStream_t s);
hipError_t launch_plan_levels(const PlanArgs& p, LevelTable* tab, hipStream_t s);
// cell counts of every stride-th point at the probe grid, and their sum of squares
hipError_t launch_probe(const float* xyz, uint32_t n, uint32_t stride, LevelTable* tab, uint32_t* count,
                        uint64_t count_room, hipStream_t s);
// counting sort of points into level l (level 0 from xyz, l > 0 from the
// level-0 array, already sorted by a finer grid); nothing when the level is absent
hipError_t launch_hist(const float* xyz, uint32_t n, const LevelTable* tab, int l, uint32_t* count, uint32_t* cell,
                       uint32_t* rank, hipStream_t s);
hipError_t launch_hist_sorted(const float4* pts, uint32_t n, const LevelTable* tab, int l, uint32_t* count,
                              uint32_t* cell, uint32_t* rank, hipStream_t s);
hipError_t launch_scatter(const float* xyz, uint32_t n, const uint32_t* cell, const uint32_t* rank,
                          const uint32_t* cell_start, float4* pts, const LevelTable* tab, int l, hipStream_t s);
hipError_t launch_scatter_sorted(const float4* in, uint32_t n, const uint32_t* cell, const uint32_t* rank,
                                 const uint32_t* cell_start, float4* out, const LevelTable* tab, int l,
                                 hipStream_t s);
hipError_t launch_search(int mode, const SearchArgs& a, int block, size_t lds_bytes, bool global_heap, bool budget,
                         hipStream_t s);
// reorder queries (and their bounds) by bound, largest first; count needs
// 2049 entries, key and rank n_max
hipError_t launch_order_by_bound(const float4* q, const float* bound, uint32_t n_max, const uint64_t* n_dev,
                                 uint32_t* count, uint32_t* key, uint32_t* rank, uint32_t* scan_part, float4* q_out,
                                 float* bound_out, hipStream_t s);
// dst[0..n) = host[0..n), the values passed as kernel arguments (no host
// memory is read by the device: capturable into a graph)
hipError_t launch_put_values(uint32_t* dst, const uint32_t* host, uint64_t n, hipStream_t s);
// order[t] = original index of sorted point t
hipError_t launch_order_from_sorted(const float4* pts, uint32_t n, uint32_t* order, hipStream_t s);
// flag[n] = 0 for the device-side n (the exclusive scan's total slot)
hipError_t launch_put_u32(uint32_t* p, const uint64_t* n_dev, uint32_t v, hipStream_t s);
// compaction of the flagged queries; *out_count = off[n] (their number)
hipError_t launch_compact_queries(const float4* q, uint32_t n_max, const uint64_t* n_dev, const uint32_t* flag,
                                  const uint32_t* off, float4* out, const float* bound, float* bound_out,
                                  uint64_t* out_count, hipStream_t s);
hipError_t launch_arange_ptr(uint32_t* nn_ptr, uint32_t n_points, uint32_t k, hipStream_t s);
hipError_t launch_compact_rows(const uint32_t* dense, uint32_t n_rows, uint32_t cap, const uint32_t* counts,
                               const uint32_t* row_ptr, uint32_t* nn, hipStream_t s);

struct SelectedArgs
{
    GridView grid;       // grid over float32 copies of the points, self queries
    const void* xyz;     // original coordinates (float or double), (n, 3)
    const double* xs;    // double input: the coordinates in the grid's sorted order, (n, 3); float: nullptr
    uint32_t n;
    uint32_t cap;        // heap capacity: max(max_knn, 2)
    uint32_t max_knn;
    double r2;           // search_radius^2 evaluated in the input precision
    double reach;        // search_radius + margin, for the cell range
    double margin;       // slack of the float grid copies' rounding, added to any distance bound
    const int32_t* fids;  // EFeatureID list
    uint32_t n_sel;
    void* out;            // (n, n_sel) in the input precision
    void* gheap_keys;     // global heap fallback: cap * n keys
    uint32_t* gheap_idx;  // cap * n indices
};
hipError_t launch_selected(bool f64, const SelectedArgs& a, int block, size_t lds, bool global_heap, hipStream_t s);
hipError_t launch_f64_to_f32(const double* in, float* out, uint64_t n, hipStream_t s);
hipError_t launch_sorted_copy_f64(const float4* pts, const double* xyz, uint32_t n, double* xs, hipStream_t s);

enum SearchMode
{
    kModeKnn = 0,       // knn_search outputs
    kModeRadius = 1,    // radius_search outputs (dense, -1 padded)
    kModeKnnFeat = 2,   // fused: CSR row of neighbours + 11 features
    kModeRadiusCsr = 3,   // radius hits as u32 dense rows + counts (for CSR compaction)
    kModeRadiusMulti = 4  // kModeRadiusCsr + the multiscale features of every row (fused)
};

// ---- launchers (pgeof_features.hip) ----
struct CsrArgs
{
    const float* xyz;
    uint32_t n_xyz;
    const uint32_t* nn;
    uint64_t n_nn;
    const uint32_t* nn_ptr;
    uint32_t n_points;
    uint32_t k_min, k_step, k_min_search;
    const uint32_t* scales;
    uint32_t n_scales;
    float* out;
    uint32_t* err;
    const uint32_t* order;  // rows in this order (a permutation of 0..n_points-1), or nullptr
};
hipError_t launch_csr_features(int mode, const CsrArgs& a, hipStream_t s);
// spatial order of the rows of `a` (Morton code of each row's first neighbour)
// in a workspace of csr_order_bytes(n_points) bytes; *order receives the
// permutation (inside the workspace)
size_t csr_order_bytes(uint32_t n_points);
hipError_t launch_csr_order(const CsrArgs& a, void* ws, uint32_t* order[1], hipStream_t s);

// ---- static scratch of the no-allocation entry point (pgeof_noalloc.hip) ----
// Device address and size of this library's file-scope scratch array on the
// current device, looked up once per device.
hipError_t noalloc_scratch(void** ptr, size_t* bytes);

// ---- profiling hooks (pgeof_capi.cpp) ----
void prof_begin(const char* name, hipStream_t s, void** token);
void prof_end(void* token, hipStream_t s);

struct ProfScope
{
    void* tok = nullptr;
    hipStream_t s;
    ProfScope(const char* name, hipStream_t st) : s(st) { prof_begin(name, st, &tok); }
    ~ProfScope() { prof_end(tok, s); }
};

}  // namespace pgeof_amd
