// Uniform-grid neighbour search for gfx950.
//
// Replaces the nanoflann KD-tree of the reference (include/nn_search.hpp:31-132):
//   1. bounding box (block reduction)
//   2. counting sort of the points by grid cell: histogram with atomics (the
//      atomic's return value is the point's rank inside its cell), exclusive
//      scan of the cell counts, scatter into a sorted float4 array holding
//      (x, y, z, bit pattern of the original index)
//   3. one thread per query, queries taken in cell order so that the lanes of
//      a wavefront scan nearly the same candidate rows; the k best candidates
//      are kept in a max-heap in LDS keyed by (squared distance, original
//      index) packed into 64 bits, i.e. ties broken by the smaller index.
//
// Squared distances use the reference's float expression
// ((dx*dx + dy*dy) + dz*dz) with dx = query - point and no contraction, so
// the neighbour lists and distances are those of an exact search with that
// metric.
#include <type_traits>

#include "pgeof_device.hpp"
#include "pgeof_internal.hpp"

namespace pgeof_amd
{

namespace
{

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int cell_coord(double v, double lo, double inv_h, int g)
{
    double t = (v - lo) * inv_h;
    t = fmin(fmax(t, 0.0), (double)(g - 1));  // NaN -> 0
    return (int)t;
}

// ---------------------------------------------------------------- bbox ----
__global__ __launch_bounds__(256) void k_bbox(const float* __restrict__ xyz, uint32_t n, float* __restrict__ partial)
{
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    {
        for (int c = 0; c < 3; ++c)
        {
            const float v = xyz[i * 3 + c];
            if (isfinite(v))
            {
                mn[c] = fminf(mn[c], v);
                mx[c] = fmaxf(mx[c], v);
            }
        }
    }
    __shared__ float s[6][256];
    for (int c = 0; c < 3; ++c)
    {
        s[c][threadIdx.x] = mn[c];
        s[3 + c][threadIdx.x] = mx[c];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; ++c)
            {
                s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + w]);
                s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_bbox_final(const float* __restrict__ partial, uint32_t np,
                                                    float* __restrict__ out6)
{
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < np; b += 256)
        for (int c = 0; c < 3; ++c)
        {
            mn[c] = fminf(mn[c], partial[b * 6 + c]);
            mx[c] = fmaxf(mx[c], partial[b * 6 + 3 + c]);
        }
    __shared__ float s[6][256];
    for (int c = 0; c < 3; ++c)
    {
        s[c][threadIdx.x] = mn[c];
        s[3 + c][threadIdx.x] = mx[c];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; ++c)
            {
                s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + w]);
                s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out6[threadIdx.x] = s[threadIdx.x][0];
}

// ---------------------------------------------------------- fill / ptr ----
// number of entries a kernel works on: n_max, or min(*n_dev + add, n_max)
// when the count is only known on the device
__device__ __forceinline__ uint64_t dyn_count(uint64_t n_max, const uint64_t* n_dev, uint64_t add)
{
    if (!n_dev) return n_max;
    const uint64_t v = *n_dev + add;
    return v < n_max ? v : n_max;
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint64_t n_max,
                                                  const uint64_t* __restrict__ n_dev, uint64_t add, uint32_t v)
{
    const uint64_t n = dyn_count(n_max, n_dev, add);
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) p[i] = v;
}

__global__ void k_put_u32(uint32_t* __restrict__ p, const uint64_t* __restrict__ idx, uint32_t v)
{
    if (threadIdx.x == 0) p[*idx] = v;
}

__global__ __launch_bounds__(256) void k_arange_ptr(uint32_t* __restrict__ p, uint32_t n_points, uint32_t k)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i <= n_points; i += (uint64_t)gridDim.x * 256)
        p[i] = (uint32_t)(i * k);
}

// ---------------------------------------------------------- histogram ----
// Level l of the table (l < 0: the kNN sizing probe) and its number of cells
// (0 when the level was not built).
__device__ __forceinline__ Level level_of(const LevelTable* tab, int l) { return l < 0 ? tab->probe : tab->lv[l]; }
__device__ __forceinline__ uint64_t cells_of(const LevelTable* tab, int l) { return l < 0 ? tab->probe_G : tab->G[l]; }

__device__ __forceinline__ uint32_t cell_id(const Level& L, double x, double y, double z)
{
    const int cx = cell_coord(x, L.lo[0], L.inv_h, L.g[0]);
    const int cy = cell_coord(y, L.lo[1], L.inv_h, L.g[1]);
    const int cz = cell_coord(z, L.lo[2], L.inv_h, L.g[2]);
    return ((uint32_t)cz * (uint32_t)L.g[1] + (uint32_t)cy) * (uint32_t)L.g[0] + (uint32_t)cx;
}

// Counting-sort histogram, 16 points per thread; rank = the point's rank in
// its cell.  Grids of few cells count per block in LDS first (one global
// atomic per touched cell and block: a few hot cells would otherwise
// serialise the atomics), larger ones count in global memory directly.
constexpr uint32_t kSmallHistCells = 8192;
__global__ __launch_bounds__(256) void k_hist(const float* __restrict__ xyz, uint32_t n,
                                              const LevelTable* __restrict__ tab, int l,
                                              uint32_t* __restrict__ count, uint32_t* __restrict__ cell,
                                              uint32_t* __restrict__ rank)
{
    __shared__ uint32_t bins[kSmallHistCells];
    const uint64_t G = cells_of(tab, l);
    if (G == 0) return;
    const Level L = level_of(tab, l);
    const bool small = G <= kSmallHistCells;
    if (small)
    {
        for (uint32_t c = threadIdx.x; c < (uint32_t)G; c += 256) bins[c] = 0;
        __syncthreads();
    }
    constexpr uint32_t kPer = 16;  // points per thread
    const uint64_t base = (uint64_t)blockIdx.x * 256 * kPer;
    uint32_t cid[kPer], lr[kPer];
#pragma unroll
    for (uint32_t u = 0; u < kPer; ++u)
    {
        const uint64_t i = base + (uint64_t)u * 256 + threadIdx.x;
        cid[u] = 0xFFFFFFFFu;
        lr[u] = 0;
        if (i < n)
        {
            cid[u] = cell_id(L, (double)xyz[i * 3 + 0], (double)xyz[i * 3 + 1], (double)xyz[i * 3 + 2]);
            // two atomics, not one through a selected pointer: that pointer
            // would be generic (a flat atomic) instead of an LDS or global one
            lr[u] = small ? atomicAdd(&bins[cid[u]], 1u) : atomicAdd(&count[cid[u]], 1u);
        }
    }
    if (small)
    {
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < (uint32_t)G; c += 256)
        {
            const uint32_t v = bins[c];
            bins[c] = v ? atomicAdd(&count[c], v) : 0u;
        }
        __syncthreads();
    }
#pragma unroll
    for (uint32_t u = 0; u < kPer; ++u)
    {
        const uint64_t i = base + (uint64_t)u * 256 + threadIdx.x;
        if (i < n)
        {
            cell[i] = cid[u];
            rank[i] = small ? bins[cid[u]] + lr[u] : lr[u];
        }
    }
}

// Histogram of points that are already sorted by a finer grid (the level-0
// array): neighbouring points mostly share a cell, so each run of equal cells
// within a wave takes one atomic for the whole run.  rank = the run's base +
// the lane's offset in the run; cell/rank are indexed by sorted position.
__global__ __launch_bounds__(256) void k_hist_sorted(const float4* __restrict__ pts, uint32_t n,
                                                     const LevelTable* __restrict__ tab, int l,
                                                     uint32_t* __restrict__ count, uint32_t* __restrict__ cell,
                                                     uint32_t* __restrict__ rank)
{
    if (cells_of(tab, l) == 0) return;
    const Level L = level_of(tab, l);
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool active = i < n;
    uint32_t cid = 0xFFFFFFFFu;
    if (active)
    {
        const float4 p = pts[i];
        cid = cell_id(L, (double)p.x, (double)p.y, (double)p.z);
    }
    const uint32_t prev = __shfl_up(cid, 1);
    const bool head = active && (lane == 0 || prev != cid);
    const unsigned long long heads = __ballot(head);
    const unsigned long long act = __ballot(active);
    // run of this lane: [h, e)
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int h = 63 - __clzll(heads & upto);
    const unsigned long long above = heads & ~upto;
    const int n_act = __popcll(act);
    const int e = above ? min(__ffsll(above) - 1, n_act) : n_act;
    uint32_t base = 0;
    if (head) base = atomicAdd(&count[cid], (uint32_t)(e - h));
    base = __shfl(base, h);
    if (active)
    {
        cell[i] = cid;
        rank[i] = base + (uint32_t)(lane - h);
    }
}

__global__ __launch_bounds__(256) void k_scatter_sorted(const float4* __restrict__ in, uint32_t n,
                                                        const uint32_t* __restrict__ cell,
                                                        const uint32_t* __restrict__ rank,
                                                        const uint32_t* __restrict__ cell_start,
                                                        float4* __restrict__ out, const LevelTable* __restrict__ tab,
                                                        int l)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n || cells_of(tab, l) == 0) return;
    out[cell_start[cell[i]] + rank[i]] = in[i];
}

// Cell counts of every `stride`-th point at the probe grid (grid sizing estimate only).
__global__ __launch_bounds__(256) void k_hist_probe(const float* __restrict__ xyz, uint32_t n, uint32_t stride,
                                                    const LevelTable* __restrict__ tab, uint32_t* __restrict__ count)
{
    const uint64_t i = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * stride;
    if (i >= n || tab->probe_G == 0) return;
    const Level L = tab->probe;
    atomicAdd(&count[cell_id(L, (double)xyz[i * 3 + 0], (double)xyz[i * 3 + 1], (double)xyz[i * 3 + 2])], 1u);
}

__global__ __launch_bounds__(256) void k_sumsq(const uint32_t* __restrict__ count, LevelTable* __restrict__ tab)
{
    const uint64_t G = tab->probe_G;
    unsigned long long* out = &tab->sumsq;
    unsigned long long acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < G; i += (uint64_t)gridDim.x * 256)
    {
        const unsigned long long c = count[i];
        acc += c * c;
    }
    __shared__ unsigned long long s[256];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, s[0]);
}

// ------------------------------------------------------ grid planning ----
// The grid of a search is sized on the device from the bounding box (and,
// for kNN, a sampled occupancy probe), so that a search queues all its
// launches without reading anything back: one thread computes the plan.

__device__ uint64_t plan_cells_for(const double ext[3], double h, int g[3])
{
    uint64_t G = 1;
    for (int a = 0; a < 3; ++a)
    {
        double c = floor(ext[a] / h) + 1.0;
        if (!(c >= 1.0)) c = 1.0;
        if (c > 1048576.0) c = 1048576.0;
        g[a] = (int)c;
        G *= (uint64_t)g[a];
    }
    return G;
}

// smallest h (up to 0.1 %) with plan_cells_for(h) <= cap
__device__ double plan_h_for_cap(const double ext[3], uint64_t cap)
{
    const double emax = fmax(ext[0], fmax(ext[1], ext[2]));
    if (!(emax > 0.0)) return 1.0;
    int g[3];
    double hi = emax * 2.0, lo = emax * 1e-7;
    if (plan_cells_for(ext, lo, g) <= cap) return lo;
    for (int it = 0; it < 80; ++it)
    {
        const double mid = sqrt(lo * hi);
        if (plan_cells_for(ext, mid, g) <= cap)
            hi = mid;
        else
            lo = mid;
        if (hi / lo < 1.001) break;
    }
    return hi;
}

__device__ void plan_level(const float bb[6], double h, Level& L, uint64_t& G)
{
    double ext[3];
    for (int a = 0; a < 3; ++a)
    {
        L.lo[a] = (double)bb[a];
        ext[a] = fmax(0.0, (double)bb[3 + a] - (double)bb[a]);
    }
    L.h = h;
    L.inv_h = 1.0 / h;
    L.cs = nullptr;
    L.pts = nullptr;
    G = plan_cells_for(ext, h, L.g);
}

__device__ void plan_extent(const float bb[6], double ext[3], double& emax)
{
    for (int a = 0; a < 3; ++a) ext[a] = fmax(0.0, (double)bb[3 + a] - (double)bb[a]);
    emax = fmax(ext[0], fmax(ext[1], ext[2]));
}

// After the bounding box: the grid-size floor, and for kNN the probe grid of
// the first guess of h (points per cell T, assuming a uniformly filled box).
__global__ void k_plan_probe(PlanArgs p, LevelTable* __restrict__ tab)
{
    if (threadIdx.x != 0) return;
    float bb[6];
    bool finite = true;
    for (int i = 0; i < 6; ++i) bb[i] = tab->bb[i];
    for (int i = 0; i < 3; ++i)
        if (!(bb[i] <= bb[3 + i])) finite = false;  // no finite point
    if (!finite)
        for (int i = 0; i < 6; ++i) bb[i] = 0.0f;
    for (int i = 0; i < 6; ++i) tab->bb[i] = bb[i];
    double ext[3], emax;
    plan_extent(bb, ext, emax);
    const double hcap = plan_h_for_cap(ext, p.cap);
    tab->hcap = hcap;
    tab->sumsq = 0;
    tab->probe_G = 0;
    tab->n_levels = 0;
    for (int l = 0; l < kMaxLevels; ++l) tab->G[l] = 0;
    if (!p.is_radius && emax > 0.0)
    {
        const double T = fmax(0.75, fmin(p.knn_occupancy * (double)p.k, p.knn_cell_points_max));
        double vol = 1.0;
        for (int a = 0; a < 3; ++a) vol *= fmax(ext[a], emax * 1e-3);
        double h1 = cbrt(vol * T / fmax((double)p.n, 1.0));
        h1 = fmax(h1, hcap);
        Level L;
        uint64_t G;
        plan_level(bb, h1, L, G);
        tab->probe = L;
        tab->probe_G = G;
        tab->T = T;
        tab->h1 = h1;
    }
}

// After the probe: level 0's cell edge (kNN: the first guess corrected by the
// measured point-weighted occupancy, assuming points on 2-D surfaces; radius:
// radius_cell_factor * r), then the pyramid levels that fit their tables.
__global__ void k_plan_levels(PlanArgs p, LevelTable* __restrict__ tab)
{
    if (threadIdx.x != 0) return;
    float bb[6];
    for (int i = 0; i < 6; ++i) bb[i] = tab->bb[i];
    double ext[3], emax;
    plan_extent(bb, ext, emax);
    const double hcap = tab->hcap;
    double h;
    if (!(emax > 0.0))
        h = 1.0;
    else if (p.is_radius)
    {
        const double r = fabs((double)p.radius);
        h = isfinite(r) && r > 0.0 ? r * p.radius_cell_factor : emax;
        h = fmax(h, hcap);
    }
    else
    {
        // sampling fraction pf: E[sum c_s^2] = pf^2 sum c^2 + pf (1 - pf) n
        const uint32_t stride = p.probe_stride;
        const double ns = (double)((p.n + stride - 1) / stride), pf = 1.0 / stride;
        const double T = tab->T, h1 = tab->h1;
        const double m1 = fmax(1.0, ((double)tab->sumsq - (1.0 - pf) * ns) / (pf * fmax(ns, 1.0)));
        h = h1;
        if (m1 > T) h = h1 * sqrt(T / m1);
        h = fmax(h, hcap);
    }
    const double ratio = fmax(2.0, fmin(8.0, p.level_ratio));
    int nl = 0;
    for (int l = 0; l < p.max_levels; ++l)
    {
        Level L;
        uint64_t G;
        plan_level(bb, h * pow(ratio, (double)l), L, G);
        if (l == 0 && G + 1 > p.cs_room[0]) plan_level(bb, hcap, L, G);  // (h >= hcap keeps G within the cap)
        if (l > 0 && (G + 1 > p.cs_room[l] || tab->G[l - 1] <= 64)) break;
        L.cs = p.cs[l];
        L.pts = p.pts[l];
        tab->lv[l] = L;
        tab->G[l] = G;
        nl = l + 1;
    }
    for (int l = nl; l < kMaxLevels; ++l)
    {
        tab->lv[l] = tab->lv[0];
        tab->G[l] = 0;
    }
    tab->n_levels = nl;
}

// --------------------------------------------------------------- scan ----
// Block-wide exclusive scan of one u32 per thread (256 threads, 4 waves).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total)
{
    __shared__ uint32_t wsum[kScanBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
    for (int d = 1; d < 64; d <<= 1)
    {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t wprefix = 0, tot = 0;
    for (int w = 0; w < kScanBlock / 64; ++w)
    {
        const uint32_t s = wsum[w];
        if (w < wave) wprefix += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wprefix + inc - v;
}

// Tiles of kScanTile entries, read as 16-byte vectors: one pass of a block
// covers 4 * kScanBlock consecutive entries, four per thread, so every
// wave-wide load is contiguous.  VEC requires 16-byte aligned arrays.
template <bool VEC>
__device__ __forceinline__ uint4 scan_load4(const uint32_t* in, uint64_t i, uint64_t n)
{
    if (VEC && i + 3 < n) return reinterpret_cast<const uint4*>(in)[i >> 2];
    uint4 v;
    v.x = i < n ? in[i] : 0u;
    v.y = i + 1 < n ? in[i + 1] : 0u;
    v.z = i + 2 < n ? in[i + 2] : 0u;
    v.w = i + 3 < n ? in[i + 3] : 0u;
    return v;
}

// The three scan kernels loop over the tiles (a block takes tiles b, b +
// grid, ...), so that a launch sized for the largest input does no empty
// work when the device-side count is smaller.
template <bool VEC>
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* __restrict__ in, uint64_t n_max,
                                                            const uint64_t* __restrict__ n_dev, uint64_t add,
                                                            uint32_t* __restrict__ partials)
{
    const uint64_t n = dyn_count(n_max, n_dev, add);
    const uint64_t nt = (n + kScanTile - 1) / kScanTile;
    for (uint64_t tile = blockIdx.x; tile < nt; tile += gridDim.x)
    {
        const uint64_t base = tile * kScanTile;
        uint32_t s = 0;
        for (int sub = 0; sub < kScanTile; sub += 4 * kScanBlock)
        {
            const uint4 v = scan_load4<VEC>(in, base + sub + 4 * threadIdx.x, n);
            s += v.x + v.y + v.z + v.w;
        }
        uint32_t total;
        block_exclusive_scan(s, total);
        if (threadIdx.x == 0) partials[tile] = total;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_partials(uint32_t* __restrict__ partials, uint64_t n_max,
                                                              const uint64_t* __restrict__ n_dev, uint64_t add)
{
    const uint32_t nt = (uint32_t)((dyn_count(n_max, n_dev, add) + kScanTile - 1) / kScanTile);
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nt; base += kScanTile)
    {
        uint32_t v[kScanItems];
        uint32_t loc = 0;
        for (int j = 0; j < kScanItems; ++j)
        {
            const uint32_t i = base + threadIdx.x * kScanItems + j;
            v[j] = i < nt ? partials[i] : 0u;
            loc += v[j];
        }
        uint32_t total;
        uint32_t ex = block_exclusive_scan(loc, total) + carry;
        for (int j = 0; j < kScanItems; ++j)
        {
            const uint32_t i = base + threadIdx.x * kScanItems + j;
            if (i < nt) partials[i] = ex;
            ex += v[j];
        }
        carry += total;
    }
}

template <bool VEC>
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const uint32_t* in, uint32_t* out, uint64_t n_max,
                                                           const uint64_t* __restrict__ n_dev, uint64_t add,
                                                           const uint32_t* __restrict__ partials)
{
    const uint64_t n = dyn_count(n_max, n_dev, add);
    const uint64_t nt = (n + kScanTile - 1) / kScanTile;
    for (uint64_t tile = blockIdx.x; tile < nt; tile += gridDim.x)
    {
        const uint64_t base = tile * kScanTile;
        uint32_t carry = partials[tile];
        for (int sub = 0; sub < kScanTile; sub += 4 * kScanBlock)
        {
            const uint64_t i = base + sub + 4 * threadIdx.x;
            const uint4 v = scan_load4<VEC>(in, i, n);  // read before the block scan syncs: in may alias out
            uint32_t total;
            const uint32_t ex = block_exclusive_scan(v.x + v.y + v.z + v.w, total) + carry;
            const uint4 o = make_uint4(ex, ex + v.x, ex + v.x + v.y, ex + v.x + v.y + v.z);
            if (VEC && i + 3 < n)
                reinterpret_cast<uint4*>(out)[i >> 2] = o;
            else
            {
                if (i < n) out[i] = o.x;
                if (i + 1 < n) out[i + 1] = o.y;
                if (i + 2 < n) out[i + 2] = o.z;
                if (i + 3 < n) out[i + 3] = o.w;
            }
            carry += total;
        }
    }
}

// ------------------------------------------------------------ scatter ----
__global__ __launch_bounds__(256) void k_scatter(const float* __restrict__ xyz, uint32_t n,
                                                 const uint32_t* __restrict__ cell, const uint32_t* __restrict__ rank,
                                                 const uint32_t* __restrict__ cell_start, float4* __restrict__ pts,
                                                 const LevelTable* __restrict__ tab, int l)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n || cells_of(tab, l) == 0) return;
    const uint32_t pos = cell_start[cell[i]] + rank[i];
    pts[pos] = make_float4(xyz[i * 3 + 0], xyz[i * 3 + 1], xyz[i * 3 + 2], __uint_as_float((uint32_t)i));
}

__global__ __launch_bounds__(256) void k_compact_rows(const uint32_t* __restrict__ dense, uint32_t n_rows,
                                                      uint32_t cap, const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ row_ptr,
                                                      uint32_t* __restrict__ nn)
{
    // 16 lanes per row (four rows per wave), lanes along the row: only the
    // filled entries are read, both sides coalesced, no 64-bit division per
    // entry; a lane issues up to four loads before its stores, and the four
    // rows of a wave wait for their counts and offsets together
    const uint32_t sub = threadIdx.x & 15u;
    const uint32_t grp = (blockIdx.x * 256u + threadIdx.x) >> 4, n_grp = gridDim.x * 16u;
    for (uint32_t r = grp; r < n_rows; r += n_grp)
    {
        const uint32_t c = min(counts[r], cap), base = row_ptr[r];
        const uint32_t* __restrict__ src = dense + (size_t)r * cap;
        uint32_t j = sub;
        for (; j + 48u < c; j += 64u)
        {
            const uint32_t v0 = src[j], v1 = src[j + 16u], v2 = src[j + 32u], v3 = src[j + 48u];
            nn[base + j] = v0;
            nn[base + j + 16u] = v1;
            nn[base + j + 32u] = v2;
            nn[base + j + 48u] = v3;
        }
        for (; j < c; j += 16u) nn[base + j] = src[j];
    }
}

// ------------------------------------------------- global-memory access ----
// Pointers that reach a kernel through a table in device memory (the levels
// of the pyramid, moved to scalar registers) lose their address space, and
// the compiler then reads them with flat loads.  A flat load also counts
// against the LDS wait counter, so every LDS access of the heaps would wait
// for the candidate loads in flight; these helpers read and write through
// global-address-space pointers instead (global_load / global_store).
#define PGEOF_GLOBAL __attribute__((address_space(1)))
typedef float pgeof_f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long pgeof_u64x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float4 gload(const float4* p, size_t i)
{
    const pgeof_f32x4 v = ((const PGEOF_GLOBAL pgeof_f32x4*)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t gload(const uint32_t* p, size_t i)
{
    return ((const PGEOF_GLOBAL uint32_t*)p)[i];
}
__device__ __forceinline__ uint64_t gload(const uint64_t* p, size_t i)
{
    return ((const PGEOF_GLOBAL unsigned long long*)p)[i];
}
// two consecutive 8-byte entries at p (16-byte aligned)
__device__ __forceinline__ ulonglong2 gload2(const uint64_t* p)
{
    const pgeof_u64x2 v = *(const PGEOF_GLOBAL pgeof_u64x2*)p;
    return make_ulonglong2(v.x, v.y);
}
__device__ __forceinline__ void gstore(uint64_t* p, size_t i, uint64_t v)
{
    ((PGEOF_GLOBAL unsigned long long*)p)[i] = v;
}

// --------------------------------------------------------------- heap ----
// Keys are 64-bit: squared-distance bits << 32 | original index, so that the
// order of keys is the order of (distance, index) and ties go to the smaller
// index.

// Binary max-heap in global memory (large k); element j at keys[j * stride].
struct Heap
{
    uint64_t* keys;
    uint32_t stride;
    uint32_t k;
    uint32_t count;  // entries held; the heap property holds once count == k

    __device__ __forceinline__ uint64_t key_at(uint32_t i) const { return keys[(size_t)i * stride]; }

    // put `key` at index i of heap [0, n) and sift it down
    __device__ __forceinline__ void sift_from(uint32_t i, uint64_t key, uint32_t n)
    {
        for (;;)
        {
            const uint32_t l = 2 * i + 1;
            if (l >= n) break;
            uint64_t ck = keys[(size_t)l * stride];
            uint32_t c = l;
            if (l + 1 < n)
            {
                const uint64_t rk = keys[(size_t)(l + 1) * stride];
                if (rk > ck)
                {
                    ck = rk;
                    c = l + 1;
                }
            }
            if (ck <= key) break;
            keys[(size_t)i * stride] = ck;
            i = c;
        }
        keys[(size_t)i * stride] = key;
    }

    __device__ __forceinline__ void heapify()
    {
        for (int i = (int)k / 2 - 1; i >= 0; --i) sift_from((uint32_t)i, keys[(size_t)i * stride], k);
    }

    // precondition: key < worst (worst = sentinel until the heap is full)
    __device__ __forceinline__ void add(uint64_t key, uint64_t& worst)
    {
        if (count < k)
        {
            keys[(size_t)count * stride] = key;
            if (++count == k)
            {
                heapify();
                worst = keys[0];
            }
        }
        else
        {
            sift_from(0, key, k);
            worst = keys[0];
        }
    }

    // pad with the sentinel, restore the heap property, sort ascending
    template <bool SHORT = false>
    __device__ __forceinline__ void finish(uint64_t sentinel)
    {
        if (count < k)
        {
            for (uint32_t j = count; j < k; ++j) keys[(size_t)j * stride] = sentinel;
            heapify();
        }
        for (uint32_t end = k - 1; end > 0; --end)
        {
            const uint64_t top_k = keys[0];
            const uint64_t last = keys[(size_t)end * stride];
            keys[(size_t)end * stride] = top_k;
            sift_from(0, last, end);
        }
    }
};

// 4-ary max-heap of k keys for one lane.  Node 0 (the largest key) and its
// four children, nodes 1..4, are held in registers; nodes 5..k-1 live in
// LDS.  The four children 4i+1..4i+4 of node i >= 1 form LDS group i - 1,
// stored as two 16-byte pairs (children 0-1 and 2-3).  Pair p of every lane
// of the block is contiguous, 16 bytes per lane, so a wave reads one pair with
// conflict-free 16-byte loads.  A sift from the root of a 20-entry heap reads
// four registers and one LDS group; the binary-heap equivalent walks four
// levels of LDS.  With TAIL (large k) only nodes 5..20 (groups 0..3, 128
// bytes per lane) are in LDS and nodes 21..k-1 in global memory as
// lane-interleaved 16-byte pairs, in the slot of the lane's hardware wave
// slot or of its lane of the launch: LDS then stays small enough for full
// occupancy whatever k is, and only sifts that reach the fourth level touch
// global memory.
template <bool TAIL>
struct Heap4
{
    uint64_t* keys;    // this lane's pair 0; pair p at keys + p * pstride
    uint32_t pstride;  // entries between consecutive pairs of one lane (2 * block size)
    uint32_t k;
    uint32_t count;    // keys inserted, up to k (after finish(): entries held, key_at past them = sent)
    uint64_t sent;     // the sentinel
    uint64_t* gkeys;   // this lane's first global pair; pair p at gkeys + p * gpstride
    size_t gpstride;   // entries between consecutive global pairs of one lane
    uint64_t rkey;     // node 0
    uint64_t n1, n2, n3, n4;  // nodes 1..4 (separate scalars, never an indexed array)

    static constexpr uint32_t kLdsTop = 20;  // with a tail: LDS holds nodes 5..20 (groups 0..3)

    __device__ __forceinline__ static bool in_lds(uint32_t i) { return !TAIL || i <= kLdsTop; }
    __device__ __forceinline__ uint32_t slot(uint32_t i) const  // LDS node
    {
        const uint32_t s = i - 5;
        return (s >> 1) * pstride + (s & 1);
    }
    // global node; 64-bit: (k - 21) / 2 * gpstride passes 2^32 for large
    // searches (10 M queries at k >= 452)
    __device__ __forceinline__ size_t gslot(uint32_t i) const
    {
        const uint32_t s = i - (kLdsTop + 1);
        return (size_t)(s >> 1) * gpstride + (s & 1);
    }
    __device__ __forceinline__ uint64_t load(uint32_t i) const  // i >= 5
    {
        return in_lds(i) ? keys[slot(i)] : gload(gkeys, gslot(i));
    }
    __device__ __forceinline__ void store(uint32_t i, uint64_t kk)  // i >= 5
    {
        if (in_lds(i))
            keys[slot(i)] = kk;
        else
            gstore(gkeys, gslot(i), kk);
    }
    // The register nodes are combined with bit masks, never indexed or
    // stored in a branch: a select chain over consecutive members would be
    // turned back into an indexed access, which lives in scratch memory.
    __device__ __forceinline__ static uint64_t mask(bool b) { return 0ull - (uint64_t)b; }
    __device__ __forceinline__ uint64_t reg(uint32_t i) const  // i in 0..4
    {
        return (rkey & mask(i == 0)) | (n1 & mask(i == 1)) | (n2 & mask(i == 2)) | (n3 & mask(i == 3)) |
               (n4 & mask(i == 4));
    }
    __device__ __forceinline__ void set_reg(uint32_t i, uint64_t kk)  // i in 0..4
    {
        const uint64_t m0 = mask(i == 0), m1 = mask(i == 1), m2 = mask(i == 2), m3 = mask(i == 3), m4 = mask(i == 4);
        rkey = (kk & m0) | (rkey & ~m0);
        n1 = (kk & m1) | (n1 & ~m1);
        n2 = (kk & m2) | (n2 & ~m2);
        n3 = (kk & m3) | (n3 & ~m3);
        n4 = (kk & m4) | (n4 & ~m4);
    }
    __device__ __forceinline__ uint64_t node(uint32_t i) const { return i <= 4 ? reg(i) : load(i); }
    // entry i of the sorted heap (after finish)
    __device__ __forceinline__ uint64_t key_at(uint32_t i) const { return i < count ? node(i) : sent; }
    __device__ __forceinline__ void put(uint32_t i, uint64_t kk)
    {
        if (i <= 4)
            set_reg(i, kk);
        else
            store(i, kk);
    }

    // largest of the children 4i+1..4i+4 (< n) of node i >= 1: group i - 1,
    // in LDS (or, with a tail, in global memory past group 3)
    __device__ __forceinline__ uint64_t max_child(uint32_t i, uint32_t n, uint32_t& c) const
    {
        const uint32_t c0 = 4 * i + 1, g = i - 1;
        ulonglong2 g01, g23;
        if (in_lds(c0))
        {
            g01 = *reinterpret_cast<const ulonglong2*>(keys + (2 * g) * pstride);
            g23 = *reinterpret_cast<const ulonglong2*>(keys + (2 * g + 1) * pstride);
        }
        else
        {
            const uint32_t gg = g - (kLdsTop - 4) / 4;
            g01 = gload2(gkeys + (size_t)(2 * gg) * gpstride);
            g23 = gload2(gkeys + (size_t)(2 * gg + 1) * gpstride);
        }
        uint64_t mk = g01.x;
        uint32_t mj = 0;
        if (c0 + 1 < n && g01.y > mk) mk = g01.y, mj = 1;
        if (c0 + 2 < n && g23.x > mk) mk = g23.x, mj = 2;
        if (c0 + 3 < n && g23.y > mk) mk = g23.y, mj = 3;
        c = c0 + mj;
        return mk;
    }

    // put `key` at node i of the heap of nodes [0, n) and sift it down:
    // node 0, then one of nodes 1..4 (registers), then memory nodes only
    __device__ __forceinline__ void sift_from(uint32_t i, uint64_t key, uint32_t n)
    {
        if (i == 0)
        {
            uint64_t mk = n > 1 ? n1 : 0;
            uint32_t mj = 0;
            if (2 < n && n2 > mk) mk = n2, mj = 1;
            if (3 < n && n3 > mk) mk = n3, mj = 2;
            if (4 < n && n4 > mk) mk = n4, mj = 3;
            const bool up = n > 1 && mk > key;
            rkey = up ? mk : key;
            if (!up) return;
            i = 1 + mj;
        }
        if (i <= 4)
        {
            uint32_t c = 0;
            const uint64_t mk = 4 * i + 1 < n ? max_child(i, n, c) : 0;
            const bool down = 4 * i + 1 < n && mk > key;
            set_reg(i, down ? mk : key);
            if (!down) return;
            i = c;
        }
        for (;;)
        {
            if (4 * i + 1 >= n) break;
            uint32_t c;
            const uint64_t mk = max_child(i, n, c);
            if (mk <= key) break;
            store(i, mk);
            i = c;
        }
        store(i, key);
    }

    // The heap starts full of the sentinel (larger than or equal to every
    // key that can be accepted), so it is valid from the start.  The first k
    // keys replace the sentinels from the last node upwards, each sifted down
    // from its own node (a bottom-up heap construction spread over the
    // insertions: most of them land on leaves and stop at once); the root
    // stays a sentinel until the k-th key.  Later keys replace the root.
    // Both cases are one sift from a computed node, with no per-lane switch
    // between code paths.  Only the register nodes are set here: a sift from
    // node s reads only nodes below s, which the fill has already written, so
    // the memory nodes the fill has not reached are never read before
    // finish(), which writes the sentinel into them (a full heap, the common
    // case, never stores them at all).
    __device__ __forceinline__ void init(uint64_t sentinel)
    {
        count = 0;
        sent = sentinel;
        rkey = n1 = n2 = n3 = n4 = sentinel;
    }

    // precondition: key < worst, worst = the root (the sentinel until k keys
    // are in)
    __device__ __forceinline__ void add(uint64_t key, uint64_t& worst)
    {
        const bool filling = count < k;
        sift_from(filling ? k - 1 - count : 0u, key, k);
        count += filling ? 1u : 0u;
        worst = count >= k ? rkey : worst;
    }

    // sort ascending (entries never filled keep the sentinel)
    // SHORT: sort rows of few keys in registers (radius searches)
    static constexpr uint32_t kShortFinish = 8;
    template <bool SHORT = false>
    __device__ __forceinline__ void finish(uint64_t sentinel)
    {
        if (SHORT && count < k && count <= kShortFinish)
        {
            // Few keys (a radius search's sparse rows): they sit in the fill's
            // nodes k-count..k-1; sort them in registers (a 19-comparator
            // network for 8) into nodes 0..count-1, and key_at() answers the
            // sentinel past them, instead of a heap sort of all k nodes that
            // sifts sentinels through LDS and global memory.
            uint64_t v[kShortFinish];
#pragma unroll
            for (uint32_t u = 0; u < kShortFinish; ++u) v[u] = u < count ? node(k - count + u) : ~0ull;
            constexpr int kNet[19][2] = {{0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6},
                                         {3, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}, {2, 4}, {3, 5},
                                         {1, 4}, {3, 6}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
            for (int c = 0; c < 19; ++c)
            {
                const uint64_t x = v[kNet[c][0]], y = v[kNet[c][1]];
                v[kNet[c][0]] = x < y ? x : y;
                v[kNet[c][1]] = x < y ? y : x;
            }
#pragma unroll
            for (uint32_t u = 0; u < kShortFinish; ++u)
                if (u < count) put(u, v[u]);
            return;
        }
        // memory nodes 5..k-1-count were never reached by the fill
        for (uint32_t i = 5; i + count < k; ++i) store(i, sentinel);
        uint32_t end = k - 1;
        for (; end > 4; --end)
        {
            const uint64_t last = load(end);
            store(end, rkey);
            sift_from(0, last, end);
        }
        for (; end > 0; --end)
        {
            const uint64_t last = reg(end);
            set_reg(end, rkey);
            sift_from(0, last, end);
        }
        count = k;
    }
};

// The hardware wave slot this wave occupies, < kHwWaveSlots: XCC (XCC_ID
// register, bits 3:0, at most kHwMaxXcc of them), shader engine (HW_ID bits
// 14:13), shader array (12), CU (11:8), SIMD (5:4) and wave buffer slot
// (3:0).  Two waves resident at the same time have different slots, so a
// slot's memory is private to the wave that holds it until the wave ends.
__device__ __forceinline__ uint32_t hw_wave_slot()
{
    constexpr int kHwIdReg = 4, kXccIdReg = 20;
    const uint32_t hw = __builtin_amdgcn_s_getreg(((15 - 1) << 11) | (0 << 6) | kHwIdReg);
    const uint32_t xcc = __builtin_amdgcn_s_getreg(((4 - 1) << 11) | (0 << 6) | kXccIdReg);
    const uint32_t wave = hw & 15u, simd = (hw >> 4) & 3u, cu = (hw >> 8) & 15u, sh = (hw >> 12) & 1u,
                   se = (hw >> 13) & 3u;
    return (((((xcc & (kHwMaxXcc - 1)) * 4u + se) * 2u + sh) * 16u + cu) * 4u + simd) * 16u + wave;
}

// A query's row of n 32-bit entries in oi (and, when od is set, the row at
// the same offset of od), 16 bytes per store: the entries up to oi's next
// 16-byte boundary one by one, then four per store, then the rest.  One
// lane writes one row and the rows of a wave's lanes are far apart, so a
// store of the wave touches as many cache lines as it has lanes: writing
// four entries per store cuts those requests fourfold.  (od not aligned
// like oi: every entry on its own.)  emit(j, i, d) gives entry j.
template <class Emit>
__device__ __forceinline__ void store_rows(uint32_t* __restrict__ oi, float* __restrict__ od, uint32_t n, Emit emit)
{
    const uintptr_t ai = reinterpret_cast<uintptr_t>(oi);
    const bool same = od == nullptr || ((ai ^ reinterpret_cast<uintptr_t>(od)) & 15u) == 0;
    const uint32_t head = same ? min((uint32_t)((16u - (ai & 15u)) & 15u) >> 2, n) : n;
    uint32_t j = 0;
    for (; j < head; ++j)
    {
        uint32_t i;
        float d;
        emit(j, i, d);
        oi[j] = i;
        if (od) od[j] = d;
    }
    for (; j + 4 <= n; j += 4)
    {
        uint32_t i0, i1, i2, i3;
        float d0, d1, d2, d3;
        emit(j, i0, d0);
        emit(j + 1, i1, d1);
        emit(j + 2, i2, d2);
        emit(j + 3, i3, d3);
        *reinterpret_cast<uint4*>(oi + j) = make_uint4(i0, i1, i2, i3);
        if (od) *reinterpret_cast<float4*>(od + j) = make_float4(d0, d1, d2, d3);
    }
    for (; j < n; ++j)
    {
        uint32_t i;
        float d;
        emit(j, i, d);
        oi[j] = i;
        if (od) od[j] = d;
    }
}

// distance from v to the slab [lo + c h, lo + (c+1) h], less a safety margin
__device__ __forceinline__ double axis_gap(double v, int c, double lo, double h, double margin)
{
    const double a = lo + (double)c * h, b = a + h;
    const double g = fmax(fmax(a - v, v - b), 0.0) - margin;
    return g > 0.0 ? g : 0.0;
}

// ------------------------------------------------------------- search ----
// One thread per query, queries in cell-sorted order.  Each query first
// picks the finest grid level whose 3x3x3 cell cube around it already holds
// k points (kNN; radius searches use level 0), so its row count stays small
// whatever the local density.  Rows are then scanned in one of three ways,
// chosen per wavefront and pass:
//   - cooperative (with the staging knob: radius searches, and the budgeted
//     kNN pass for k >= 32 or always): when all
//     lanes of the wave use the same level and the union of their rows is
//     compact, the wave walks that union in lockstep, stages each row's
//     point slice into LDS with coalesced loads and every lane scans its own
//     sub-range from LDS;
//   - per lane in step (budgeted kNN pass, radius searches without
//     staging): each lane walks its own rows,
//     the wave opening rows r..r+2 of a plane of every lane together and
//     each lane scanning their pieces as one stream of candidates;
//   - per lane with a cursor (unbudgeted pass): each lane opens its next row
//     as soon as its current one ends.
constexpr int kStage = 128;  // float4 points staged per wave per chunk

// Wavefront reductions over the active lanes (ROCm device library, DPP based).
extern "C" __device__ uint32_t __ockl_wfred_min_u32(uint32_t);
extern "C" __device__ uint32_t __ockl_wfred_max_u32(uint32_t);
extern "C" __device__ int32_t __ockl_wfred_min_i32(int32_t);
extern "C" __device__ int32_t __ockl_wfred_max_i32(int32_t);

__device__ __forceinline__ int wave_min_i(int v) { return __ockl_wfred_min_i32(v); }
__device__ __forceinline__ int wave_max_i(int v) { return __ockl_wfred_max_i32(v); }
__device__ __forceinline__ uint32_t wave_min_u(uint32_t v) { return __ockl_wfred_min_u32(v); }
__device__ __forceinline__ uint32_t wave_max_u(uint32_t v) { return __ockl_wfred_max_u32(v); }

// LDS visibility between the lanes of one wavefront
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct Box
{
    int x0, x1, y0, y1, z0, z1;  // inclusive cell ranges; empty when x0 > x1
};

__device__ __forceinline__ Box empty_box() { return Box{1, 0, 1, 0, 1, 0}; }

__device__ __forceinline__ Box cube_box(int cx, int cy, int cz, int n, const Level& G)
{
    return Box{max(cx - n, 0), min(cx + n, G.g[0] - 1), max(cy - n, 0), min(cy + n, G.g[1] - 1), max(cz - n, 0),
               min(cz + n, G.g[2] - 1)};
}

// i-th element of the order c, c-1, c+1, c-2, c+2, ... when the range [lo, hi]
// starts at c (values outside [lo, hi] are skipped by the caller)
__device__ __forceinline__ int near_index(int i, int c, int lo, int hi)
{
    if (c <= lo) return lo + i;  // plain ascending order
    c = min(c, hi);
    const int d = (i + 1) >> 1;
    return (i & 1) ? c - d : c + d;
}

// t-th element (t in [0, hi - lo]) of the order c, c-1, c+1, c-2, c+2, ...
// over [lo, hi] with c in [lo, hi]: alternating while both sides last, then
// the longer side outwards
__device__ __forceinline__ int nth_near(int t, int c, int lo, int hi)
{
    const int below = c - lo, above = hi - c, m = min(below, above);
    if (t <= 2 * m)
    {
        const int d = (t + 1) >> 1;
        return (t & 1) ? c - d : c + d;
    }
    const int r = t - m;
    return below > above ? c - r : c + r;
}

// Level l of the pyramid, read from the device-side table (a per-lane level
// is a vector load into vector registers).
__device__ __forceinline__ Level pick_level(const SearchArgs& a, int l) { return a.tab->lv[l]; }

// A level every lane of the wave uses, moved to scalar registers: the table
// is written by other kernels, so the compiler cannot load it with scalar
// loads by itself, and the level's 18 dwords would otherwise occupy vector
// registers for the whole search.
__device__ __forceinline__ Level uniform_level(const LevelTable* tab, int l)
{
    static_assert(sizeof(Level) % 4 == 0, "Level is moved as dwords");
    union
    {
        Level L;
        uint32_t u[sizeof(Level) / 4];
    } b;
    b.L = tab->lv[l];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(Level) / 4); ++i) b.u[i] = __builtin_amdgcn_readfirstlane(b.u[i]);
    return b.L;
}

#ifndef PGEOF_SCAN_UNROLL
#define PGEOF_SCAN_UNROLL 4
#endif
constexpr uint32_t kScanUnroll = PGEOF_SCAN_UNROLL;  // candidates loaded together per lane
// unbudgeted pass (per-lane cursor): candidates loaded together per lane
// (10 M points, k = 20: 4 -> 3.5 ms, 6 -> 3.35, 8 -> 3.35 with three times
// the spill of 6; k = 50: 8.4, 7.9-8.2, 7.8-8.0)
constexpr uint32_t kCursorUnroll = 6;
// budgeted kNN pass: rows of a plane scanned as one stream per lane (3 rows:
// 6.3 -> 5.8 ms per 10 M queries at k = 20; 2 rows 6.0, 4 rows 5.9)
constexpr int kRowsPerStream = 3;

// where a search kernel keeps its heaps
constexpr int kHeapLds = 0;     // 4-ary, registers + LDS
constexpr int kHeapGlobal = 1;  // binary, global memory (fallback)
constexpr int kHeapTail = 2;    // 4-ary, registers + LDS + global tail past node 20

// The search of query slot t (sorted position) by one lane; every lane of the
// wave calls it together (inactive lanes take part in the wave-wide steps).
template <int MODE, int WAVES, int HEAP, bool BUDGET>
__device__ __forceinline__ void search_slot(const SearchArgs& a, const uint32_t t, const bool active,
                                            unsigned char* smem, const uint32_t tid)
{
    constexpr bool RADIUS = (MODE == kModeRadius || MODE == kModeRadiusCsr || MODE == kModeRadiusMulti);
    constexpr int BLOCK = WAVES * 64;

    const int lane = tid & 63, wid = tid >> 6;
    const int n_levels = __builtin_amdgcn_readfirstlane(a.tab->n_levels);
    float4* stage = reinterpret_cast<float4*>(smem) + wid * kStage;
    unsigned char* hbase = smem + (size_t)WAVES * kStage * sizeof(float4);

    // heaps: 4-ary with nodes 0..4 in registers and ceil((k - 5) / 4) groups
    // of four keys per lane in LDS; global-memory fallback for large k: binary
    constexpr bool GLOBAL_HEAP = HEAP == kHeapGlobal;
    using HeapT = typename std::conditional<GLOBAL_HEAP, Heap, Heap4<HEAP == kHeapTail>>::type;
    HeapT heap;
    if constexpr (GLOBAL_HEAP)
    {
        // inactive lanes never touch their (non-existent) global heap
        heap.keys = a.gheap_keys + (active ? t : 0);
        heap.stride = a.n_query;
    }
    else
    {
        heap.keys = reinterpret_cast<uint64_t*>(hbase) + (size_t)tid * 2;
        heap.pstride = 2 * BLOCK;
        if constexpr (HEAP == kHeapTail)
        {
            // nodes past 20 in global memory as lane-interleaved pairs: one
            // slot per hardware wave slot (a wave's pairs are contiguous and
            // the few MB of resident slots stay in cache) or, for small
            // searches, one slot per lane of the launch
            if (a.heap_tail == 2)
            {
                const uint32_t w = hw_wave_slot();
                heap.gkeys = a.gheap_keys + (((size_t)w * a.tail_pairs) * 64 + (size_t)lane) * 2;
                heap.gpstride = 128;
            }
            else
            {
                heap.gkeys = a.gheap_keys + ((size_t)blockIdx.x * BLOCK + tid) * 2;
                heap.gpstride = (size_t)2 * gridDim.x * (size_t)BLOCK;
            }
        }
    }
    heap.k = a.k;

    const float4 q4 = active ? a.queries[t] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t qorig = __float_as_uint(q4.w);
    const uint64_t sentinel = RADIUS ? ((uint64_t)__float_as_uint(a.r2) << 32) : ~0ull;
    uint64_t worst = sentinel;
    if constexpr (GLOBAL_HEAP)
        heap.count = 0;
    else
        heap.init(sentinel);

    bool live = active && isfinite(q4.x) && isfinite(q4.y) && isfinite(q4.z);
    // known upper bound of the k-th squared distance (from an earlier pass), or +inf
    const float qbound = (!RADIUS && a.query_bound && active) ? a.query_bound[t] : INFINITY;
    const bool bounded = qbound < INFINITY;
    if (BUDGET && a.base_level >= n_levels)
    {
        // no such level: the query goes on to the next pass unchanged
        if (active)
        {
            a.defer_flag[t] = 1u;
            a.defer_bound[t] = qbound;
        }
        return;
    }
    bool deferred = false;  // BUDGET: query handed over to the unbudgeted pass
    const float qxf = q4.x, qyf = q4.y, qzf = q4.z;
    const double qx = qxf, qy = qyf, qz = qzf;

    // ---- level: finest one whose 3x3x3 cube holds k points ----
    int level = 0;
    if (!RADIUS && !BUDGET && live && bounded)
    {
        const double rb = sqrt((double)qbound) * 1.00001;
        while (level < n_levels - 1 && rb > 2.0 * pick_level(a, level).h) ++level;
    }
    else if (!RADIUS && !BUDGET && live)
    {
        for (; level < n_levels - 1; ++level)
        {
            const Level L = pick_level(a, level);
            const int lx = cell_coord(qx, L.lo[0], L.inv_h, L.g[0]);
            const int ly = cell_coord(qy, L.lo[1], L.inv_h, L.g[1]);
            const int lz = cell_coord(qz, L.lo[2], L.inv_h, L.g[2]);
            const int x0 = max(lx - 1, 0), x1 = min(lx + 1, L.g[0] - 1);
            uint32_t count = 0;
            for (int z = max(lz - 1, 0); z <= min(lz + 1, L.g[2] - 1); ++z)
                for (int y = max(ly - 1, 0); y <= min(ly + 1, L.g[1] - 1); ++y)
                {
                    const uint32_t base = ((uint32_t)z * (uint32_t)L.g[1] + (uint32_t)y) * (uint32_t)L.g[0];
                    count += gload(L.cs, base + (uint32_t)x1 + 1) - gload(L.cs, base + (uint32_t)x0);
                }
            if (count >= a.k) break;
        }
    }
    const Level G = (RADIUS || BUDGET) ? uniform_level(a.tab, a.base_level) : pick_level(a, level);
    const float4* __restrict__ pts = G.pts;
    const uint32_t* __restrict__ cs = G.cs;
    const int G0 = G.g[0], G1 = G.g[1], G2 = G.g[2];
    const uint32_t gx = (uint32_t)G0, gy = (uint32_t)G1;
    const int cx = live ? cell_coord(qx, G.lo[0], G.inv_h, G0) : 0;
    const int cy = live ? cell_coord(qy, G.lo[1], G.inv_h, G1) : 0;
    const int cz = live ? cell_coord(qz, G.lo[2], G.inv_h, G2) : 0;
    const double h = G.h, margin = G.h * (double)a.margin;

    // squared pruning distance: the current k-th distance, or the known
    // bound while the heap is not full, with slack for rounding
    auto wlim = [&]() -> double {
        float w = worst == ~0ull ? INFINITY : __uint_as_float((uint32_t)(worst >> 32));
        w = fminf(w, qbound);
        if (!(w < INFINITY)) return INFINITY;
        return (double)w * 1.00001 + 1e-37;
    };

    // POSKEY (budgeted fused pass with heap tails, k > 21): keys carry the
    // candidate's position in the level's sorted array instead of its
    // original index, so that the epilogue gathers the k neighbours from the
    // spatially sorted points rather than from the caller's array in index
    // order (10 M points, k = 50: budgeted pass 19.2 -> 17.0 ms; at k = 20
    // the tie bookkeeping costs what the gathers save).  Ties in distance
    // are then ordered by position, not by index: any exact tie that could
    // matter (at the k-th entry, or between two entries of the row) sends
    // the query to the unbudgeted pass, which keys by index.
    constexpr bool POSKEY = BUDGET && MODE == kModeKnnFeat && HEAP == kHeapTail;
    bool tie = false;
    auto consider = [&](const float4& p, uint32_t pos) {
        const float d2 = sqdist32(qxf, qyf, qzf, p.x, p.y, p.z);
        const uint32_t lo = POSKEY ? pos : __float_as_uint(p.w);
        const uint64_t key = ((uint64_t)__float_as_uint(d2) << 32) | (uint64_t)lo;
        if constexpr (POSKEY)
        {
            // a candidate with the worst entry's distance: rejected or
            // replacing it, the choice between the two is a tie
            const uint32_t wd = (uint32_t)(worst >> 32);
            tie = tie || __float_as_uint(d2) == wd;
            if (key < worst)
            {
                heap.add(key, worst);
                // the entry now at the root has the evicted one's distance
                tie = tie || (wd != 0xFFFFFFFFu && (uint32_t)(worst >> 32) == wd);
            }
        }
        else if (key < worst)
            heap.add(key, worst);
    };


    // Candidates are loaded kScanUnroll at a time (indices clamped to the
    // range, so every load is in bounds) and then considered in order: the
    // loads overlap instead of each waiting for the previous one.

    // this lane's parts [s0, e0) and [s1, e1) of the staged chunk [c, c + cnt)
    auto scan_staged = [&](uint32_t s0, uint32_t e0, uint32_t s1, uint32_t e1, uint32_t c, uint32_t cnt) {
        for (int pc = 0; pc < 2; ++pc)
        {
            const uint32_t lo = max(pc ? s1 : s0, c), hi = min(pc ? e1 : e0, c + cnt);
            for (uint32_t j = lo; j < hi; j += kScanUnroll)
            {
                float4 p[kScanUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kScanUnroll; ++u) p[u] = stage[min(j + u, hi - 1) - c];
#pragma unroll
                for (uint32_t u = 0; u < kScanUnroll; ++u)
                    if (j + u < hi) consider(p[u], j + u);
            }
        }
    };

    // this lane's x pieces [p.x, p.y] and [p.z, p.w] of row (y, z) of box B
    // minus box C (a piece is empty when its start exceeds its end); returned
    // by value so that the pieces stay in registers
    auto row_pieces = [&](const Box& B, const Box& C, bool hasC, bool tighten, double dz2, int y, int z) -> int4 {
        int xa = B.x0, xb = B.x1;
        if (tighten)
        {
            const double wl = wlim();
            const double dy = axis_gap(qy, y, G.lo[1], h, margin);
            const double rem = wl - dz2 - dy * dy;
            if (rem < 0.0) xb = xa - 1;
            else if (rem < INFINITY)
            {
                const double xr = sqrt(rem) + margin;
                xa = max(xa, cell_coord(qx - xr, G.lo[0], G.inv_h, G0));
                xb = min(xb, cell_coord(qx + xr, G.lo[0], G.inv_h, G0));
            }
        }
        const bool split = hasC && y >= C.y0 && y <= C.y1 && z >= C.z0 && z <= C.z1;
        int4 p;
        p.x = xa;
        p.y = split ? min(xb, C.x0 - 1) : xb;
        p.z = split ? max(xa, C.x1 + 1) : 1;
        p.w = split ? xb : 0;
        return p;
    };

    // Scan the cells of box B minus box C (per lane; empty boxes allowed).
    // With `tighten`, each row's x-range is cut to the ball of the lane's
    // current worst distance.
    auto scan_pass = [&](const Box& B, const Box& C, bool tighten) {
        const bool has = B.x0 <= B.x1 && B.y0 <= B.y1 && B.z0 <= B.z1;
        const bool hasC = C.x0 <= C.x1 && C.y0 <= C.y1 && C.z0 <= C.z1;
        const int Z0 = wave_min_i(has ? B.z0 : INT_MAX), Z1 = wave_max_i(has ? B.z1 : INT_MIN);
        const int Y0 = wave_min_i(has ? B.y0 : INT_MAX), Y1 = wave_max_i(has ? B.y1 : INT_MIN);
        if (Z0 > Z1) return;  // no lane has work
        const int lmin = wave_min_i(has ? level : INT_MAX), lmax = wave_max_i(has ? level : INT_MIN);
        const int my_rows = has ? (B.z1 - B.z0 + 1) * (B.y1 - B.y0 + 1) : 0;
        const int max_rows = wave_max_i(my_rows);
        const long union_rows = (long)(Z1 - Z0 + 1) * (long)(Y1 - Y0 + 1);
        // The unbudgeted kNN pass always scans per lane: its queries are
        // sparse and far apart, so shared staging costs more than it saves
        // (measured).  The budgeted pass stages only when asked to: its lanes
        // scan each other's rows in lockstep either way, and per-lane loads
        // from L2 measure faster than staging the rows' union.
        const bool stage_here = RADIUS ? a.staging >= 1 : (BUDGET && a.staging >= 2);
        const bool coop = stage_here && (lmin == lmax) && union_rows <= 2L * max_rows + 8;
        if (coop)
        {
            // every lane stages from the working level's array, also lanes
            // that sit on another level and have no row in this pass
            const float4* __restrict__ wpts = (RADIUS || BUDGET) ? G.pts : pick_level(a, lmin).pts;
            // rows nearest to the wave's centre first: close candidates fill
            // the heaps early and most later ones fail the threshold test
            const int zc = a.near_first ? (Z0 + Z1) / 2 : Z0, yc = a.near_first ? (Y0 + Y1) / 2 : Y0;
            const int nz = Z1 - Z0 + 1, ny = Y1 - Y0 + 1;
            for (int iz = 0; iz < 2 * nz; ++iz)
            {
                const int z = near_index(iz, zc, Z0, Z1);
                if (z < Z0 || z > Z1) continue;
                const bool zin = has && z >= B.z0 && z <= B.z1;
                double dz2 = 0.0;
                if (tighten && zin)
                {
                    const double dz = axis_gap(qz, z, G.lo[2], h, margin);
                    dz2 = dz * dz;
                }
                for (int iy = 0; iy < 2 * ny; ++iy)
                {
                    const int y = near_index(iy, yc, Y0, Y1);
                    if (y < Y0 || y > Y1) continue;
                    int4 rp = make_int4(1, 0, 1, 0);
                    if (zin && y >= B.y0 && y <= B.y1) rp = row_pieces(B, C, hasC, tighten, dz2, y, z);
                    const int a0 = rp.x, b0 = rp.y, a1 = rp.z, b1 = rp.w;
                    const uint32_t base = ((uint32_t)z * gy + (uint32_t)y) * gx;
                    uint32_t s0 = 0xFFFFFFFFu, e0 = 0, s1 = 0xFFFFFFFFu, e1 = 0;
                    if (a0 <= b0)
                    {
                        s0 = gload(cs, base + (uint32_t)a0);
                        e0 = gload(cs, base + (uint32_t)b0 + 1);
                    }
                    if (a1 <= b1)
                    {
                        s1 = gload(cs, base + (uint32_t)a1);
                        e1 = gload(cs, base + (uint32_t)b1 + 1);
                    }
                    const uint32_t S = wave_min_u(min(s0 < e0 ? s0 : 0xFFFFFFFFu, s1 < e1 ? s1 : 0xFFFFFFFFu));
                    const uint32_t E = wave_max_u(max(s0 < e0 ? e0 : 0u, s1 < e1 ? e1 : 0u));
                    for (uint32_t c = S; c < E; c += kStage)
                    {
                        const uint32_t cnt = min((uint32_t)kStage, E - c);
#pragma unroll
                        for (int u = 0; u < kStage / 64; ++u)
                        {
                            const uint32_t i = (uint32_t)(u * 64 + lane);
                            if (i < cnt) stage[i] = gload(wpts, c + i);
                        }
                        wave_sync();
                        scan_staged(s0, e0, s1, e1, c, cnt);
                        wave_sync();
                    }
                }
            }
        }
        else if (has && (BUDGET || RADIUS))
        {
            // budgeted kNN pass and radius searches: the lanes walk their
            // rows in step (nearest first), a row's setup shared by the wave
            const int zc = a.near_first ? min(max(cz, B.z0), B.z1) : B.z0;
            const int yc = a.near_first ? min(max(cy, B.y0), B.y1) : B.y0;
            for (int iz = 0; iz <= B.z1 - B.z0; ++iz)
            {
                const int z = nth_near(iz, zc, B.z0, B.z1);
                double dz2 = 0.0;
                if (tighten)
                {
                    const double dz = axis_gap(qz, z, G.lo[2], h, margin);
                    dz2 = dz * dz;
                    if (dz2 > wlim()) continue;
                }
                // kRowsPerStream rows of the plane (two pieces each) form one
                // stream of candidates per lane: the wave steps over its
                // busiest lane's sum of those pieces instead of each piece's
                // busiest lane, and the rows' cell starts load together
                constexpr int kR = kRowsPerStream;
                for (int iy = 0; iy <= B.y1 - B.y0; iy += kR)
                {
                    uint32_t rs[2 * kR], rl[2 * kR];
#pragma unroll
                    for (int h2 = 0; h2 < kR; ++h2)
                    {
                        rs[2 * h2] = rl[2 * h2] = rs[2 * h2 + 1] = rl[2 * h2 + 1] = 0;
                        if (iy + h2 > B.y1 - B.y0) continue;
                        const int y = nth_near(iy + h2, yc, B.y0, B.y1);
                        const int4 rp = row_pieces(B, C, hasC, tighten, dz2, y, z);
                        const uint32_t base = ((uint32_t)z * gy + (uint32_t)y) * gx;
                        if (rp.x <= rp.y)
                        {
                            rs[2 * h2] = gload(cs, base + (uint32_t)rp.x);
                            rl[2 * h2] = gload(cs, base + (uint32_t)rp.y + 1) - rs[2 * h2];
                        }
                        if (rp.z <= rp.w)
                        {
                            rs[2 * h2 + 1] = gload(cs, base + (uint32_t)rp.z);
                            rl[2 * h2 + 1] = gload(cs, base + (uint32_t)rp.w + 1) - rs[2 * h2 + 1];
                        }
                    }
                    // stream position -> sorted position: cum[r] is the end of
                    // range r in the stream, off[r] = its start in pts minus its
                    // start in the stream (u32 wrap-around is fine)
                    uint32_t cum[2 * kR], off[2 * kR], tot = 0;
#pragma unroll
                    for (int r = 0; r < 2 * kR; ++r)
                    {
                        off[r] = rs[r] - tot;
                        tot += rl[r];
                        cum[r] = tot;
                    }
                    for (uint32_t c = 0; c < tot; c += kScanUnroll)
                    {
                        float4 p[kScanUnroll];
                        uint32_t pp[kScanUnroll];
#pragma unroll
                        for (uint32_t u = 0; u < kScanUnroll; ++u)
                        {
                            const uint32_t cc = min(c + u, tot - 1);
                            uint32_t o = off[2 * kR - 1];
#pragma unroll
                            for (int r = 2 * kR - 2; r >= 0; --r) o = cc < cum[r] ? off[r] : o;
                            pp[u] = cc + o;
                            p[u] = gload(pts, pp[u]);
                        }
#pragma unroll
                        for (uint32_t u = 0; u < kScanUnroll; ++u)
                            if (c + u < tot) consider(p[u], pp[u]);
                    }
                }
            }
        }
        else if (has)
        {
            // Unbudgeted pass: each lane walks its own rows (nearest first)
            // with a cursor and its candidates four per step.  A lane that
            // finishes a row opens its next one within the same step, so the
            // lanes of a wave do not wait for each other at row ends: the wave
            // takes as many steps as its busiest lane, not the sum over rows
            // of the busiest lane per row.  (In the budgeted pass the rows
            // are short and alike, and sharing each row's setup wins.)
            const int zc = a.near_first ? min(max(cz, B.z0), B.z1) : B.z0;
            const int yc = a.near_first ? min(max(cy, B.y0), B.y1) : B.y0;
            const int nz = B.z1 - B.z0 + 1, ny = B.y1 - B.y0 + 1;
            int iz = 0, iy = 0;  // next row to open: row iy (near order) of plane iz
            int z = nth_near(0, zc, B.z0, B.z1);
            double dz2 = 0.0;
            if (tighten)
            {
                const double dz = axis_gap(qz, z, G.lo[2], h, margin);
                dz2 = dz * dz;
            }
            bool rows = true;  // rows left to open
            uint32_t pos = 0, end = 0, ps = 0, pe = 0;  // current piece, pending second piece
            while (__ballot(rows || pos < end || ps < pe) != 0)
            {
                if (pos >= end)
                {
                    if (ps < pe)
                    {
                        pos = ps;
                        end = pe;
                        ps = pe = 0;
                    }
                    else
                    {
                        // open the next row with a non-empty x range; every
                        // pass moves the cursor by one row or one plane
                        while (rows)
                        {
                            if (iy >= ny || (tighten && dz2 > wlim()))
                            {
                                iy = 0;
                                if (++iz >= nz)
                                {
                                    rows = false;
                                    break;
                                }
                                z = nth_near(iz, zc, B.z0, B.z1);
                                if (tighten)
                                {
                                    const double dz = axis_gap(qz, z, G.lo[2], h, margin);
                                    dz2 = dz * dz;
                                }
                                continue;
                            }
                            const int y = nth_near(iy++, yc, B.y0, B.y1);
                            const int4 rp = row_pieces(B, C, hasC, tighten, dz2, y, z);
                            const bool p0 = rp.x <= rp.y, p1 = rp.z <= rp.w;
                            if (!p0 && !p1) continue;
                            const uint32_t base = ((uint32_t)z * gy + (uint32_t)y) * gx;
                            const uint32_t a0 = (uint32_t)(p0 ? rp.x : rp.z), b0 = (uint32_t)(p0 ? rp.y : rp.w);
                            pos = gload(cs, base + a0);
                            end = gload(cs, base + b0 + 1);
                            if (p0 && p1)
                            {
                                ps = gload(cs, base + (uint32_t)rp.z);
                                pe = gload(cs, base + (uint32_t)rp.w + 1);
                            }
                            break;
                        }
                    }
                }
                if (pos < end)
                {
                    float4 p[kCursorUnroll];
#pragma unroll
                    for (uint32_t u = 0; u < kCursorUnroll; ++u) p[u] = gload(pts, min(pos + u, end - 1));
#pragma unroll
                    for (uint32_t u = 0; u < kCursorUnroll; ++u)
                        if (pos + u < end) consider(p[u], pos + u);
                    pos = min(pos + kCursorUnroll, end);
                }
            }
        }
    };

    // Phase 1 (kNN): cubic shells around the query cell until the heap holds
    // k candidates.
    int n_done = -1;
    bool covered = false;
    if (!RADIUS)
    {
        // a query with a known bound goes straight to the ball of that bound
        bool need = live && !bounded;
        for (int n = BUDGET ? 1 : 0; __ballot(need) != 0; ++n)
        {
            const Box B = need ? cube_box(cx, cy, cz, n, G) : empty_box();
            const Box C = (need && n > (BUDGET ? 1 : 0)) ? cube_box(cx, cy, cz, n - 1, G) : empty_box();
            scan_pass(B, C, true);
            if (need)
            {
                n_done = n;
                covered = (cx - n <= 0) && (cx + n >= G0 - 1) && (cy - n <= 0) && (cy + n >= G1 - 1) &&
                          (cz - n <= 0) && (cz + n >= G2 - 1);
                if (worst != sentinel || covered) need = false;
                if (BUDGET && need && n >= a.budget_shells)
                {
                    // sparse neighbourhood: leave it to the next pass
                    need = false;
                    deferred = true;
                    live = false;
                }
            }
        }
    }

    // Phase 2: every row meeting the ball of the current k-th distance (kNN)
    // or of the search radius, minus the phase-1 cube.
    {
        Box B = empty_box();
        if (live && !covered)
        {
            const double wl = wlim();
            if (wl < INFINITY)
            {
                const double R = sqrt(wl) + margin;
                B = Box{cell_coord(qx - R, G.lo[0], G.inv_h, G0), cell_coord(qx + R, G.lo[0], G.inv_h, G0),
                        cell_coord(qy - R, G.lo[1], G.inv_h, G1), cell_coord(qy + R, G.lo[1], G.inv_h, G1),
                        cell_coord(qz - R, G.lo[2], G.inv_h, G2), cell_coord(qz + R, G.lo[2], G.inv_h, G2)};
            }
            else
                B = Box{0, G0 - 1, 0, G1 - 1, 0, G2 - 1};
            const int bc = a.budget_cells;
            if (BUDGET && (B.x0 < cx - bc || B.x1 > cx + bc || B.y0 < cy - bc || B.y1 > cy + bc ||
                           B.z0 < cz - bc || B.z1 > cz + bc))
            {
                B = empty_box();
                deferred = true;
                live = false;
            }
        }
        const Box C = n_done >= 0 ? cube_box(cx, cy, cz, n_done, G) : empty_box();
        scan_pass(B, C, true);
    }

    if (!active) return;
    if (BUDGET)
    {
        if constexpr (POSKEY)
        {
            if (!deferred)
            {
                // two entries of the row with one distance: their order (and
                // with a tie at the k-th, the choice) follows positions here
                heap.finish(sentinel);
                uint32_t prev = 0xFFFFFFFFu;
                for (uint32_t j = 0; j < a.k; ++j)
                {
                    const uint32_t d = (uint32_t)(heap.key_at(j) >> 32);
                    tie = tie || d == prev;
                    prev = d;
                }
                deferred = tie;
            }
        }
        a.defer_flag[t] = deferred ? 1u : 0u;
        if (deferred)
        {
            // the k-th distance so far (heap full) bounds the final one
            const float wd = worst == ~0ull ? INFINITY : __uint_as_float((uint32_t)(worst >> 32));
            a.defer_bound[t] = fminf(wd, qbound);
            return;
        }
    }
    // keys held (radius searches: the hits, at most k; they come first in
    // the sorted heap, the sentinels after them)
    const uint32_t hits = heap.count;
    if constexpr (!POSKEY) heap.template finish<RADIUS>(sentinel);
    const uint32_t k = a.k;

    if (MODE == kModeKnn)
    {
        uint32_t* oi = reinterpret_cast<uint32_t*>(a.out_idx) + (size_t)qorig * k;
        float* od = a.out_d2 ? a.out_d2 + (size_t)qorig * k : nullptr;
        store_rows(oi, od, k, [&](uint32_t j, uint32_t& i, float& d) {
            const uint64_t key = heap.key_at(j);
            i = (uint32_t)key;
            d = key == ~0ull ? INFINITY : __uint_as_float((uint32_t)(key >> 32));
        });
    }
    else if (MODE == kModeRadius || MODE == kModeRadiusCsr || MODE == kModeRadiusMulti)
    {
        const uint32_t cnt = hits;
        if (MODE == kModeRadius)
        {
            uint32_t* oi = reinterpret_cast<uint32_t*>(a.out_idx) + (size_t)qorig * k;
            float* od = a.out_d2 + (size_t)qorig * k;
            // -1 / 0 past the hits, and nanoflann's RKNNResultSet::init
            // leaves r^2 in the last slot when it is not a hit
            store_rows(oi, od, k, [&](uint32_t j, uint32_t& i, float& d) {
                const uint64_t key = j < cnt ? heap.key_at(j) : sentinel;
                i = j < cnt ? (uint32_t)key : 0xFFFFFFFFu;
                d = j < cnt ? __uint_as_float((uint32_t)(key >> 32)) : (j == k - 1 ? a.r2 : 0.0f);
            });
        }
        else
        {
            uint32_t* oi = reinterpret_cast<uint32_t*>(a.out_idx) + (size_t)qorig * k;
            store_rows(oi, nullptr, cnt, [&](uint32_t j, uint32_t& i, float& d) {
                i = (uint32_t)heap.key_at(j);
                d = 0.0f;
            });
        }
        if (a.out_count) a.out_count[qorig] = cnt;
        if (MODE == kModeRadiusMulti)
        {
            // compute_features_multiscale over this row (include/pgeof.hpp:
            // 159-211): the features of its first k_s hits for every scale
            // k_s <= cnt, in one walk of the sorted hits (the same shifted
            // FP64 moments and eigensolve as the CSR kernel, so the values
            // equal the two-call path's bit for bit); later scales stay 0.
            const uint32_t S = a.n_scales;
            float* of = a.out_feat + (size_t)qorig * S * 11;
            uint32_t walk = 0;
            for (uint32_t si = 0; si < S; ++si)
            {
                if (a.scales[si] > cnt) break;
                walk = a.scales[si];
            }
            uint32_t si = 0;
            if (walk)
            {
                // the hits are read back from the row just written (this
                // lane's own stores): the heap is dead during the walk
                const uint32_t* __restrict__ row = reinterpret_cast<const uint32_t*>(a.out_idx) + (size_t)qorig * k;
                const float3* __restrict__ xyz3 = reinterpret_cast<const float3*>(a.data);
                const float3 p0 = xyz3[row[0]];
                const double ox = p0.x, oy = p0.y, oz = p0.z;
                Moments m;
                mom_zero(m);
                for (uint32_t j = 0; j < walk; ++j)
                {
                    const float3 p = xyz3[row[j]];
                    mom_add(m, (double)p.x - ox, (double)p.y - oy, (double)p.z - oz);
                    while (si < S && a.scales[si] == j + 1)
                    {
                        double c[6];
                        mom_cov(m, (double)(j + 1), c);
                        const Eig E = eig_full(c);
                        const Pca32 P = to_pca32(E);
                        float f[11];
                        features11(P, f);
                        store_row11(of + (size_t)si * 11, f);
                        ++si;
                    }
                }
            }
            for (uint32_t c = si * 11; c < S * 11; ++c) of[c] = 0.0f;
        }
    }
    else  // kModeKnnFeat
    {
        uint32_t* oi = reinterpret_cast<uint32_t*>(a.out_idx) + (size_t)qorig * k;
        float* od = a.out_d2 ? a.out_d2 + (size_t)qorig * k : nullptr;
        float f[11];
        if (live)
        {
            // Neighbour coordinates are gathered from the caller's xyz by
            // original index (one 12-byte load each), four at a time, the next
            // four issued before the current ones are accumulated.  Rows of
            // the index (and distance) output are written four entries per
            // store when k is a multiple of 4 and the arrays are 16-byte
            // aligned (every row is then aligned too), two per store when k
            // is even and the arrays 8-byte aligned (the rows of a wave's
            // lanes are far apart, so a store reaches as many cache lines as
            // the wave has lanes: fewer, wider stores make fewer requests).
            Moments m;
            mom_zero(m);
            const float3* __restrict__ xyz3 = reinterpret_cast<const float3*>(a.data);
            // neighbour of a key: its coordinates and (bits of) original index,
            // from the sorted points by position (POSKEY) or from xyz by index
            auto nbr = [&](uint64_t key) -> float4 {
                if constexpr (POSKEY)
                    return gload(pts, (uint32_t)key);
                else
                {
                    const float3 q3 = xyz3[(uint32_t)key];
                    return make_float4(q3.x, q3.y, q3.z, __uint_as_float((uint32_t)key));
                }
            };
            constexpr uint32_t kChunk = 4;
            const uintptr_t aln = reinterpret_cast<uintptr_t>(a.out_idx) | reinterpret_cast<uintptr_t>(a.out_d2);
            const bool vec_out = (k & 3u) == 0 && (aln & 15u) == 0;
            // k even, rows 8-byte aligned: 16 bytes per store from each row's
            // first 16-byte boundary when the two arrays are aligned alike (a
            // row 8 bytes past one stores its first two entries alone, then
            // carries two entries of every chunk into the next chunk's
            // store), else 8 bytes per store
            const bool pair_out = !vec_out && (k & 1u) == 0 && (aln & 7u) == 0;
            const bool carry_ok =
                pair_out && (od == nullptr || ((reinterpret_cast<uintptr_t>(oi) ^ reinterpret_cast<uintptr_t>(od)) & 15u) == 0);
            const bool shifted = carry_ok && (reinterpret_cast<uintptr_t>(oi) & 15u) != 0;
            uint32_t ci0 = 0, ci1 = 0;
            float cd0 = 0.0f, cd1 = 0.0f;
            bool carry = false;
            uint64_t kn[kChunk];
            float4 pn[kChunk];
#pragma unroll
            for (uint32_t u = 0; u < kChunk; ++u)
            {
                kn[u] = heap.key_at(u < k ? u : 0);
                pn[u] = nbr(kn[u]);
            }
            const double ox = pn[0].x, oy = pn[0].y, oz = pn[0].z;
            for (uint32_t j0 = 0; j0 < k; j0 += kChunk)
            {
                uint64_t kc[kChunk];
                float4 pc[kChunk];
#pragma unroll
                for (uint32_t u = 0; u < kChunk; ++u)
                {
                    kc[u] = kn[u];
                    pc[u] = pn[u];
                }
                if (j0 + kChunk < k)
                {
#pragma unroll
                    for (uint32_t u = 0; u < kChunk; ++u)
                    {
                        const uint32_t j = j0 + kChunk + u;
                        kn[u] = heap.key_at(j < k ? j : 0);
                        pn[u] = nbr(kn[u]);
                    }
                }
                if (vec_out)
                {
                    *reinterpret_cast<uint4*>(oi + j0) =
                        make_uint4(__float_as_uint(pc[0].w), __float_as_uint(pc[1].w), __float_as_uint(pc[2].w),
                                   __float_as_uint(pc[3].w));
                    if (od)
                        *reinterpret_cast<float4*>(od + j0) = make_float4(
                            __uint_as_float((uint32_t)(kc[0] >> 32)), __uint_as_float((uint32_t)(kc[1] >> 32)),
                            __uint_as_float((uint32_t)(kc[2] >> 32)), __uint_as_float((uint32_t)(kc[3] >> 32)));
                }
                else if (carry_ok)
                {
                    // k even: entries j0, j0 + 1 exist, and j0 + 2, j0 + 3 together
                    const uint32_t i0 = __float_as_uint(pc[0].w), i1 = __float_as_uint(pc[1].w);
                    const uint32_t i2 = __float_as_uint(pc[2].w), i3 = __float_as_uint(pc[3].w);
                    const float e0 = __uint_as_float((uint32_t)(kc[0] >> 32)), e1 = __uint_as_float((uint32_t)(kc[1] >> 32));
                    const float e2 = __uint_as_float((uint32_t)(kc[2] >> 32)), e3 = __uint_as_float((uint32_t)(kc[3] >> 32));
                    const bool full = j0 + 3 < k;
                    if (!shifted)
                    {
                        if (full)
                        {
                            *reinterpret_cast<uint4*>(oi + j0) = make_uint4(i0, i1, i2, i3);
                            if (od) *reinterpret_cast<float4*>(od + j0) = make_float4(e0, e1, e2, e3);
                        }
                        else
                        {
                            *reinterpret_cast<uint2*>(oi + j0) = make_uint2(i0, i1);
                            if (od) *reinterpret_cast<float2*>(od + j0) = make_float2(e0, e1);
                        }
                    }
                    else
                    {
                        if (j0 == 0)
                        {
                            *reinterpret_cast<uint2*>(oi) = make_uint2(i0, i1);
                            if (od) *reinterpret_cast<float2*>(od) = make_float2(e0, e1);
                        }
                        else
                        {
                            *reinterpret_cast<uint4*>(oi + j0 - 2) = make_uint4(ci0, ci1, i0, i1);
                            if (od) *reinterpret_cast<float4*>(od + j0 - 2) = make_float4(cd0, cd1, e0, e1);
                        }
                        ci0 = i2, ci1 = i3, cd0 = e2, cd1 = e3;
                        carry = full;
                    }
                }
                else if (pair_out)
                {
                    // k even: entries j0, j0 + 1 exist, and j0 + 2, j0 + 3 together
#pragma unroll
                    for (uint32_t u = 0; u < kChunk; u += 2)
                        if (j0 + u < k)
                        {
                            *reinterpret_cast<uint2*>(oi + j0 + u) =
                                make_uint2(__float_as_uint(pc[u].w), __float_as_uint(pc[u + 1].w));
                            if (od)
                                *reinterpret_cast<float2*>(od + j0 + u) =
                                    make_float2(__uint_as_float((uint32_t)(kc[u] >> 32)),
                                                __uint_as_float((uint32_t)(kc[u + 1] >> 32)));
                        }
                }
                else
                {
#pragma unroll
                    for (uint32_t u = 0; u < kChunk; ++u)
                        if (j0 + u < k)
                        {
                            oi[j0 + u] = __float_as_uint(pc[u].w);
                            if (od) od[j0 + u] = __uint_as_float((uint32_t)(kc[u] >> 32));
                        }
                }
#pragma unroll
                for (uint32_t u = 0; u < kChunk; ++u)
                    if (j0 + u < k) mom_add(m, (double)pc[u].x - ox, (double)pc[u].y - oy, (double)pc[u].z - oz);
            }
            if (carry)
            {
                // the row's last two entries (k a multiple of 4, shifted row)
                *reinterpret_cast<uint2*>(oi + k - 2) = make_uint2(ci0, ci1);
                if (od) *reinterpret_cast<float2*>(od + k - 2) = make_float2(cd0, cd1);
            }
            double c[6];
            mom_cov(m, (double)k, c);
            const Eig E = eig_full(c);
            const Pca32 P = to_pca32(E);
            features11(P, f);
        }
        else
        {
            for (uint32_t j = 0; j < k; ++j)
            {
                oi[j] = 0xFFFFFFFFu;
                if (od) od[j] = INFINITY;
            }
            for (int c = 0; c < 11; ++c) f[c] = 0.0f;
        }
        store_row11(a.out_feat + (size_t)qorig * 11, f);
    }
}

// One lane per query.  With a device-side query count (the deferred pass)
// the launch is sized for the largest possible count and the blocks past the
// count exit.  Block b takes queries xcd_logical_block(b) * BLOCK ...:
// consecutive (spatially adjacent) blocks of work share an XCD's L2.
template <int MODE, int WAVES, int HEAP, bool BUDGET>
__global__ __launch_bounds__(WAVES * 64, 4) void k_search(SearchArgs a)
{
    constexpr int BLOCK = WAVES * 64;
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t nq = a.n_query, nb = gridDim.x;
    if (a.n_query_dev)
    {
        const uint64_t c = *a.n_query_dev;
        nq = c < (uint64_t)a.n_query ? (uint32_t)c : a.n_query;
        nb = (nq + BLOCK - 1) / BLOCK;
    }
    // (a loop over several blocks of work per block would hold loop-invariant
    // per-lane state through the whole search and spill: blocks past the
    // work exit at once instead)
    if (blockIdx.x >= nb) return;
    const uint32_t t = xcd_logical_block(blockIdx.x, nb) * BLOCK + threadIdx.x;
    search_slot<MODE, WAVES, HEAP, BUDGET>(a, t, t < nq, smem, threadIdx.x);
}

// ----------------------------------------------------------- selected ----
// compute_features_selected (include/pgeof.hpp:325-375): radius neighbourhood
// in the input precision T, the min(found, max_knn) nearest hits (ties by
// index), rows with fewer than 2 hits stay 0, then the EFeatureID list of
// include/pca.hpp:212-295 evaluated in T.
template <typename T>
struct PairHeap
{
    T* key;
    uint32_t* idx;
    uint32_t stride;
    uint32_t k;
    __device__ __forceinline__ static bool less(T a, uint32_t ia, T b, uint32_t ib)
    {
        return a < b || (a == b && ia < ib);
    }
    __device__ __forceinline__ void sift_root(T kk, uint32_t ii, uint32_t n)
    {
        uint32_t i = 0;
        for (;;)
        {
            const uint32_t l = 2 * i + 1;
            if (l >= n) break;
            T ck = key[(size_t)l * stride];
            uint32_t ci = idx[(size_t)l * stride];
            uint32_t c = l;
            if (l + 1 < n)
            {
                const T rk = key[(size_t)(l + 1) * stride];
                const uint32_t ri = idx[(size_t)(l + 1) * stride];
                if (less(ck, ci, rk, ri))
                {
                    ck = rk;
                    ci = ri;
                    c = l + 1;
                }
            }
            if (!less(kk, ii, ck, ci)) break;
            key[(size_t)i * stride] = ck;
            idx[(size_t)i * stride] = ci;
            i = c;
        }
        key[(size_t)i * stride] = kk;
        idx[(size_t)i * stride] = ii;
    }
    __device__ __forceinline__ void sort_ascending()
    {
        for (uint32_t end = k - 1; end > 0; --end)
        {
            const T tk = key[0];
            const uint32_t ti = idx[0];
            const T lk = key[(size_t)end * stride];
            const uint32_t li = idx[(size_t)end * stride];
            key[(size_t)end * stride] = tk;
            idx[(size_t)end * stride] = ti;
            sift_root(lk, li, end);
        }
    }
};

template <typename T>
__device__ __forceinline__ T sqdist_t(const T* q, const T* p)
{
#pragma clang fp contract(off)
    const T dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    T d = dx * dx;
    d = d + dy * dy;
    d = d + dz * dz;
    return d;
}

template <typename T>
struct PcaT
{
    T val[3];
    T v[3][3];
};

template <typename T>
__device__ __forceinline__ PcaT<T> to_pca_t(const Eig& E)
{
    PcaT<T> P;
    for (int i = 0; i < 3; ++i)
    {
        P.val[i] = (T)E.l[i] > (T)0 ? (T)E.l[i] : (T)0;
        for (int j = 0; j < 3; ++j) P.v[i][j] = (T)E.v[i][j];
    }
    bool flip = P.v[2][2] < (T)0;
    if (P.v[2][2] == (T)0) flip = (P.v[2][1] < (T)0) || (P.v[2][1] == (T)0 && P.v[2][0] < (T)0);
    if (flip)
        for (int j = 0; j < 3; ++j) P.v[2][j] = -P.v[2][j];
    return P;
}

template <typename T>
__device__ __forceinline__ T selected_feature(const PcaT<T>& P, int fid)
{
#pragma clang fp contract(off)
    const T eps = (T)1e-3;
    const T val0 = sqrt(P.val[0]), val1 = sqrt(P.val[1]), val2 = sqrt(P.val[2]);
    const T fact = (T)1.0 / (val0 + eps);
    switch (fid)
    {
        case 0: return (val0 - val1) * fact;
        case 1: return (val1 - val2) * fact;
        case 2: return val2 * fact;
        case 3:
            if (val0 > (T)0)
            {
                T u[3];
                for (int c = 0; c < 3; ++c)
                    u[c] = P.val[0] * fabs(P.v[0][c]) + P.val[1] * fabs(P.v[1][c]) + P.val[2] * fabs(P.v[2][c]);
                return u[2] / sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            }
            return (T)0;
        case 4: return P.v[2][0];
        case 5: return P.v[2][1];
        case 6: return P.v[2][2];
        case 7: return val0;
        case 8: return sqrt(val0 * val1 + (T)1e-6f);  // pca.hpp:252: float literal in both builds
        case 9: return pow(val0 * val1 * val2 + (T)1e-9, (T)1.0 / (T)3.0);
        case 10: return val2 / (val0 + val1 + val2 + eps);
        case 12: return (T)1.0 - fabs(P.v[2][2]);
        case 13:
        {
            const T sum = P.val[0] + P.val[1] + P.val[2] + eps;
            const T e0 = P.val[0] / sum, e1 = P.val[1] / sum, e2 = P.val[2] / sum;
            return -e0 * log(e0 + eps) - e1 * log(e1 + eps) - e2 * log(e2 + eps);
        }
        default: return (T)0;
    }
}

template <typename T, int BLOCK, bool GLOBAL_HEAP>
__global__ __launch_bounds__(BLOCK) void k_selected(SelectedArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t t = lb * BLOCK + threadIdx.x;
    if (t >= a.n) return;
    const uint32_t cap = a.cap;
    PairHeap<T> heap;
    if (GLOBAL_HEAP)
    {
        heap.key = reinterpret_cast<T*>(a.gheap_keys) + t;
        heap.idx = a.gheap_idx + t;
        heap.stride = a.n;
    }
    else
    {
        heap.key = reinterpret_cast<T*>(smem) + threadIdx.x;
        heap.idx = reinterpret_cast<uint32_t*>(smem + (size_t)cap * BLOCK * sizeof(T)) + threadIdx.x;
        heap.stride = BLOCK;
    }
    heap.k = cap;
    const T r2 = (T)a.r2;
    for (uint32_t j = 0; j < cap; ++j)
    {
        heap.key[(size_t)j * heap.stride] = r2;
        heap.idx[(size_t)j * heap.stride] = 0xFFFFFFFFu;
    }
    const T* __restrict__ X = reinterpret_cast<const T*>(a.xyz);
    const float4* __restrict__ pts = a.grid.pts;
    // candidate coordinates in the input precision, read in the grid's
    // sorted order: float input is the sorted float4 array itself, double
    // input a sorted copy (a.xs); the original index comes from pts[j].w
    const double* __restrict__ xs = a.xs;
    auto coords = [&](uint32_t j, const float4& pj, T c[3]) {
        if constexpr (sizeof(T) == 8)
        {
            c[0] = xs[(size_t)j * 3 + 0];
            c[1] = xs[(size_t)j * 3 + 1];
            c[2] = xs[(size_t)j * 3 + 2];
        }
        else
        {
            c[0] = pj.x;
            c[1] = pj.y;
            c[2] = pj.z;
        }
    };
    const float4 qf = pts[t];
    const uint32_t qi = __float_as_uint(qf.w);
    T q[3];
    coords(t, qf, q);
    T worst_k = r2;
    uint32_t worst_i = 0xFFFFFFFFu;
    const GridParams& gp = a.grid.p;
    const bool finite = isfinite((double)q[0]) && isfinite((double)q[1]) && isfinite((double)q[2]);
    if (finite)
    {
        const double R = a.reach;
        const double fx = qf.x, fy = qf.y, fz = qf.z;
        const int x0 = cell_coord(fx - R, gp.lo[0], gp.inv_h, gp.g[0]);
        const int x1 = cell_coord(fx + R, gp.lo[0], gp.inv_h, gp.g[0]);
        const int y0 = cell_coord(fy - R, gp.lo[1], gp.inv_h, gp.g[1]);
        const int y1 = cell_coord(fy + R, gp.lo[1], gp.inv_h, gp.g[1]);
        const int z0 = cell_coord(fz - R, gp.lo[2], gp.inv_h, gp.g[2]);
        const int z1 = cell_coord(fz + R, gp.lo[2], gp.inv_h, gp.g[2]);
        const int yc = min(max(cell_coord(fy, gp.lo[1], gp.inv_h, gp.g[1]), y0), y1);
        const int zc = min(max(cell_coord(fz, gp.lo[2], gp.inv_h, gp.g[2]), z0), z1);
        // rows nearest first, each cut to the sphere of the current bound (in
        // the float grid's space: the radius, or the heap's worst distance
        // once it is full, plus the rounding margin of the float copies)
        auto bound2 = [&]() -> double {
            const double b = worst_i == 0xFFFFFFFFu ? R : sqrt((double)worst_k * 1.00001) + a.margin;
            return (b < R ? b : R) * (b < R ? b : R);
        };
        const double gm = gp.h * 1e-6;  // cell boxes widened against rounding of the cell coordinates
        for (int iz = 0; iz <= z1 - z0; ++iz)
        {
            const int z = nth_near(iz, zc, z0, z1);
            const double dz = axis_gap(fz, z, gp.lo[2], gp.h, gm);
            if (dz * dz > bound2()) continue;
            for (int iy = 0; iy <= y1 - y0; ++iy)
            {
                const int y = nth_near(iy, yc, y0, y1);
                const double dy = axis_gap(fy, y, gp.lo[1], gp.h, gm);
                const double rem = bound2() - dz * dz - dy * dy;
                if (rem < 0.0) continue;
                const double xr = sqrt(rem) + gm;
                const int xa = max(x0, cell_coord(fx - xr, gp.lo[0], gp.inv_h, gp.g[0]));
                const int xb = min(x1, cell_coord(fx + xr, gp.lo[0], gp.inv_h, gp.g[0]));
                if (xa > xb) continue;
                const uint32_t base = ((uint32_t)z * (uint32_t)gp.g[1] + (uint32_t)y) * (uint32_t)gp.g[0];
                const uint32_t s = a.grid.cell_start[base + (uint32_t)xa], e = a.grid.cell_start[base + (uint32_t)xb + 1];
                // four candidates loaded together, then considered in order
                for (uint32_t j = s; j < e; j += 4)
                {
                    float4 pj[4];
                    T pc[4][3];
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u)
                    {
                        const uint32_t jj = min(j + u, e - 1);
                        pj[u] = pts[jj];
                        coords(jj, pj[u], pc[u]);
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u)
                    {
                        if (j + u >= e) break;
                        const uint32_t pi = __float_as_uint(pj[u].w);
                        const T d = sqdist_t<T>(q, pc[u]);
                        if (d < r2 && PairHeap<T>::less(d, pi, worst_k, worst_i))
                        {
                            heap.sift_root(d, pi, cap);
                            worst_k = heap.key[0];
                            worst_i = heap.idx[0];
                        }
                    }
                }
            }
        }
    }
    heap.sort_ascending();
    uint32_t cnt = 0;
    for (uint32_t j = 0; j < cap; ++j)
        if (heap.idx[(size_t)j * heap.stride] != 0xFFFFFFFFu) ++cnt;
    T* out = reinterpret_cast<T*>(a.out) + (size_t)qi * a.n_sel;
    if (cnt < 2 || a.max_knn == 0)
    {
        for (uint32_t s = 0; s < a.n_sel; ++s) out[s] = (T)0;
        return;
    }
    const uint32_t m = cnt < a.max_knn ? cnt : a.max_knn;
    Moments mo;
    mom_zero(mo);
    double o[3] = {0.0, 0.0, 0.0};
    for (uint32_t j = 0; j < m; ++j)
    {
        const uint32_t pi = heap.idx[(size_t)j * heap.stride];
        const double px = (double)X[(size_t)pi * 3 + 0], py = (double)X[(size_t)pi * 3 + 1],
                     pz = (double)X[(size_t)pi * 3 + 2];
        if (j == 0) o[0] = px, o[1] = py, o[2] = pz;
        mom_add(mo, px - o[0], py - o[1], pz - o[2]);
    }
    double c[6];
    mom_cov(mo, (double)m, c);
    const Eig E = eig_full(c);
    const PcaT<T> P = to_pca_t<T>(E);
    for (uint32_t s = 0; s < a.n_sel; ++s) out[s] = selected_feature<T>(P, a.fids[s]);
}

template <typename T, int BLOCK>
hipError_t launch_selected_b(const SelectedArgs& a, size_t lds, bool gh, hipStream_t s)
{
    const uint32_t nb = (a.n + BLOCK - 1) / BLOCK;
    if (nb == 0) return hipSuccess;
    if (gh)
        k_selected<T, BLOCK, true><<<nb, BLOCK, 0, s>>>(a);
    else
        k_selected<T, BLOCK, false><<<nb, BLOCK, lds, s>>>(a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_selected_t(const SelectedArgs& a, int block, size_t lds, bool gh, hipStream_t s)
{
    switch (block)
    {
        case 64: return launch_selected_b<T, 64>(a, lds, gh, s);
        case 128: return launch_selected_b<T, 128>(a, lds, gh, s);
        default: return launch_selected_b<T, 256>(a, lds, gh, s);
    }
}

// xs[j] = xyz[original index of sorted point j] (the grid's order), doubles
__global__ __launch_bounds__(256) void k_sorted_copy_f64(const float4* __restrict__ pts,
                                                         const double* __restrict__ xyz, uint32_t n,
                                                         double* __restrict__ xs)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= n) return;
    const uint32_t pi = __float_as_uint(pts[j].w);
    xs[(size_t)j * 3 + 0] = xyz[(size_t)pi * 3 + 0];
    xs[(size_t)j * 3 + 1] = xyz[(size_t)pi * 3 + 1];
    xs[(size_t)j * 3 + 2] = xyz[(size_t)pi * 3 + 2];
}

__global__ __launch_bounds__(256) void k_f64_to_f32(const double* __restrict__ in, float* __restrict__ out, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        out[i] = (float)in[i];
}

// A search whose query count is only known on the device (n_query_dev) is
// launched for the largest count, n_query: every query slot has a block, and
// the blocks past the device-side count exit at once.
template <int MODE, int WAVES, bool BUDGET>
hipError_t launch_search_mb(const SearchArgs& a, size_t lds, bool gh, hipStream_t s)
{
    constexpr int BLOCK = WAVES * 64;
    const uint32_t nb = (a.n_query + BLOCK - 1) / BLOCK;
    if (nb == 0) return hipSuccess;
    if (a.heap_tail && (gh || !a.gheap_keys || (a.heap_tail == 2 && a.tail_pairs == 0))) return hipErrorInvalidValue;
    if (gh)
        k_search<MODE, WAVES, kHeapGlobal, BUDGET><<<nb, BLOCK, (size_t)WAVES * kStage * sizeof(float4), s>>>(a);
    else if (a.heap_tail)
        k_search<MODE, WAVES, kHeapTail, BUDGET><<<nb, BLOCK, lds, s>>>(a);
    else
        k_search<MODE, WAVES, kHeapLds, BUDGET><<<nb, BLOCK, lds, s>>>(a);
    return hipGetLastError();
}

template <int MODE, bool BUDGET>
hipError_t launch_search_m(const SearchArgs& a, int block, size_t lds, bool gh, hipStream_t s)
{
    switch (block)
    {
        case 64: return launch_search_mb<MODE, 1, BUDGET>(a, lds, gh, s);
        case 128: return launch_search_mb<MODE, 2, BUDGET>(a, lds, gh, s);
        default: return launch_search_mb<MODE, 4, BUDGET>(a, lds, gh, s);
    }
}

// The flagged queries (flag[i] = 1) of the first n, moved to off[i] (the
// exclusive scan of the flags over n + 1 entries, flag[n] = 0); *out_count
// receives off[n], their number.
__global__ __launch_bounds__(256) void k_compact_queries(const float4* __restrict__ q, uint32_t n_max,
                                                         const uint64_t* __restrict__ n_dev,
                                                         const uint32_t* __restrict__ flag,
                                                         const uint32_t* __restrict__ off, float4* __restrict__ out,
                                                         const float* __restrict__ bound, float* __restrict__ bound_out,
                                                         uint64_t* __restrict__ out_count)
{
    const uint64_t n = dyn_count(n_max, n_dev, 0);
    if (blockIdx.x == 0 && threadIdx.x == 0) *out_count = off[n];
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u)
        if (flag[i])
        {
            out[off[i]] = q[i];
            if (bound) bound_out[off[i]] = bound[i];
        }
}

inline uint32_t blocks_for(uint64_t n, uint32_t per, uint32_t cap)
{
    uint64_t b = (n + per - 1) / per;
    if (b < 1) b = 1;
    return (uint32_t)(b > cap ? cap : b);
}

}  // namespace

// ---------------------------------------------------------- launchers ----
hipError_t launch_bbox(const float* xyz, uint32_t n, float* partial, uint32_t n_partial, float* out6, hipStream_t s)
{
    ProfScope ps("bbox", s);
    k_bbox<<<n_partial, 256, 0, s>>>(xyz, n, partial);
    k_bbox_final<<<1, 256, 0, s>>>(partial, n_partial, out6);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s)
{
    return launch_fill_u32_dyn(p, n, nullptr, 0, v, s);
}

hipError_t launch_fill_u32_dyn(uint32_t* p, uint64_t n_max, const uint64_t* n_dev, uint64_t add, uint32_t v,
                               hipStream_t s)
{
    if (n_max == 0) return hipSuccess;
    k_fill_u32<<<blocks_for(n_max, 256, n_dev ? 4096 : 65536), 256, 0, s>>>(p, n_max, n_dev, add, v);
    return hipGetLastError();
}

// Small host arrays travel to the device as kernel arguments (16 values per
// launch), so that a call that uploads them can be captured into a graph.
struct U32x16
{
    uint32_t v[16];
};

__global__ void k_put_values(uint32_t* __restrict__ dst, U32x16 vals, uint32_t count)
{
    if (threadIdx.x != 0) return;
#pragma unroll
    for (uint32_t u = 0; u < 16; ++u)
        if (u < count) dst[u] = vals.v[u];
}

// order[t] = original index of sorted point t (the spatial order of a self
// search's rows)
__global__ __launch_bounds__(256) void k_order_from_sorted(const float4* __restrict__ pts, uint32_t n,
                                                           uint32_t* __restrict__ order)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) order[i] = __float_as_uint(pts[i].w);
}

hipError_t launch_put_values(uint32_t* dst, const uint32_t* host, uint64_t n, hipStream_t s)
{
    for (uint64_t b = 0; b < n; b += 16)
    {
        U32x16 v{};
        const uint32_t c = (uint32_t)std::min<uint64_t>(16, n - b);
        for (uint32_t u = 0; u < c; ++u) v.v[u] = host[b + u];
        k_put_values<<<1, 64, 0, s>>>(dst + b, v, c);
    }
    return hipGetLastError();
}

hipError_t launch_order_from_sorted(const float4* pts, uint32_t n, uint32_t* order, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_order_from_sorted<<<(n + 255) / 256, 256, 0, s>>>(pts, n, order);
    return hipGetLastError();
}

hipError_t launch_put_u32(uint32_t* p, const uint64_t* n_dev, uint32_t v, hipStream_t s)
{
    k_put_u32<<<1, 64, 0, s>>>(p, n_dev, v);
    return hipGetLastError();
}

hipError_t launch_plan_probe(const PlanArgs& p, LevelTable* tab, hipStream_t s)
{
    k_plan_probe<<<1, 64, 0, s>>>(p, tab);
    return hipGetLastError();
}

hipError_t launch_plan_levels(const PlanArgs& p, LevelTable* tab, hipStream_t s)
{
    k_plan_levels<<<1, 64, 0, s>>>(p, tab);
    return hipGetLastError();
}

hipError_t launch_arange_ptr(uint32_t* nn_ptr, uint32_t n_points, uint32_t k, hipStream_t s)
{
    k_arange_ptr<<<blocks_for((uint64_t)n_points + 1, 256, 65536), 256, 0, s>>>(nn_ptr, n_points, k);
    return hipGetLastError();
}

hipError_t launch_hist(const float* xyz, uint32_t n, const LevelTable* tab, int l, uint32_t* count, uint32_t* cell,
                       uint32_t* rank, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    ProfScope ps("grid_histogram", s);
    k_hist<<<(n + 256 * 16 - 1) / (256 * 16), 256, 0, s>>>(xyz, n, tab, l, count, cell, rank);
    return hipGetLastError();
}

// ---- work order of the deferred queries ----
// A deferred query's search cost grows with its bound of the k-th distance,
// and a wave runs as long as its slowest lane: the unbudgeted pass therefore
// takes the deferred queries ordered by bound, largest first (+inf, unknown
// bound, first of all).  Bucket = the bound's float bits >> 20 (exponent and
// three mantissa bits), reversed.
constexpr uint32_t kBoundBuckets = 2048;

__device__ __forceinline__ uint32_t bound_bucket(float b)
{
    const uint32_t bits = (__float_as_uint(b) & 0x7FFFFFFFu) >> 20;
    return (kBoundBuckets - 1) - min(bits, kBoundBuckets - 1);
}

__global__ __launch_bounds__(256) void k_bound_hist(const float* __restrict__ bound, uint32_t n_max,
                                                    const uint64_t* __restrict__ n_dev,
                                                    uint32_t* __restrict__ count, uint32_t* __restrict__ key,
                                                    uint32_t* __restrict__ rank)
{
    const uint64_t n = dyn_count(n_max, n_dev, 0);
    if ((uint64_t)blockIdx.x * 256 * 16 >= n) return;
    __shared__ uint32_t bins[kBoundBuckets];
    for (uint32_t c = threadIdx.x; c < kBoundBuckets; c += 256) bins[c] = 0;
    __syncthreads();
    constexpr uint32_t kPer = 16;
    const uint64_t base = (uint64_t)blockIdx.x * 256 * kPer;
    uint32_t b[kPer], lr[kPer];
#pragma unroll
    for (uint32_t u = 0; u < kPer; ++u)
    {
        const uint64_t i = base + (uint64_t)u * 256 + threadIdx.x;
        b[u] = 0;
        lr[u] = 0;
        if (i < n)
        {
            b[u] = bound_bucket(bound[i]);
            lr[u] = atomicAdd(&bins[b[u]], 1u);
        }
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < kBoundBuckets; c += 256)
    {
        const uint32_t v = bins[c];
        bins[c] = v ? atomicAdd(&count[c], v) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < kPer; ++u)
    {
        const uint64_t i = base + (uint64_t)u * 256 + threadIdx.x;
        if (i < n)
        {
            key[i] = b[u];
            rank[i] = bins[b[u]] + lr[u];
        }
    }
}

__global__ __launch_bounds__(256) void k_bound_scatter(const float4* __restrict__ q, const float* __restrict__ bound,
                                                       uint32_t n_max, const uint64_t* __restrict__ n_dev,
                                                       const uint32_t* __restrict__ key,
                                                       const uint32_t* __restrict__ rank,
                                                       const uint32_t* __restrict__ start, float4* __restrict__ q_out,
                                                       float* __restrict__ bound_out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= dyn_count(n_max, n_dev, 0)) return;
    const uint32_t p = start[key[i]] + rank[i];
    q_out[p] = q[i];
    bound_out[p] = bound[i];
}

hipError_t launch_order_by_bound(const float4* q, const float* bound, uint32_t n_max, const uint64_t* n_dev,
                                 uint32_t* count, uint32_t* key, uint32_t* rank, uint32_t* scan_part, float4* q_out,
                                 float* bound_out, hipStream_t s)
{
    if (n_max == 0) return hipSuccess;
    ProfScope ps("knn_defer_order", s);
    hipError_t e = launch_fill_u32(count, kBoundBuckets + 1, 0u, s);
    if (e != hipSuccess) return e;
    k_bound_hist<<<(n_max + 256 * 16 - 1) / (256 * 16), 256, 0, s>>>(bound, n_max, n_dev, count, key, rank);
    e = launch_scan_exclusive(count, count, kBoundBuckets + 1, scan_part, s);
    if (e != hipSuccess) return e;
    k_bound_scatter<<<(n_max + 255) / 256, 256, 0, s>>>(q, bound, n_max, n_dev, key, rank, count, q_out, bound_out);
    return hipGetLastError();
}

hipError_t launch_compact_queries(const float4* q, uint32_t n_max, const uint64_t* n_dev, const uint32_t* flag,
                                  const uint32_t* off, float4* out, const float* bound, float* bound_out,
                                  uint64_t* out_count, hipStream_t s)
{
    k_compact_queries<<<blocks_for(n_max, 256, 16384), 256, 0, s>>>(q, n_max, n_dev, flag, off, out, bound,
                                                                     bound_out, out_count);
    return hipGetLastError();
}

size_t scan_partials_needed(uint64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

hipError_t launch_scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* partials, hipStream_t s)
{
    return launch_scan_exclusive_dyn(in, out, n, nullptr, 0, partials, s);
}

hipError_t launch_scan_exclusive_dyn(const uint32_t* in, uint32_t* out, uint64_t n_max, const uint64_t* n_dev,
                                     uint64_t add, uint32_t* partials, hipStream_t s)
{
    if (n_max == 0) return hipSuccess;
    ProfScope ps("grid_scan", s);
    const uint64_t nt = (n_max + kScanTile - 1) / kScanTile;
    const uint32_t nb = (uint32_t)std::min<uint64_t>(nt, 2048);
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    if (vec)
        k_scan_reduce<true><<<nb, kScanBlock, 0, s>>>(in, n_max, n_dev, add, partials);
    else
        k_scan_reduce<false><<<nb, kScanBlock, 0, s>>>(in, n_max, n_dev, add, partials);
    k_scan_partials<<<1, kScanBlock, 0, s>>>(partials, n_max, n_dev, add);
    if (vec)
        k_scan_apply<true><<<nb, kScanBlock, 0, s>>>(in, out, n_max, n_dev, add, partials);
    else
        k_scan_apply<false><<<nb, kScanBlock, 0, s>>>(in, out, n_max, n_dev, add, partials);
    return hipGetLastError();
}

hipError_t launch_probe(const float* xyz, uint32_t n, uint32_t stride, LevelTable* tab, uint32_t* count,
                        uint64_t count_room, hipStream_t s)
{
    const uint32_t ns = (n + stride - 1) / stride;
    if (ns == 0) return hipSuccess;
    ProfScope ps("grid_probe", s);
    hipError_t e = launch_fill_u32_dyn(count, count_room, &tab->probe_G, 0, 0u, s);
    if (e != hipSuccess) return e;
    k_hist_probe<<<(ns + 255) / 256, 256, 0, s>>>(xyz, n, stride, tab, count);
    k_sumsq<<<blocks_for(count_room, 256, 2048), 256, 0, s>>>(count, tab);
    return hipGetLastError();
}

hipError_t launch_hist_sorted(const float4* pts, uint32_t n, const LevelTable* tab, int l, uint32_t* count,
                              uint32_t* cell, uint32_t* rank, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    ProfScope ps("grid_histogram_sorted", s);
    k_hist_sorted<<<(n + 255) / 256, 256, 0, s>>>(pts, n, tab, l, count, cell, rank);
    return hipGetLastError();
}

hipError_t launch_scatter_sorted(const float4* in, uint32_t n, const uint32_t* cell, const uint32_t* rank,
                                 const uint32_t* cell_start, float4* out, const LevelTable* tab, int l,
                                 hipStream_t s)
{
    if (n == 0) return hipSuccess;
    ProfScope ps("grid_scatter", s);
    k_scatter_sorted<<<(n + 255) / 256, 256, 0, s>>>(in, n, cell, rank, cell_start, out, tab, l);
    return hipGetLastError();
}

hipError_t launch_scatter(const float* xyz, uint32_t n, const uint32_t* cell, const uint32_t* rank,
                          const uint32_t* cell_start, float4* pts, const LevelTable* tab, int l, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    ProfScope ps("grid_scatter", s);
    k_scatter<<<(n + 255) / 256, 256, 0, s>>>(xyz, n, cell, rank, cell_start, pts, tab, l);
    return hipGetLastError();
}

hipError_t launch_compact_rows(const uint32_t* dense, uint32_t n_rows, uint32_t cap, const uint32_t* counts,
                               const uint32_t* row_ptr, uint32_t* nn, hipStream_t s)
{
    if ((uint64_t)n_rows * cap == 0) return hipSuccess;
    ProfScope ps("radius_compact", s);
    k_compact_rows<<<blocks_for((uint64_t)n_rows * 16, 256, 65536), 256, 0, s>>>(dense, n_rows, cap, counts,
                                                                                row_ptr, nn);
    return hipGetLastError();
}

hipError_t launch_selected(bool f64, const SelectedArgs& a, int block, size_t lds, bool gh, hipStream_t s)
{
    ProfScope ps(f64 ? "features_selected_f64" : "features_selected_f32", s);
    return f64 ? launch_selected_t<double>(a, block, lds, gh, s) : launch_selected_t<float>(a, block, lds, gh, s);
}

hipError_t launch_sorted_copy_f64(const float4* pts, const double* xyz, uint32_t n, double* xs, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_sorted_copy_f64<<<(n + 255) / 256, 256, 0, s>>>(pts, xyz, n, xs);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double* in, float* out, uint64_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_f64_to_f32<<<blocks_for(n, 256, 65536), 256, 0, s>>>(in, out, n);
    return hipGetLastError();
}

hipError_t launch_search(int mode, const SearchArgs& a, int block, size_t lds, bool gh, bool budget, hipStream_t s)
{
    static const char* names[5] = {"knn_search", "radius_search", "knn_features_fused", "radius_search_csr",
                                   "radius_multiscale_fused"};
    static const char* names_b[5] = {"knn_search_dense", "radius_search", "knn_features_fused_dense",
                                     "radius_search_csr", "radius_multiscale_fused"};
    if (mode < 0 || mode > 4) return hipErrorInvalidValue;
    ProfScope ps(budget ? names_b[mode] : names[mode], s);
    if (budget)
    {
        switch (mode)
        {
            case kModeKnn: return launch_search_m<kModeKnn, true>(a, block, lds, gh, s);
            case kModeKnnFeat: return launch_search_m<kModeKnnFeat, true>(a, block, lds, gh, s);
            default: return hipErrorInvalidValue;
        }
    }
    switch (mode)
    {
        case kModeKnn: return launch_search_m<kModeKnn, false>(a, block, lds, gh, s);
        case kModeRadius: return launch_search_m<kModeRadius, false>(a, block, lds, gh, s);
        case kModeKnnFeat: return launch_search_m<kModeKnnFeat, false>(a, block, lds, gh, s);
        case kModeRadiusMulti: return launch_search_m<kModeRadiusMulti, false>(a, block, lds, gh, s);
        default: return launch_search_m<kModeRadiusCsr, false>(a, block, lds, gh, s);
    }
}

}  // namespace pgeof_amd
