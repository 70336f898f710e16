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
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) p[i] = v;
}

__global__ __launch_bounds__(256) void k_arange_ptr(uint32_t* __restrict__ p, uint32_t n_points, uint32_t k)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i <= n_points; i += (uint64_t)gridDim.x * 256)
        p[i] = (uint32_t)(i * k);
}

// ---------------------------------------------------------- histogram ----
__global__ __launch_bounds__(256) void k_hist(const float* __restrict__ xyz, uint32_t n, GridParams gp,
                                              uint32_t* __restrict__ count, uint32_t* __restrict__ cell,
                                              uint32_t* __restrict__ rank)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int cx = cell_coord((double)xyz[i * 3 + 0], gp.lo[0], gp.inv_h, gp.g[0]);
    const int cy = cell_coord((double)xyz[i * 3 + 1], gp.lo[1], gp.inv_h, gp.g[1]);
    const int cz = cell_coord((double)xyz[i * 3 + 2], gp.lo[2], gp.inv_h, gp.g[2]);
    const uint32_t cid = ((uint32_t)cz * (uint32_t)gp.g[1] + (uint32_t)cy) * (uint32_t)gp.g[0] + (uint32_t)cx;
    const uint32_t r = atomicAdd(&count[cid], 1u);
    if (cell) cell[i] = cid;
    if (rank) rank[i] = r;
}

__global__ __launch_bounds__(256) void k_sumsq(const uint32_t* __restrict__ count, uint64_t G,
                                               unsigned long long* __restrict__ out)
{
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

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                            uint32_t* __restrict__ partials)
{
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; ++j)
    {
        const uint64_t i = base + (uint64_t)j * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t total;
    block_exclusive_scan(s, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_partials(uint32_t* __restrict__ partials, uint32_t nt)
{
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

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const uint32_t* in, uint32_t* out, uint64_t n,
                                                           const uint32_t* __restrict__ partials)
{
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t loc = 0;
    for (int j = 0; j < kScanItems; ++j)
    {
        const uint64_t i = base + j;
        v[j] = i < n ? in[i] : 0u;
        loc += v[j];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(loc, total) + partials[blockIdx.x];
    for (int j = 0; j < kScanItems; ++j)
    {
        const uint64_t i = base + j;
        if (i < n) out[i] = ex;
        ex += v[j];
    }
}

// ------------------------------------------------------------ scatter ----
__global__ __launch_bounds__(256) void k_scatter(const float* __restrict__ xyz, uint32_t n,
                                                 const uint32_t* __restrict__ cell, const uint32_t* __restrict__ rank,
                                                 const uint32_t* __restrict__ cell_start, float4* __restrict__ pts)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t pos = cell_start[cell[i]] + rank[i];
    pts[pos] = make_float4(xyz[i * 3 + 0], xyz[i * 3 + 1], xyz[i * 3 + 2], __uint_as_float((uint32_t)i));
}

__global__ __launch_bounds__(256) void k_compact_rows(const uint32_t* __restrict__ dense, uint32_t n_rows,
                                                      uint32_t cap, const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ row_ptr,
                                                      uint32_t* __restrict__ nn)
{
    const uint64_t total = (uint64_t)n_rows * cap;
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256)
    {
        const uint32_t r = (uint32_t)(e / cap), j = (uint32_t)(e % cap);
        if (j < counts[r]) nn[row_ptr[r] + j] = dense[e];
    }
}

// --------------------------------------------------------------- heap ----
// Max-heap of 64-bit keys (squared-distance bits << 32 | original index);
// element j lives at keys[j * stride].  With POS, the sorted-array position
// of each entry travels with its key.
template <bool POS>
struct Heap
{
    uint64_t* keys;
    uint32_t* pos;
    uint32_t stride;
    uint32_t k;

    __device__ __forceinline__ void init(uint64_t sentinel)
    {
        for (uint32_t j = 0; j < k; ++j)
        {
            keys[j * stride] = sentinel;
            if (POS) pos[j * stride] = 0xFFFFFFFFu;
        }
    }
    __device__ __forceinline__ uint64_t top() const { return keys[0]; }

    // put `key` at the root of heap [0, n) and restore the heap property
    __device__ __forceinline__ void sift_root(uint64_t key, uint32_t p, uint32_t n)
    {
        uint32_t i = 0;
        for (;;)
        {
            const uint32_t l = 2 * i + 1;
            if (l >= n) break;
            uint64_t ck = keys[l * stride];
            uint32_t c = l;
            if (l + 1 < n)
            {
                const uint64_t rk = keys[(l + 1) * stride];
                if (rk > ck)
                {
                    ck = rk;
                    c = l + 1;
                }
            }
            if (ck <= key) break;
            keys[i * stride] = ck;
            if (POS) pos[i * stride] = pos[c * stride];
            i = c;
        }
        keys[i * stride] = key;
        if (POS) pos[i * stride] = p;
    }

    __device__ __forceinline__ void sort_ascending()
    {
        for (uint32_t end = k - 1; end > 0; --end)
        {
            const uint64_t top_k = keys[0];
            const uint64_t last = keys[end * stride];
            uint32_t tp = 0, lp = 0;
            if (POS)
            {
                tp = pos[0];
                lp = pos[end * stride];
            }
            keys[end * stride] = top_k;
            if (POS) pos[end * stride] = tp;
            sift_root(last, lp, end);
        }
    }
};

// distance from v to the slab [lo + c h, lo + (c+1) h], less a safety margin
__device__ __forceinline__ double axis_gap(double v, int c, double lo, double h, double margin)
{
    const double a = lo + (double)c * h, b = a + h;
    const double g = fmax(fmax(a - v, v - b), 0.0) - margin;
    return g > 0.0 ? g : 0.0;
}

// ------------------------------------------------------------- search ----
template <int MODE, int BLOCK, bool GLOBAL_HEAP>
__global__ __launch_bounds__(BLOCK) void k_search(SearchArgs a)
{
    constexpr bool POS = (MODE == kModeKnnFeat);
    constexpr bool RADIUS = (MODE == kModeRadius || MODE == kModeRadiusCsr);
    extern __shared__ __align__(16) unsigned char smem[];

    const uint32_t lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t t = lb * BLOCK + threadIdx.x;
    if (t >= a.n_query) return;

    Heap<POS> heap;
    if (GLOBAL_HEAP)
    {
        heap.keys = a.gheap_keys + t;
        heap.pos = a.gheap_pos + t;
        heap.stride = a.n_query;
    }
    else
    {
        heap.keys = reinterpret_cast<uint64_t*>(smem) + threadIdx.x;
        heap.pos = reinterpret_cast<uint32_t*>(smem + (size_t)a.k * BLOCK * sizeof(uint64_t)) + threadIdx.x;
        heap.stride = BLOCK;
    }
    heap.k = a.k;

    const float4 q4 = a.queries[t];
    const uint32_t qorig = __float_as_uint(q4.w);
    const uint64_t sentinel = RADIUS ? ((uint64_t)__float_as_uint(a.r2) << 32) : ~0ull;
    heap.init(sentinel);
    uint64_t worst = sentinel;

    const GridParams& gp = a.grid.p;
    const float4* __restrict__ pts = a.grid.pts;
    const uint32_t* __restrict__ cs = a.grid.cell_start;
    const bool finite = isfinite(q4.x) && isfinite(q4.y) && isfinite(q4.z);

    if (finite)
    {
        const float qxf = q4.x, qyf = q4.y, qzf = q4.z;
        auto scan = [&](uint32_t s, uint32_t e) {
            for (uint32_t j = s; j < e; ++j)
            {
                const float4 p = pts[j];
                const float d2 = sqdist32(qxf, qyf, qzf, p.x, p.y, p.z);
                const uint64_t key = ((uint64_t)__float_as_uint(d2) << 32) | (uint64_t)__float_as_uint(p.w);
                if (key < worst)
                {
                    heap.sift_root(key, j, heap.k);
                    worst = heap.top();
                }
            }
        };
        const uint32_t gx = (uint32_t)gp.g[0], gy = (uint32_t)gp.g[1];
        auto row = [&](int y, int z, int x0, int x1) {
            const uint32_t base = ((uint32_t)z * gy + (uint32_t)y) * gx;
            scan(cs[base + (uint32_t)x0], cs[base + (uint32_t)x1 + 1]);
        };
        const double qx = qxf, qy = qyf, qz = qzf;
        const int G0 = gp.g[0], G1 = gp.g[1], G2 = gp.g[2];
        const int cx = cell_coord(qx, gp.lo[0], gp.inv_h, G0);
        const int cy = cell_coord(qy, gp.lo[1], gp.inv_h, G1);
        const int cz = cell_coord(qz, gp.lo[2], gp.inv_h, G2);

        // Phase 1 (kNN): scan cubic shells around the query's cell until k
        // candidates are held.
        int n_done = -1;
        bool covered = false;
        if (!RADIUS)
        {
            for (int n = 0;; ++n)
            {
                const int z0 = max(cz - n, 0), z1 = min(cz + n, G2 - 1);
                const int y0 = max(cy - n, 0), y1 = min(cy + n, G1 - 1);
                const int x0 = max(cx - n, 0), x1 = min(cx + n, G0 - 1);
                for (int z = z0; z <= z1; ++z)
                {
                    const bool ze = (z == cz - n) || (z == cz + n);
                    for (int y = y0; y <= y1; ++y)
                    {
                        const bool ye = (y == cy - n) || (y == cy + n);
                        if (ze || ye)
                            row(y, z, x0, x1);
                        else
                        {
                            if (cx - n >= 0) row(y, z, cx - n, cx - n);
                            if (cx + n < G0) row(y, z, cx + n, cx + n);
                        }
                    }
                }
                n_done = n;
                covered = (cx - n <= 0) && (cx + n >= G0 - 1) && (cy - n <= 0) && (cy + n >= G1 - 1) &&
                          (cz - n <= 0) && (cz + n >= G2 - 1);
                if (worst != sentinel || covered) break;
            }
        }

        // Phase 2: every cell row that intersects the ball of the current
        // k-th distance (or of the search radius), minus the phase-1 cube.
        if (!covered)
        {
            const double h = gp.h, margin = (double)a.margin;
            auto wlim = [&]() -> double {
                if (worst == ~0ull) return INFINITY;
                const float w = __uint_as_float((uint32_t)(worst >> 32));
                if (!(w < INFINITY)) return INFINITY;
                return (double)w * 1.00001 + 1e-37;
            };
            double wl = wlim();
            int tx0 = 0, tx1 = G0 - 1, ty0 = 0, ty1 = G1 - 1, tz0 = 0, tz1 = G2 - 1;
            if (wl < INFINITY)
            {
                const double R = sqrt(wl) + margin;
                tx0 = cell_coord(qx - R, gp.lo[0], gp.inv_h, G0);
                tx1 = cell_coord(qx + R, gp.lo[0], gp.inv_h, G0);
                ty0 = cell_coord(qy - R, gp.lo[1], gp.inv_h, G1);
                ty1 = cell_coord(qy + R, gp.lo[1], gp.inv_h, G1);
                tz0 = cell_coord(qz - R, gp.lo[2], gp.inv_h, G2);
                tz1 = cell_coord(qz + R, gp.lo[2], gp.inv_h, G2);
            }
            const int ix0 = n_done >= 0 ? max(cx - n_done, 0) : 1, ix1 = n_done >= 0 ? min(cx + n_done, G0 - 1) : 0;
            for (int z = tz0; z <= tz1; ++z)
            {
                const double dz = axis_gap(qz, z, gp.lo[2], h, margin);
                wl = wlim();
                if (dz * dz > wl) continue;
                const bool zin = n_done >= 0 && abs(z - cz) <= n_done;
                for (int y = ty0; y <= ty1; ++y)
                {
                    const double dy = axis_gap(qy, y, gp.lo[1], h, margin);
                    wl = wlim();
                    const double rem = wl - dz * dz - dy * dy;
                    if (rem < 0.0) continue;
                    int xa = tx0, xb = tx1;
                    if (rem < INFINITY)
                    {
                        const double xr = sqrt(rem) + margin;
                        xa = max(xa, cell_coord(qx - xr, gp.lo[0], gp.inv_h, G0));
                        xb = min(xb, cell_coord(qx + xr, gp.lo[0], gp.inv_h, G0));
                    }
                    if (xa > xb) continue;
                    if (zin && abs(y - cy) <= n_done)
                    {
                        if (xa < ix0) row(y, z, xa, min(xb, ix0 - 1));
                        if (xb > ix1) row(y, z, max(xa, ix1 + 1), xb);
                    }
                    else
                        row(y, z, xa, xb);
                }
            }
        }
    }

    heap.sort_ascending();
    const uint32_t k = a.k;

    if (MODE == kModeKnn)
    {
        uint32_t* oi = reinterpret_cast<uint32_t*>(a.out_idx) + (size_t)qorig * k;
        float* od = a.out_d2 ? a.out_d2 + (size_t)qorig * k : nullptr;
        for (uint32_t j = 0; j < k; ++j)
        {
            const uint64_t key = heap.keys[j * heap.stride];
            oi[j] = (uint32_t)key;
            if (od) od[j] = key == ~0ull ? INFINITY : __uint_as_float((uint32_t)(key >> 32));
        }
    }
    else if (MODE == kModeRadius || MODE == kModeRadiusCsr)
    {
        uint32_t cnt = 0;
        if (MODE == kModeRadius)
        {
            int32_t* oi = reinterpret_cast<int32_t*>(a.out_idx) + (size_t)qorig * k;
            float* od = a.out_d2 + (size_t)qorig * k;
            for (uint32_t j = 0; j < k; ++j)
            {
                const uint64_t key = heap.keys[j * heap.stride];
                if (key != sentinel)
                {
                    oi[j] = (int32_t)(uint32_t)key;
                    od[j] = __uint_as_float((uint32_t)(key >> 32));
                    ++cnt;
                }
                else
                {
                    oi[j] = -1;
                    od[j] = 0.0f;
                }
            }
            // nanoflann's RKNNResultSet::init stores r^2 in the last slot
            if (cnt < k) od[k - 1] = a.r2;
        }
        else
        {
            uint32_t* oi = reinterpret_cast<uint32_t*>(a.out_idx) + (size_t)qorig * k;
            for (uint32_t j = 0; j < k; ++j)
            {
                const uint64_t key = heap.keys[j * heap.stride];
                if (key != sentinel) oi[cnt++] = (uint32_t)key;
            }
        }
        if (a.out_count) a.out_count[qorig] = cnt;
    }
    else  // kModeKnnFeat
    {
        uint32_t* oi = reinterpret_cast<uint32_t*>(a.out_idx) + (size_t)qorig * k;
        float* od = a.out_d2 ? a.out_d2 + (size_t)qorig * k : nullptr;
        float f[11];
        if (finite)
        {
            Moments m;
            mom_zero(m);
            double ox = 0.0, oy = 0.0, oz = 0.0;
            for (uint32_t j = 0; j < k; ++j)
            {
                const uint64_t key = heap.keys[j * heap.stride];
                const uint32_t p = heap.pos[j * heap.stride];
                const float4 pt = pts[p];
                if (j == 0) ox = pt.x, oy = pt.y, oz = pt.z;
                mom_add(m, (double)pt.x - ox, (double)pt.y - oy, (double)pt.z - oz);
                oi[j] = (uint32_t)key;
                if (od) od[j] = __uint_as_float((uint32_t)(key >> 32));
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
        float* of = a.out_feat + (size_t)qorig * 11;
        for (int c = 0; c < 11; ++c) of[c] = f[c];
    }
}

template <int MODE, int BLOCK>
hipError_t launch_search_mb(const SearchArgs& a, size_t lds, bool gh, hipStream_t s)
{
    const uint32_t nb = (a.n_query + BLOCK - 1) / BLOCK;
    if (nb == 0) return hipSuccess;
    if (gh)
        k_search<MODE, BLOCK, true><<<nb, BLOCK, 0, s>>>(a);
    else
        k_search<MODE, BLOCK, false><<<nb, BLOCK, lds, s>>>(a);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_search_m(const SearchArgs& a, int block, size_t lds, bool gh, hipStream_t s)
{
    switch (block)
    {
        case 64: return launch_search_mb<MODE, 64>(a, lds, gh, s);
        case 128: return launch_search_mb<MODE, 128>(a, lds, gh, s);
        default: return launch_search_mb<MODE, 256>(a, lds, gh, s);
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
    if (n == 0) return hipSuccess;
    k_fill_u32<<<blocks_for(n, 256, 65536), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_arange_ptr(uint32_t* nn_ptr, uint32_t n_points, uint32_t k, hipStream_t s)
{
    k_arange_ptr<<<blocks_for((uint64_t)n_points + 1, 256, 65536), 256, 0, s>>>(nn_ptr, n_points, k);
    return hipGetLastError();
}

hipError_t launch_hist(const float* xyz, uint32_t n, const GridParams& gp, uint32_t* count, uint32_t* cell,
                       uint32_t* rank, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    ProfScope ps("grid_histogram", s);
    k_hist<<<(n + 255) / 256, 256, 0, s>>>(xyz, n, gp, count, cell, rank);
    return hipGetLastError();
}

hipError_t launch_sumsq(const uint32_t* count, uint64_t G, unsigned long long* out, hipStream_t s)
{
    k_sumsq<<<blocks_for(G, 256, 2048), 256, 0, s>>>(count, G, out);
    return hipGetLastError();
}

size_t scan_partials_needed(uint64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

hipError_t launch_scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* partials, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    ProfScope ps("grid_scan", s);
    const uint32_t nt = (uint32_t)((n + kScanTile - 1) / kScanTile);
    k_scan_reduce<<<nt, kScanBlock, 0, s>>>(in, n, partials);
    k_scan_partials<<<1, kScanBlock, 0, s>>>(partials, nt);
    k_scan_apply<<<nt, kScanBlock, 0, s>>>(in, out, n, partials);
    return hipGetLastError();
}

hipError_t launch_scatter(const float* xyz, uint32_t n, const uint32_t* cell, const uint32_t* rank,
                          const uint32_t* cell_start, float4* pts, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    ProfScope ps("grid_scatter", s);
    k_scatter<<<(n + 255) / 256, 256, 0, s>>>(xyz, n, cell, rank, cell_start, pts);
    return hipGetLastError();
}

hipError_t launch_compact_rows(const uint32_t* dense, uint32_t n_rows, uint32_t cap, const uint32_t* counts,
                               const uint32_t* row_ptr, uint32_t* nn, hipStream_t s)
{
    if ((uint64_t)n_rows * cap == 0) return hipSuccess;
    ProfScope ps("radius_compact", s);
    k_compact_rows<<<blocks_for((uint64_t)n_rows * cap, 256, 65536), 256, 0, s>>>(dense, n_rows, cap, counts,
                                                                                 row_ptr, nn);
    return hipGetLastError();
}

hipError_t launch_search(int mode, const SearchArgs& a, int block, size_t lds, bool gh, hipStream_t s)
{
    static const char* names[4] = {"knn_search", "radius_search", "knn_features_fused", "radius_search_csr"};
    ProfScope ps(names[mode & 3], s);
    switch (mode)
    {
        case kModeKnn: return launch_search_m<kModeKnn>(a, block, lds, gh, s);
        case kModeRadius: return launch_search_m<kModeRadius>(a, block, lds, gh, s);
        case kModeKnnFeat: return launch_search_m<kModeKnnFeat>(a, block, lds, gh, s);
        default: return launch_search_m<kModeRadiusCsr>(a, block, lds, gh, s);
    }
}

}  // namespace pgeof_amd
