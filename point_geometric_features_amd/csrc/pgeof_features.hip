// Per-point geometric features over a CSR neighbourhood list, for gfx950.
//
// One thread per point walks its CSR row once, accumulating FP64 moments of
// the neighbours shifted by the row's first neighbour.  Because the moments
// of every prefix of the row are available during that single walk:
//   - plain      (include/pgeof.hpp:75-117)  one eigensolve at k_nn
//   - multiscale (include/pgeof.hpp:159-211) one eigensolve per scale, no
//                re-gathering of the prefix
//   - optimal    (include/pgeof.hpp:243-310) eigenvalues at every candidate k
//                (Weinmann et al. 2015 eigentropy), full PCA only at k_opt
// Every neighbour index and every nn_ptr entry is bounds-checked; a bad value
// zeroes the row and raises a bit in the error word instead of reading out of
// bounds (the reference leaves this undefined, pca.hpp:103-104).
#include <algorithm>

#include "pgeof_device.hpp"
#include "pgeof_internal.hpp"

namespace pgeof_amd
{

namespace
{

enum
{
    kCsrPlain = 0,
    kCsrMulti = 1,
    kCsrOptimal = 2
};

__device__ __forceinline__ void features_of(const Moments& m, uint32_t k, float f[11])
{
    double c[6];
    mom_cov(m, (double)k, c);
    const Eig E = eig_full(c);
    const Pca32 P = to_pca32(E);
    features11(P, f);
}

// Neighbours gathered per chunk: the plain kernel's rows are short and its
// per-neighbour work light, so it takes whole rows of up to 32 at once; the
// multiscale and optimal kernels evaluate features or eigentropies inside
// the loop and keep their registers for that (measured best: 32 and 8).
#ifndef PGEOF_CSR_CHUNK
#define PGEOF_CSR_CHUNK 32
#endif
#ifndef PGEOF_CSR_WAVES_PER_SIMD
#define PGEOF_CSR_WAVES_PER_SIMD 1
#endif
constexpr uint32_t kCsrChunkPlain = PGEOF_CSR_CHUNK;
constexpr uint32_t kCsrChunkScales = 8;

template <int MODE>
__global__ __launch_bounds__(256, PGEOF_CSR_WAVES_PER_SIMD) void k_csr(CsrArgs a)
{
    // consecutive blocks of rows (neighbours in space when an order is
    // given) run on the same XCD and share its L2
    const uint32_t t = xcd_logical_block(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
    if (t >= a.n_points) return;
    // rows are visited in the spatial order of their points when an
    // order is given, so that a wave's gathers share cache lines
    const uint32_t i = a.order ? a.order[t] : t;
    const uint32_t b = a.nn_ptr[i], e = a.nn_ptr[i + 1];
    const uint32_t nf = MODE == kCsrOptimal ? 12u : 11u;
    const uint32_t n_out = MODE == kCsrMulti ? a.n_scales * 11u : nf;
    float* __restrict__ out = a.out + (size_t)i * n_out;

    if (e < b || (uint64_t)e > a.n_nn)
    {
        atomicOr(a.err, kErrNnPtr);
        for (uint32_t c = 0; c < n_out; ++c) out[c] = 0.0f;
        return;
    }
    const uint32_t k_nn = e - b;
    const uint32_t* __restrict__ row = a.nn + b;
    const float* __restrict__ xyz = a.xyz;

    // number of neighbours this row needs to walk
    uint32_t walk = 0;
    uint32_t k0 = 0;
    if (MODE == kCsrPlain)
        walk = (k_nn >= a.k_min) ? k_nn : 0;
    else if (MODE == kCsrMulti)
    {
        for (uint32_t s = 0; s < a.n_scales; ++s)
        {
            if (a.scales[s] > k_nn) break;
            walk = a.scales[s];
        }
    }
    else
    {
        if (k_nn >= a.k_min && k_nn >= a.k_min_search)
        {
            k0 = max(a.k_min, a.k_min_search);
            k0 = min(k0, k_nn);
            walk = k_nn;
        }
    }

    float f[12];
    for (int c = 0; c < 12; ++c) f[c] = 0.0f;
    if (walk == 0)
    {
        for (uint32_t c = 0; c < n_out; ++c) out[c] = 0.0f;
        return;
    }

    // One pass over the row: indices are checked as they are read (a bad one
    // is replaced by 0 so every gather stays in bounds, and the row is zeroed
    // at the end); coordinates are gathered as one 12-byte load each, a chunk
    // of neighbours at a time.
    if (a.n_xyz == 0)  // every index is out of range and there is nothing to gather
    {
        atomicOr(a.err, kErrNnIndex);
        for (uint32_t c = 0; c < n_out; ++c) out[c] = 0.0f;
        return;
    }
    const float3* __restrict__ xyz3 = reinterpret_cast<const float3*>(xyz);
    bool bad = row[0] >= a.n_xyz;
    const uint32_t i0 = bad ? 0u : row[0];
    const float3 p0 = xyz3[i0];
    const double ox = p0.x, oy = p0.y, oz = p0.z;
    Moments m, best;
    mom_zero(m);
    mom_zero(best);
    uint32_t s = 0;
    float h_best = 1.0f;
    uint32_t k_opt = k_nn;
    // The indices of the next chunk are loaded before the current chunk is
    // accumulated, so their loads overlap the gathers and the FP64 work
    // (10 M rows of 20: 5.5 ms with chunks of 4 and no look-ahead, 3.9 ms
    // with one chunk of 32).
    constexpr uint32_t kChunk = MODE == kCsrPlain ? kCsrChunkPlain : kCsrChunkScales;
    uint32_t nid[kChunk];
    // a row that starts on a 16-byte boundary (every row of a kNN CSR with k
    // a multiple of 4) is read four indices per load
    const bool row16 = (reinterpret_cast<uintptr_t>(row) & 15u) == 0;
    auto load_idx = [&](uint32_t j0) {
        static_assert(kChunk % 4 == 0, "chunks are read in groups of four");
#pragma unroll
        for (uint32_t u = 0; u < kChunk; u += 4)
        {
            uint32_t v[4];
            if (row16 && j0 + u + 3 < walk)
            {
                const uint4 q = *reinterpret_cast<const uint4*>(row + j0 + u);
                v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
            }
            else
            {
#pragma unroll
                for (uint32_t w = 0; w < 4; ++w) v[w] = (j0 + u + w < walk) ? row[j0 + u + w] : i0;
            }
#pragma unroll
            for (uint32_t w = 0; w < 4; ++w)
            {
                bad |= v[w] >= a.n_xyz;
                nid[u + w] = v[w] < a.n_xyz ? v[w] : 0u;
            }
        }
    };
    load_idx(0);
    for (uint32_t j0 = 0; j0 < walk; j0 += kChunk)
    {
        float px[kChunk], py[kChunk], pz[kChunk];
#pragma unroll
        for (uint32_t u = 0; u < kChunk; ++u)
        {
            const float3 p = xyz3[nid[u]];
            px[u] = p.x;
            py[u] = p.y;
            pz[u] = p.z;
        }
        if (j0 + kChunk < walk) load_idx(j0 + kChunk);
#pragma unroll
        for (uint32_t u = 0; u < kChunk; ++u)
        {
            if (j0 + u >= walk) break;
            mom_add(m, (double)px[u] - ox, (double)py[u] - oy, (double)pz[u] - oz);
            const uint32_t cnt = j0 + u + 1;
            if (MODE == kCsrMulti)
            {
                while (s < a.n_scales && a.scales[s] == cnt)
                {
                    features_of(m, cnt, f);
                    store_row11(out + (size_t)s * 11, f);
                    ++s;
                }
            }
            else if (MODE == kCsrOptimal)
            {
                if (cnt >= k0 && (cnt == k0 || cnt % a.k_step == 0 || cnt == k_nn))
                {
                    double c6[6], l[3];
                    mom_cov(m, (double)cnt, c6);
                    eig_vals_only(c6, l);
                    const float hval = eigentropy32(l);
                    if (cnt == k0 || hval < h_best)
                    {
                        h_best = hval;
                        k_opt = cnt;
                        best = m;
                    }
                }
            }
        }
    }
    if (bad)
    {
        atomicOr(a.err, kErrNnIndex);
        for (uint32_t c = 0; c < n_out; ++c) out[c] = 0.0f;
        return;
    }
    if (MODE == kCsrPlain)
    {
        features_of(m, walk, f);
        store_row11(out, f);
    }
    else if (MODE == kCsrMulti)
    {
        for (uint32_t c = s * 11; c < n_out; ++c) out[c] = 0.0f;
    }
    else
    {
        // rows of 12 floats are 16-byte aligned: three 16-byte stores
        features_of(best, k_opt, f);
        float4* o4 = reinterpret_cast<float4*>(out);
        o4[0] = make_float4(f[0], f[1], f[2], f[3]);
        o4[1] = make_float4(f[4], f[5], f[6], f[7]);
        o4[2] = make_float4(f[8], f[9], f[10], (float)k_opt);
    }
}

// ---- spatial order of CSR rows ----
// Key of a row: a Morton-like code of kOrderBits bits, in the bounding box of
// xyz (read from device memory, no host round trip), of the row's own point
// (row i of a pgeof CSR is point i, whose first neighbour in a search of the
// cloud against itself is the point itself), read as a stream; a row past the
// end of xyz takes its first neighbour instead, and an empty or invalid one 0
// (the feature kernel reports it).  The bits are handed out one at a time to
// the axis whose cells are currently the longest, so the finest cells are
// about cubic whatever the box's shape: a flat scene (200 m x 200 m x 20 m)
// gets most bits in x and y.  Equal bits per axis would leave the ground of
// such a scene in one z slice, with hundreds of rows per bucket and the
// histogram's atomics serialised on them.
//
// The rows are sorted by key in two counting passes, most significant bits
// first, with no global atomics: rows are ranked within blocks of
// kOrderRows by the key's high kOrderBits - kLoBits bits (LDS atomics), the
// blocks' counts per high bucket scanned bucket-major give every row its
// place among the rows of its high bucket, and one workgroup per high
// bucket then sorts those rows by the low kLoBits bits in LDS.  (A single
// pass over 2^21 buckets took one returning global atomic per row: 0.45 ms
// of the order's 0.76 ms per 10 M rows.)
constexpr int kOrderBits = 20;  // 2^20 buckets
constexpr int kLoBits = 10;
constexpr uint32_t kHiBins = 1u << (kOrderBits - kLoBits), kLoBins = 1u << kLoBits;
constexpr uint32_t kOrderBlock = 1024;  // threads of the two LDS passes (== kLoBins)
constexpr uint32_t kOrderPer = 4;       // rows per thread of the first pass
constexpr uint32_t kOrderRows = kOrderBlock * kOrderPer;
constexpr int kRankBits = 12;  // rank of a row within its block and high bucket (< kOrderRows)
static_assert(kLoBins == kOrderBlock && kOrderRows == (1u << kRankBits) && kOrderBits + kRankBits <= 32,
              "row-order packing");

__device__ __forceinline__ int longest_axis(const float cell[3])
{
    return cell[0] >= cell[1] ? (cell[0] >= cell[2] ? 0 : 2) : (cell[1] >= cell[2] ? 1 : 2);
}

// Where each axis' bits go in the key, from the box alone (the same for
// every row): bit b handed out (coarsest first) is key bit kOrderBits-1-b,
// and an axis' bits are handed out from its coordinate's most significant
// one down, so the key is a bit deposit of each cell coordinate into its
// axis' mask.
struct KeyPlan
{
    float lo[3], ext[3];
    int bits[3];
    uint32_t mask[3];
};

__device__ __forceinline__ KeyPlan key_plan(const float* __restrict__ bbox)
{
    KeyPlan P;
    float cell[3];
    for (int ax = 0; ax < 3; ++ax)
    {
        P.lo[ax] = bbox[ax];
        P.ext[ax] = fmaxf(bbox[3 + ax] - bbox[ax], 0.0f);
        cell[ax] = P.ext[ax];
        P.bits[ax] = 0;
        P.mask[ax] = 0;
    }
    for (int b = 0; b < kOrderBits; ++b)
    {
        const int ax = longest_axis(cell);
        ++P.bits[ax];
        P.mask[ax] |= 1u << (kOrderBits - 1 - b);
        cell[ax] *= 0.5f;
    }
    return P;
}

// the low bits of v, lowest first, into the set bits of mask, lowest first
__device__ __forceinline__ uint32_t deposit(uint32_t v, uint32_t mask)
{
    uint32_t r = 0;
    for (uint32_t m = mask; m; m &= m - 1u)
    {
        if (v & 1u) r |= m & (0u - m);
        v >>= 1;
    }
    return r;
}

__device__ __forceinline__ uint32_t row_key(const float* __restrict__ xyz, uint32_t n_xyz,
                                            const uint32_t* __restrict__ nn, uint64_t n_nn,
                                            const uint32_t* __restrict__ nn_ptr, const KeyPlan& P, uint32_t i)
{
    uint32_t j = i;
    if (i >= n_xyz)
    {
        const uint32_t b = nn_ptr[i], e = nn_ptr[i + 1];
        j = (b < e && (uint64_t)e <= n_nn) ? nn[b] : n_xyz;
    }
    if (j >= n_xyz) return 0;
    uint32_t k = 0;
    for (int ax = 0; ax < 3; ++ax)
    {
        const float side = (float)(1u << P.bits[ax]);
        float f = P.ext[ax] > 0.0f ? (xyz[(size_t)j * 3 + ax] - P.lo[ax]) / P.ext[ax] * side : 0.0f;
        f = fminf(fmaxf(f, 0.0f), side - 1.0f);  // also maps NaN to 0
        k |= deposit((uint32_t)f, P.mask[ax]);
    }
    return k;
}

// First pass: every row's key and rank among the rows of its block with the
// same high bucket, packed as (key << kRankBits) | rank; the block's count
// per high bucket at bcount[bucket * n_blocks + block] (one more entry, the
// last, is 0: its exclusive scan is the number of rows).
__global__ __launch_bounds__(kOrderBlock) void k_row_keys(const float* __restrict__ xyz, uint32_t n_xyz,
                                                          const uint32_t* __restrict__ nn, uint64_t n_nn,
                                                          const uint32_t* __restrict__ nn_ptr, uint32_t n_points,
                                                          const float* __restrict__ bbox, uint32_t* __restrict__ kr,
                                                          uint32_t* __restrict__ bcount, uint32_t n_blocks)
{
    __shared__ uint32_t hist[kHiBins];
    for (uint32_t b = threadIdx.x; b < kHiBins; b += kOrderBlock) hist[b] = 0;
    const KeyPlan P = key_plan(bbox);
    __syncthreads();
#pragma unroll
    for (uint32_t t = 0; t < kOrderPer; ++t)
    {
        const uint32_t i = blockIdx.x * kOrderRows + t * kOrderBlock + threadIdx.x;
        if (i < n_points)
        {
            const uint32_t k = row_key(xyz, n_xyz, nn, n_nn, nn_ptr, P, i);
            const uint32_t r = atomicAdd(&hist[k >> kLoBits], 1u);
            kr[i] = (k << kRankBits) | r;
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kHiBins; b += kOrderBlock) bcount[(size_t)b * n_blocks + blockIdx.x] = hist[b];
    if (blockIdx.x == 0 && threadIdx.x == 0) bcount[(size_t)kHiBins * n_blocks] = 0;
}

// Second pass: each row to its place among the rows of its high bucket
// (bucket-major offsets), as (row, low bits of its key).
__global__ __launch_bounds__(256) void k_row_hi_scatter(const uint32_t* __restrict__ kr, uint32_t n,
                                                        const uint32_t* __restrict__ boff, uint32_t n_blocks,
                                                        uint2* __restrict__ tmp)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = kr[i], k = v >> kRankBits, r = v & ((1u << kRankBits) - 1u);
    const uint32_t pos = boff[(size_t)(k >> kLoBits) * n_blocks + i / kOrderRows] + r;
    tmp[pos] = make_uint2(i, k & (kLoBins - 1u));
}

// Third pass: one workgroup per high bucket counts its rows' low bits in
// LDS, scans the counts and places every row (the rank within a low bucket
// is the order of the LDS atomics: any order within a bucket will do).
__global__ __launch_bounds__(kOrderBlock) void k_row_lo_sort(const uint2* __restrict__ tmp,
                                                             const uint32_t* __restrict__ boff, uint32_t n_blocks,
                                                             uint32_t* __restrict__ order)
{
    __shared__ uint32_t h[kLoBins];
    __shared__ uint32_t wsum[kOrderBlock / 64];
    const uint32_t s = boff[(size_t)blockIdx.x * n_blocks], e = boff[(size_t)(blockIdx.x + 1) * n_blocks];
    if (s >= e) return;  // (uniform over the workgroup)
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    h[tid] = 0;
    __syncthreads();
    for (uint32_t x = s + tid; x < e; x += kOrderBlock) atomicAdd(&h[tmp[x].y], 1u);
    __syncthreads();
    // exclusive scan of the kLoBins counts, one per thread
    const uint32_t v = h[tid];
    uint32_t incl = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1)
    {
        const uint32_t t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    if (w == 0)
    {
        const uint32_t ws = lane < kOrderBlock / 64 ? wsum[lane] : 0u;
        uint32_t wi = ws;
#pragma unroll
        for (uint32_t d = 1; d < kOrderBlock / 64; d <<= 1)
        {
            const uint32_t t = __shfl_up(wi, d);
            if (lane >= d) wi += t;
        }
        if (lane < kOrderBlock / 64) wsum[lane] = wi - ws;
    }
    __syncthreads();
    h[tid] = incl - v + wsum[w];
    __syncthreads();
    for (uint32_t x = s + tid; x < e; x += kOrderBlock)
    {
        const uint2 r = tmp[x];
        order[s + atomicAdd(&h[r.y], 1u)] = r.x;
    }
}

uint32_t order_blocks(uint32_t n_points) { return (n_points + kOrderRows - 1) / kOrderRows; }

}  // namespace

size_t csr_order_bytes(uint32_t n_points)
{
    const uint64_t counts = (uint64_t)kHiBins * order_blocks(n_points) + 1;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    return al(1024 * 6 * 4) + al(16 * 4) + al((size_t)n_points * 4) + al((size_t)n_points * 8) + al(counts * 4) +
           al(scan_partials_needed(counts) * 4);
}

hipError_t launch_csr_order(const CsrArgs& a, void* ws, uint32_t* order_out_ptr_holder[1], hipStream_t s)
{
    const uint32_t n = a.n_points;
    const uint32_t nb = order_blocks(n);
    const uint64_t counts = (uint64_t)kHiBins * nb + 1;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    char* p = (char*)ws;
    float* bb_part = (float*)p;
    p += al(1024 * 6 * 4);
    float* bb = (float*)p;
    p += al(16 * 4);
    // the packed keys, then (once the second pass has read them) the order
    uint32_t* kr = (uint32_t*)p;
    uint32_t* order = kr;
    p += al((size_t)n * 4);
    uint2* tmp = (uint2*)p;
    p += al((size_t)n * 8);
    uint32_t* bcount = (uint32_t*)p;
    p += al(counts * 4);
    uint32_t* part = (uint32_t*)p;
    ProfScope ps("csr_row_order", s);
    const uint32_t nbb = std::max<uint32_t>(1, std::min<uint32_t>(1024, (a.n_xyz + 255) / 256));
    hipError_t e = launch_bbox(a.xyz, a.n_xyz, bb_part, nbb, bb, s);
    if (e != hipSuccess) return e;
    k_row_keys<<<nb, kOrderBlock, 0, s>>>(a.xyz, a.n_xyz, a.nn, a.n_nn, a.nn_ptr, n, bb, kr, bcount, nb);
    e = launch_scan_exclusive(bcount, bcount, counts, part, s);
    if (e != hipSuccess) return e;
    k_row_hi_scatter<<<(n + 255) / 256, 256, 0, s>>>(kr, n, bcount, nb, tmp);
    k_row_lo_sort<<<kHiBins, kOrderBlock, 0, s>>>(tmp, bcount, nb, order);
    order_out_ptr_holder[0] = order;
    return hipGetLastError();
}

hipError_t launch_csr_features(int mode, const CsrArgs& a, hipStream_t s)
{
    if (a.n_points == 0) return hipSuccess;
    const uint32_t nb = (a.n_points + 255) / 256;
    switch (mode)
    {
        case kCsrPlain:
        {
            ProfScope ps("csr_features", s);
            k_csr<kCsrPlain><<<nb, 256, 0, s>>>(a);
            break;
        }
        case kCsrMulti:
        {
            ProfScope ps("csr_features_multiscale", s);
            k_csr<kCsrMulti><<<nb, 256, 0, s>>>(a);
            break;
        }
        default:
        {
            ProfScope ps("csr_features_optimal", s);
            k_csr<kCsrOptimal><<<nb, 256, 0, s>>>(a);
            break;
        }
    }
    return hipGetLastError();
}

}  // namespace pgeof_amd
