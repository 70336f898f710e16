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

template <int MODE>
__global__ __launch_bounds__(256) void k_csr(CsrArgs a)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n_points) return;
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
    // at the end); coordinates are gathered as one 12-byte load each, four
    // neighbours at a time.
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
    constexpr uint32_t kChunk = 4;
    for (uint32_t j0 = 0; j0 < walk; j0 += kChunk)
    {
        float px[kChunk], py[kChunk], pz[kChunk];
#pragma unroll
        for (uint32_t u = 0; u < kChunk; ++u)
        {
            uint32_t idx = (j0 + u < walk) ? row[j0 + u] : i0;
            bad |= idx >= a.n_xyz;
            idx = idx < a.n_xyz ? idx : 0u;
            const float3 p = xyz3[idx];
            px[u] = p.x;
            py[u] = p.y;
            pz[u] = p.z;
        }
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
                    float* o = out + (size_t)s * 11;
                    for (int c = 0; c < 11; ++c) o[c] = f[c];
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
        for (int c = 0; c < 11; ++c) out[c] = f[c];
    }
    else if (MODE == kCsrMulti)
    {
        for (uint32_t c = s * 11; c < n_out; ++c) out[c] = 0.0f;
    }
    else
    {
        features_of(best, k_opt, f);
        for (int c = 0; c < 11; ++c) out[c] = f[c];
        out[11] = (float)k_opt;
    }
}

}  // namespace

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
