// CPU check of the search kernels' 4-ary heap (Heap4 in pgeof_search.hip).
// tests/test_cpu_heap.py copies the struct out of the HIP source into the
// marked line below, compiles this file with g++ and runs it.
//
// Many "lanes" share one LDS-like buffer (pairs of all lanes of a block
// interleaved) and one global-tail buffer (pairs of all lanes of the launch
// interleaved, or the pairs of each wave's 64 lanes interleaved and the
// waves' tails one after another), exactly as the kernel lays them out.  Both buffers start
// filled with poison (small, medium and huge keys), so a read of a node the
// heap has not written shows up as a wrong result.  Each lane offers a random
// number of random keys (often fewer than k, so partially filled heaps are
// covered), under the kernel's precondition key < worst; after finish() the
// heap must hold the k smallest keys in ascending order, padded with the
// sentinel.  Guard words after both buffers must stay untouched.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#define __device__
#define __forceinline__ inline
struct ulonglong2
{
    uint64_t x, y;
};
static ulonglong2 make_ulonglong2(uint64_t a, uint64_t b) { return {a, b}; }
using std::min;
// host forms of the kernel's global-memory accessors
static uint64_t gload(const uint64_t* p, size_t i) { return p[i]; }
static ulonglong2 gload2(const uint64_t* p) { return {p[0], p[1]}; }
static void gstore(uint64_t* p, size_t i, uint64_t v) { p[i] = v; }

// @HEAP4@

static uint64_t poison(size_t q)
{
    switch (q % 3)
    {
        case 0: return 0ull;
        case 1: return 3ull << 32;
        default: return ~0ull - 1;
    }
}

// WAVE_SLOTS: the tail of lane l of wave w (64 lanes) at pairs
// ((w * tail_pairs + p) * 64 + l) * 2, as with one tail per hardware wave
// slot; otherwise pair p of lane t at (p * lanes + t) * 2.  SHORT: the
// radius searches' finish (rows of at most 8 keys sorted in registers), with
// 0..11 keys per lane so that rows below, at and above that size occur.
template <bool TAIL, bool WAVE_SLOTS = false, bool SHORT = false>
static int check(std::mt19937_64& rng)
{
    constexpr uint64_t kGuard = 0x5a5a5a5a5a5a5a5aull;
    constexpr uint32_t kBlock = 64, kBlocks = 3, kLanes = kBlock * kBlocks;
    int bad = 0;
    for (uint32_t k = 1; k <= 130; ++k)
    {
        if (TAIL && k <= 21) continue;  // the kernel uses the tail only for k > 21
        const uint32_t groups = k > 5 ? (k - 2) / 4 : 0;
        const uint32_t lds_groups = TAIL ? 4 : groups;
        const uint32_t tail_pairs = TAIL ? 2 * ((k - 21 + 3) / 4) : 0;
        std::vector<uint64_t> lds((size_t)kBlocks * lds_groups * 4 * kBlock + 1);
        std::vector<uint64_t> glob((size_t)tail_pairs * 2 * kLanes + 1);
        for (size_t q = 0; q + 1 < lds.size(); ++q) lds[q] = poison(q);
        for (size_t q = 0; q + 1 < glob.size(); ++q) glob[q] = poison(q + 1);
        lds.back() = glob.back() = kGuard;

        const uint64_t sentinel = (rng() & 1) ? ~0ull : (0x3f000000ull << 32);
        std::vector<Heap4<TAIL>> heap(kLanes);
        std::vector<uint64_t> worst(kLanes, sentinel);
        std::vector<std::vector<uint64_t>> offered(kLanes);
        for (uint32_t t = 0; t < kLanes; ++t)
        {
            const uint32_t blk = t / kBlock, th = t % kBlock;
            heap[t].keys = lds.data() + (size_t)blk * lds_groups * 4 * kBlock + th * 2;
            heap[t].pstride = 2 * kBlock;
            heap[t].k = k;
            if (WAVE_SLOTS)
            {
                heap[t].gkeys = glob.data() + (((size_t)(t / 64) * tail_pairs) * 64 + t % 64) * 2;
                heap[t].gpstride = 128;
            }
            else
            {
                heap[t].gkeys = glob.data() + (size_t)t * 2;
                heap[t].gpstride = 2 * kLanes;
            }
            heap[t].init(sentinel);
        }
        const int n = SHORT ? (int)(rng() % 12) : (int)(rng() % 300);
        for (int j = 0; j < n; ++j)
            for (uint32_t t = 0; t < kLanes; ++t)
            {
                const uint64_t d = sentinel == ~0ull ? rng() % 100000 : rng() % 0x3f000000ull;
                const uint64_t key = (d << 32) | ((uint64_t)j * kLanes + t);
                offered[t].push_back(key);
                if (key < worst[t]) heap[t].add(key, worst[t]);
            }
        for (uint32_t t = 0; t < kLanes; ++t)
        {
            heap[t].template finish<SHORT>(sentinel);
            std::vector<uint64_t> want;
            for (uint64_t x : offered[t])
                if (x < sentinel) want.push_back(x);
            std::sort(want.begin(), want.end());
            want.resize(std::min<size_t>(want.size(), k));
            want.resize(k, sentinel);
            for (uint32_t i = 0; i < k; ++i)
                if (heap[t].key_at(i) != want[i])
                {
                    if (bad < 10)
                        std::printf("mismatch: tail=%d slots=%d short=%d k=%u lane=%u node=%u\n", (int)TAIL,
                                    (int)WAVE_SLOTS, (int)SHORT, k, t, i);
                    ++bad;
                    break;
                }
        }
        if (lds.back() != kGuard || glob.back() != kGuard)
        {
            std::printf("overrun: tail=%d k=%u\n", (int)TAIL, k);
            ++bad;
        }
    }
    return bad;
}

// Offsets of global tail nodes for launches with billions of lanes: the
// node offset (i - 21) / 2 * gpstride must not wrap at 2^32.
static int check_tail_offsets()
{
    int bad = 0;
    Heap4<true> h{};
    for (uint64_t lanes : {1ull << 30, (1ull << 31) - 64, 10000000ull, 4294967040ull})
    {
        h.gpstride = 2 * lanes;
        for (uint32_t i = 21; i < 2000; ++i)
        {
            const uint64_t s = i - 21;
            const uint64_t want = (s >> 1) * (2 * lanes) + (s & 1);
            if ((uint64_t)h.gslot(i) != want)
            {
                if (bad < 10) std::printf("tail offset: lanes=%llu node=%u\n", (unsigned long long)lanes, i);
                ++bad;
            }
        }
    }
    return bad;
}

int main()
{
    std::mt19937_64 rng(7);
    const int bad = check<false>(rng) + check<true>(rng) + check<true, true>(rng) + check<false, false, true>(rng) +
                    check<true, false, true>(rng) + check<true, true, true>(rng) + check_tail_offsets();
    std::printf("heap4 check: %d mismatches\n", bad);
    return bad != 0;
}
