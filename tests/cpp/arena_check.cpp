// CPU check of the workspace arena (Arena in pgeof_internal.hpp).
// tests/test_cpu_heap.py copies the struct out of the header into the marked
// line below, compiles this file with g++ and runs it.
//
// A search carves its workspace with Arena::take; kernels then use 64-bit
// atomics (the occupancy probe's sum of squares) and 16-byte vector loads on
// the regions.  The host entry points place the workspace after the call's
// outputs, so its base is only as aligned as the outputs' length makes it
// (4 bytes after an odd number of floats).  Every region must nevertheless
// start on a 256-byte boundary of the address space, and the regions must not
// overlap or leave the buffer.
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <vector>

// @ARENA@

int main()
{
    std::vector<char> buf(1 << 20);
    int bad = 0, cases = 0;
    for (size_t skew = 0; skew < 512; skew += 4)  // every 4-byte-aligned base offset
    {
        Arena ar{buf.data() + skew, buf.size() - skew - 4096, 0};
        char* prev_end = ar.base;
        const size_t counts[] = {6144, 16, 8, 640001, 1280066, 11429 * 3, 102861, 1, 7, 33};
        for (size_t i = 0; i < sizeof(counts) / sizeof(counts[0]); ++i)
        {
            // alternate element types as the search's carve does
            char* p;
            size_t bytes;
            if (i % 3 == 0)
            {
                p = reinterpret_cast<char*>(ar.take<float>(counts[i] % 5000 + 1));
                bytes = (counts[i] % 5000 + 1) * 4;
            }
            else if (i % 3 == 1)
            {
                p = reinterpret_cast<char*>(ar.take<unsigned long long>(counts[i] % 700 + 1));
                bytes = (counts[i] % 700 + 1) * 8;
            }
            else
            {
                p = reinterpret_cast<char*>(ar.take<uint32_t>(counts[i] % 9000 + 1));
                bytes = (counts[i] % 9000 + 1) * 4;
            }
            ++cases;
            if (reinterpret_cast<uintptr_t>(p) % 256 != 0)
            {
                ++bad;
                std::printf("skew %zu region %zu: address %% 256 = %zu\n", skew, i,
                            (size_t)(reinterpret_cast<uintptr_t>(p) % 256));
            }
            if (p < prev_end || p + bytes > buf.data() + buf.size())
            {
                ++bad;
                std::printf("skew %zu region %zu: overlaps or leaves the buffer\n", skew, i);
            }
            if ((size_t)(p - ar.base) > ar.off)
            {
                ++bad;
                std::printf("skew %zu region %zu: offset beyond the arena's end\n", skew, i);
            }
            prev_end = p + bytes;
        }
    }
    std::printf("%d cases, %d failures\n", cases, bad);
    return bad ? 1 : 0;
}
