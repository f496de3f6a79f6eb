// Host-side check of the radix scatter destination (gs3d_internal.h
// scatter_delta / scatter_index, the functions the k_radix_pass kernels call):
// for tiles whose digit runs start anywhere below 2^32 - 1, every position's
// index must equal the exact 64-bit destination global_start + excl +
// (j - local_start). Prints "ok <cases>" or the first mismatch and exits 1.
#include <cstdio>
#include <cstdint>

#include "gs3d_internal.h"

int main() {
    const uint64_t kMaxKeys = 0xFFFFFFFFull;
    // destination starts around the int32 and uint32 limits and small ones
    const uint64_t starts[] = {0ull, 1ull, 4095ull, 0x7FFFF000ull, 0x7FFFFFFFull, 0x80000000ull,
                               0x80000001ull, 0xC0000000ull, 0xFFFF0000ull, kMaxKeys - 16384ull};
    const uint32_t local_starts[] = {0u, 1u, 255u, 8191u, 12000u, 16383u};
    const uint32_t runs[] = {1u, 7u, 64u, 4096u};
    uint64_t cases = 0;
    for (uint64_t dst : starts) {
        for (uint32_t ls : local_starts) {
            for (uint32_t run : runs) {
                if (ls + (uint64_t)run > 16384u || dst + run > kMaxKeys) continue;
                // split the destination into a digit start and a predecessor count
                const uint32_t excl = (uint32_t)(dst / 3u);
                const uint32_t gstart = (uint32_t)(dst - excl);
                const uint32_t delta = gs3d::scatter_delta(gstart, excl, ls);
                for (uint32_t j = ls; j < ls + run; j++) {
                    const uint64_t want = dst + (j - ls);
                    const size_t got = gs3d::scatter_index(delta, j);
                    cases++;
                    if ((uint64_t)got != want) {
                        std::printf("mismatch: dst %llu local_start %u j %u: got %llu want %llu\n",
                                    (unsigned long long)dst, ls, j, (unsigned long long)got,
                                    (unsigned long long)want);
                        return 1;
                    }
                }
            }
        }
    }
    std::printf("ok %llu\n", (unsigned long long)cases);
    return 0;
}
