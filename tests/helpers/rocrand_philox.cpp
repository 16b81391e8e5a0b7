// Prints Philox4x32-10 words from AMD's open-source rocRAND engine
// (rocrand_philox4x32_10.h) for counters read from stdin, so the oracle's
// Philox can be cross-checked against an independent implementation.
// Input lines: c0 c1 c2 c3 k0 k1 (hex).  Output lines: w0 w1 w2 w3 (hex).
// The engine is seeded with key (k0, k1); subsequence (c3:c2) and skipping
// (c1:c0) blocks of four words place its counter at (c0, c1, c2, c3).
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

int main()
{
    unsigned c[4], k[2];
    while (std::scanf("%x %x %x %x %x %x", &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) == 6) {
        const unsigned long long seed = ((unsigned long long)k[1] << 32) | k[0];
        const unsigned long long subseq = ((unsigned long long)c[3] << 32) | c[2];
        unsigned long long blocks = ((unsigned long long)c[1] << 32) | c[0];
        rocrand_device::philox4x32_10_engine eng(seed, subseq, 0);
        while (blocks) {
            const unsigned long long step = blocks > (1ull << 60) ? (1ull << 60) : blocks;
            eng.discard(step * 4ull);
            blocks -= step;
        }
        unsigned w0 = eng(), w1 = eng(), w2 = eng(), w3 = eng();
        std::printf("%08x %08x %08x %08x\n", w0, w1, w2, w3);
    }
    return 0;
}
