// Writes Philox4x32-10 vectors from rocRAND's open-source philox4x32_10
// engine (rocrand/rocrand_philox4x32_10.h), compiled for the host.  The engine
// seeded with key k, subsequence (c.z | c.w << 32) and offset 4 (c.x | c.y << 32)
// returns the ten-round block of counter c under key k (c.y < 2^30, so the
// offset 4 (c.x | c.y << 32) does not overflow 64 bits).
#include <cstdio>
#include <cstdint>
#include <rocrand/rocrand_philox4x32_10.h>

static uint64_t sm_state = 0x243F6A8885A308D3ull;
static uint32_t next32() {  // splitmix64, only to spread the counters and keys
    uint64_t z = (sm_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)((z ^ (z >> 31)) >> 16);
}

int main() {
    std::printf("[\n");
    const int n = 64;
    for (int i = 0; i < n; ++i) {
        uint32_t c[4], k[2];
        for (int j = 0; j < 4; ++j) c[j] = (i < 4) ? (i == 1 ? 0xFFFFFFFFu : (uint32_t)(i * j)) : next32();
        for (int j = 0; j < 2; ++j) k[j] = (i < 4) ? (i == 1 ? 0xFFFFFFFFu : (uint32_t)i) : next32();
        c[1] &= 0x3FFFFFFFu;
        const unsigned long long seed = ((unsigned long long)k[1] << 32) | k[0];
        const unsigned long long sub = ((unsigned long long)c[3] << 32) | c[2];
        const unsigned long long off = 4ull * (((unsigned long long)c[1] << 32) | c[0]);
        rocrand_device::philox4x32_10_engine eng(seed, sub, off);
        uint32_t r[4];
        for (int j = 0; j < 4; ++j) r[j] = eng.next();
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    c[0], c[1], c[2], c[3], k[0], k[1], r[0], r[1], r[2], r[3], i + 1 < n ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
