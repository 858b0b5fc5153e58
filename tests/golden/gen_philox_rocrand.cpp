// Writes tests/golden/philox_rocrand.json: Philox4x32-10 blocks computed by
// rocRAND's open-source engine (rocrand_philox4x32_10.h, host side), used to
// pin the oracle's and the kernels' Philox independently of this build.
// Build and run: python tests/golden/gen_philox_rocrand.py
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdint>
#include <cstdio>

struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

static uint64_t sm(uint64_t& s)
{
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int main()
{
    Probe p;
    uint64_t st = 20261203ull;
    std::printf("{\"source\": \"rocrand_philox4x32_10.h ten_rounds\", \"vectors\": [\n");
    for (int q = 0; q < 64; ++q) {
        uint4 c;
        uint2 k;
        if (q == 0) { c = {0u, 0u, 0u, 0u}; k = {0u, 0u}; }
        else if (q == 1) { c = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}; k = {0xFFFFFFFFu, 0xFFFFFFFFu}; }
        else {
            const uint64_t a = sm(st), b = sm(st), e = sm(st);
            c = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
            k = {(unsigned)e, (unsigned)(e >> 32)};
        }
        const uint4 r = p.block(c, k);
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n", c.x, c.y, c.z,
                    c.w, k.x, k.y, r.x, r.y, r.z, r.w, q == 63 ? "" : ",");
    }
    std::printf("]}\n");
    return 0;
}
