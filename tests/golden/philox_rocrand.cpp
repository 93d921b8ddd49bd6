// Writes Philox4x32-10 outputs computed by rocRAND's engine
// (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h, ten_rounds) for
// counters/keys from a fixed LCG, as JSON on stdout. Built and run by
// make_philox_rocrand.sh; the JSON is the committed fixture.
#include <cstdint>
#include <cstdio>

#include <rocrand/rocrand_philox4x32_10.h>

struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main()
{
    Probe p;
    uint64_t s = 0x9E3779B97F4A7C15ull;
    auto next = [&s]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (uint32_t)(s >> 32);
    };
    std::printf("[\n");
    for (int i = 0; i < 64; ++i) {
        uint4 c = {next(), next(), next(), next()};
        uint2 k = {next(), next()};
        if (i == 0) c = {0u, 0u, 0u, 0u}, k = {0u, 0u};
        if (i == 1) c = {1u, 7u, 3u, 0xFFFFFFF0u}, k = {0u, 0u};
        const uint4 o = p.run(c, k);
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n", c.x, c.y, c.z,
                    c.w, k.x, k.y, o.x, o.y, o.z, o.w, i == 63 ? "" : ",");
    }
    std::printf("]\n");
    return 0;
}
