// Writes Philox4x32-10 vectors computed by rocRAND's open-source engine
// (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h), host side only, as JSON.
// The counters/keys are a fixed pseudo-random sweep plus the edge patterns.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

struct Engine : rocrand_device::philox4x32_10_engine {
    uint4 bijection(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

static uint32_t lcg(uint64_t &s) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(s >> 33);
}

int main() {
    Engine e;
    uint64_t s = 12345;
    std::printf("[\n");
    const int N = 64;
    for (int t = 0; t < N; t++) {
        uint4 c;
        uint2 k;
        if (t == 0) { c = {0, 0, 0, 0}; k = {0, 0}; }
        else if (t == 1) { c = {~0u, ~0u, ~0u, ~0u}; k = {~0u, ~0u}; }
        else { c = {lcg(s), lcg(s), lcg(s), lcg(s)}; k = {lcg(s), lcg(s)}; }
        uint4 o = e.bijection(c, k);
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    c.x, c.y, c.z, c.w, k.x, k.y, o.x, o.y, o.z, o.w, t + 1 < N ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
