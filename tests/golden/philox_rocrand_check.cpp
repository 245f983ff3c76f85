// Cross-check of the Philox4x32-10 known-answer vectors against rocRAND's
// host-callable engine (an implementation independent of this repository).
// Build + run (host only, no GPU needed):
//   hipcc -O1 -o /tmp/philox_rocrand_check tests/golden/philox_rocrand_check.cpp && /tmp/philox_rocrand_check
// Prints one line per vector: ctr0..3 key0..1 -> out0..3 (hex).
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

struct probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main()
{
    const unsigned v[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {1u, 2u, 3u, 4u, 7u, 0u},
        {17u, 123456u, (6u << 24) | (1u << 16), 0u, 42u, 4095u},
    };
    probe p;
    for (const auto& x : v) {
        uint4 c = {x[0], x[1], x[2], x[3]};
        uint2 k = {x[4], x[5]};
        uint4 o = p.run(c, k);
        std::printf("%08x %08x %08x %08x %08x %08x -> %08x %08x %08x %08x\n", x[0], x[1], x[2], x[3],
                    x[4], x[5], o.x, o.y, o.z, o.w);
    }
    return 0;
}
