// Writes Philox4x32-10 known-answer vectors computed by rocRAND's host-callable
// engine (third-party implementation of Salmon et al., SC'11), used to pin the
// oracle's Philox.  Build and run: tests/golden/make_philox_vectors.sh
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

// ten_rounds() is a protected member; expose it for known-answer output.
struct Engine : rocrand_device::philox4x32_10_engine
{
    uint4 rounds(uint4 ctr, uint2 key) { return ten_rounds(ctr, key); }
};

int main()
{
    Engine eng;
    const uint32_t cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {1u, 0u, 0u, 0u, 0u, 0u},
        {0u, 1u, 0u, 0u, 0u, 0u},
        {0u, 0u, 1u, 0u, 0u, 0u},
        {0u, 0u, 0u, 1u, 0u, 0u},
        {0u, 0u, 0u, 0u, 1u, 0u},
        {0u, 0u, 0u, 0u, 0u, 1u},
        {12345u, 999u, 0u, 1u, 0x2545f491u, 0x4f6cdd1du},
        {7u, 1000000u, 3u, 2u, 42u, 0u},
        {0x80000000u, 0x7fffffffu, 0xdeadbeefu, 0xcafef00du, 0x01234567u, 0x89abcdefu},
    };
    std::printf("[\n");
    const int ncase = sizeof(cases) / sizeof(cases[0]);
    for (int i = 0; i < ncase; ++i) {
        const uint32_t* c = cases[i];
        uint4 ctr = {c[0], c[1], c[2], c[3]};
        uint2 key = {c[4], c[5]};
        uint4 r = eng.rounds(ctr, key);
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    c[0], c[1], c[2], c[3], c[4], c[5], r.x, r.y, r.z, r.w, i + 1 < ncase ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
