// Writes Philox4x32-10 known-answer vectors computed by AMD's open-source
// rocRAND implementation (an implementation independent of this project's),
// so the project's Philox is pinned against it.
// Build & run:  hipcc -O1 tests/golden/make_philox_kat.cpp -o /tmp/kat && /tmp/kat > tests/golden/philox_kat.txt
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

// ten_rounds() is a protected member: expose it through a derived class.
struct kat_engine : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 ctr, uint2 key) { return this->ten_rounds(ctr, key); }
};

int main() {
    kat_engine eng;
    const uint32_t cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {1u, 0u, 0u, 0u, 0u, 0u},
        {0u, 7u, 12345u, 0u, 0xdeadbeefu, 0x01234567u},
        {41u, 1000000u, 1048575u, 0u, 20240601u, 0u},
    };
    // counter words, key words, output words
    for (const auto& c : cases) {
        uint4 ctr = {c[0], c[1], c[2], c[3]};
        uint2 key = {c[4], c[5]};
        uint4 o = eng.block(ctr, key);
        std::printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n",
                    c[0], c[1], c[2], c[3], c[4], c[5], o.x, o.y, o.z, o.w);
    }
    return 0;
}
