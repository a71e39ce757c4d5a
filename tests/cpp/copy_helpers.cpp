// Unit check of the host copies of the staged transfers (CopyHelpers,
// dxmclib_amd/csrc/hostcopy.hpp): lengths below, at and around the part
// threshold and the 8 MiB staging chunk, odd lengths and unaligned
// addresses, repeated copies on the same helpers, and copies without
// helpers. Prints "ok" and exits 0, or names the failing case and exits 1.
#include "../../dxmclib_amd/csrc/hostcopy.hpp"

#include <cstdint>
#include <cstdio>
#include <vector>

int main()
{
    using dxmc_backend::CopyHelpers;
    constexpr std::size_t MiB = std::size_t { 1 } << 20;
    const std::size_t lengths[] = { 0, 1, 7, 4 * MiB - 1, 4 * MiB, 4 * MiB + 3, 8 * MiB - 1, 8 * MiB, 8 * MiB + 1, 13 * MiB + 5 };
    const std::size_t maxLen = 13 * MiB + 5 + 64;
    std::vector<unsigned char> src(maxLen), dst(maxLen);
    std::uint64_t x = 0x9E3779B97F4A7C15ULL;
    for (auto& b : src) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        b = static_cast<unsigned char>(x);
    }
    CopyHelpers helpers;
    for (int round = 0; round < 3; ++round) {
        if (round == 0 && !helpers.start()) {
            std::printf("helpers did not start\n");
            return 1;
        }
        if (round == 2)
            helpers.stop(); // copies on the calling thread alone
        for (const std::size_t len : lengths)
            for (const std::size_t so : { 0, 3 })
                for (const std::size_t doff : { 0, 5 }) {
                    std::fill(dst.begin(), dst.end(), 0xA5);
                    helpers.copy(dst.data() + doff, src.data() + so, len);
                    for (std::size_t i = 0; i < maxLen; ++i) {
                        const bool inside = i >= doff && i < doff + len;
                        const unsigned char want = inside ? src[i - doff + so] : 0xA5;
                        if (dst[i] != want) {
                            std::printf("round %d len %zu src+%zu dst+%zu: byte %zu differs\n", round, len, so, doff, i);
                            return 1;
                        }
                    }
                }
    }
    std::printf("ok\n");
    return 0;
}
