// Host check of the kernels' exact division (make_fastdiv / fast_div of
// lb-cube_amd/csrc/lbc_launch.hpp, the header's own code): for every divisor
// read from stdin, every dividend at a remainder boundary (k d - 1, k d,
// k d + d - 1 for k spread over [0, 2^32 / d]), the top of the range and a
// pseudo-random sample are divided and compared with 64-bit division.
// Prints "ok <divisors> <dividends>" or the first failure.
#include <cstdint>
#include <cstdio>

#include "../lb-cube_amd/csrc/lbc_launch.hpp"

int main() {
    unsigned long long d64;
    uint64_t divisors = 0, checked = 0, state = 0x9E3779B97F4A7C15ull;
    while (std::scanf("%llu", &d64) == 1) {
        const uint32_t d = static_cast<uint32_t>(d64);
        const lbc::FastDiv f = lbc::make_fastdiv(d);
        auto check = [&](uint64_t n) {
            if (n > 0xffffffffull) return true;
            ++checked;
            const uint32_t q = lbc::fast_div(static_cast<uint32_t>(n), f);
            if (q != n / d) {
                std::printf("FAIL d=%u n=%llu got %u want %llu\n", d, static_cast<unsigned long long>(n), q,
                            static_cast<unsigned long long>(n / d));
                return false;
            }
            return true;
        };
        const uint64_t kmax = 0xffffffffull / d + 1;
        const uint64_t stride = kmax > 4096 ? kmax / 4096 : 1;
        for (uint64_t k = 0; k <= kmax; k += stride)
            if (!check(k * d) || (k > 0 && !check(k * d - 1)) || !check(k * d + d - 1)) return 1;
        for (uint64_t k = kmax > 64 ? kmax - 64 : 0; k <= kmax; ++k)
            if (!check(k * d) || (k > 0 && !check(k * d - 1)) || !check(k * d + d - 1)) return 1;
        for (int r = 0; r < 4096; ++r) {
            state ^= state << 13;
            state ^= state >> 7;
            state ^= state << 17;
            if (!check(state & 0xffffffffull)) return 1;
        }
        if (!check(0xffffffffull)) return 1;
        ++divisors;
    }
    std::printf("ok %llu %llu\n", static_cast<unsigned long long>(divisors), static_cast<unsigned long long>(checked));
    return 0;
}
