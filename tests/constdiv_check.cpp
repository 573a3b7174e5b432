// Host check of lbc::div_const<R, D> (lb-cube_amd/csrc/lbc_constdiv.hpp, the
// header's own code) against the hardware's IEEE division x / D:
//   float:  every finite bit pattern (normal, subnormal and zero), D = 3 and 6;
//   double: every exponent with pseudo-random significands, the significands
//           next to 1, 1.5 and 2 in every binade, the subnormals near zero and
//           the largest finite values, D = 3 and 6.
// Signed zeros are compared by value.  Prints "ok <floats> <doubles>" or the
// first mismatch.
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "../lb-cube_amd/csrc/lbc_constdiv.hpp"

namespace {

template <class R, class U> bool same(R a, R b) {
    if (a == R(0) && b == R(0)) return true;
    U ua, ub;
    std::memcpy(&ua, &a, sizeof(R));
    std::memcpy(&ub, &b, sizeof(R));
    return ua == ub;
}

std::atomic<bool> failed{false};

template <int D> void float_range(uint64_t lo, uint64_t hi) {
    for (uint64_t bits = lo; bits < hi && !failed.load(std::memory_order_relaxed); ++bits) {
        const uint32_t b = static_cast<uint32_t>(bits);
        if ((b & 0x7f800000u) == 0x7f800000u) continue;  // inf and NaN
        float x;
        std::memcpy(&x, &b, 4);
        const float want = x / static_cast<float>(D);
        const float got = lbc::div_const<float, D>(x);
        if (!same<float, uint32_t>(got, want)) {
            if (!failed.exchange(true))
                std::printf("FAIL float D=%d x=%a got %a want %a\n", D, static_cast<double>(x),
                            static_cast<double>(got), static_cast<double>(want));
            return;
        }
    }
}

template <int D> uint64_t check_floats(unsigned threads) {
    const uint64_t total = 1ull << 32;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back(float_range<D>, total * t / threads, total * (t + 1) / threads);
    for (auto& th : pool) th.join();
    return total - 2 * (1ull << 23);  // finite patterns
}

uint64_t state = 0x9E3779B97F4A7C15ull;
uint64_t next_random() {
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return state;
}

template <int D> bool check_double(double x, uint64_t& count) {
    ++count;
    const double want = x / static_cast<double>(D);
    const double got = lbc::div_const<double, D>(x);
    if (!same<double, uint64_t>(got, want)) {
        std::printf("FAIL double D=%d x=%a got %a want %a\n", D, x, got, want);
        return false;
    }
    return true;
}

template <int D> bool check_doubles(uint64_t& count) {
    for (uint64_t e = 0; e < 2047; ++e) {  // biased exponent 0 (subnormals) .. 2046
        for (int r = 0; r < 20000; ++r) {
            const uint64_t bits = (e << 52) | (next_random() & ((1ull << 52) - 1));
            double x;
            std::memcpy(&x, &bits, 8);
            if (!check_double<D>(x, count) || !check_double<D>(-x, count)) return false;
        }
        for (uint64_t m : {0ull, 1ull << 51}) {  // significands next to 1, 1.5 (and 2 from below)
            for (int k = -64; k <= 64; ++k) {
                const uint64_t sig = (m + static_cast<uint64_t>(k)) & ((1ull << 52) - 1);
                const uint64_t bits = (e << 52) | sig;
                double x;
                std::memcpy(&x, &bits, 8);
                if (!check_double<D>(x, count)) return false;
            }
        }
    }
    for (uint64_t bits = 0; bits < (1u << 20); ++bits) {  // the smallest subnormals
        double x;
        std::memcpy(&x, &bits, 8);
        if (!check_double<D>(x, count)) return false;
    }
    for (uint64_t k = 0; k < (1u << 20); ++k) {  // the largest finite values
        const uint64_t bits = 0x7fefffffffffffffull - k;
        double x;
        std::memcpy(&x, &bits, 8);
        if (!check_double<D>(x, count)) return false;
    }
    return true;
}

}  // namespace

int main() {
    const unsigned threads = std::thread::hardware_concurrency() ? std::min(16u, std::thread::hardware_concurrency())
                                                                 : 4u;
    uint64_t floats = check_floats<3>(threads);
    if (failed) return 1;
    floats += check_floats<6>(threads);
    if (failed) return 1;
    uint64_t doubles = 0;
    if (!check_doubles<3>(doubles) || !check_doubles<6>(doubles)) return 1;
    std::printf("ok %llu %llu\n", static_cast<unsigned long long>(floats), static_cast<unsigned long long>(doubles));
    return 0;
}
