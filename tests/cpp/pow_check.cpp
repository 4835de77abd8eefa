// Host build of the device step-controller pow (vectorizedadjoint_amd/csrc/va_pow.hpp),
// compared with the C library's pow.  Built and run by tests/test_pow_controller.py.
//
//   pow_check random <count> <seed>   random arguments in [2^-40, 2^120] against each
//                                     controller exponent -1/q (q = 2..8) and random
//                                     exponents in [-4, 4]
//   pow_check file <path>             "x y" pairs, one per line (hex floats)
// Prints "checked N mismatches M"; exits 1 on a mismatch.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "va_pow.hpp"

namespace {

uint64_t g_state = 0;

uint64_t next_u64()
{
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double unit() { return (double)(next_u64() >> 11) * 0x1p-53; }

long g_bad = 0, g_checked = 0;

void check(double x, double y)
{
    const double a = va::powref::pow_ctrl(x, y);
    const double b = std::pow(x, y);
    ++g_checked;
    if (std::memcmp(&a, &b, sizeof a) != 0) {
        if (g_bad < 5) std::printf("mismatch x=%a y=%a device=%a libm=%a\n", x, y, a, b);
        ++g_bad;
    }
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc >= 4 && std::strcmp(argv[1], "random") == 0) {
        const long count = std::atol(argv[2]);
        g_state = std::strtoull(argv[3], nullptr, 10);
        for (long k = 0; k < count; ++k) {
            const double x = std::exp2(-40.0 + 160.0 * unit());
            for (int q = 2; q <= 8; ++q) check(x, -1.0 / (double)q);
            check(x, -4.0 + 8.0 * unit());
        }
        check(__builtin_huge_val(), -0.2);
    } else if (argc >= 3 && std::strcmp(argv[1], "file") == 0) {
        FILE *f = std::fopen(argv[2], "r");
        if (!f) return 2;
        char xs[64], ys[64];
        while (std::fscanf(f, "%63s %63s", xs, ys) == 2) check(std::strtod(xs, nullptr), std::strtod(ys, nullptr));
        std::fclose(f);
    } else {
        std::fprintf(stderr, "usage: pow_check random <count> <seed> | file <path>\n");
        return 2;
    }
    std::printf("checked %ld mismatches %ld\n", g_checked, g_bad);
    return g_bad ? 1 : 0;
}
