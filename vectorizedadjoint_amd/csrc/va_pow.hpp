// va_pow.hpp -- pow(x, y) with the rounding of the host C library's pow, for the
// step controller.  Compiles as HIP device code and as plain host C++.
//
// Boost.odeint's default_step_adjuster (called by the reference's controlled
// loop, lib/include/detail/runge_kutta.hpp:104-107) scales the step by
// std::pow(err, -1/q).  On Linux that is glibc's pow (2.28 and later), whose
// result is not always the correctly rounded power: on the controller's
// arguments about 8 in 10^4 calls differ from it by one ulp.  One ulp in a step
// size changes the rest of an instance's accept/reject history, so a device
// pow that rounded differently would change the reference's step sequence on
// several percent of the instances.  This file restates the published
// algorithm that glibc's pow uses (Szabolcs Nagy's log/exp pair, also in ARM's
// optimized-routines) in the operation order of its x86-64 FMA build, where
// the compiler contracts every a*b+c into one fused multiply-add:
//
//   log:  x = 2^k z with z in [0x1.69555p-1, 0x1.69555p0);
//         log x = k ln2 + log c + log1p(r),  r = z (1/c) - 1  (exact: 1/c has
//         few bits), c from a 128-entry table indexed by z's leading mantissa
//         bits; log1p(r) - r + r^2/2 by a polynomial in r; the sum carried as a
//         head hi and a tail lo.
//   mul:  y log x as ehi + elo, elo holding the rounding error of y*hi.
//   exp:  exp(ehi + elo) = 2^(j/128) exp(r'), j = round(128 ehi / ln2),
//         r' the reduced argument; 2^(j/128) = H (1 + T) from a 128-entry
//         table; exp(r') - 1 by a degree-5 polynomial; result scale + scale*tmp.
//
// Tables: va_pow_tables.hpp (generated from their definitions by
// gen_va_pow_tables.py).  The polynomial coefficients and ln2 splits are the
// published ones.  tests/test_pow_controller.py checks this code, built on
// the host, against the host's libm pow bit for bit.
//
// Domain: x > 0 normal or +inf (pow(+inf, y < 0) = +0), y negative with
// |y log x| < 512 -- the controller's error ratios in [5^-8, 1e300] and
// exponents -1/q.  The C library's special-case paths (zero, subnormal,
// negative or NaN x, huge or tiny exponents) are not restated.
#pragma once

#include <math.h>
#include <stdint.h>

#include "va_pow_tables.hpp"

#if defined(__HIPCC__)
#define VA_POW_HD __host__ __device__ __forceinline__
#else
#define VA_POW_HD inline
#endif

namespace va {
namespace powref {

VA_POW_HD double as_double(uint64_t u) { return __builtin_bit_cast(double, u); }
VA_POW_HD uint64_t as_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

// log1p polynomial, scaled by the evaluation (p0 = -1/2)
constexpr double P0 = -0x1p-1;
constexpr double P1 = 0x1.555555555556p-2 * -2;
constexpr double P2 = -0x1.0000000000006p-2 * -2;
constexpr double P3 = 0x1.999999959554ep-3 * 4;
constexpr double P4 = -0x1.555555529a47ap-3 * 4;
constexpr double P5 = 0x1.2495b9b4845e9p-3 * -8;
constexpr double P6 = -0x1.0002b8b263fc3p-3 * -8;
constexpr double LN2_HI = 0x1.62e42fefa3800p-1;
constexpr double LN2_LO = 0x1.ef35793c76730p-45;
// exp: 128/ln2, -ln2/128 split, rounding shift, polynomial
constexpr double INV_LN2_128 = 0x1.71547652b82fep0 * 128;
constexpr double NEG_LN2_HI_128 = -0x1.62e42fefa0000p-8;
constexpr double NEG_LN2_LO_128 = -0x1.cf79abc9e3b3ap-47;
constexpr double SHIFT = 0x1.8p52;
constexpr double E2 = 0x1.ffffffffffdbdp-2;
constexpr double E3 = 0x1.555555555543cp-3;
constexpr double E4 = 0x1.55555cf172b91p-5;
constexpr double E5 = 0x1.1111167a4d017p-7;
constexpr uint64_t OFF = 0x3fe6955500000000ull;

// Table access.  GlobalTables reads the constant tables; a caller may supply
// another accessor with the same two functions (the kernels read LDS copies).
struct GlobalTables {
    VA_POW_HD static void log_entry(int i, double &invc, double &logc, double &logctail)
    {
        invc = powtab::LOG[i][0];
        logc = powtab::LOG[i][1];
        logctail = powtab::LOG[i][2];
    }
    VA_POW_HD static void exp_entry(uint64_t idx, uint64_t &tail_bits, uint64_t &scale_bits)
    {
        tail_bits = powtab::EXP[idx];
        scale_bits = powtab::EXP[idx + 1];
    }
};

// log(x) as hi + *lo for x > 0 normal (x given by its bits)
template <class T = GlobalTables>
VA_POW_HD double log_hilo(uint64_t ix, double *lo)
{
    const uint64_t tmp = ix - OFF;
    const int i = (int)((tmp >> 45) & 127);
    const int k = (int)((int64_t)tmp >> 52);
    const double z = as_double(ix - (tmp & (0xfffull << 52)));
    const double kd = (double)k;
    double invc, logc, logctail;
    T::log_entry(i, invc, logc, logctail);
    const double r = fma(z, invc, -1.0);
    const double t1 = fma(kd, LN2_HI, logc);
    const double t2 = t1 + r;
    const double lo1 = fma(kd, LN2_LO, logctail);
    const double lo2 = t1 - t2 + r;
    const double ar = P0 * r;
    const double ar2 = r * ar;
    const double ar3 = r * ar2;
    const double hi = t2 + ar2;
    const double lo3 = fma(ar, r, -ar2);
    const double lo4 = t2 - hi + ar2;
    const double q = fma(ar2, fma(ar2, fma(r, P6, P5), fma(r, P4, P3)), fma(r, P2, P1));
    const double l = fma(ar3, q, lo1 + lo2 + lo3 + lo4);
    const double y = hi + l;
    *lo = hi - y + l;
    return y;
}

// exp(x + xtail) for |x| in [2^-54, 512)
template <class T = GlobalTables>
VA_POW_HD double exp_hilo(double x, double xtail)
{
    double kd = fma(INV_LN2_128, x, SHIFT);
    const uint64_t ki = as_bits(kd);
    kd = kd - SHIFT;
    double r = fma(kd, NEG_LN2_LO_128, fma(kd, NEG_LN2_HI_128, x));
    r = r + xtail;
    const uint64_t idx = 2 * (ki & 127);
    const uint64_t top = ki << 45;
    uint64_t tail_bits, scale_bits;
    T::exp_entry(idx, tail_bits, scale_bits);
    const double tail = as_double(tail_bits);
    const uint64_t sbits = scale_bits + top;
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, E5, E4), fma(r2, fma(r, E3, E2), tail + r));
    const double scale = as_double(sbits);
    return fma(scale, tmp, scale);
}

// pow(x, y) on the domain stated above
template <class T = GlobalTables>
VA_POW_HD double pow_ctrl(double x, double y)
{
    if (x == __builtin_huge_val()) return 0.0;
    double lo;
    const double hi = log_hilo<T>(as_bits(x), &lo);
    const double ehi = y * hi;
    const double elo = fma(y, lo, fma(y, hi, -ehi));
    return exp_hilo<T>(ehi, elo);
}

}  // namespace powref
}  // namespace va
