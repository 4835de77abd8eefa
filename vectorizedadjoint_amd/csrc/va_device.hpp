// va_device.hpp -- device building blocks of the batched RK forward/adjoint path
// for gfx950: Butcher tableaus, the step-controller pow, the built-in systems
// (right-hand side and vector-Jacobian products) and the explicit RK attempt.
//
// Rounding sequence: all HIP sources are compiled with -ffp-contract=off and
// every multiply-add of the step arithmetic is an explicit fma(), so the IEEE
// operation sequence is exactly the one written here.  It is the same sequence
// the CPU oracle uses, which makes the forward pass (step sizes, accept/reject
// decisions, y(T)) reproducible bit for bit on the host.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace va {

#define VA_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------------------
// Butcher tableaus.  Published coefficients: classic RK4; Cash & Karp (1990);
// Dormand & Prince (1980); Fehlberg (1968, NASA TR R-287).  Row layout follows
// lib/include/ButcherTable.hpp:88-120,154-186,210-242 (row 0 is zero); db is
// b - b_embedded rounded as odeint 1.74 rounds it.  DOPRI5 (unsupported in the
// reference, ButcherTable.hpp:247-250) carries its FSAL stage as stage 7.
// ---------------------------------------------------------------------------
struct Euler {
    static constexpr int S = 1, SF = 1, ORDER = 1, ERR_ORDER = 0;
    static constexpr bool FSAL = false, HAS_ERR = false;
    static constexpr double A[1][1] = {{0.0}};
    static constexpr double B[1] = {1.0};
    static constexpr double DB[1] = {0.0};
    static constexpr double C[1] = {0.0};
};

struct RK4 {
    static constexpr int S = 4, SF = 4, ORDER = 4, ERR_ORDER = 0;
    static constexpr bool FSAL = false, HAS_ERR = false;
    static constexpr double A[4][4] = {{0, 0, 0, 0},
                                       {1.0 / 2.0, 0, 0, 0},
                                       {0, 1.0 / 2.0, 0, 0},
                                       {0, 0, 1.0, 0}};
    static constexpr double B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    static constexpr double DB[4] = {0, 0, 0, 0};
    static constexpr double C[4] = {0, 1.0 / 2.0, 1.0 / 2.0, 1.0};
};

struct CashKarp54 {
    static constexpr int S = 6, SF = 6, ORDER = 5, ERR_ORDER = 4;
    static constexpr bool FSAL = false, HAS_ERR = true;
    static constexpr double A[6][6] = {
        {0, 0, 0, 0, 0, 0},
        {1.0 / 5.0, 0, 0, 0, 0, 0},
        {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0},
        {3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0, 0, 0, 0},
        {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0, 0, 0},
        {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0, 0}};
    static constexpr double B[6] = {37.0 / 378.0, 0, 250.0 / 621.0, 125.0 / 594.0, 0,
                                    512.0 / 1771.0};
    static constexpr double DB[6] = {37.0 / 378.0 - 2825.0 / 27648.0,
                                     0,
                                     250.0 / 621.0 - 18575.0 / 48384.0,
                                     125.0 / 594.0 - 13525.0 / 55296.0,
                                     -277.0 / 14336.0,
                                     512.0 / 1771.0 - 1.0 / 4.0};
    static constexpr double C[6] = {0, 1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0};
};

struct Dopri5 {
    // S = 6 solution stages; SF = 7 evaluated stages (stage 7 = f(y_{n+1}), FSAL).
    static constexpr int S = 6, SF = 7, ORDER = 5, ERR_ORDER = 4;
    static constexpr bool FSAL = true, HAS_ERR = true;
    static constexpr double A[7][7] = {
        {0, 0, 0, 0, 0, 0, 0},
        {1.0 / 5.0, 0, 0, 0, 0, 0, 0},
        {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0, 0},
        {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0, 0},
        {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0, 0},
        {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0, 0},
        {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0}};
    static constexpr double B[7] = {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0,
                                    -2187.0 / 6784.0, 11.0 / 84.0, 0};
    static constexpr double DB[7] = {35.0 / 384.0 - 5179.0 / 57600.0,
                                     0,
                                     500.0 / 1113.0 - 7571.0 / 16695.0,
                                     125.0 / 192.0 - 393.0 / 640.0,
                                     -2187.0 / 6784.0 - -92097.0 / 339200.0,
                                     11.0 / 84.0 - 187.0 / 2100.0,
                                     -1.0 / 40.0};
    static constexpr double C[7] = {0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
};

struct Fehlberg78 {
    static constexpr int S = 13, SF = 13, ORDER = 8, ERR_ORDER = 7;
    static constexpr bool FSAL = false, HAS_ERR = true;
    static constexpr double A[13][13] = {
        {0},
        {2.0 / 27.0},
        {1.0 / 36.0, 1.0 / 12.0},
        {1.0 / 24.0, 0, 1.0 / 8.0},
        {5.0 / 12.0, 0, -25.0 / 16.0, 25.0 / 16.0},
        {1.0 / 20.0, 0, 0, 1.0 / 4.0, 1.0 / 5.0},
        {-25.0 / 108.0, 0, 0, 125.0 / 108.0, -65.0 / 27.0, 125.0 / 54.0},
        {31.0 / 300.0, 0, 0, 0, 61.0 / 225.0, -2.0 / 9.0, 13.0 / 900.0},
        {2.0, 0, 0, -53.0 / 6.0, 704.0 / 45.0, -107.0 / 9.0, 67.0 / 90.0, 3.0},
        {-91.0 / 108.0, 0, 0, 23.0 / 108.0, -976.0 / 135.0, 311.0 / 54.0, -19.0 / 60.0,
         17.0 / 6.0, -1.0 / 12.0},
        {2383.0 / 4100.0, 0, 0, -341.0 / 164.0, 4496.0 / 1025.0, -301.0 / 82.0, 2133.0 / 4100.0,
         45.0 / 82.0, 45.0 / 164.0, 18.0 / 41.0},
        {3.0 / 205.0, 0, 0, 0, 0, -6.0 / 41.0, -3.0 / 205.0, -3.0 / 41.0, 3.0 / 41.0, 6.0 / 41.0,
         0},
        {-1777.0 / 4100.0, 0, 0, -341.0 / 164.0, 4496.0 / 1025.0, -289.0 / 82.0, 2193.0 / 4100.0,
         51.0 / 82.0, 33.0 / 164.0, 12.0 / 41.0, 0, 1.0}};
    static constexpr double B[13] = {0, 0, 0, 0, 0, 34.0 / 105.0, 9.0 / 35.0, 9.0 / 35.0,
                                     9.0 / 280.0, 9.0 / 280.0, 0, 41.0 / 840.0, 41.0 / 840.0};
    static constexpr double DB[13] = {0.0 - 41.0 / 840.0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                      0.0 - 41.0 / 840.0, 41.0 / 840.0, 41.0 / 840.0};
    static constexpr double C[13] = {0, 2.0 / 27.0, 1.0 / 9.0, 1.0 / 6.0, 5.0 / 12.0, 1.0 / 2.0,
                                     5.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0, 1.0, 0.0, 1.0};
};

// ---------------------------------------------------------------------------
// Step-controller root x^(-1/q), q = 2..8 (the adjuster raises the error ratio
// to -1/q).  Own method:
//   x = m 2^e, m in [1,2), e = q k + j with 0 <= j < q:  x^(-1/q) = 2^-k (m 2^j)^(-1/q);
//   start  y = P_q(m - 1.5) 2^(-j/q), P_q the degree-5 Chebyshev interpolant of
//          m^(-1/q) on [1,2] (relative error <= 1.4e-5);
//   refine two Newton steps for y^-q = m 2^j: y += (y/q)(1 - m 2^j y^q);
//   scale  by 2^-k, exactly.
// IEEE basic operations, fma and integer operations on the representation
// only, so the GPU and the host obtain the same bits.  The dependent chain is
// about 20 FP64 operations deep, against about 60 for exp(y log x) with an
// IEEE division inside the log.  Within one ulp of the exact root.
// ---------------------------------------------------------------------------
template <int Q>
struct RootCoef;
template <>
struct RootCoef<2> {
    static constexpr double p[6] = {0x1.a20d0b2f16c1fp-1, -0x1.16b40de0b165ap-2,
                                    0x1.1559ba6062540p-3, -0x1.33fdf4c87b433p-4,
                                    0x1.a0f6e83f40934p-5, -0x1.f69d70fc899e8p-6};
    static constexpr double s[2] = {0x1p+0, 0x1.6a09e667f3bcdp-1};
};
template <>
struct RootCoef<3> {
    static constexpr double p[6] = {0x1.bf46914f5b764p-1, -0x1.8d95060abdbffp-3,
                                    0x1.5ffe431f2a20ep-4, -0x1.6cc93eb91b616p-5,
                                    0x1.d18071c6f18d9p-6, -0x1.0e87006cbb389p-6};
    static constexpr double s[3] = {0x1p+0, 0x1.965fea53d6e3dp-1, 0x1.428a2f98d728bp-1};
};
template <>
struct RootCoef<4> {
    static constexpr double p[6] = {0x1.cea555e5f371ap-1, -0x1.346eefc659ea0p-3,
                                    0x1.00192f0f33fe4p-4, -0x1.ffd60027966f0p-6,
                                    0x1.3cd1d65f2357fp-6, -0x1.6962c4875515fp-7};
    static constexpr double s[4] = {0x1p+0, 0x1.ae89f995ad3adp-1, 0x1.6a09e667f3bcdp-1,
                                    0x1.306fe0a31b715p-1};
};
template <>
struct RootCoef<5> {
    static constexpr double p[6] = {0x1.d81ed544cf8d8p-1, -0x1.f79979e0f7a07p-4,
                                    0x1.918282e8ac699p-5, -0x1.884dc61ebd241p-6,
                                    0x1.dcc723482118fp-7, -0x1.0cd24fbb9fc7fp-7};
    static constexpr double s[5] = {0x1p+0, 0x1.bdb8cdadbe120p-1, 0x1.8406003b2ae5cp-1,
                                    0x1.51cb453b9536cp-1, 0x1.2611186bae675p-1};
};
template <>
struct RootCoef<6> {
    static constexpr double p[6] = {0x1.de8b600813da4p-1, -0x1.a9605d403f50fp-4,
                                    0x1.49c4eebfec5aap-5, -0x1.3d52ab9af5948p-6,
                                    0x1.7ce3ef31baabcp-7, -0x1.aa352530fcd10p-8};
    static constexpr double s[6] = {0x1p+0, 0x1.c823e074ec129p-1, 0x1.965fea53d6e3dp-1,
                                    0x1.6a09e667f3bcdp-1, 0x1.428a2f98d728bp-1,
                                    0x1.1f59ac3c7d6c0p-1};
};
template <>
struct RootCoef<7> {
    static constexpr double p[6] = {0x1.e32fbc32c3ccbp-1, -0x1.702525e04166bp-4,
                                    0x1.179a8cef4f954p-5, -0x1.0a17a5e275aafp-6,
                                    0x1.3c8ea867dbd19p-7, -0x1.60416b7d9015bp-8};
    static constexpr double s[7] = {0x1p+0, 0x1.cfbb031a741a5p-1, 0x1.a402feeb9c533p-1,
                                    0x1.7c6a1f29e2ce6p-1, 0x1.588cea3f093bep-1,
                                    0x1.381147622f886p-1, 0x1.1aa59c4115e7dp-1};
};
template <>
struct RootCoef<8> {
    static constexpr double p[6] = {0x1.e6b29102ed063p-1, -0x1.4477b7dbbc595p-4,
                                    0x1.e5320a8f3a376p-6, -0x1.c9e63649fd2a0p-7,
                                    0x1.0e8b165b61159p-7, -0x1.2bcadec4f337fp-8};
    static constexpr double s[8] = {0x1p+0, 0x1.d5818dcfba487p-1, 0x1.ae89f995ad3adp-1,
                                    0x1.8ace5422aa0dbp-1, 0x1.6a09e667f3bcdp-1,
                                    0x1.4bfdad5362a27p-1, 0x1.306fe0a31b715p-1,
                                    0x1.172b83c7d517bp-1};
};

// y^Q by a fixed sequence of products
template <int Q>
VA_HD double root_ipow(double y)
{
    const double y2 = y * y;
    if constexpr (Q == 2) return y2;
    else if constexpr (Q == 3) return y2 * y;
    else if constexpr (Q == 4) return y2 * y2;
    else if constexpr (Q == 5) return (y2 * y2) * y;
    else if constexpr (Q == 6) { const double y3 = y2 * y; return y3 * y3; }
    else if constexpr (Q == 7) { const double y3 = y2 * y, y4 = y2 * y2; return y4 * y3; }
    else { const double y4 = y2 * y2; return y4 * y4; }
}

// 2^(-j/Q) for 0 <= j < Q by selects (no indexed load)
template <int Q>
VA_HD double root_scale(int j)
{
    double s = RootCoef<Q>::s[0];
#pragma unroll
    for (int i = 1; i < Q; ++i) s = (j == i) ? RootCoef<Q>::s[i] : s;
    return s;
}

VA_HD double bits_double(uint64_t b) { return __builtin_bit_cast(double, b); }

// P_Q(u) by Estrin's scheme
template <int Q>
VA_HD double root_poly(double u, double u2, double u4)
{
    using R = RootCoef<Q>;
    const double p01 = fma(R::p[1], u, R::p[0]);
    const double p23 = fma(R::p[3], u, R::p[2]);
    const double p45 = fma(R::p[5], u, R::p[4]);
    return fma(p45, u4, fma(p23, u2, p01));
}

// x^(-1/QA) if use_a, else x^(-1/QB).  Both candidates' integer splits,
// polynomials and scale constants are formed unconditionally and the lane's
// one is selected (no branch, no table load), so lanes that pick different
// orders execute one instruction stream.
template <int QA, int QB>
VA_HD double ctrl_root(double x, bool use_a)
{
    static_assert(QA >= 2 && QA <= 8 && QB >= 2 && QB <= 8, "root order out of range");
    const bool sub = x < 2.2250738585072014e-308;  // subnormal (or <= 0: replaced below)
    const double xs = sub ? x * 0x1p64 : x;
    const uint64_t bits = __builtin_bit_cast(uint64_t, xs);
    const int e = (int)(bits >> 52) - 1023 - (sub ? 64 : 0);
    const uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    const int ea = e + QA * 2048, eb = e + QB * 2048;  // > 0 for every finite x > 0
    const int ka = ea / QA - 2048, kb = eb / QB - 2048;
    const int ja = ea - (ea / QA) * QA, jb = eb - (eb / QB) * QB;
    const int k = use_a ? ka : kb;
    const int j = use_a ? ja : jb;
    const double m = bits_double(mb);
    const double xr = bits_double(mb + ((uint64_t)(uint32_t)j << 52));  // m 2^j, exact
    const double u = m - 1.5, u2 = u * u, u4 = u2 * u2;
    const double ya = root_poly<QA>(u, u2, u4) * root_scale<QA>(ja);
    const double yb = root_poly<QB>(u, u2, u4) * root_scale<QB>(jb);
    double y = use_a ? ya : yb;
    const double iq = use_a ? 1.0 / (double)QA : 1.0 / (double)QB;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double yqa = root_ipow<QA>(y), yqb = root_ipow<QB>(y);
        const double r = fma(-xr, use_a ? yqa : yqb, 1.0);
        y = fma(y * iq, r, y);
    }
    double res = y * bits_double((uint64_t)(uint32_t)(1023 - k) << 52);
    res = (x == __builtin_huge_val()) ? 0.0 : res;
    res = (x == 0.0) ? __builtin_huge_val() : res;
    res = (x < 0.0) ? __builtin_nan("") : res;
    return (x != x) ? x : res;
}

// ---------------------------------------------------------------------------
// odeint helpers and controller (Boost.odeint 1.74 default_error_checker,
// default_step_adjuster with max_dt = 0, less_with_sign; restated).
// ---------------------------------------------------------------------------
#define VA_DBL_EPSILON 2.220446049250313080847e-16

VA_HD bool less_with_sign(double t1, double t2, double dt)
{
    return (dt > 0.0) ? (t2 - t1 > VA_DBL_EPSILON) : (t1 - t2 > VA_DBL_EPSILON);
}

VA_HD bool less_eq_with_sign(double t1, double t2, double dt)
{
    return (dt > 0.0) ? (t1 - t2 <= VA_DBL_EPSILON) : (t2 - t1 <= VA_DBL_EPSILON);
}

template <int Q>
struct FivePow {
    static constexpr double inv()
    {
        double d = 1.0;
        for (int i = 0; i < Q; ++i) d = d * 5.0;
        return 1.0 / d;
    }
};

// Accept/reject decision and the next step size (default_step_adjuster with
// max_dt = 0):
//   reject (err > 1):   h * max(0.9 err^(-1/(ERR_ORDER-1)), 0.2)
//   accept, err < 0.5:  h * 0.9 max(5^-ORDER, err)^(-1/ORDER)
//   accept otherwise:   h
// with ONE root evaluation: its argument and order are selected first, so lanes
// of a wave that take different decisions share one instruction stream.  When
// every active lane of the wave takes the same decision (a wave-uniform branch)
// only that order's root is formed; otherwise both orders' candidates are
// formed and each lane selects its own.  The lane's value is the same either way.
template <class Tab>
__device__ __forceinline__ bool control_step(double h, double err, double &h_next)
{
    constexpr double lo = FivePow<Tab::ORDER>::inv();
    // fixed-step tableaus (no error estimate) never reach the controller
    constexpr int QI = Tab::ORDER >= 2 ? Tab::ORDER : 2;
    constexpr int QD = Tab::ERR_ORDER >= 3 ? Tab::ERR_ORDER - 1 : 2;
    const bool accept = !(err > 1.0);
    const double base = accept ? ((lo < err) ? err : lo) : err;
    const uint64_t all = __builtin_amdgcn_ballot_w64(true);
    const uint64_t acc = __builtin_amdgcn_ballot_w64(accept);
    double root;
    if (acc == all) root = ctrl_root<QI, QI>(base, true);
    else if (acc == 0) root = ctrl_root<QD, QD>(base, false);
    else root = ctrl_root<QI, QD>(base, accept);
    const double f = 0.9 * root;
    if (!accept) {
        h_next = h * ((f < 0.2) ? 0.2 : f);
    } else {
        h_next = (err < 0.5) ? h * f : h;
    }
    return accept;
}

// ---------------------------------------------------------------------------
// Systems.  rhs(p, x, f); vjp(p, x, w, wJx, pbar) accumulates
// pbar += w^T df/dp.  Formulas follow examples/VanDerPol/main.cpp:41-42,
// examples/HarmonicOscillator/main.cpp:27-28 and
// examples/GeneralizedLotkaVolterra/main.cpp:109-118 (accumulation starts at
// r_i: g_i = fma(A_i,N-1, x_N-1, ... fma(A_i0, x_0, r_i))).
// ---------------------------------------------------------------------------
struct VanDerPol {
    static constexpr int N = 2, P = 1;
    static constexpr int FLOP_RHS = 5, FLOP_VJP = 13;
    struct Params {
        double mu;
    };
    __device__ static void load(Params &q, const double *p, int64_t B, int64_t b, double)
    {
        q.mu = p[b];
    }
    VA_HD static void rhs(const Params &q, const double *x, double *f)
    {
        double x0 = x[0], x1 = x[1];
        f[0] = x1;
        f[1] = q.mu * ((1.0 - x0 * x0) * x1 - x0);
    }
    VA_HD static void vjp(const Params &q, const double *x, const double *w, double *wjx,
                          double *pbar)
    {
        double x0 = x[0], x1 = x[1];
        double wm = w[1] * q.mu;
        double qq = 1.0 - x0 * x0;
        wjx[0] = wm * fma(-2.0 * x0, x1, -1.0);
        wjx[1] = fma(wm, qq, w[0]);
        pbar[0] = pbar[0] + w[1] * fma(qq, x1, -x0);
    }
};

struct Oscillator {  // k and c both differentiable parameters
    static constexpr int N = 2, P = 2;
    static constexpr int FLOP_RHS = 3, FLOP_VJP = 7;
    struct Params {
        double k, c;
    };
    __device__ static void load(Params &q, const double *p, int64_t B, int64_t b, double)
    {
        q.k = p[b];
        q.c = p[B + b];
    }
    VA_HD static void rhs(const Params &q, const double *x, double *f)
    {
        f[0] = x[1];
        f[1] = -q.k * x[0] - q.c * x[1];
    }
    VA_HD static void vjp(const Params &q, const double *x, const double *w, double *wjx,
                          double *pbar)
    {
        wjx[0] = -(w[1] * q.k);
        wjx[1] = w[0] - w[1] * q.c;
        pbar[0] = pbar[0] + -(w[1] * x[0]);
        pbar[1] = pbar[1] + -(w[1] * x[1]);
    }
};

struct OscillatorFixedK {  // reference form: k a constant of the functor, mu the parameter
    static constexpr int N = 2, P = 1;
    static constexpr int FLOP_RHS = 3, FLOP_VJP = 5;
    struct Params {
        double k, mu;
    };
    __device__ static void load(Params &q, const double *p, int64_t B, int64_t b, double kconst)
    {
        q.k = kconst;
        q.mu = p[b];
    }
    VA_HD static void rhs(const Params &q, const double *x, double *f)
    {
        f[0] = x[1];
        f[1] = -q.k * x[0] - q.mu * x[1];
    }
    VA_HD static void vjp(const Params &q, const double *x, const double *w, double *wjx,
                          double *pbar)
    {
        wjx[0] = -(w[1] * q.k);
        wjx[1] = w[0] - w[1] * q.mu;
        pbar[0] = pbar[0] + -(w[1] * x[1]);
    }
};

template <int NN>
struct LotkaVolterra {
    static constexpr int N = NN, P = NN + NN * NN;
    static constexpr int FLOP_RHS = 2 * NN * NN + NN;
    static constexpr int FLOP_VJP = 4 * NN * NN + 3 * NN;
    struct Params {
        double r[NN];
        double A[NN][NN];
    };
    __device__ static void load(Params &q, const double *p, int64_t B, int64_t b, double)
    {
#pragma unroll
        for (int i = 0; i < NN; ++i) q.r[i] = p[(int64_t)i * B + b];
#pragma unroll
        for (int i = 0; i < NN; ++i)
#pragma unroll
            for (int j = 0; j < NN; ++j) q.A[i][j] = p[(int64_t)(NN * (i + 1) + j) * B + b];
    }
    // g_i = r_i + sum_j A_ij x_j as two interleaved partial sums (even j from
    // r_i, odd j from 0), added at the end.
    VA_HD static double growth(const Params &q, const double *x, int i)
    {
        double ge = q.r[i], go = 0.0;
#pragma unroll
        for (int j = 0; j < NN; j += 2) ge = fma(q.A[i][j], x[j], ge);
#pragma unroll
        for (int j = 1; j < NN; j += 2) go = fma(q.A[i][j], x[j], go);
        return ge + go;
    }
    VA_HD static void rhs(const Params &q, const double *x, double *f)
    {
#pragma unroll
        for (int i = 0; i < NN; ++i) f[i] = x[i] * growth(q, x, i);
    }
    // pbar layout: r_bar[NN] then A_bar row-major (same as the parameter vector).
    VA_HD static void vjp(const Params &q, const double *x, const double *w, double *wjx,
                          double *pbar)
    {
        double u[NN];
#pragma unroll
        for (int i = 0; i < NN; ++i) u[i] = w[i] * x[i];
#pragma unroll
        for (int j = 0; j < NN; ++j) {
            double ae = w[j] * growth(q, x, j), ao = 0.0;
#pragma unroll
            for (int i = 0; i < NN; i += 2) ae = fma(u[i], q.A[i][j], ae);
#pragma unroll
            for (int i = 1; i < NN; i += 2) ao = fma(u[i], q.A[i][j], ao);
            wjx[j] = ae + ao;
        }
#pragma unroll
        for (int i = 0; i < NN; ++i) {
            pbar[i] = pbar[i] + u[i];
#pragma unroll
            for (int j = 0; j < NN; ++j)
                pbar[NN * (i + 1) + j] = fma(u[i], x[j], pbar[NN * (i + 1) + j]);
        }
    }
};

// ---------------------------------------------------------------------------
// Explicit RK attempt from (y, K[0] = f(y)):
//   X_m  = y + sum_{j<m, a_mj != 0} (a_mj h) K_j   (fma chain in j order)
//   y'   = y + sum_{b_j != 0} (b_j h) K_j
//   xerr = 0 + sum_{db_j != 0} (db_j h) K_j
// ---------------------------------------------------------------------------
template <class Sys, class Tab>
struct RK {
    static constexpr int N = Sys::N;

    __device__ __forceinline__ static void stage_input(int m, double h, const double *y,
                                                       const double (*K)[N], double *X)
    {
#pragma unroll
        for (int i = 0; i < N; ++i) X[i] = y[i];
#pragma unroll
        for (int j = 0; j < Tab::SF; ++j) {
            if (j < m && Tab::A[m][j] != 0.0) {
                const double ah = Tab::A[m][j] * h;
#pragma unroll
                for (int i = 0; i < N; ++i) X[i] = fma(ah, K[j][i], X[i]);
            }
        }
    }

    __device__ __forceinline__ static void attempt(const typename Sys::Params &q, const double *y,
                                                   double (*K)[N], double h, double *ynew,
                                                   double *xerr)
    {
#pragma unroll
        for (int m = 1; m < Tab::SF; ++m) {
            double X[N];
            stage_input(m, h, y, K, X);
            Sys::rhs(q, X, K[m]);
        }
#pragma unroll
        for (int i = 0; i < N; ++i) ynew[i] = y[i];
#pragma unroll
        for (int j = 0; j < Tab::S; ++j) {
            if (Tab::B[j] != 0.0) {
                const double bh = Tab::B[j] * h;
#pragma unroll
                for (int i = 0; i < N; ++i) ynew[i] = fma(bh, K[j][i], ynew[i]);
            }
        }
        if constexpr (Tab::HAS_ERR) {
#pragma unroll
            for (int i = 0; i < N; ++i) xerr[i] = 0.0;
#pragma unroll
            for (int j = 0; j < Tab::SF; ++j) {
                if (Tab::DB[j] != 0.0) {
                    const double dh = Tab::DB[j] * h;
#pragma unroll
                    for (int i = 0; i < N; ++i) xerr[i] = fma(dh, K[j][i], xerr[i]);
                }
            }
        }
    }

    // default_error_checker::error with norm_inf (std::max semantics)
    __device__ __forceinline__ static double error_norm(double atol, double rtol, double h,
                                                        const double *y, const double *dydt,
                                                        const double *xerr)
    {
        double m = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double den = fma(rtol, fma(h, fabs(dydt[i]), fabs(y[i])), atol);
            double e = fabs(xerr[i]) / den;
            m = (m < e) ? e : m;
        }
        return m;
    }
};

}  // namespace va
