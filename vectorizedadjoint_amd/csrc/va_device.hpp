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
// Step-controller pow (x >= 0).  Own implementation from textbook methods:
//   log x = e ln2 + 2 atanh(s), s = (m-1)/(m+1), m in [sqrt(1/2), sqrt(2)),
//   exp t = 2^k exp(r), r = t - k ln2 (two-constant Cody-Waite split), Taylor to
//   degree 13.  Only IEEE basic operations, fma, frexp, ldexp and rint, so the
//   result is the same bits on the GPU and on the host.
// ---------------------------------------------------------------------------
#define VA_LN2_HI 0x1.62e42fee00000p-1
#define VA_LN2_LO 0x1.a39ef35793c76p-33
#define VA_INV_LN2 0x1.71547652b82fep+0
#define VA_SQRT_HALF 0x1.6a09e667f3bcdp-1

VA_HD double log_pos(double x)
{
    int e;
    double m = frexp(x, &e);
    if (m < VA_SQRT_HALF) {
        m = m * 2.0;
        e = e - 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    // q = 1/3 + z/5 + ... + z^10/23 by Estrin's scheme (short dependency chains:
    // a dependent FP64 FMA costs ~32 cycles on gfx950)
    double z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
    double p01 = fma(1.0 / 5.0, z, 1.0 / 3.0);
    double p23 = fma(1.0 / 9.0, z, 1.0 / 7.0);
    double p45 = fma(1.0 / 13.0, z, 1.0 / 11.0);
    double p67 = fma(1.0 / 17.0, z, 1.0 / 15.0);
    double p89 = fma(1.0 / 21.0, z, 1.0 / 19.0);
    double p03 = fma(p23, z2, p01);
    double p47 = fma(p67, z2, p45);
    double p8a = fma(1.0 / 23.0, z2, p89);
    double p07 = fma(p47, z4, p03);
    double q = fma(p8a, z8, p07);
    double zq = z * q;
    double s2 = s + s;
    double lm = fma(s2, zq, s2);
    double de = (double)e;
    return fma(de, VA_LN2_HI, fma(de, VA_LN2_LO, lm));
}

VA_HD double exp_(double t)
{
    if (t != t) return t;
    if (t > 709.8) return __builtin_huge_val();
    if (t < -745.2) return 0.0;
    double k = rint(t * VA_INV_LN2);
    double r = fma(-k, VA_LN2_HI, t);
    r = fma(-k, VA_LN2_LO, r);
    // sum_{i=0}^{13} r^i / i! by Estrin's scheme
    double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
    double e01 = fma(1.0, r, 1.0);
    double e23 = fma(1.0 / 6.0, r, 0.5);
    double e45 = fma(1.0 / 120.0, r, 1.0 / 24.0);
    double e67 = fma(1.0 / 5040.0, r, 1.0 / 720.0);
    double e89 = fma(1.0 / 362880.0, r, 1.0 / 40320.0);
    double eab = fma(1.0 / 39916800.0, r, 1.0 / 3628800.0);
    double ecd = fma(1.0 / 6227020800.0, r, 1.0 / 479001600.0);
    double e03 = fma(e23, r2, e01);
    double e47 = fma(e67, r2, e45);
    double e8b = fma(eab, r2, e89);
    double e07 = fma(e47, r4, e03);
    double e8d = fma(ecd, r4, e8b);
    double q = fma(e8d, r8, e07);
    return ldexp(q, (int)k);
}

VA_HD double ctrl_pow(double x, double y)
{
    if (x != x || y != y) return x + y;
    if (y == 0.0) return 1.0;
    if (x == 0.0) return y < 0.0 ? __builtin_huge_val() : 0.0;
    if (x < 0.0) return __builtin_nan("");
    if (x == __builtin_huge_val()) return y < 0.0 ? 0.0 : __builtin_huge_val();
    return exp_(y * log_pos(x));
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

template <class Tab>
VA_HD double decrease_step(double h, double err)
{
    double f = 0.9 * ctrl_pow(err, -1.0 / (double)(Tab::ERR_ORDER - 1));
    f = (f < 0.2) ? 0.2 : f;
    return h * f;
}

template <class Tab>
VA_HD double increase_step(double h, double err)
{
    if (err < 0.5) {
        constexpr double lo = FivePow<Tab::ORDER>::inv();
        err = (lo < err) ? err : lo;
        h = h * (0.9 * ctrl_pow(err, -1.0 / (double)Tab::ORDER));
    }
    return h;
}

// Accept/reject decision and the next step size with ONE pow evaluation:
// decrease_step and increase_step above, with the pow argument and exponent
// selected first.  Lanes of a wave that take different decisions then share a
// single pow instead of executing one per branch; every value is computed with
// the same operations as in the two functions (bit-identical).
template <class Tab>
VA_HD bool control_step(double h, double err, double &h_next)
{
    constexpr double lo = FivePow<Tab::ORDER>::inv();
    constexpr double e_dec = -1.0 / (double)(Tab::ERR_ORDER - 1);
    constexpr double e_inc = -1.0 / (double)Tab::ORDER;
    const bool accept = !(err > 1.0);
    const double base = accept ? ((lo < err) ? err : lo) : err;
    const double f = 0.9 * ctrl_pow(base, accept ? e_inc : e_dec);
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
            for (int j = 0; j < NN; ++j) pbar[NN * (i + 1) + j] = pbar[NN * (i + 1) + j] + u[i] * x[j];
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
