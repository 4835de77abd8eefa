/*
 * va_oracle.c -- CPU restatement of VectorizedAdjoint's hot path (test
 * infrastructure; see va_oracle.h for the scope and the file:line map).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -mfma).
 */
#define _GNU_SOURCE
#include "va_oracle.h"

/* ORC_MUTANT selects a deliberately wrong variant for the mutation tests
 * (tests/test_oracle_mutants.py builds them into a temporary directory);
 * 0, the only value liboracle.so is built with, is the restatement itself. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* Operation sequence.  By default every step follows the IEEE operation order
 * of the reference's own arithmetic: Boost.odeint 1.74's scale_sum stage sums
 * and rel_error norm with each product rounded (the reference is built without
 * FMA contraction: -mavx2, no -mfma, examples/GeneralizedLotkaVolterra/build/
 * CMakeFiles/lotka.dir/flags.make:9), the example functors' loop orders, the
 * C library's pow in the step adjuster, and the reverse sweep of
 * lib/include/detail/backpropagation.hpp:83-158 with h = t[n+1] - t[n].
 * ORC_FUSED=1 selects an alternative sequence -- fused multiply-adds, split
 * growth sums, a Newton root in place of pow, the stored step size in the
 * reverse -- kept only so that tests can measure how far results move when
 * nothing but the rounding sequence changes. */
#ifndef ORC_FUSED
#define ORC_FUSED 0
#endif
/* ORC_FUSED_REVERSE: the alternative sequence in the reverse sweep only (stage
 * recompute, growth rates, Kbar and parameter accumulation, stored step), with
 * the reference's forward -- so the gradients' sensitivity to the reverse
 * sweep's rounding alone can be measured. */
#ifndef ORC_FUSED_REVERSE
#define ORC_FUSED_REVERSE ORC_FUSED
#endif
#if ORC_FUSED
#define ORC_MAD(a, b, c) fma((a), (b), (c))
#define STORED_STEP(t_next, t, h) (h)
#else
#define ORC_MAD(a, b, c) ((c) + (a) * (b))
/* the trajectory keeps h_n = t_{n+1} - t_n, the reference's GetDt
 * (StateStorage.hpp:22), which its reverse uses */
#define STORED_STEP(t_next, t, h) ((t_next) - (t))
#endif

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Controller root x^(-1/q), q = 2..8 (the step adjuster raises the error    */
/* ratio to -1/q).  Own method:                                              */
/*   x = m 2^e, m in [1,2), e = q k + j with 0 <= j < q, so                  */
/*   x^(-1/q) = 2^-k (m 2^j)^(-1/q);                                         */
/*   start  y = P_q(m - 1.5) 2^(-j/q), P_q the degree-5 Chebyshev            */
/*          interpolant of m^(-1/q) on [1,2] (relative error <= 1.4e-5);     */
/*   refine two Newton steps for y^-q = m 2^j:  y += (y/q)(1 - m 2^j y^q),   */
/*          each taking the relative error d to about (q+1)/2 d^2;           */
/*   scale  by 2^-k, exactly.                                                */
/* Only IEEE basic operations, fma and integer operations on the binary      */
/* representation are used, so the HIP kernels (which use the same sequence) */
/* obtain the same bits.  The result is within about one ulp of the exact    */
/* root; it differs from pow(x, -1.0/q) by that plus the rounding of the     */
/* exponent -1.0/q, a relative |ln x| 2^-54 (under 1.5 ulp for x <= 1e6).    */
/* ------------------------------------------------------------------------ */
static const double ROOT_P[7][6] = {
    /* q = 2 */ {0x1.a20d0b2f16c1fp-1, -0x1.16b40de0b165ap-2, 0x1.1559ba6062540p-3,
                 -0x1.33fdf4c87b433p-4, 0x1.a0f6e83f40934p-5, -0x1.f69d70fc899e8p-6},
    /* q = 3 */ {0x1.bf46914f5b764p-1, -0x1.8d95060abdbffp-3, 0x1.5ffe431f2a20ep-4,
                 -0x1.6cc93eb91b616p-5, 0x1.d18071c6f18d9p-6, -0x1.0e87006cbb389p-6},
    /* q = 4 */ {0x1.cea555e5f371ap-1, -0x1.346eefc659ea0p-3, 0x1.00192f0f33fe4p-4,
                 -0x1.ffd60027966f0p-6, 0x1.3cd1d65f2357fp-6, -0x1.6962c4875515fp-7},
    /* q = 5 */ {0x1.d81ed544cf8d8p-1, -0x1.f79979e0f7a07p-4, 0x1.918282e8ac699p-5,
                 -0x1.884dc61ebd241p-6, 0x1.dcc723482118fp-7, -0x1.0cd24fbb9fc7fp-7},
    /* q = 6 */ {0x1.de8b600813da4p-1, -0x1.a9605d403f50fp-4, 0x1.49c4eebfec5aap-5,
                 -0x1.3d52ab9af5948p-6, 0x1.7ce3ef31baabcp-7, -0x1.aa352530fcd10p-8},
    /* q = 7 */ {0x1.e32fbc32c3ccbp-1, -0x1.702525e04166bp-4, 0x1.179a8cef4f954p-5,
                 -0x1.0a17a5e275aafp-6, 0x1.3c8ea867dbd19p-7, -0x1.60416b7d9015bp-8},
    /* q = 8 */ {0x1.e6b29102ed063p-1, -0x1.4477b7dbbc595p-4, 0x1.e5320a8f3a376p-6,
                 -0x1.c9e63649fd2a0p-7, 0x1.0e8b165b61159p-7, -0x1.2bcadec4f337fp-8},
};

/* 2^(-j/q), correctly rounded */
static const double ROOT_S[7][8] = {
    /* q = 2 */ {0x1p+0, 0x1.6a09e667f3bcdp-1},
    /* q = 3 */ {0x1p+0, 0x1.965fea53d6e3dp-1, 0x1.428a2f98d728bp-1},
    /* q = 4 */ {0x1p+0, 0x1.ae89f995ad3adp-1, 0x1.6a09e667f3bcdp-1, 0x1.306fe0a31b715p-1},
    /* q = 5 */ {0x1p+0, 0x1.bdb8cdadbe120p-1, 0x1.8406003b2ae5cp-1, 0x1.51cb453b9536cp-1,
                 0x1.2611186bae675p-1},
    /* q = 6 */ {0x1p+0, 0x1.c823e074ec129p-1, 0x1.965fea53d6e3dp-1, 0x1.6a09e667f3bcdp-1,
                 0x1.428a2f98d728bp-1, 0x1.1f59ac3c7d6c0p-1},
    /* q = 7 */ {0x1p+0, 0x1.cfbb031a741a5p-1, 0x1.a402feeb9c533p-1, 0x1.7c6a1f29e2ce6p-1,
                 0x1.588cea3f093bep-1, 0x1.381147622f886p-1, 0x1.1aa59c4115e7dp-1},
    /* q = 8 */ {0x1p+0, 0x1.d5818dcfba487p-1, 0x1.ae89f995ad3adp-1, 0x1.8ace5422aa0dbp-1,
                 0x1.6a09e667f3bcdp-1, 0x1.4bfdad5362a27p-1, 0x1.306fe0a31b715p-1,
                 0x1.172b83c7d517bp-1},
};

/* y^q by a fixed sequence of products */
static double root_ipow(double y, int q)
{
    double y2 = y * y;
    switch (q) {
    case 2: return y2;
    case 3: return y2 * y;
    case 4: return y2 * y2;
    case 5: { double y4 = y2 * y2; return y4 * y; }
    case 6: { double y3 = y2 * y; return y3 * y3; }
    case 7: { double y3 = y2 * y, y4 = y2 * y2; return y4 * y3; }
    default: { double y4 = y2 * y2; return y4 * y4; }
    }
}

static double bits_to_double(uint64_t b)
{
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

double orc_root(double x, int q)
{
    if (q < 2 || q > 8) return NAN;
    if (x != x) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return INFINITY;
    if (isinf(x)) return 0.0;
    int eadj = 0;
    if (x < DBL_MIN) { /* subnormal: scale into the normal range */
        x = x * 0x1p64;
        eadj = -64;
    }
    uint64_t bits;
    memcpy(&bits, &x, sizeof bits);
    int e = (int)(bits >> 52) - 1023 + eadj;
    uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    int ee = e + q * 2048; /* > 0, so / and % are floor division */
    int k = ee / q - 2048;
    int j = ee - (ee / q) * q;
    double m = bits_to_double(mb);
    double xr = bits_to_double(mb + ((uint64_t)j << 52)); /* m 2^j, exact */
    const double *c = ROOT_P[q - 2];
    double u = m - 1.5, u2 = u * u, u4 = u2 * u2;
    double p01 = fma(c[1], u, c[0]);
    double p23 = fma(c[3], u, c[2]);
    double p45 = fma(c[5], u, c[4]);
    double y = fma(p45, u4, fma(p23, u2, p01)) * ROOT_S[q - 2][j];
    double iq = 1.0 / (double)q;
    for (int it = 0; it < 2; ++it) {
        double r = fma(-xr, root_ipow(y, q), 1.0);
        y = fma(y * iq, r, y);
    }
    return y * bits_to_double((uint64_t)(1023 - k) << 52);
}

/* ------------------------------------------------------------------------ */
/* Butcher tableaus (ButcherTable.hpp:39-251 plus DOPRI5, which the         */
/* reference leaves unsupported at ButcherTable.hpp:247-250).  Coefficients  */
/* are the published ones: classic RK4; Cash & Karp (1990); Dormand & Prince */
/* (1980); Fehlberg (1968, NASA TR R-287).  db = b - b_embedded is rounded   */
/* the way odeint 1.74 rounds it (difference of two rounded quotients).      */
/* ------------------------------------------------------------------------ */
static void tab_clear(orc_tableau *t)
{
    memset(t, 0, sizeof(*t));
}

int orc_get_tableau(int id, orc_tableau *t)
{
    tab_clear(t);
    switch (id) {
    case ORC_TAB_EULER:
        t->s = t->s_fwd = 1;
        t->order = 1;
        t->b[0] = 1.0;
        return 0;
    case ORC_TAB_RK4:
        t->s = t->s_fwd = 4;
        t->order = 4;
        t->a[1][0] = 1.0 / 2.0;
        t->a[2][1] = 1.0 / 2.0;
        t->a[3][2] = 1.0;
        t->b[0] = 1.0 / 6.0; t->b[1] = 1.0 / 3.0; t->b[2] = 1.0 / 3.0; t->b[3] = 1.0 / 6.0;
        t->c[1] = 1.0 / 2.0; t->c[2] = 1.0 / 2.0; t->c[3] = 1.0;
        return 0;
    case ORC_TAB_CASH_KARP54:
        t->s = t->s_fwd = 6;
        t->order = 5;
        t->err_order = 4;
        t->has_err = 1;
        t->a[1][0] = 1.0 / 5.0;
        t->a[2][0] = 3.0 / 40.0; t->a[2][1] = 9.0 / 40.0;
        t->a[3][0] = 3.0 / 10.0; t->a[3][1] = -9.0 / 10.0; t->a[3][2] = 6.0 / 5.0;
        t->a[4][0] = -11.0 / 54.0; t->a[4][1] = 5.0 / 2.0; t->a[4][2] = -70.0 / 27.0;
        t->a[4][3] = 35.0 / 27.0;
        t->a[5][0] = 1631.0 / 55296.0; t->a[5][1] = 175.0 / 512.0; t->a[5][2] = 575.0 / 13824.0;
        t->a[5][3] = 44275.0 / 110592.0; t->a[5][4] = 253.0 / 4096.0;
        t->b[0] = 37.0 / 378.0; t->b[2] = 250.0 / 621.0; t->b[3] = 125.0 / 594.0;
        t->b[5] = 512.0 / 1771.0;
        t->db[0] = 37.0 / 378.0 - 2825.0 / 27648.0;
        t->db[2] = 250.0 / 621.0 - 18575.0 / 48384.0;
        t->db[3] = 125.0 / 594.0 - 13525.0 / 55296.0;
        t->db[4] = -277.0 / 14336.0;
        t->db[5] = 512.0 / 1771.0 - 1.0 / 4.0;
        t->c[1] = 1.0 / 5.0; t->c[2] = 3.0 / 10.0; t->c[3] = 3.0 / 5.0; t->c[4] = 1.0;
        t->c[5] = 7.0 / 8.0;
        return 0;
    case ORC_TAB_DOPRI5:
        t->s = 6;
        t->s_fwd = 7;
        t->order = 5;
        t->err_order = 4;
        t->has_err = 1;
        t->fsal = 1;
        t->a[1][0] = 1.0 / 5.0;
        t->a[2][0] = 3.0 / 40.0; t->a[2][1] = 9.0 / 40.0;
        t->a[3][0] = 44.0 / 45.0; t->a[3][1] = -56.0 / 15.0; t->a[3][2] = 32.0 / 9.0;
        t->a[4][0] = 19372.0 / 6561.0; t->a[4][1] = -25360.0 / 2187.0;
        t->a[4][2] = 64448.0 / 6561.0; t->a[4][3] = -212.0 / 729.0;
        t->a[5][0] = 9017.0 / 3168.0; t->a[5][1] = -355.0 / 33.0; t->a[5][2] = 46732.0 / 5247.0;
        t->a[5][3] = 49.0 / 176.0; t->a[5][4] = -5103.0 / 18656.0;
        t->a[6][0] = 35.0 / 384.0; t->a[6][2] = 500.0 / 1113.0; t->a[6][3] = 125.0 / 192.0;
        t->a[6][4] = -2187.0 / 6784.0; t->a[6][5] = 11.0 / 84.0;
        t->b[0] = 35.0 / 384.0; t->b[2] = 500.0 / 1113.0; t->b[3] = 125.0 / 192.0;
        t->b[4] = -2187.0 / 6784.0; t->b[5] = 11.0 / 84.0;
        t->db[0] = 35.0 / 384.0 - 5179.0 / 57600.0;
        t->db[2] = 500.0 / 1113.0 - 7571.0 / 16695.0;
        t->db[3] = 125.0 / 192.0 - 393.0 / 640.0;
        t->db[4] = -2187.0 / 6784.0 - -92097.0 / 339200.0;
        t->db[5] = 11.0 / 84.0 - 187.0 / 2100.0;
        t->db[6] = -1.0 / 40.0;
        t->c[1] = 1.0 / 5.0; t->c[2] = 3.0 / 10.0; t->c[3] = 4.0 / 5.0; t->c[4] = 8.0 / 9.0;
        t->c[5] = 1.0; t->c[6] = 1.0;
        return 0;
    case ORC_TAB_FEHLBERG78:
        t->s = t->s_fwd = 13;
        t->order = 8;
        t->err_order = 7;
        t->has_err = 1;
        t->a[1][0] = 2.0 / 27.0;
        t->a[2][0] = 1.0 / 36.0; t->a[2][1] = 1.0 / 12.0;
        t->a[3][0] = 1.0 / 24.0; t->a[3][2] = 1.0 / 8.0;
        t->a[4][0] = 5.0 / 12.0; t->a[4][2] = -25.0 / 16.0; t->a[4][3] = 25.0 / 16.0;
        t->a[5][0] = 1.0 / 20.0; t->a[5][3] = 1.0 / 4.0; t->a[5][4] = 1.0 / 5.0;
        t->a[6][0] = -25.0 / 108.0; t->a[6][3] = 125.0 / 108.0; t->a[6][4] = -65.0 / 27.0;
        t->a[6][5] = 125.0 / 54.0;
        t->a[7][0] = 31.0 / 300.0; t->a[7][4] = 61.0 / 225.0; t->a[7][5] = -2.0 / 9.0;
        t->a[7][6] = 13.0 / 900.0;
        t->a[8][0] = 2.0; t->a[8][3] = -53.0 / 6.0; t->a[8][4] = 704.0 / 45.0;
        t->a[8][5] = -107.0 / 9.0; t->a[8][6] = 67.0 / 90.0; t->a[8][7] = 3.0;
        t->a[9][0] = -91.0 / 108.0; t->a[9][3] = 23.0 / 108.0; t->a[9][4] = -976.0 / 135.0;
        t->a[9][5] = 311.0 / 54.0; t->a[9][6] = -19.0 / 60.0; t->a[9][7] = 17.0 / 6.0;
        t->a[9][8] = -1.0 / 12.0;
        t->a[10][0] = 2383.0 / 4100.0; t->a[10][3] = -341.0 / 164.0; t->a[10][4] = 4496.0 / 1025.0;
        t->a[10][5] = -301.0 / 82.0; t->a[10][6] = 2133.0 / 4100.0; t->a[10][7] = 45.0 / 82.0;
        t->a[10][8] = 45.0 / 164.0; t->a[10][9] = 18.0 / 41.0;
        t->a[11][0] = 3.0 / 205.0; t->a[11][5] = -6.0 / 41.0; t->a[11][6] = -3.0 / 205.0;
        t->a[11][7] = -3.0 / 41.0; t->a[11][8] = 3.0 / 41.0; t->a[11][9] = 6.0 / 41.0;
        t->a[12][0] = -1777.0 / 4100.0; t->a[12][3] = -341.0 / 164.0; t->a[12][4] = 4496.0 / 1025.0;
        t->a[12][5] = -289.0 / 82.0; t->a[12][6] = 2193.0 / 4100.0; t->a[12][7] = 51.0 / 82.0;
        t->a[12][8] = 33.0 / 164.0; t->a[12][9] = 12.0 / 41.0; t->a[12][11] = 1.0;
        t->b[5] = 34.0 / 105.0; t->b[6] = 9.0 / 35.0; t->b[7] = 9.0 / 35.0; t->b[8] = 9.0 / 280.0;
        t->b[9] = 9.0 / 280.0; t->b[11] = 41.0 / 840.0; t->b[12] = 41.0 / 840.0;
        t->db[0] = 0.0 - 41.0 / 840.0;
        t->db[10] = 0.0 - 41.0 / 840.0;
        t->db[11] = 41.0 / 840.0;
        t->db[12] = 41.0 / 840.0;
        t->c[1] = 2.0 / 27.0; t->c[2] = 1.0 / 9.0; t->c[3] = 1.0 / 6.0; t->c[4] = 5.0 / 12.0;
        t->c[5] = 1.0 / 2.0; t->c[6] = 5.0 / 6.0; t->c[7] = 1.0 / 6.0; t->c[8] = 2.0 / 3.0;
        t->c[9] = 1.0 / 3.0; t->c[10] = 1.0; t->c[11] = 0.0; t->c[12] = 1.0;
        return 0;
    default:
        return -1;
    }
}

/* ------------------------------------------------------------------------ */
/* Systems: right-hand sides as the reference examples define them, and     */
/* their hand-derived vector-Jacobian products (these replace the AADC tape, */
/* AadData.hpp:124-171,291-330).                                             */
/* ------------------------------------------------------------------------ */

/* Lotka-Volterra growth rate g_i = r_i + sum_j A_ij x_j.  Reference order
 * (examples/GeneralizedLotkaVolterra/main.cpp:112-117): sum = 0.0, then
 * sum += A_ij x_j for j = 0..N-1, each product rounded, then r_i + sum.
 * ORC_FUSED: two interleaved fma chains (even j from r_i, odd j from 0). */
static double glv_growth(int n, const double *Ai, double r, const double *x, int fused)
{
    if (fused) {
        double ge = r, go = 0.0;
        for (int j = 0; j < n; j += 2) ge = fma(Ai[j], x[j], ge);
        for (int j = 1; j < n; j += 2) go = fma(Ai[j], x[j], go);
        return ge + go;
    }
    double sum = 0.0;
    for (int j = 0; j < n; ++j) sum = sum + Ai[j] * x[j];
    return r + sum;
}

static orc_user_rhs g_user_rhs = NULL;
static orc_user_vjp g_user_vjp = NULL;

void orc_set_user_system(orc_user_rhs rhs, orc_user_vjp vjp)
{
    g_user_rhs = rhs;
    g_user_vjp = vjp;
}

static void rhs_v(int system, int n, const double *x, const double *p, double t, double K,
                  double *f, int fused)
{
    (void)t;
    if (system == ORC_SYS_USER) {
        if (g_user_rhs) g_user_rhs(x, p, f);
        else for (int i = 0; i < n; ++i) f[i] = NAN;
        return;
    }
    switch (system) {
    case ORC_SYS_VAN_DER_POL: { /* examples/VanDerPol/main.cpp:41-42 */
        double x0 = x[0], x1 = x[1];
        f[0] = x1;
        f[1] = p[0] * ((1.0 - x0 * x0) * x1 - x0);
        break;
    }
    case ORC_SYS_OSCILLATOR: /* examples/HarmonicOscillator/main.cpp:27-28 with k a parameter */
        f[0] = x[1];
        f[1] = -p[0] * x[0] - p[1] * x[1];
        break;
    case ORC_SYS_OSCILLATOR_FIXED_K: /* examples/HarmonicOscillator/main.cpp:27-28 */
        f[0] = x[1];
        f[1] = -K * x[0] - p[0] * x[1];
        break;
    case ORC_SYS_LOTKA_VOLTERRA:
    case ORC_SYS_LOTKA_VOLTERRA_SHARED: /* examples/GeneralizedLotkaVolterra/main.cpp:109-118 */
        for (int i = 0; i < n; ++i)
            f[i] = x[i] * glv_growth(n, p + (size_t)n * (i + 1), p[i], x, fused);
        break;
    default:
        for (int i = 0; i < n; ++i) f[i] = NAN;
    }
}

void orc_rhs(int system, int n, const double *x, const double *p, double t, double K,
             double *f)
{
    rhs_v(system, n, x, p, t, K, f, ORC_FUSED);
}

static void vjp_v(int system, int n, int np, const double *x, const double *p, double t,
                  double K, const double *w, double *wJx, double *wJp, int fused)
{
    if (system == ORC_SYS_USER) {
        if (g_user_vjp) g_user_vjp(x, p, w, wJx, wJp);
        else for (int i = 0; i < n; ++i) wJx[i] = NAN;
        return;
    }
    (void)t;
    (void)np;
    switch (system) {
    case ORC_SYS_VAN_DER_POL: {
        double x0 = x[0], x1 = x[1];
        double wm = w[1] * p[0];
        double q = 1.0 - x0 * x0;
#if ORC_MUTANT == 3 /* planted sign error, test builds only */
        wJx[0] = wm * fma(2.0 * x0, x1, -1.0);
#else
        wJx[0] = wm * fma(-2.0 * x0, x1, -1.0);
#endif
        wJx[1] = fma(wm, q, w[0]);
        wJp[0] = w[1] * fma(q, x1, -x0);
        break;
    }
    case ORC_SYS_OSCILLATOR:
        wJx[0] = -(w[1] * p[0]);
        wJx[1] = w[0] - w[1] * p[1];
        wJp[0] = -(w[1] * x[0]);
        wJp[1] = -(w[1] * x[1]);
        break;
    case ORC_SYS_OSCILLATOR_FIXED_K:
        wJx[0] = -(w[1] * K);
        wJx[1] = w[0] - w[1] * p[0];
        wJp[0] = -(w[1] * x[1]);
        break;
    case ORC_SYS_LOTKA_VOLTERRA:
    case ORC_SYS_LOTKA_VOLTERRA_SHARED: {
        /* f_i = x_i g_i, g_i = r_i + sum_j A_ij x_j
         * wJx_j = w_j g_j + sum_i (w_i x_i) A_ij;  r_bar_i = w_i x_i;  A_bar_ij = w_i x_i x_j
         * The sum over i is split like the growth sum: even i starting from
         * w_j g_j, odd i starting from 0, added at the end. */
        double u[ORC_MAX_N];
        for (int i = 0; i < n; ++i) u[i] = w[i] * x[i];
        for (int j = 0; j < n; ++j) {
            double g = glv_growth(n, p + (size_t)n * (j + 1), p[j], x, fused);
            double ae = w[j] * g, ao = 0.0;
#if ORC_MUTANT == 1 /* planted transposed operand, test builds only */
            for (int i = 0; i < n; i += 2) ae = fma(u[i], p[(size_t)n * (j + 1) + i], ae);
            for (int i = 1; i < n; i += 2) ao = fma(u[i], p[(size_t)n * (j + 1) + i], ao);
#else
            for (int i = 0; i < n; i += 2) ae = fma(u[i], p[(size_t)n * (i + 1) + j], ae);
            for (int i = 1; i < n; i += 2) ao = fma(u[i], p[(size_t)n * (i + 1) + j], ao);
#endif
            wJx[j] = ae + ao;
        }
        for (int i = 0; i < n; ++i) {
            wJp[i] = u[i];
            double *Ab = wJp + (size_t)n * (i + 1);
            for (int j = 0; j < n; ++j) Ab[j] = u[i] * x[j];
        }
        break;
    }
    default:
        for (int i = 0; i < n; ++i) wJx[i] = NAN;
    }
}

void orc_vjp(int system, int n, int np, const double *x, const double *p, double t, double K,
             const double *w, double *wJx, double *wJp)
{
    vjp_v(system, n, np, x, p, t, K, w, wJx, wJp, ORC_FUSED);
}

/* ------------------------------------------------------------------------ */
/* odeint helpers (Boost 1.74 util/detail/less_with_sign.hpp, restated)     */
/* ------------------------------------------------------------------------ */
static int less_with_sign(double t1, double t2, double dt)
{
    if (dt > 0.0) return t2 - t1 > DBL_EPSILON;
    return t1 - t2 > DBL_EPSILON;
}

static int less_eq_with_sign(double t1, double t2, double dt)
{
    if (dt > 0.0) return t1 - t2 <= DBL_EPSILON;
    return t2 - t1 <= DBL_EPSILON;
}

/* Stage sums (Boost.odeint 1.74 generic_rk_algorithm with scale_sumN, and the
 * hand-written runge_kutta_dopri5::do_step_impl; third party, restated):
 *   X_m  = y + (a_m0 h) K_0 + (a_m1 h) K_1 + ...    left to right, products rounded
 *   y'   = y + (b_0 h) K_0 + ...
 *   xerr = (db_0 h) K_0 + (db_1 h) K_1 + ...        from the first product
 * Terms with a zero coefficient add a zero and are skipped (same value).
 * The HIP kernels use the same sequence.  ORC_FUSED: one fma per term and
 * xerr started from 0.                                                        */
static void stage_input_v(const orc_tableau *tb, int m, int n, double h, const double *y,
                          double K[][ORC_MAX_N], double *X, int fused)
{
    for (int i = 0; i < n; ++i) X[i] = y[i];
    for (int j = 0; j < m; ++j) {
#if ORC_MUTANT == 4 /* planted off-by-one stage index, test builds only */
        double a = tb->a[m][j];
        int jj = j > 0 ? j - 1 : j;
        if (a == 0.0) continue;
        double ah = a * h;
        for (int i = 0; i < n; ++i) X[i] = fused ? fma(ah, K[jj][i], X[i]) : X[i] + ah * K[jj][i];
#else
        double a = tb->a[m][j];
        if (a == 0.0) continue;
        double ah = a * h;
        for (int i = 0; i < n; ++i) X[i] = fused ? fma(ah, K[j][i], X[i]) : X[i] + ah * K[j][i];
#endif
    }
}

static void stage_input(const orc_tableau *tb, int m, int n, double h, const double *y,
                        double K[][ORC_MAX_N], double *X)
{
    stage_input_v(tb, m, n, h, y, K, X, ORC_FUSED);
}

typedef struct {
    double K[ORC_MAX_STAGES][ORC_MAX_N];
    double X[ORC_MAX_N];
    double ynew[ORC_MAX_N];
    double xerr[ORC_MAX_N];
} orc_work;

/* One attempt from (y, K0 = f(y)) with step h: fills w->K[1..s_fwd-1],
 * w->ynew and, when the tableau has an embedded pair, w->xerr. */
static int attempt(const orc_problem *pb, const orc_tableau *tb, const double *y,
                   const double *p, double t, double h, orc_work *w)
{
    int n = pb->n;
    int evals = 0;
    for (int m = 1; m < tb->s_fwd; ++m) {
        stage_input(tb, m, n, h, y, w->K, w->X);
        double tm = t + tb->c[m] * h;
        orc_rhs(pb->system, n, w->X, p, tm, pb->sys_const, w->K[m]);
        ++evals;
    }
    for (int i = 0; i < n; ++i) w->ynew[i] = y[i];
    for (int j = 0; j < tb->s; ++j) {
        double bb = tb->b[j];
        if (bb == 0.0) continue;
        double bh = bb * h;
        for (int i = 0; i < n; ++i) w->ynew[i] = ORC_MAD(bh, w->K[j][i], w->ynew[i]);
    }
    if (tb->has_err) {
        int first = 1;
        for (int i = 0; i < n; ++i) w->xerr[i] = 0.0;
        for (int j = 0; j < tb->s_fwd; ++j) {
            double d = tb->db[j];
            if (d == 0.0) continue;
            double dh = d * h;
#if ORC_FUSED
            for (int i = 0; i < n; ++i) w->xerr[i] = fma(dh, w->K[j][i], w->xerr[i]);
#else
            if (first)
                for (int i = 0; i < n; ++i) w->xerr[i] = dh * w->K[j][i];
            else
                for (int i = 0; i < n; ++i) w->xerr[i] = w->xerr[i] + dh * w->K[j][i];
#endif
            first = 0;
        }
        (void)first;
    }
    return evals;
}

void orc_rk_step(const orc_problem *pb, const double *y, const double *p, double t, double h,
                 double *out, double *xerr)
{
    orc_tableau tb;
    orc_work w;
    orc_get_tableau(pb->tableau, &tb);
    orc_rhs(pb->system, pb->n, y, p, t, pb->sys_const, w.K[0]);
    attempt(pb, &tb, y, p, t, h, &w);
    for (int i = 0; i < pb->n; ++i) out[i] = w.ynew[i];
    if (xerr)
        for (int i = 0; i < pb->n; ++i) xerr[i] = tb.has_err ? w.xerr[i] : 0.0;
}

/* default_error_checker::error (a_x = a_dxdt = 1) with rel_error
 * |xerr_i| / (atol + rtol (|y_i| + h |dydt_i|)), each operation rounded, and
 * norm_inf with std::max semantics (a NaN component never replaces the running
 * maximum). */
static double error_norm(int n, double atol, double rtol, double h, const double *y,
                         const double *dydt, const double *xerr)
{
    double m = 0.0;
    for (int i = 0; i < n; ++i) {
#if ORC_FUSED
        double den = fma(rtol, fma(h, fabs(dydt[i]), fabs(y[i])), atol);
#else
        double den = atol + rtol * (fabs(y[i]) + h * fabs(dydt[i]));
#endif
        double e = fabs(xerr[i]) / den;
        m = (m < e) ? e : m;
    }
    return m;
}

/* err^(-1/q): std::pow(err, -1.0/q) as Boost's step adjuster calls it (the C
 * library's pow, not always correctly rounded); ORC_FUSED: the Newton root. */
static double ctrl_pow(double err, int q)
{
#if ORC_FUSED
    return orc_root(err, q);
#else
    return pow(err, -1.0 / (double)q);
#endif
}

/* default_step_adjuster (max_dt = 0) */
static double decrease_step(double h, double err, int err_order)
{
#if ORC_MUTANT == 5 /* planted wrong exponent, test builds only */
    double f = 0.9 * ctrl_pow(err, err_order);
#else
    double f = 0.9 * ctrl_pow(err, err_order - 1); /* pow(err, -1/(err_order-1)) */
#endif
    f = (f < 0.2) ? 0.2 : f;
    return h * f;
}

static double min_err_for_order(int q)
{
    double d = 1.0;
    for (int i = 0; i < q; ++i) d = d * 5.0;
    return 1.0 / d; /* pow(5, -q), exact quotient correctly rounded */
}

static double increase_step(double h, double err, int order)
{
    if (err < 0.5) {
        double lo = min_err_for_order(order);
        err = (lo < err) ? err : lo;
        h = h * (0.9 * ctrl_pow(err, order)); /* pow(err, -1/order) */
    }
    return h;
}

int orc_traj_alloc(orc_trajectory *tr, int64_t cap, int n)
{
    tr->cap = cap;
    tr->T = 0;
    tr->t = (double *)malloc(sizeof(double) * (size_t)cap);
    tr->h = (double *)calloc((size_t)cap, sizeof(double));
    tr->y = (double *)malloc(sizeof(double) * (size_t)cap * (size_t)n);
    return (tr->t && tr->h && tr->y) ? 0 : -1;
}

void orc_traj_free(orc_trajectory *tr)
{
    free(tr->t);
    free(tr->h);
    free(tr->y);
    tr->t = tr->h = tr->y = NULL;
    tr->cap = tr->T = 0;
}

static int record(orc_trajectory *tr, int n, double t, const double *y)
{
    if (!tr) return 0;
    if (tr->T >= tr->cap) return -1;
    tr->t[tr->T] = t;
    tr->h[tr->T] = 0.0;
    memcpy(tr->y + (size_t)tr->T * n, y, sizeof(double) * (size_t)n);
    tr->T += 1;
    return 0;
}

/* Digest of an accepted step sequence: a 64-bit mix over the bit patterns of
 * (t_n, h_n) for every stored point n (h = 0 on the final point).  Equal step
 * histories give equal digests; the device computes the same function
 * (va_step_digest). */
static uint64_t digest_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t orc_step_digest(const orc_trajectory *tr)
{
    uint64_t d = 0x9E3779B97F4A7C15ull;
    for (int64_t k = 0; k < tr->T; ++k) {
        uint64_t bt, bh;
        memcpy(&bt, &tr->t[k], 8);
        memcpy(&bh, &tr->h[k], 8);
        d = digest_mix(d ^ bt);
        d = digest_mix(d ^ bh);
    }
    return d;
}

static int all_finite(int n, const double *y)
{
    for (int i = 0; i < n; ++i)
        if (!isfinite(y[i])) return 0;
    return 1;
}

static int64_t forward_impl(const orc_problem *pb, const orc_obs *ob, double *y, const double *p,
                            orc_trajectory *tr, orc_stats *st, double *log, int64_t log_cap,
                            int64_t *log_len, int32_t *obs_row, double *Jobs);

int64_t orc_forward(const orc_problem *pb, double *y, const double *p, orc_trajectory *tr,
                    orc_stats *st, double *log, int64_t log_cap, int64_t *log_len)
{
    return forward_impl(pb, NULL, y, p, tr, st, log, log_cap, log_len, NULL, NULL);
}

int64_t orc_forward_obs(const orc_problem *pb, const orc_obs *ob, double *y, const double *p,
                        orc_trajectory *tr, orc_stats *st, int32_t *obs_row, double *J)
{
    return forward_impl(pb, ob, y, p, tr, st, NULL, 0, NULL, obs_row, J);
}

/* Forward solve; with observations (ob != NULL, controlled loop only) the loop
 * runs segment by segment to each observation time and then to t1.  A step
 * that would pass the next observation time is shortened to land on it, as at
 * t1 (detail/runge_kutta.hpp:98-101).  After such a shortened step is accepted
 * the next step size is the larger of the controller's proposal and the step
 * size before shortening -- Boost.odeint 1.74 integrate_times for controlled
 * steppers does the same (third party, restated).  At each observation time
 * the running cost J += sum_i (y_i - y_obs_k,i)^2 and obs_row[k] = the index
 * of the stored point (the accepted steps so far). */
static int64_t forward_impl(const orc_problem *pb, const orc_obs *ob, double *y, const double *p,
                            orc_trajectory *tr, orc_stats *st, double *log, int64_t log_cap,
                            int64_t *log_len, int32_t *obs_row, double *Jobs)
{
    orc_tableau tb;
    orc_work w;
    int n = pb->n;
    orc_stats s0 = {0, 0, 0, ORC_OK};
    int64_t nlog = 0;
    if (tr) tr->T = 0;
    const int nobs = ob ? ob->k : 0;
    double Jacc = 0.0;
    if (Jobs) *Jobs = 0.0;
    if (orc_get_tableau(pb->tableau, &tb) != 0 || n <= 0 || n > ORC_MAX_N ||
        (pb->controlled && !tb.has_err) || (nobs > 0 && !pb->controlled)) {
        s0.status = ORC_BAD_INPUT;
        if (st) *st = s0;
        return 0;
    }
    double t = pb->t0, t1 = pb->t1, dt = pb->dt0;
    int64_t count = 0;

    if (!pb->controlled) {
        /* detail/runge_kutta.hpp:36-72 (stepper_tag) */
        int64_t step = 0;
        while (less_eq_with_sign(t + dt, t1, dt)) {
            if (record(tr, n, t, y) != 0 || (pb->max_steps > 0 && step >= pb->max_steps)) {
                s0.status = ORC_MAX_STEPS;
                goto done;
            }
            orc_rhs(pb->system, n, y, p, t, pb->sys_const, w.K[0]);
            s0.rhs_evals += 1 + attempt(pb, &tb, y, p, t, dt, &w);
            memcpy(y, w.ynew, sizeof(double) * (size_t)n);
            ++step;
            s0.accepted = (int32_t)step;
            const double tn = pb->t0 + (double)step * dt;
            if (tr) tr->h[tr->T - 1] = STORED_STEP(tn, t, dt);
            t = tn;
        }
        if (record(tr, n, t, y) != 0) s0.status = ORC_MAX_STEPS;
        count = step;
        goto done;
    }

    /* detail/runge_kutta.hpp:74-118 (controlled_stepper_tag) */
    int have_k0 = 0;
    for (int kobs = 0;; ++kobs) {
    const int is_obs = kobs < nobs;
    const double target = is_obs ? ob->t[kobs] : t1;
    while (less_with_sign(t, target, dt)) {
        if (record(tr, n, t, y) != 0 || (pb->max_steps > 0 && count >= pb->max_steps)) {
            s0.status = ORC_MAX_STEPS;
            goto done;
        }
        int clipped = 0;
        const double dt_keep = dt;
        if (less_with_sign(target, t + dt, dt)) {
            dt = target - t;
            clipped = is_obs;
        }
        int tries = 0;
        int ok = 0;
        /* odeint evaluates f(y) on every try of a non-FSAL stepper; the value
         * does not change between tries, so it is evaluated once per step.  A
         * FSAL stepper (DOPRI5) evaluates it on its first call only and then
         * carries f(y_{n+1}) over from the last stage. */
        if (!tb.fsal || !have_k0) {
            orc_rhs(pb->system, n, y, p, t, pb->sys_const, w.K[0]);
            s0.rhs_evals += 1;
            have_k0 = 1;
        }
        do {
            s0.rhs_evals += attempt(pb, &tb, y, p, t, dt, &w);
            double err = error_norm(n, pb->abs_tol, pb->rel_tol, dt, y, w.K[0], w.xerr);
            double t_at = t, h_at = dt;
            if (err > 1.0) {
                dt = decrease_step(dt, err, tb.err_order);
                s0.rejected += 1;
                clipped = 0;
                ok = 0;
            } else {
                const double tn = t + dt;
                if (tr) tr->h[tr->T - 1] = STORED_STEP(tn, t, dt);
                t = tn;
                dt = increase_step(dt, err, tb.order);
                if (clipped && fabs(dt_keep) > fabs(dt)) dt = dt_keep;
                memcpy(y, w.ynew, sizeof(double) * (size_t)n);
                if (tb.fsal) memcpy(w.K[0], w.K[tb.s_fwd - 1], sizeof(double) * (size_t)n);
                ok = 1;
            }
            if (log && nlog < log_cap) {
                log[4 * nlog + 0] = t_at;
                log[4 * nlog + 1] = h_at;
                log[4 * nlog + 2] = err;
                log[4 * nlog + 3] = (double)ok;
                ++nlog;
            }
            ++tries; /* failed_step_checker: throws on the try after 500 */
            if (tries > pb->max_attempts) {
                if (ok) s0.accepted = (int32_t)(count + 1);
                s0.status = ORC_MAX_FAIL;
                goto done;
            }
        } while (!ok);
        ++count;
        s0.accepted = (int32_t)count;
    }
    if (!is_obs) break;
    if (obs_row) obs_row[kobs] = (int32_t)count;
    for (int i = 0; i < n; ++i) {
        const double d = y[i] - ob->y[(size_t)kobs * n + i];
        Jacc = fma(d, d, Jacc);
    }
    }
    if (record(tr, n, t, y) != 0) s0.status = ORC_MAX_STEPS;
    if (Jobs) *Jobs = Jacc;

done:
    if (s0.status == ORC_OK && !all_finite(n, y)) s0.status = ORC_NONFINITE;
    if (st) *st = s0;
    if (log_len) *log_len = nlog;
    return count;
}

void orc_replay(const orc_problem *pb, const double *y0, const double *p,
                const orc_trajectory *tr, double *yT)
{
    orc_tableau tb;
    orc_work w;
    int n = pb->n;
    double y[ORC_MAX_N];
    orc_get_tableau(pb->tableau, &tb);
    memcpy(y, y0, sizeof(double) * (size_t)n);
    for (int64_t k = 0; k + 1 < tr->T; ++k) {
        double h = tr->h[k];
        orc_rhs(pb->system, n, y, p, tr->t[k], pb->sys_const, w.K[0]);
        attempt(pb, &tb, y, p, tr->t[k], h, &w);
        memcpy(y, w.ynew, sizeof(double) * (size_t)n);
    }
    memcpy(yT, y, sizeof(double) * (size_t)n);
}

/* ------------------------------------------------------------------------ */
/* Discrete adjoint, detail/backpropagation.hpp:83-158 (one seed).          */
/* ------------------------------------------------------------------------ */
typedef struct {
    double K[ORC_MAX_STAGES][ORC_MAX_N];
    double X[ORC_MAX_STAGES][ORC_MAX_N];
    double Kb[ORC_MAX_STAGES][ORC_MAX_N];
    double ybar[ORC_MAX_N];
    double xi[ORC_MAX_N];
} orc_rwork;

/* pbar += the stage's parameter VJP, term by term (alphabar[k] += wJalpha[k],
 * detail/backpropagation.hpp:220-223: each VJP entry is a rounded value).
 * ORC_FUSED: for Lotka-Volterra with per-instance A the entries A_ij get
 * u_i x_j with one fused multiply-add each (wjp holds u_i in its first n
 * entries). */
static void accumulate_param_bar(int system, int n, int np, const double *x, const double *wjp,
                                 double *pbar)
{
    if (ORC_FUSED_REVERSE && system == ORC_SYS_LOTKA_VOLTERRA) {
        for (int i = 0; i < n; ++i) {
            pbar[i] += wjp[i];
            double *Ab = pbar + (size_t)n * (i + 1);
            for (int j = 0; j < n; ++j) Ab[j] = fma(wjp[i], x[j], Ab[j]);
        }
        return;
    }
    for (int k = 0; k < np; ++k) pbar[k] += wjp[k];
}

static void back_prop_step(const orc_problem *pb, const orc_tableau *tb, const double *p,
                           const orc_trajectory *tr, int64_t nidx, double *lambda,
                           double *pbar, double *wjp, orc_rwork *w)
{
    int n = pb->n, np = pb->np, s = tb->s;
    double tn = tr->t[nidx];
    /* The reference takes h = t[n+1] - t[n] (Driver::GetDt, StateStorage.hpp:22),
     * which can differ from the accepted step by an ulp of t.  ORC_FUSED takes
     * the stored accepted step instead. */
#if ORC_FUSED
    double h = tr->h[nidx];
#else
    double h = tr->t[nidx + 1] - tr->t[nidx];
#endif
    const double *yn = tr->y + (size_t)nidx * n;

    /* compute_intermediate_states (:24-64): recompute K_m from y_n */
    for (int m = 0; m < s; ++m) {
        stage_input_v(tb, m, n, h, yn, w->K, w->X[m], ORC_FUSED_REVERSE);
        rhs_v(pb->system, n, w->X[m], p, tn + tb->c[m] * h, pb->sys_const, w->K[m],
              ORC_FUSED_REVERSE);
    }
    /* seed (:104-112): ybar = lambda_{n+1}; Kbar_m = (b_m h) lambda_{n+1} */
    for (int i = 0; i < n; ++i) w->ybar[i] = lambda[i];
    for (int m = 0; m < s; ++m) {
        double bh = tb->b[m] * h;
        for (int i = 0; i < n; ++i) w->Kb[m][i] = bh * lambda[i];
    }
    /* stages s..1 (:126-153): ybar += xi; Kbar_j += (xi a_mj) h, in that
     * order of products (:210-216) */
    for (int m = s - 1; m >= 0; --m) {
        vjp_v(pb->system, n, np, w->X[m], p, tn + tb->c[m] * h, pb->sys_const, w->Kb[m], w->xi,
              wjp, ORC_FUSED_REVERSE);
        for (int i = 0; i < n; ++i) w->ybar[i] += w->xi[i];
        for (int j = 0; j < m; ++j) {
            double a = tb->a[m][j];
            if (a == 0.0) continue;
#if ORC_MUTANT == 2 /* planted dropped term, test builds only */
            (void)a;
#elif ORC_FUSED_REVERSE
            double ah = a * h;
            for (int i = 0; i < n; ++i) w->Kb[j][i] = fma(ah, w->xi[i], w->Kb[j][i]);
#else
            for (int i = 0; i < n; ++i) w->Kb[j][i] = w->Kb[j][i] + (w->xi[i] * a) * h;
#endif
        }
        accumulate_param_bar(pb->system, n, np, w->X[m], wjp, pbar);
    }
    for (int i = 0; i < n; ++i) lambda[i] = w->ybar[i];
}

void orc_backward(const orc_problem *pb, const double *p, const orc_trajectory *tr,
                  double *lambda, double *mu)
{
    orc_tableau tb;
    orc_get_tableau(pb->tableau, &tb);
    int np = pb->np;
    orc_rwork *w = (orc_rwork *)malloc(sizeof(orc_rwork));
    double *pbar = (double *)calloc((size_t)np, sizeof(double));
    double *wjp = (double *)calloc((size_t)np, sizeof(double));
    /* back_prop (:231-254): n = T-2 ... 0 */
    for (int64_t k = tr->T - 2; k >= 0; --k) back_prop_step(pb, &tb, p, tr, k, lambda, pbar, wjp, w);
    for (int k = 0; k < np; ++k) mu[k] += pbar[k];
    free(wjp);
    free(pbar);
    free(w);
}

/* lambda += dJ_k/dy = 2 (y_n - y_obs_k) for every observation taken at point n. */
static int obs_jumps(const orc_obs *ob, int n, const int32_t *obs_row, int kobs, int64_t point,
                     const double *yn, double *lambda)
{
    while (kobs >= 0 && obs_row[kobs] == point) {
        for (int i = 0; i < n; ++i)
#if ORC_MUTANT == 6 /* planted sign error, test builds only */
            lambda[i] = fma(-2.0, yn[i] - ob->y[(size_t)kobs * n + i], lambda[i]);
#else
            lambda[i] = fma(2.0, yn[i] - ob->y[(size_t)kobs * n + i], lambda[i]);
#endif
        --kobs;
    }
    return kobs;
}

void orc_backward_obs(const orc_problem *pb, const orc_obs *ob, const double *p,
                      const orc_trajectory *tr, const int32_t *obs_row, double *lambda,
                      double *mu)
{
    orc_tableau tb;
    orc_get_tableau(pb->tableau, &tb);
    int n = pb->n, np = pb->np;
    orc_rwork *w = (orc_rwork *)malloc(sizeof(orc_rwork));
    double *pbar = (double *)calloc((size_t)np, sizeof(double));
    double *wjp = (double *)calloc((size_t)np, sizeof(double));
    int kobs = ob->k - 1;
    kobs = obs_jumps(ob, n, obs_row, kobs, tr->T - 1, tr->y + (size_t)(tr->T - 1) * n, lambda);
    for (int64_t k = tr->T - 2; k >= 0; --k) {
        back_prop_step(pb, &tb, p, tr, k, lambda, pbar, wjp, w);
        kobs = obs_jumps(ob, n, obs_row, kobs, k, tr->y + (size_t)k * n, lambda);
    }
    for (int k = 0; k < np; ++k) mu[k] += pbar[k];
    free(wjp);
    free(pbar);
    free(w);
}

void orc_adjoint_solve(const orc_problem *pb, const double *p, const orc_trajectory *tr,
                       int nout, double *lambda, double *mu)
{
    for (int o = 0; o < nout; ++o)
        orc_backward(pb, p, tr, lambda + (size_t)o * pb->n, mu + (size_t)o * pb->np);
}

void orc_sensitivity_matrix(const orc_problem *pb, const double *p, const orc_trajectory *tr,
                            double *S)
{
    int n = pb->n, np = pb->np;
    double lam[ORC_MAX_N];
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) lam[j] = (i == j) ? 1.0 : 0.0;
        double *row = S + (size_t)i * np;
        for (int k = 0; k < np; ++k) row[k] = 0.0;
        orc_backward(pb, p, tr, lam, row);
    }
}

/* ------------------------------------------------------------------------ */
/* Batched driver (SoA in/out), pthreads over instances.                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    const orc_problem *pb;
    int64_t B, b_begin, b_end;
    const double *y0, *p, *p_shared;
    int cost;
    const double *cost_w;
    double *J, *yT, *dJdy0, *dJdp;
    int32_t *n_acc, *n_rej, *status;
    uint64_t *digest;
    int64_t next; /* shared work counter */
    pthread_mutex_t mtx;
    double *shared_acc;   /* sum of dJ/dA over instances */
    int np_full;
    const orc_obs *ob;    /* observation cost instead of the terminal cost */
} batch_ctx;

static void *batch_worker(void *arg)
{
    batch_ctx *c = (batch_ctx *)arg;
    const orc_problem *pb = c->pb;
    int n = pb->n;
    int shared = pb->system == ORC_SYS_LOTKA_VOLTERRA_SHARED;
    int npf = c->np_full;
    orc_problem lp = *pb;
    lp.np = npf;
    orc_trajectory tr;
    int64_t cap = (pb->max_steps > 0 ? pb->max_steps : 100000) + 1;
    if (orc_traj_alloc(&tr, cap, n) != 0) return NULL;
    double *pp = (double *)malloc(sizeof(double) * (size_t)npf);
    double *mu = (double *)malloc(sizeof(double) * (size_t)npf);
    double *acc = shared ? (double *)calloc((size_t)n * n, sizeof(double)) : NULL;
    double y[ORC_MAX_N], lam[ORC_MAX_N];
    int32_t *obs_row = c->ob ? (int32_t *)malloc(sizeof(int32_t) * (size_t)(c->ob->k + 1)) : NULL;
    const int64_t chunk = 8;
    for (;;) {
        int64_t b0 = __atomic_fetch_add(&c->next, chunk, __ATOMIC_RELAXED);
        if (b0 >= c->b_end) break;
        int64_t b1 = b0 + chunk < c->b_end ? b0 + chunk : c->b_end;
        for (int64_t b = b0; b < b1; ++b) {
            int64_t B = c->B;
            for (int i = 0; i < n; ++i) y[i] = c->y0[(size_t)i * B + b];
            if (shared) {
                for (int i = 0; i < n; ++i) pp[i] = c->p[(size_t)i * B + b];
                memcpy(pp + n, c->p_shared, sizeof(double) * (size_t)n * n);
            } else {
                for (int k = 0; k < npf; ++k) pp[k] = c->p[(size_t)k * B + b];
            }
            orc_stats st;
            double Jo = 0.0;
            if (c->ob)
                orc_forward_obs(&lp, c->ob, y, pp, &tr, &st, obs_row, &Jo);
            else
                orc_forward(&lp, y, pp, &tr, &st, NULL, 0, NULL);
            if (c->n_acc) c->n_acc[b] = st.accepted;
            if (c->n_rej) c->n_rej[b] = st.rejected;
            if (c->status) c->status[b] = st.status;
            if (c->digest) c->digest[b] = orc_step_digest(&tr);
            int okf = st.status == ORC_OK;
            double Jv = 0.0;
            for (int i = 0; i < n && !c->ob; ++i) {
                switch (c->cost) {
                case ORC_COST_SUM: Jv += y[i]; lam[i] = 1.0; break;
                case ORC_COST_SQUARED_NORM: Jv = fma(y[i], y[i], Jv); lam[i] = 2.0 * y[i]; break;
                default: Jv = fma(c->cost_w[i], y[i], Jv); lam[i] = c->cost_w[i]; break;
                }
            }
            if (c->ob) {
                Jv = Jo;
                for (int i = 0; i < n; ++i) lam[i] = 0.0;
            }
            if (c->J) c->J[b] = okf ? Jv : NAN;
            if (c->yT)
                for (int i = 0; i < n; ++i) c->yT[(size_t)i * B + b] = y[i];
            for (int k = 0; k < npf; ++k) mu[k] = 0.0;
            if (okf && c->ob)
                orc_backward_obs(&lp, c->ob, pp, &tr, obs_row, lam, mu);
            else if (okf)
                orc_backward(&lp, pp, &tr, lam, mu);
            if (c->dJdy0)
                for (int i = 0; i < n; ++i) c->dJdy0[(size_t)i * B + b] = okf ? lam[i] : NAN;
            if (c->dJdp) {
                int npi = shared ? n : npf;
                for (int k = 0; k < npi; ++k) c->dJdp[(size_t)k * B + b] = okf ? mu[k] : NAN;
            }
            if (shared && okf)
                for (int k = 0; k < n * n; ++k) acc[k] += mu[n + k];
        }
    }
    if (shared && c->shared_acc) {
        pthread_mutex_lock(&c->mtx);
        for (int k = 0; k < n * n; ++k) c->shared_acc[k] += acc[k];
        pthread_mutex_unlock(&c->mtx);
    }
    free(obs_row);
    free(acc);
    free(mu);
    free(pp);
    orc_traj_free(&tr);
    return NULL;
}

static int solve_batch_impl(const orc_problem *pb, const orc_obs *ob, int64_t B, int64_t b_begin,
                            int64_t b_count, const double *y0, const double *p,
                            const double *p_shared, int cost, const double *cost_w, double *J,
                            double *yT, double *dJdy0, double *dJdp, double *dJdp_shared,
                            int32_t *n_acc, int32_t *n_rej, int32_t *status, uint64_t *digest,
                            int nthreads);

int orc_solve_batch(const orc_problem *pb, int64_t B, int64_t b_begin, int64_t b_count,
                    const double *y0, const double *p, const double *p_shared,
                    int cost, const double *cost_w,
                    double *J, double *yT, double *dJdy0, double *dJdp, double *dJdp_shared,
                    int32_t *n_acc, int32_t *n_rej, int32_t *status, int nthreads)
{
    return solve_batch_impl(pb, NULL, B, b_begin, b_count, y0, p, p_shared, cost, cost_w, J, yT,
                            dJdy0, dJdp, dJdp_shared, n_acc, n_rej, status, NULL, nthreads);
}

int orc_solve_batch_digest(const orc_problem *pb, int64_t B, int64_t b_begin, int64_t b_count,
                           const double *y0, const double *p, const double *p_shared, int cost,
                           const double *cost_w, double *J, double *yT, double *dJdy0,
                           double *dJdp, double *dJdp_shared, int32_t *n_acc, int32_t *n_rej,
                           int32_t *status, uint64_t *digest, int nthreads)
{
    return solve_batch_impl(pb, NULL, B, b_begin, b_count, y0, p, p_shared, cost, cost_w, J, yT,
                            dJdy0, dJdp, dJdp_shared, n_acc, n_rej, status, digest, nthreads);
}

int orc_solve_batch_obs(const orc_problem *pb, const orc_obs *ob, int64_t B, int64_t b_begin,
                        int64_t b_count, const double *y0, const double *p, double *J,
                        double *yT, double *dJdy0, double *dJdp, int32_t *n_acc,
                        int32_t *n_rej, int32_t *status, int nthreads)
{
    if (!ob || ob->k < 1 || !pb || !pb->controlled ||
        pb->system == ORC_SYS_LOTKA_VOLTERRA_SHARED)
        return -1;
    return solve_batch_impl(pb, ob, B, b_begin, b_count, y0, p, NULL, 0, NULL, J, yT, dJdy0, dJdp,
                            NULL, n_acc, n_rej, status, NULL, nthreads);
}

static int solve_batch_impl(const orc_problem *pb, const orc_obs *ob, int64_t B, int64_t b_begin,
                            int64_t b_count, const double *y0, const double *p,
                            const double *p_shared, int cost, const double *cost_w, double *J,
                            double *yT, double *dJdy0, double *dJdp, double *dJdp_shared,
                            int32_t *n_acc, int32_t *n_rej, int32_t *status, uint64_t *digest,
                            int nthreads)
{
    if (!pb || pb->n <= 0 || pb->n > ORC_MAX_N || b_begin < 0 || b_begin + b_count > B) return -1;
    if (!ob && cost == ORC_COST_LINEAR && !cost_w) return -1;
    int shared = pb->system == ORC_SYS_LOTKA_VOLTERRA_SHARED;
    if (shared && !p_shared) return -1;
    batch_ctx c;
    memset(&c, 0, sizeof(c));
    c.pb = pb;
    c.B = B;
    c.b_begin = b_begin;
    c.b_end = b_begin + b_count;
    c.y0 = y0;
    c.p = p;
    c.p_shared = p_shared;
    c.cost = cost;
    c.cost_w = cost_w;
    c.J = J;
    c.yT = yT;
    c.dJdy0 = dJdy0;
    c.dJdp = dJdp;
    c.n_acc = n_acc;
    c.n_rej = n_rej;
    c.status = status;
    c.digest = digest;
    c.next = b_begin;
    c.np_full = shared ? pb->n + pb->n * pb->n : pb->np;
    c.ob = ob;
    pthread_mutex_init(&c.mtx, NULL);
    double *acc = NULL;
    if (shared) {
        acc = (double *)calloc((size_t)pb->n * pb->n, sizeof(double));
        c.shared_acc = acc;
    }
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    int started = 0;
    for (int i = 1; i < nthreads; ++i)
        if (pthread_create(&th[i], NULL, batch_worker, &c) == 0) ++started;
        else break;
    batch_worker(&c);
    for (int i = 1; i <= started; ++i) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&c.mtx);
    if (shared) {
        if (dJdp_shared)
            for (int k = 0; k < pb->n * pb->n; ++k) dJdp_shared[k] = acc[k];
        free(acc);
    }
    return 0;
}
