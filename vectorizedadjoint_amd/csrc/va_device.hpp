// va_device.hpp -- device building blocks of the batched RK forward/adjoint path
// for gfx950: Butcher tableaus, the step controller, the built-in systems
// (right-hand side and vector-Jacobian products) and the explicit RK attempt.
//
// Rounding sequence: all HIP sources are compiled with -ffp-contract=off, so
// the IEEE operation sequence is exactly the one written here.  The forward
// step follows the reference's own arithmetic operation for operation:
// Boost.odeint 1.74's stage sums and error norm with every product rounded
// (the reference is built without FMA contraction), the example functors'
// loop orders, and the C library's pow in the step adjuster (va_pow.hpp).  So
// step sizes, accept/reject decisions and y(T) equal the reference's bit for
// bit.  The CPU oracle uses the same sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "va_pow.hpp"

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

// The controller pow's tables in LDS (5 KB per block): every kernel that runs the
// controller copies them in with load_ctrl_tables() at its start (a block-wide
// barrier follows), so the two table reads on each attempt's dependent chain are
// LDS reads instead of cache reads.
struct CtrlTablesLds {
    double log[128][3];
    uint64_t exp[256];
};

__device__ __forceinline__ CtrlTablesLds &ctrl_lds()
{
    __shared__ CtrlTablesLds tables;
    return tables;
}

// Call at the top of the kernel, by every thread of the block.
__device__ __forceinline__ void load_ctrl_tables()
{
    CtrlTablesLds &t = ctrl_lds();
    for (int k = threadIdx.x; k < 128 * 3; k += blockDim.x)
        (&t.log[0][0])[k] = (&powtab::LOG[0][0])[k];
    for (int k = threadIdx.x; k < 256; k += blockDim.x) t.exp[k] = powtab::EXP[k];
    __syncthreads();
}

struct LdsTables {
    __device__ static void log_entry(int i, double &invc, double &logc, double &logctail)
    {
        const CtrlTablesLds &t = ctrl_lds();
        invc = t.log[i][0];
        logc = t.log[i][1];
        logctail = t.log[i][2];
    }
    __device__ static void exp_entry(uint64_t idx, uint64_t &tail_bits, uint64_t &scale_bits)
    {
        const CtrlTablesLds &t = ctrl_lds();
        tail_bits = t.exp[idx];
        scale_bits = t.exp[idx + 1];
    }
};

// Accept/reject decision and the next step size (default_step_adjuster with
// max_dt = 0):
//   reject (err > 1):   h * max(0.9 pow(err, -1/(ERR_ORDER-1)), 0.2)
//   accept, err < 0.5:  h * (0.9 pow(max(5^-ORDER, err), -1/ORDER))
//   accept otherwise:   h
// One pow evaluation whose argument and exponent are selected per lane first,
// so lanes of a wave that take different decisions share one instruction
// stream (tables read through PowTables: LDS copies by default, which the
// kernel must have loaded).  The exponents are the rounded quotients -1.0/q, as Boost forms them;
// pow(5.0, -q) is 1/5^q exactly rounded, which the C library returns for the
// orders here (q <= 8).
template <class Tab, class PowTables = LdsTables>
__device__ __forceinline__ bool control_step(double h, double err, double &h_next)
{
    constexpr double lo = FivePow<Tab::ORDER>::inv();
    constexpr double y_inc = -1.0 / (double)Tab::ORDER;
    constexpr double y_dec = -1.0 / (double)(Tab::ERR_ORDER - 1);
    const bool accept = !(err > 1.0);
    const double base = accept ? ((lo < err) ? err : lo) : err;
    const double f = 0.9 * powref::pow_ctrl<PowTables>(base, accept ? y_inc : y_dec);
    if (!accept) {
        h_next = h * ((f < 0.2) ? 0.2 : f);
    } else {
        h_next = (err < 0.5) ? h * f : h;
    }
    return accept;
}

// Step size stored for the reverse sweep: t_{n+1} - t_n, which the reference's
// reverse takes as its h (Driver::GetDt, lib/include/StateStorage.hpp:22); it can
// differ from the accepted step by an ulp of t.
VA_HD double stored_step(double t_next, double t) { return t_next - t; }

// ---------------------------------------------------------------------------
// Systems.  rhs(p, x, f); vjp(p, x, w, wJx, pbar) accumulates
// pbar += w^T df/dp term by term, each term a rounded value (alphabar[k] +=
// wJalpha[k], detail/backpropagation.hpp:220-223).  Right-hand sides follow the
// example functors operation for operation: examples/VanDerPol/main.cpp:41-42,
// examples/HarmonicOscillator/main.cpp:27-28, examples/GeneralizedLotkaVolterra/
// main.cpp:111-118 (sum = 0; sum += A_ij x_j for j = 0..N-1; x_i (r_i + sum)).
// The VJPs are hand-written in place of the reference's AADC tape.
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
    // g_i = r_i + sum_j A_ij x_j in the functor's order
    VA_HD static double growth(const Params &q, const double *x, int i)
    {
        double sum = 0.0;
#pragma unroll
        for (int j = 0; j < NN; ++j) sum = sum + q.A[i][j] * x[j];
        return q.r[i] + sum;
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
                pbar[NN * (i + 1) + j] = pbar[NN * (i + 1) + j] + u[i] * x[j];
        }
    }
};

// ---------------------------------------------------------------------------
// Explicit RK attempt from (y, K[0] = f(y)), Boost.odeint 1.74's sums
// (generic_rk_algorithm with scale_sumN, runge_kutta_dopri5::do_step_impl):
//   X_m  = y + (a_m0 h) K_0 + (a_m1 h) K_1 + ...   left to right, products rounded
//   y'   = y + (b_0 h) K_0 + ...
//   xerr = (db_0 h) K_0 + (db_1 h) K_1 + ...       from the first product
// Terms with a zero coefficient add a zero and are skipped (same value).
// ---------------------------------------------------------------------------
template <class Tab>
struct ErrFirst {  // index of the first nonzero embedded-error weight
    static constexpr int value()
    {
        for (int j = 0; j < Tab::SF; ++j)
            if (Tab::DB[j] != 0.0) return j;
        return Tab::SF;
    }
};

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
                for (int i = 0; i < N; ++i) X[i] = X[i] + ah * K[j][i];
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
                for (int i = 0; i < N; ++i) ynew[i] = ynew[i] + bh * K[j][i];
            }
        }
        if constexpr (Tab::HAS_ERR) {
            constexpr int j0 = ErrFirst<Tab>::value();
            const double d0 = Tab::DB[j0] * h;
#pragma unroll
            for (int i = 0; i < N; ++i) xerr[i] = d0 * K[j0][i];
#pragma unroll
            for (int j = j0 + 1; j < Tab::SF; ++j) {
                if (Tab::DB[j] != 0.0) {
                    const double dh = Tab::DB[j] * h;
#pragma unroll
                    for (int i = 0; i < N; ++i) xerr[i] = xerr[i] + dh * K[j][i];
                }
            }
        }
    }

    // default_error_checker::error: rel_error |xerr_i| / (atol + rtol (|y_i| +
    // h |dydt_i|)), each operation rounded; norm_inf with std::max semantics
    __device__ __forceinline__ static double error_norm(double atol, double rtol, double h,
                                                        const double *y, const double *dydt,
                                                        const double *xerr)
    {
        double m = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double den = atol + rtol * (fabs(y[i]) + h * fabs(dydt[i]));
            const double e = fabs(xerr[i]) / den;
            m = (m < e) ? e : m;
        }
        return m;
    }
};

}  // namespace va
