// va_glv_rows.hpp -- generalized Lotka-Volterra forward and reverse sweeps with one
// instance per group of N lanes ("row per lane"), gfx950.
//
// Layout.  A wave64 holds G = 64/N instance groups (N = 10: 6 groups, 60 busy
// lanes).  Lane i of a group owns state component i, row i of A (for the
// growth rates g_i = r_i + sum_j A_ij x_j), column i of A (for the adjoint
// term sum_k u_k A_ki) and, in the reverse sweep, column i of dJ/dA (sum of
// u_k X_i: the u-vector is gathered for the adjoint term anyway, X_i is the
// lane's own).  Stage vectors are exchanged
// through a per-wave LDS slab: every lane writes its component, the group
// reads the whole vector back as broadcast 16-byte loads.  Nothing crosses
// waves, so there is no block barrier; LDS operations of one wave complete in
// program order.
//
// Compared with one lane per instance this keeps A and dJ/dA in registers
// (3N doubles per lane instead of 2N^2) and cuts the instruction count of one
// instance-step by ~N, which is what bounds the long tail of instances that
// need ~10x the median number of steps.
//
// Every value is formed by the same operations in the same order as in the
// one-lane kernels and the CPU oracle (forward: the reference's sums, growth
// j = 0..N-1 from 0 then r_i + sum; adjoint: k = 0..N-1 from w_i g_i in two
// interleaved fma chains; accumulations in stage order), so results are
// bit-identical.
//
// Reference path: lib/include/detail/runge_kutta.hpp:74-118 (forward loop),
// lib/include/detail/backpropagation.hpp:83-158,231-254 (reverse step/loop),
// examples/GeneralizedLotkaVolterra/main.cpp:105-119 (system, parameter layout).
#pragma once

#include <type_traits>
#include <utility>

#include "va_kernels_impl.hpp"

namespace va {

template <int NN>
struct GlvRows {
    static constexpr int G = NN;          // lanes per instance
    static constexpr int GPW = 64 / NN;   // instance groups per wave
    static constexpr int F = NN + 2;      // trajectory row (t, h, y[N])
    static constexpr int WAVES = 4;       // block of 256 threads
    static_assert(NN >= 2 && NN <= 32, "row-per-lane layout needs 2 <= N <= 32");
};

// Broadcast-gather of a group's vector from the wave's LDS slab.
template <int NN>
__device__ __forceinline__ void gather_vec(const double *slab, int gbase, double (&out)[NN])
{
    if constexpr (NN % 2 == 0) {
        const double2 *s2 = reinterpret_cast<const double2 *>(slab + gbase);
#pragma unroll
        for (int j = 0; j < NN / 2; ++j) {
            const double2 v = s2[j];
            out[2 * j] = v.x;
            out[2 * j + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < NN; ++j) out[j] = slab[gbase + j];
    }
}

// v_max_f64 without the quieting of its operands that fmax() adds for values
// the compiler cannot prove quiet (LDS loads and DPP moves): the error ratios
// are results of arithmetic, so a NaN among them is already quiet, and
// v_max_f64 returns the other operand for a quiet NaN.
__device__ __forceinline__ double max_quiet(double a, double b)
{
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Maximum of a group's error ratios.  Every value is >= +0 or a quiet NaN; the
// maximum is exact, ignores NaN operands like the sequential (m < e) ? e : m
// fold, and is order-independent on such values, so a balanced tree gives the
// same result with a dependency depth of log2(N); the fold starts at 0, which
// an all-NaN group returns.
template <int NN>
__device__ __forceinline__ double group_max(const double (&v)[NN])
{
    double t[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) t[j] = v[j];
#pragma unroll
    for (int w = 1; w < NN; w *= 2)
#pragma unroll
        for (int j = 0; j + w < NN; j += 2 * w) t[j] = max_quiet(t[j], t[j + w]);
    return max_quiet(t[0], 0.0);
}

// Long-running instances set the job's critical path: a wave holding one gets
// SIMD issue priority so it runs near its lone-wave latency while other waves
// fill the remaining issue slots.  (The one-instance-per-wave long kernel runs
// at the highest priority throughout; persistent kernels running beside it,
// i.e. with a queue split, leave their priority at 0.)
__device__ __forceinline__ void prioritise_long(bool long_running)
{
    if (__builtin_amdgcn_ballot_w64(long_running) != 0)
        __builtin_amdgcn_s_setprio(2);
    else
        __builtin_amdgcn_s_setprio(0);
}

constexpr int VA_LONG_ATTEMPTS = 160;  // ~2x the median attempts of an instance

// Wait for this wave's outstanding vector-memory operations.  Placed at the end
// of a refill, it retires the instance's parameter loads there; otherwise the
// compiler cannot tell at the loop head whether a load is pending and waits
// for every outstanding operation -- including the trajectory stores, which
// count in the same in-order counter -- on each trip.  (gfx9 s_waitcnt
// encoding: vmcnt 0, expcnt and lgkmcnt left at their maxima.)
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

template <int NN>
__device__ __forceinline__ void exchange(double *slab, int lane, int gbase, double v,
                                         double (&out)[NN])
{
    slab[lane] = v;
    __builtin_amdgcn_wave_barrier();
    gather_vec<NN>(slab, gbase, out);
    __builtin_amdgcn_wave_barrier();
}

// Two exchanges through two slabs with one pair of wave barriers, so the
// arithmetic on both results can be interleaved by the scheduler.
template <int NN>
__device__ __forceinline__ void exchange2(double *slab_a, double *slab_b, int lane, int gbase,
                                          double va, double (&out_a)[NN], double vb,
                                          double (&out_b)[NN])
{
    slab_a[lane] = va;
    slab_b[lane] = vb;
    __builtin_amdgcn_wave_barrier();
    gather_vec<NN>(slab_a, gbase, out_a);
    gather_vec<NN>(slab_b, gbase, out_b);
    __builtin_amdgcn_wave_barrier();
}

// 64-bit DPP move (llvm.amdgcn.update.dpp at i64; the clang builtin takes 32-bit
// operands only).  With control ROW_NEWBCAST + j (0x150 + j) every lane of a
// 16-lane row receives lane j of that row in one v_mov_b64_dpp.
extern "C" __device__ long va_update_dpp_i64(long old, long src, int ctrl, int row_mask,
                                             int bank_mask, bool bound_ctrl)
    __asm("llvm.amdgcn.update.dpp.i64");

template <int J>
__device__ __forceinline__ double row_bcast(double x)
{
    return __builtin_bit_cast(
        double, va_update_dpp_i64(0, __builtin_bit_cast(long, x), 0x150 + J, 0xF, 0xF, true));
}

template <int NN, int... J>
__device__ __forceinline__ void bcast_row0(double x, double (&out)[NN],
                                           std::integer_sequence<int, J...>)
{
    ((out[J] = row_bcast<J>(x)), ...);
}

// Group 0 of a wave (lanes 0..N-1) running alone: each lane's value reaches the
// group without an LDS round trip -- by row broadcasts when the group lies in
// the first 16-lane row (one instruction per value), else through scalar
// registers (v_readlane, two per value).
template <int NN>
__device__ __forceinline__ void bcast_group0(double x, double (&out)[NN])
{
    if constexpr (NN <= 16) {
        bcast_row0<NN>(x, out, std::make_integer_sequence<int, NN>{});
    } else {
        const int lo = __double2loint(x), hi = __double2hiint(x);
#pragma unroll
        for (int j = 0; j < NN; ++j)
            out[j] = __hiloint2double(__builtin_amdgcn_readlane(hi, j),
                                      __builtin_amdgcn_readlane(lo, j));
    }
}

template <bool OneShot, int NN>
__device__ __forceinline__ void xchg(double *slab, int lane, int gbase, double x,
                                     double (&out)[NN])
{
    if constexpr (OneShot) bcast_group0<NN>(x, out);
    else exchange<NN>(slab, lane, gbase, x, out);
}

// g_i = r_i + sum_j A_ij x_j in the reference functor's order
// (examples/GeneralizedLotkaVolterra/main.cpp:112-117): sum = 0, sum += A_ij x_j
// for j = 0..N-1 with each product rounded, then r_i + sum.
//
// The sum starts at the first product instead of at 0.0 + p_0: one add fewer
// per row, with the same result when r is taken as r + 0.0 (growth_rate()).
// 0.0 + p equals p unless p = -0, and a sum of rounded products started at
// +0.0 is never -0, so the two sums differ only when every product is -0 (+0
// against -0); then r + (+0) and r' + (-0) with r' = r + 0.0 agree for every r,
// and where they do not, r' + s = r + s as well.
template <int NN>
__device__ __forceinline__ double growth_row(const double (&Arow)[NN], double r0,
                                             const double (&x)[NN])
{
    double sum = Arow[0] * x[0];
#pragma unroll
    for (int j = 1; j < NN; ++j) sum = sum + Arow[j] * x[j];
    return r0 + sum;
}

// The growth rate as growth_row takes it: r + 0.0 (-0 becomes +0, every other
// value is unchanged).
__device__ __forceinline__ double growth_rate(double r) { return r + 0.0; }

// Parameters and dJ/dp of the row kernels.  The persistent kernels over the
// bulk of the batch (IM) read the parameters instance-major, [B][P] (the
// instance's 8(N + N^2) bytes contiguous: they take instances in a
// data-dependent order, and [P][B] reads would fetch a cache line per value),
// and store each instance's dJ/dp sums instance-major as they are,
// [nout][B][P], marking the instance in pbar_set; the host then forms
// 0.0 + sum (with a cost) or mu + sum (without one, adjointSolve adds to mu) in
// [nout][P][B].  The long-instance kernels, a few hundred instances that must
// not wait for those copies, read [P][B] and update mu themselves.
template <int NN, bool IM, class Args>
__device__ __forceinline__ double param(const Args &a, int64_t k, int64_t b)
{
    constexpr int64_t P = NN + NN * NN;
    if constexpr (IM) return a.p[b * P + k];
    else return a.p[k * a.sp.B + b];
}

template <int NN, bool IM>
__device__ __forceinline__ void put_pbar(const RevArgs &a, int64_t o, int64_t k, int64_t b,
                                         double x)
{
    constexpr int64_t P = NN + NN * NN;
    if constexpr (IM) {
        a.mu[(o * a.sp.B + b) * P + k] = x;
    } else {
        double *m = a.mu + (o * P + k) * a.sp.B + b;
        *m = a.cost ? 0.0 + x : *m + x;
    }
}

// A failed instance's dJ/dp output: NaN.
template <int NN, bool IM>
__device__ __forceinline__ void put_pbar_nan(const RevArgs &a, int64_t o, int64_t k, int64_t b)
{
    constexpr int64_t P = NN + NN * NN;
    if constexpr (IM) a.mu[(o * a.sp.B + b) * P + k] = __builtin_nan("");
    else a.mu[(o * P + k) * a.sp.B + b] = __builtin_nan("");
}

// xi_j = w_j g_j + sum_k u_k A_kj, split the same way.
template <int NN>
__device__ __forceinline__ double adjoint_col(const double (&Acol)[NN], double wg,
                                              const double (&u)[NN])
{
    double ae = wg, ao = 0.0;
#pragma unroll
    for (int k = 0; k < NN; k += 2) ae = fma(u[k], Acol[k], ae);
#pragma unroll
    for (int k = 1; k < NN; k += 2) ao = fma(u[k], Acol[k], ao);
    return ae + ao;
}

// Where a group's instances come from: the persistent kernels take queue
// positions [lo, hi) from an atomic counter (mapped through `order` when given;
// `split`/`part` select the positions below or from *split); OneShot bodies
// run the single instance `one` on group 0 of the wave.
struct QueueRange {
    int64_t lo, hi;
};

__device__ __forceinline__ QueueRange queue_range(int64_t B, const int32_t *split, int part)
{
    QueueRange q{0, B};
    if (split) {
        const int64_t k = *split;
        if (part == 0) q.hi = k < B ? k : B;
        else q.lo = k < B ? k : B;
    }
    return q;
}

// The tableau's coefficients as values in registers.  With Opaque the compiler
// sees them as unknown values formed once per kernel, so it keeps them in
// vector registers instead of forming each 64-bit constant in scalar registers
// on every use (two scalar moves each).  The values are the tableau's.
template <class Tab, bool Opaque>
struct TabRegs {
    double a[Tab::SF][Tab::SF], b[Tab::S], db[Tab::SF];
    __device__ __forceinline__ TabRegs()
    {
#pragma unroll
        for (int m = 0; m < Tab::SF; ++m)
#pragma unroll
            for (int j = 0; j < Tab::SF; ++j) {
                a[m][j] = Tab::A[m][j];
                if constexpr (Opaque)
                    if (j < m && Tab::A[m][j] != 0.0) asm volatile("" : "+v"(a[m][j]));
            }
#pragma unroll
        for (int j = 0; j < Tab::S; ++j) {
            b[j] = Tab::B[j];
            if constexpr (Opaque)
                if (Tab::B[j] != 0.0) asm volatile("" : "+v"(b[j]));
        }
#pragma unroll
        for (int j = 0; j < Tab::SF; ++j) {
            db[j] = Tab::HAS_ERR ? Tab::DB[j] : 0.0;
            if constexpr (Opaque && Tab::HAS_ERR)
                if (Tab::DB[j] != 0.0) asm volatile("" : "+v"(db[j]));
        }
    }
};

// ---------------------------------------------------------------------------
// Forward sweep (persistent, refill per instance group)
// ---------------------------------------------------------------------------
template <int NN, class Tab, bool Adaptive, bool OneShot>
__device__ __forceinline__ void glv_rows_fwd_body(const FwdArgs &a, int64_t one)
{
    using L = GlvRows<NN>;
    constexpr int F = L::F;
    __shared__ double2 slab2[L::WAVES][32];
    __shared__ double2 slabb2[L::WAVES][32];
    __shared__ int gidx[L::WAVES][L::GPW];

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int grp = lane / NN, i = lane - grp * NN;
    if (OneShot ? grp != 0 : grp >= L::GPW) return;
    double *slab = reinterpret_cast<double *>(slab2[wv]);
    double *slabb = reinterpret_cast<double *>(slabb2[wv]);
    const int gbase = grp * NN;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    // For tableaus without FSAL the next step's f(y_{n+1}) is computed with the
    // error norm of the attempt (it does not depend on the step decision), so it
    // is off the accept path's critical chain; it is the same arithmetic as
    // evaluating it at the start of the next step.
    constexpr bool kSpec = Adaptive && !Tab::FSAL;

    double Arow[NN], r = 0.0, y = 0.0, K[Tab::SF], ynew = 0.0, xerr = 0.0;
    double t = 0.0, dt = 0.0;
    int64_t b = 0;
    int32_t count = 0, rej = 0, tries = 0, st = LANE_OK;
    bool new_step = true, have_k0 = false, need = true, taken = false;
    double *rows = nullptr;
    double v[NN];
    const QueueRange q = queue_range(B, a.split, a.part);
    const TabRegs<Tab, OneShot> tc;

    for (;;) {
        if constexpr (OneShot) __builtin_amdgcn_s_setprio(3);
        else if (!a.split) prioritise_long(!need && count + rej > VA_LONG_ATTEMPTS);
        if (need) {
            if constexpr (OneShot) {
                if (taken) break;
                taken = true;
                b = one;
            } else {
                if (i == 0) gidx[wv][grp] = atomicAdd(a.counter, 1);
                __builtin_amdgcn_wave_barrier();
                b = q.lo + gidx[wv][grp];
                __builtin_amdgcn_wave_barrier();
                if (b >= q.hi) break;
                if (a.order) b = a.order[b];
            }
            r = growth_rate(param<NN, !OneShot>(a, i, b));
#pragma unroll
            for (int j = 0; j < NN; ++j) Arow[j] = param<NN, !OneShot>(a, NN * (i + 1) + j, b);
            y = a.y[(int64_t)i * B + b];
            t = sp.t0;
            dt = sp.dt0;
            count = rej = tries = 0;
            st = LANE_OK;
            new_step = true;
            have_k0 = false;
            need = false;
            rows = a.traj + (size_t)b * (size_t)(sp.cap + 1) * F;
            wait_vmem();
        }
        bool finish = false;
        if (new_step) {
            const bool more = Adaptive ? less_with_sign(t, sp.t1, dt)
                                       : less_eq_with_sign(t + sp.dt0, sp.t1, sp.dt0);
            if (!more) {
                finish = true;
            } else if (count >= sp.cap) {
                st = LANE_MAX_STEPS;
                finish = true;
            } else {
                double *row = rows + (size_t)count * F;
                row[0] = t;
                row[2 + i] = y;
                if (Adaptive && less_with_sign(sp.t1, t + dt, dt)) dt = sp.t1 - t;
                if ((!Tab::FSAL && !kSpec) || !have_k0) {
                    xchg<OneShot, NN>(slab, lane, gbase, y, v);
                    K[0] = y * growth_row<NN>(Arow, r, v);
                    have_k0 = true;
                }
                tries = 0;
                new_step = false;
            }
        }
        if (!finish) {
            const double h = Adaptive ? dt : sp.dt0;
#pragma unroll
            for (int m = 1; m < Tab::SF; ++m) {
                double X = y;
#pragma unroll
                for (int j = 0; j < Tab::SF; ++j)
                    if (j < m && Tab::A[m][j] != 0.0) X = X + (tc.a[m][j] * h) * K[j];
                xchg<OneShot, NN>(slab, lane, gbase, X, v);
                K[m] = X * growth_row<NN>(Arow, r, v);
            }
            ynew = y;
#pragma unroll
            for (int j = 0; j < Tab::S; ++j)
                if (Tab::B[j] != 0.0) ynew = ynew + (tc.b[j] * h) * K[j];
            if constexpr (Adaptive) {
                constexpr int j0 = ErrFirst<Tab>::value();
                xerr = (tc.db[j0] * h) * K[j0];
#pragma unroll
                for (int j = j0 + 1; j < Tab::SF; ++j)
                    if (Tab::DB[j] != 0.0) xerr = xerr + (tc.db[j] * h) * K[j];
                const double den = sp.atol + sp.rtol * (fabs(y) + h * fabs(K[0]));
                double k0n = 0.0, ve[NN];
                if constexpr (kSpec) {
                    if constexpr (OneShot) {
                        bcast_group0<NN>(ynew, v);
                        bcast_group0<NN>(fabs(xerr) / den, ve);
                    } else {
                        exchange2<NN>(slab, slabb, lane, gbase, ynew, v, fabs(xerr) / den, ve);
                    }
                    k0n = ynew * growth_row<NN>(Arow, r, v);
                } else {
                    xchg<OneShot, NN>(slabb, lane, gbase, fabs(xerr) / den, ve);
                }
                const double err = group_max<NN>(ve);
                double dt_next;
                if (!control_step<Tab>(dt, err, dt_next)) {
                    ++rej;
                } else {
                    const double tn = t + dt;
                    rows[(size_t)count * F + 1] = stored_step(tn, t);
                    t = tn;
                    y = ynew;
                    if constexpr (Tab::FSAL) K[0] = K[Tab::SF - 1];
                    if constexpr (kSpec) K[0] = k0n;
                    ++count;
                    new_step = true;
                }
                dt = dt_next;
                ++tries;
                if (tries > sp.max_attempts) {
                    st = LANE_MAX_FAIL;
                    finish = true;
                }
            } else {
                const double tn = sp.t0 + (double)(count + 1) * sp.dt0;
                rows[(size_t)count * F + 1] = stored_step(tn, t);
                y = ynew;
                if constexpr (Tab::FSAL) K[0] = K[Tab::SF - 1];
                ++count;
                t = tn;
                new_step = true;
            }
        }
        if (finish) {
            if (st == LANE_OK) {
                double *row = rows + (size_t)count * F;
                if (i == 0) {
                    row[0] = t;
                    row[1] = 0.0;
                }
                row[2 + i] = y;
                xchg<OneShot, NN>(slab, lane, gbase, fabs(y), v);
                bool fin = true;
#pragma unroll
                for (int j = 0; j < NN; ++j) fin = fin && (v[j] <= 1.7976931348623157e308);
                if (!fin) st = LANE_NONFINITE;
            }
            a.y[(int64_t)i * B + b] = y;
            if (i == 0) {
                a.n_acc[b] = count;
                a.n_rej[b] = rej;
                a.status[b] = st;
                atomicMax(a.max_steps_seen, count);
            }
            need = true;
        }
    }
}

template <int NN, class Tab, bool Adaptive>
__global__ void __launch_bounds__(256) k_glv_rows_fwd(FwdArgs a)
{
    if constexpr (Adaptive) load_ctrl_tables();
    glv_rows_fwd_body<NN, Tab, Adaptive, false>(a, 0);
}

// ---------------------------------------------------------------------------
// Forward sweep with two rows per lane (N even): N/2 lanes per instance, so a
// wave holds 64/(N/2) instances (12 for N = 10) instead of 64/N (6).  Each
// lane forms both of its rows with the operations of glv_rows_fwd_body, and
// the work every lane of a group repeats -- the step control, the controller's
// pow, the products (a_mj h), the exchanges and the error maximum -- is spread
// over twice as many instances.  Same values as glv_rows_fwd_body.
// ---------------------------------------------------------------------------
template <int NN>
struct GlvRows2 {
    static_assert(NN % 2 == 0, "two rows per lane need an even N");
    static constexpr int H = NN / 2;      // lanes per instance
    static constexpr int GPW = 64 / H;    // instance groups per wave
    static constexpr int WAVES = 4;       // block of 256 threads
};

// The group's N-vector from the two values of each of its lanes (row 2l and
// 2l + 1 in lane l), through the wave's LDS slab.
template <int NN>
__device__ __forceinline__ void exchange_pairs(double *slab, int gbase, int li, double v0,
                                               double v1, double (&out)[NN])
{
    reinterpret_cast<double2 *>(slab + gbase)[li] = make_double2(v0, v1);
    __builtin_amdgcn_wave_barrier();
    gather_vec<NN>(slab, gbase, out);
    __builtin_amdgcn_wave_barrier();
}

template <int NN>
__device__ __forceinline__ void exchange_pairs2(double *slab_a, double *slab_b, int gbase, int li,
                                                double a0, double a1, double (&out_a)[NN],
                                                double b0, double b1, double (&out_b)[NN])
{
    reinterpret_cast<double2 *>(slab_a + gbase)[li] = make_double2(a0, a1);
    reinterpret_cast<double2 *>(slab_b + gbase)[li] = make_double2(b0, b1);
    __builtin_amdgcn_wave_barrier();
    gather_vec<NN>(slab_a, gbase, out_a);
    gather_vec<NN>(slab_b, gbase, out_b);
    __builtin_amdgcn_wave_barrier();
}

template <int NN, class Tab, bool Adaptive>
__device__ __forceinline__ void glv_rows2_fwd_body(const FwdArgs &a)
{
    using L = GlvRows2<NN>;
    constexpr int F = NN + 2, SF = Tab::SF;
    __shared__ double2 slab2[L::WAVES][L::GPW * L::H];
    __shared__ double2 slabb2[L::WAVES][L::GPW * L::H];
    __shared__ int gidx[L::WAVES][L::GPW];

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int grp = lane / L::H, li = lane - grp * L::H;
    if (grp >= L::GPW) return;
    const int i0 = 2 * li;  // this lane's rows: i0 and i0 + 1
    double *slab = reinterpret_cast<double *>(slab2[wv]);
    double *slabb = reinterpret_cast<double *>(slabb2[wv]);
    const int gbase = grp * NN;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    constexpr bool kSpec = Adaptive && !Tab::FSAL;

    double A[2][NN], r[2], y[2], K[SF][2], ynew[2], xerr[2];
    double t = 0.0, dt = 0.0;
    int64_t b = 0;
    int32_t count = 0, rej = 0, tries = 0, st = LANE_OK;
    bool new_step = true, have_k0 = false, need = true;
    double *rows = nullptr;
    double v[NN];
    const QueueRange q = queue_range(B, a.split, a.part);
    const TabRegs<Tab, false> tc;

    for (;;) {
        if (!a.split) prioritise_long(!need && count + rej > VA_LONG_ATTEMPTS);
        if (need) {
            if (li == 0) gidx[wv][grp] = atomicAdd(a.counter, 1);
            __builtin_amdgcn_wave_barrier();
            b = q.lo + gidx[wv][grp];
            __builtin_amdgcn_wave_barrier();
            if (b >= q.hi) break;
            if (a.order) b = a.order[b];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int i = i0 + e;
                r[e] = growth_rate(param<NN, true>(a, i, b));
#pragma unroll
                for (int j = 0; j < NN; ++j) A[e][j] = param<NN, true>(a, NN * (i + 1) + j, b);
                y[e] = a.y[(int64_t)i * B + b];
            }
            t = sp.t0;
            dt = sp.dt0;
            count = rej = tries = 0;
            st = LANE_OK;
            new_step = true;
            have_k0 = false;
            need = false;
            rows = a.traj + (size_t)b * (size_t)(sp.cap + 1) * F;
            wait_vmem();
        }
        bool finish = false;
        if (new_step) {
            const bool more = Adaptive ? less_with_sign(t, sp.t1, dt)
                                       : less_eq_with_sign(t + sp.dt0, sp.t1, sp.dt0);
            if (!more) {
                finish = true;
            } else if (count >= sp.cap) {
                st = LANE_MAX_STEPS;
                finish = true;
            } else {
                double *row = rows + (size_t)count * F;
                row[0] = t;
                *reinterpret_cast<double2 *>(row + 2 + i0) = make_double2(y[0], y[1]);
                if (Adaptive && less_with_sign(sp.t1, t + dt, dt)) dt = sp.t1 - t;
                if ((!Tab::FSAL && !kSpec) || !have_k0) {
                    exchange_pairs<NN>(slab, gbase, li, y[0], y[1], v);
#pragma unroll
                    for (int e = 0; e < 2; ++e) K[0][e] = y[e] * growth_row<NN>(A[e], r[e], v);
                    have_k0 = true;
                }
                tries = 0;
                new_step = false;
            }
        }
        if (!finish) {
            const double h = Adaptive ? dt : sp.dt0;
#pragma unroll
            for (int m = 1; m < SF; ++m) {
                double X[2] = {y[0], y[1]};
#pragma unroll
                for (int j = 0; j < SF; ++j)
                    if (j < m && Tab::A[m][j] != 0.0) {
                        const double ah = tc.a[m][j] * h;
#pragma unroll
                        for (int e = 0; e < 2; ++e) X[e] = X[e] + ah * K[j][e];
                    }
                exchange_pairs<NN>(slab, gbase, li, X[0], X[1], v);
#pragma unroll
                for (int e = 0; e < 2; ++e) K[m][e] = X[e] * growth_row<NN>(A[e], r[e], v);
            }
#pragma unroll
            for (int e = 0; e < 2; ++e) ynew[e] = y[e];
#pragma unroll
            for (int j = 0; j < Tab::S; ++j)
                if (Tab::B[j] != 0.0) {
                    const double bh = tc.b[j] * h;
#pragma unroll
                    for (int e = 0; e < 2; ++e) ynew[e] = ynew[e] + bh * K[j][e];
                }
            if constexpr (Adaptive) {
                constexpr int j0 = ErrFirst<Tab>::value();
                const double dh0 = tc.db[j0] * h;
#pragma unroll
                for (int e = 0; e < 2; ++e) xerr[e] = dh0 * K[j0][e];
#pragma unroll
                for (int j = j0 + 1; j < SF; ++j)
                    if (Tab::DB[j] != 0.0) {
                        const double dh = tc.db[j] * h;
#pragma unroll
                        for (int e = 0; e < 2; ++e) xerr[e] = xerr[e] + dh * K[j][e];
                    }
                double ratio[2];
#pragma unroll
                for (int e = 0; e < 2; ++e)
                    ratio[e] = fabs(xerr[e]) / (sp.atol + sp.rtol * (fabs(y[e]) + h * fabs(K[0][e])));
                double k0n[2] = {0.0, 0.0}, ve[NN];
                if constexpr (kSpec) {
                    exchange_pairs2<NN>(slab, slabb, gbase, li, ynew[0], ynew[1], v, ratio[0],
                                        ratio[1], ve);
#pragma unroll
                    for (int e = 0; e < 2; ++e) k0n[e] = ynew[e] * growth_row<NN>(A[e], r[e], v);
                } else {
                    exchange_pairs<NN>(slabb, gbase, li, ratio[0], ratio[1], ve);
                }
                const double err = group_max<NN>(ve);
                double dt_next;
                if (!control_step<Tab>(dt, err, dt_next)) {
                    ++rej;
                } else {
                    const double tn = t + dt;
                    rows[(size_t)count * F + 1] = stored_step(tn, t);
                    t = tn;
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        y[e] = ynew[e];
                        if constexpr (Tab::FSAL) K[0][e] = K[SF - 1][e];
                        if constexpr (kSpec) K[0][e] = k0n[e];
                    }
                    ++count;
                    new_step = true;
                }
                dt = dt_next;
                ++tries;
                if (tries > sp.max_attempts) {
                    st = LANE_MAX_FAIL;
                    finish = true;
                }
            } else {
                const double tn = sp.t0 + (double)(count + 1) * sp.dt0;
                rows[(size_t)count * F + 1] = stored_step(tn, t);
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    y[e] = ynew[e];
                    if constexpr (Tab::FSAL) K[0][e] = K[SF - 1][e];
                }
                ++count;
                t = tn;
                new_step = true;
            }
        }
        if (finish) {
            if (st == LANE_OK) {
                double *row = rows + (size_t)count * F;
                if (li == 0) {
                    row[0] = t;
                    row[1] = 0.0;
                }
                *reinterpret_cast<double2 *>(row + 2 + i0) = make_double2(y[0], y[1]);
                exchange_pairs<NN>(slab, gbase, li, fabs(y[0]), fabs(y[1]), v);
                bool fin = true;
#pragma unroll
                for (int j = 0; j < NN; ++j) fin = fin && (v[j] <= 1.7976931348623157e308);
                if (!fin) st = LANE_NONFINITE;
            }
#pragma unroll
            for (int e = 0; e < 2; ++e) a.y[(int64_t)(i0 + e) * B + b] = y[e];
            if (li == 0) {
                a.n_acc[b] = count;
                a.n_rej[b] = rej;
                a.status[b] = st;
                atomicMax(a.max_steps_seen, count);
            }
            need = true;
        }
    }
}

template <int NN, class Tab, bool Adaptive>
__global__ void __launch_bounds__(256) k_glv_rows2_fwd(FwdArgs a)
{
    if constexpr (Adaptive) load_ctrl_tables();
    glv_rows2_fwd_body<NN, Tab, Adaptive>(a);
}



// ---------------------------------------------------------------------------
// Reverse sweep (persistent, refill per instance group), software-pipelined.
//
// The reverse of step n needs the stage inputs X_m and growth rates g_m of
// step n, recomputed from the stored (h_n, y_n); that recompute does not
// depend on the adjoint.  So each sweep iteration recomputes step n-1 while it
// reverses step n: stage i of the recompute runs beside backward stage S-1-i,
// the two share one LDS exchange round (two writes, two broadcast gathers: the
// X-vector and the u-vector), and their dependent chains interleave.  Each
// lane keeps its own recomputed X_m and g_m in registers until the backward of
// the next iteration reads them: the slot that backward stage S-1-i frees is
// where recomputed stage i goes, so the slot of stage m is m or S-1-m by the
// parity of the step.  The first iteration of an
// instance (or of a further seed) only recomputes.  Every value is formed by
// the same operations in the same order as in the oracle.
// ---------------------------------------------------------------------------
template <int NN, class Tab, bool OneShot>
__device__ __forceinline__ void glv_rows_rev_body(const RevArgs &a, int64_t one)
{
    using L = GlvRows<NN>;
    constexpr int F = L::F, S = Tab::S, P = NN + NN * NN;
    // per wave: one exchange slab for a stage-input vector, one for the
    // adjoint products u
    __shared__ double2 xsl[L::WAVES][32];
    __shared__ double2 usl2[L::WAVES][32];
    __shared__ int gidx[L::WAVES][L::GPW];

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int grp = lane / NN, i = lane - grp * NN;
    if (OneShot ? grp != 0 : grp >= L::GPW) return;
    const int gbase = grp * NN;
    double *uslab = reinterpret_cast<double *>(usl2[wv]);
    double *xslab = reinterpret_cast<double *>(xsl[wv]);
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const double qnan = __builtin_nan("");

    // Abar: column i of dJ/dA (entries A_ki, k = 0..N-1)
    double Arow[NN], Acol[NN], Abar[NN], r = 0.0, rbar = 0.0, lam = 0.0;
    double Kn[S], Kb[S];
    // the lane's own stage inputs X_m and growth rates g_m of the last
    // recomputed step, in slot m or S-1-m by the parity of that step
    double Xs[S], Gs[S];
    // row nr's (h, y_i) of this lane, loaded one iteration ahead of its use
    double h_nxt = 0.0, y_nxt = 0.0, h_back = 0.0;
    int64_t b = 0;
    int32_t o = 0, nr = -1;
    int par = 0;  // slot parity of the step being reversed
    bool need = true, back = false, taken = false;
    const double *rows = nullptr;
    double v[NN];
    const QueueRange q = queue_range(B, a.split, a.part);

    for (;;) {
        if constexpr (OneShot) __builtin_amdgcn_s_setprio(3);
        else if (!a.split) prioritise_long(!need && nr > VA_LONG_ATTEMPTS);
        if (need) {
            if constexpr (OneShot) {
                if (taken) break;
                taken = true;
                b = one;
            } else {
                if (i == 0) gidx[wv][grp] = atomicAdd(a.counter, 1);
                __builtin_amdgcn_wave_barrier();
                b = q.lo + gidx[wv][grp];
                __builtin_amdgcn_wave_barrier();
                if (b >= q.hi) break;
                if (a.order) b = a.order[b];
            }
            if (a.status[b] != LANE_OK) {
                const int nseed = a.cost ? 1 : a.nout;
                for (int oo = 0; oo < nseed; ++oo) {
                    a.lambda[((int64_t)oo * NN + i) * B + b] = qnan;
                    put_pbar_nan<NN, !OneShot>(a, oo, i, b);
                    for (int j = 0; j < NN; ++j)
                        put_pbar_nan<NN, !OneShot>(a, oo, NN * (i + 1) + j, b);
                }
                if (!OneShot && a.pbar_set && i == 0) a.pbar_set[b] = 1;
                if (a.cost && i == 0) a.J[b] = qnan;
                continue;
            }
            r = growth_rate(param<NN, !OneShot>(a, i, b));
#pragma unroll
            for (int j = 0; j < NN; ++j) {
                Arow[j] = param<NN, !OneShot>(a, NN * (i + 1) + j, b);
                Acol[j] = param<NN, !OneShot>(a, NN * (j + 1) + i, b);
            }
            rows = a.traj + (size_t)b * (size_t)(sp.cap + 1) * F;
            o = 0;
            if (a.cost) {
                const double yT = a.yT[(int64_t)i * B + b];
                // terminal cost: J = fold over i in order, lambda_i from y_i(T)
                xchg<OneShot, NN>(uslab, lane, gbase, yT, v);
                double J = 0.0;
#pragma unroll
                for (int j = 0; j < NN; ++j) {
                    if (a.cost == COST_SUM) J = J + v[j];
                    else if (a.cost == COST_SQUARED_NORM) J = fma(v[j], v[j], J);
                    else J = fma(a.cost_w[j], v[j], J);
                }
                if (a.cost == COST_SUM) lam = 1.0;
                else if (a.cost == COST_SQUARED_NORM) lam = 2.0 * yT;
                else lam = a.cost_w[i];
                if (i == 0) a.J[b] = J;
            } else {
                lam = a.lambda[(int64_t)i * B + b];
            }
            rbar = 0.0;
#pragma unroll
            for (int j = 0; j < NN; ++j) Abar[j] = 0.0;
            nr = a.n_acc[b] - 1;
            back = false;
            if (nr >= 0) {
                h_nxt = rows[(size_t)nr * F + 1];
                y_nxt = rows[(size_t)nr * F + 2 + i];
            }
            need = false;
            wait_vmem();
        }
        if (nr >= 0 || back) {
            // recompute step nr (stage inputs from (h_nr, y_nr), backpropagation.hpp:24-64)
            // beside the reverse of step nr + 1 (:104-153)
            const double hr = h_nxt, yr = y_nxt;
            if (nr >= 1) {
                const double *prev = rows + (size_t)(nr - 1) * F;
                h_nxt = prev[1];
                y_nxt = prev[2 + i];
            }
            // Stage ii of step nr goes to the register slot that stage S-1-ii of
            // step nr + 1 frees in the same iteration: slot ii when step nr has
            // parity 0, S-1-ii when it has parity 1 (the parity alternates from
            // step to step, so both mappings are fixed at compile time).
            auto step = [&](auto parity) {
                constexpr int Q = decltype(parity)::value;
                if (!back) {
#pragma unroll
                    for (int m = 0; m < S; ++m) {
                        double x = yr;
#pragma unroll
                        for (int j = 0; j < S; ++j)
                            if (j < m && Tab::A[m][j] != 0.0) x = x + (Tab::A[m][j] * hr) * Kn[j];
                        if constexpr (OneShot) bcast_group0<NN>(x, v);
                        else exchange<NN>(xslab, lane, gbase, x, v);
                        const double g = growth_row<NN>(Arow, r, v);
                        const int sl = Q ? S - 1 - m : m;
                        Xs[sl] = x;
                        Gs[sl] = g;
                        Kn[m] = x * g;
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < S; ++m) Kb[m] = (Tab::B[m] * h_back) * lam;
                    double ybar = lam;
#pragma unroll
                    for (int ii = 0; ii < S; ++ii) {
                        const int m = S - 1 - ii;
                        const int sl = Q ? S - 1 - ii : ii;
                        const double w = Kb[m];
                        const double xown = Xs[sl];
                        const double u = w * xown;
                        const double wg = w * Gs[sl];
                        double x = yr;
#pragma unroll
                        for (int j = 0; j < S; ++j)
                            if (j < ii && Tab::A[ii][j] != 0.0) x = x + (Tab::A[ii][j] * hr) * Kn[j];
                        double vu[NN];
                        if constexpr (OneShot) {
                            bcast_group0<NN>(x, v);
                            bcast_group0<NN>(u, vu);
                        } else {
                            exchange2<NN>(xslab, uslab, lane, gbase, x, v, u, vu);
                        }
                        const double g = growth_row<NN>(Arow, r, v);
                        Xs[sl] = x;
                        Gs[sl] = g;
                        Kn[ii] = x * g;
                        const double xi = adjoint_col<NN>(Acol, wg, vu);
                        rbar = rbar + u;
                        // column i of dJ/dA: dA_ki += u_k X_i, from the gathered u-vector
                        // and the lane's own X (no X-vector gather), each product
                        // rounded before it is added
#pragma unroll
                        for (int k = 0; k < NN; ++k) Abar[k] = Abar[k] + vu[k] * xown;
                        ybar = ybar + xi;
                        // Kbar_j += (xi a_mj) h (detail/backpropagation.hpp:215)
#pragma unroll
                        for (int j = 0; j < S; ++j)
                            if (j < m && Tab::A[m][j] != 0.0) Kb[j] = Kb[j] + (xi * Tab::A[m][j]) * h_back;
                    }
                    lam = ybar;
                }
            };
            const int pr = par ^ 1;  // parity of step nr
            if (pr) step(std::integral_constant<int, 1>{});
            else step(std::integral_constant<int, 0>{});
            h_back = hr;
            par = pr;
            back = nr >= 0;
            --nr;
        }
        if (!back) {
            const int64_t lo = (int64_t)o * NN;
            a.lambda[(lo + i) * B + b] = lam;
            put_pbar<NN, !OneShot>(a, o, i, b, rbar);
#pragma unroll
            for (int k = 0; k < NN; ++k) put_pbar<NN, !OneShot>(a, o, NN * (k + 1) + i, b, Abar[k]);
            if (!OneShot && a.pbar_set && i == 0) a.pbar_set[b] = 1;
            ++o;
            if (a.cost || o >= a.nout) {
                need = true;
            } else {
                lam = a.lambda[((int64_t)o * NN + i) * B + b];
                rbar = 0.0;
#pragma unroll
                for (int j = 0; j < NN; ++j) Abar[j] = 0.0;
                nr = a.n_acc[b] - 1;
                if (nr >= 0) {
                    h_nxt = rows[(size_t)nr * F + 1];
                    y_nxt = rows[(size_t)nr * F + 2 + i];
                }
                wait_vmem();
            }
        }
    }
}

template <int NN, class Tab>
__global__ void __launch_bounds__(256) k_glv_rows_rev(RevArgs a)
{
    glv_rows_rev_body<NN, Tab, false>(a, 0);
}

// ---------------------------------------------------------------------------
// Instances expected to run long: queue positions [0, *split) of the forward
// order, one block of three waves per instance.  Wave 0 runs the instance's
// forward and then at once the backward of its reverse, so it never waits for
// another instance; the rest of the batch (positions from *split) runs in the
// persistent kernels on another stream meanwhile.  During the reverse, wave 1
// forms the stage values of each stored step -- the stage inputs X_m and the
// growth rates g_m from (h_n, y_n) with h_n = t_{n+1} - t_n, the operations of
// the recompute in glv_rows_rev_body -- a few steps ahead of wave 0 into a ring
// of LDS slots, and wave 2 accumulates dJ/dr and dJ/dA from the adjoint
// products u that wave 0 leaves in the same slot, so wave 0's dependent chain
// holds the adjoint recursion alone (backpropagation.hpp:104-153).  The waves
// meet through three counters in LDS: the steps staged (wave 1), the steps
// whose products are written (wave 0) and the steps accumulated (wave 2, which
// frees the slot; it cannot pass wave 0, so wave 0 needs no count of its own
// to reuse a slot).
// ---------------------------------------------------------------------------
constexpr int VA_STAGE_RING = 8;         // steps staged ahead of the backward
constexpr int VA_SPIN_LIMIT = 1 << 22;   // polls of a ring counter (about 0.1 s) before giving up
// LDS a long-instance block reserves beyond its own use (its own is ~13 KB), so
// that no two such blocks share a CU's 160 KB: each long forward then has a
// SIMD without another long wave (two long waves on one SIMD each ran ~40%
// slower), and the persistent kernels' blocks (<= 10 KB) still fit beside it.
constexpr int VA_LONG_LDS_PAD = 68 * 1024;

template <int NN, class Tab>
struct StageRing {
    double2 xg[VA_STAGE_RING][Tab::S][NN];  // (X_m,i, g_m,i) of a staged step
    double u[VA_STAGE_RING][Tab::S][NN];    // its adjoint products u_m,i = w_m,i X_m,i
    double h[VA_STAGE_RING];                // its h_n
    int avail;                              // steps staged (wave 1)
    int uavail;                             // steps whose products wave 0 has written
    int done;                               // steps wave 2 has accumulated: slots free
    int abort;                              // a wait gave up: every wave stops
};

// The ring's counters are relaxed workgroup-scope atomics (single LDS reads and
// writes, no wait for their completion); the fences around them order the
// slot accesses.
__device__ __forceinline__ int ring_load(int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void ring_store(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Wait until *ctr > k.  Every wait ends: it returns false when the other wave
// gave up, and gives up itself (setting abort) after VA_SPIN_LIMIT polls.
template <class Ring>
__device__ __forceinline__ bool wait_above(Ring &R, int *ctr, int k)
{
    for (int spins = 0; spins < VA_SPIN_LIMIT; ++spins) {
        if (ring_load(ctr) > k) {
            // the other wave's slot accesses before its count come first
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            return true;
        }
        if (ring_load(&R.abort)) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    ring_store(&R.abort, 1);
    return false;
}

// Wave 1: stage values of steps n_acc-1 .. 0 into the ring.
template <int NN, class Tab>
__device__ __forceinline__ void glv_long_stage(const RevArgs &a, int64_t b, StageRing<NN, Tab> &R)
{
    constexpr int S = Tab::S, F = NN + 2;
    const int i = threadIdx.x & 63;
    if (i >= NN) return;
    __builtin_amdgcn_s_setprio(3);
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int32_t nacc = a.n_acc[b];
    if (a.status[b] != LANE_OK || nacc <= 0) return;
    const double r = growth_rate(param<NN, false>(a, i, b));
    double Arow[NN], v[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) Arow[j] = param<NN, false>(a, NN * (i + 1) + j, b);
    const double *rows = a.traj + (size_t)b * (size_t)(sp.cap + 1) * F;
    const TabRegs<Tab, true> tc;
    // (h_n, y_n,i) of the next two steps, loaded ahead of use
    int n = nacc - 1;
    double h0 = rows[(size_t)n * F + 1], y0 = rows[(size_t)n * F + 2 + i], h1 = 0.0, y1 = 0.0;
    if (n >= 1) {
        h1 = rows[(size_t)(n - 1) * F + 1];
        y1 = rows[(size_t)(n - 1) * F + 2 + i];
    }
    for (int k = 0; k < nacc; ++k, --n) {
        const double h = h0, y = y0;
        h0 = h1;
        y0 = y1;
        if (n >= 2) {
            h1 = rows[(size_t)(n - 2) * F + 1];
            y1 = rows[(size_t)(n - 2) * F + 2 + i];
        }
        // the slot of step k - RING must have been accumulated by wave 2 (and so
        // read by wave 0 before that)
        if (k >= VA_STAGE_RING && !wait_above(R, &R.done, k - VA_STAGE_RING)) return;
        const int slot = k % VA_STAGE_RING;
        double Kn[S];
#pragma unroll
        for (int m = 0; m < S; ++m) {
            double x = y;
#pragma unroll
            for (int j = 0; j < S; ++j)
                if (j < m && Tab::A[m][j] != 0.0) x = x + (tc.a[m][j] * h) * Kn[j];
            bcast_group0<NN>(x, v);
            const double g = growth_row<NN>(Arow, r, v);
            Kn[m] = x * g;
            R.xg[slot][m][i] = make_double2(x, g);
        }
        R.h[slot] = h;
        // the slot's writes complete before the count that publishes them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        ring_store(&R.avail, k + 1);
    }
}

// Wave 0 after the forward: terminal cost, then the adjoint recursion of steps
// n_acc-1 .. 0 from the staged values (the backward of glv_rows_rev_body),
// leaving each stage's products u in the step's slot for wave 2.
template <int NN, class Tab>
__device__ __forceinline__ void glv_long_backward(const RevArgs &a, int64_t b,
                                                  StageRing<NN, Tab> &R)
{
    constexpr int S = Tab::S, P = NN + NN * NN;
    const int i = threadIdx.x & 63;
    if (i >= NN) return;
    __builtin_amdgcn_s_setprio(3);
    const int64_t B = a.sp.B;
    const double qnan = __builtin_nan("");
    if (a.status[b] != LANE_OK) {
        a.lambda[(int64_t)i * B + b] = qnan;
        put_pbar_nan<NN, false>(a, 0, i, b);
        for (int j = 0; j < NN; ++j) put_pbar_nan<NN, false>(a, 0, NN * (i + 1) + j, b);
        if (a.cost) a.J[b] = qnan;
        return;
    }
    const TabRegs<Tab, true> tc;
    double Acol[NN], v[NN], vu[NN], lam;
#pragma unroll
    for (int j = 0; j < NN; ++j) Acol[j] = param<NN, false>(a, NN * (j + 1) + i, b);
    if (a.cost) {
        const double yT = a.yT[(int64_t)i * B + b];
        // terminal cost: J = fold over i in order, lambda_i from y_i(T)
        bcast_group0<NN>(yT, v);
        double J = 0.0;
#pragma unroll
        for (int j = 0; j < NN; ++j) {
            if (a.cost == COST_SUM) J = J + v[j];
            else if (a.cost == COST_SQUARED_NORM) J = fma(v[j], v[j], J);
            else J = fma(a.cost_w[j], v[j], J);
        }
        if (a.cost == COST_SUM) lam = 1.0;
        else if (a.cost == COST_SQUARED_NORM) lam = 2.0 * yT;
        else lam = a.cost_w[i];
        a.J[b] = J;
    } else {
        lam = a.lambda[(int64_t)i * B + b];
    }
    const int32_t nacc = a.n_acc[b];
    for (int k = 0; k < nacc; ++k) {
        if (!wait_above(R, &R.avail, k)) {
            lam = qnan;
            break;
        }
        const int slot = k % VA_STAGE_RING;
        double X[S], G[S];
#pragma unroll
        for (int m = 0; m < S; ++m) {
            const double2 xg = R.xg[slot][m][i];
            X[m] = xg.x;
            G[m] = xg.y;
        }
        const double h = R.h[slot];
        double Kb[S];
#pragma unroll
        for (int m = 0; m < S; ++m) Kb[m] = (tc.b[m] * h) * lam;
        double ybar = lam;
#pragma unroll
        for (int ii = 0; ii < S; ++ii) {
            const int m = S - 1 - ii;
            const double w = Kb[m];
            const double xown = X[m];
            const double u = w * xown;
            const double wg = w * G[m];
            R.u[slot][m][i] = u;
            bcast_group0<NN>(u, vu);
            const double xi = adjoint_col<NN>(Acol, wg, vu);
            ybar = ybar + xi;
            // Kbar_j += (xi a_mj) h (detail/backpropagation.hpp:215)
#pragma unroll
            for (int j = 0; j < S; ++j)
                if (j < m && Tab::A[m][j] != 0.0) Kb[j] = Kb[j] + (xi * tc.a[m][j]) * h;
        }
        lam = ybar;
        // the step's products are written before the count that publishes them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        ring_store(&R.uavail, k + 1);
    }
    a.lambda[(int64_t)i * B + b] = lam;
}

// Wave 2: dJ/dr_i and column i of dJ/dA from the products u of each step wave 0
// has reversed, in wave 0's stage order (what glv_rows_rev_body accumulates
// beside its recursion, the same operations in the same order).
template <int NN, class Tab>
__device__ __forceinline__ void glv_long_accumulate(const RevArgs &a, int64_t b,
                                                    StageRing<NN, Tab> &R)
{
    constexpr int S = Tab::S;
    const int i = threadIdx.x & 63;
    if (i >= NN) return;
    __builtin_amdgcn_s_setprio(3);
    // wave 0 writes the outputs of an instance that did not end OK
    if (a.status[b] != LANE_OK) return;
    const double qnan = __builtin_nan("");
    double Abar[NN], vu[NN], rbar = 0.0;
#pragma unroll
    for (int j = 0; j < NN; ++j) Abar[j] = 0.0;
    const int32_t nacc = a.n_acc[b];
    for (int k = 0; k < nacc; ++k) {
        if (!wait_above(R, &R.uavail, k)) {
            rbar = qnan;
#pragma unroll
            for (int j = 0; j < NN; ++j) Abar[j] = qnan;
            break;
        }
        const int slot = k % VA_STAGE_RING;
#pragma unroll
        for (int ii = 0; ii < S; ++ii) {
            const int m = S - 1 - ii;
#pragma unroll
            for (int kk = 0; kk < NN; ++kk) vu[kk] = R.u[slot][m][kk];
            const double xown = R.xg[slot][m][i].x;
            const double u = R.u[slot][m][i];
            rbar = rbar + u;
            // column i of dJ/dA: dA_ki += u_k X_i, each product rounded before it is added
#pragma unroll
            for (int kk = 0; kk < NN; ++kk) Abar[kk] = Abar[kk] + vu[kk] * xown;
        }
        // the slot's values are read: release it to waves 0 and 1
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        ring_store(&R.done, k + 1);
    }
    put_pbar<NN, false>(a, 0, i, b, rbar);
#pragma unroll
    for (int kk = 0; kk < NN; ++kk) put_pbar<NN, false>(a, 0, NN * (kk + 1) + i, b, Abar[kk]);
}

template <int NN, class Tab>
__global__ void __launch_bounds__(192) k_glv_rows_long(FwdArgs fa, RevArgs ra)
{
    __shared__ StageRing<NN, Tab> ring;
    __shared__ double lds_pad[VA_LONG_LDS_PAD / 8];
    if (fa.sp.B < 0) lds_pad[threadIdx.x] = 0.0;  // never: keeps the reservation
    load_ctrl_tables();
    const int64_t w = blockIdx.x;
    if (w >= *fa.split || w >= fa.sp.B) return;  // the whole block
    const int64_t b = fa.order[w];
    if (threadIdx.x == 0) {
        ring.avail = 0;
        ring.uavail = 0;
        ring.done = 0;
        ring.abort = 0;
    }
    if (threadIdx.x < 64) glv_rows_fwd_body<NN, Tab, true, true>(fa, b);
    // the forward's rows, y(T), counts and status become visible to every wave
    // (the barrier orders the block's memory operations at workgroup scope)
    __syncthreads();
    if (threadIdx.x < 64) glv_long_backward<NN, Tab>(ra, b, ring);
    else if (threadIdx.x < 128) glv_long_stage<NN, Tab>(ra, b, ring);
    else glv_long_accumulate<NN, Tab>(ra, b, ring);
}

// The same instances in a forward alone (va_forward) and in a reverse alone
// (va_adjoint, positions [0, *split) of the order of the last forward).  The
// forward runs one wave per instance, four to a block: with one long forward
// per CU (the long kernel's LDS reservation) the bulk forward beside it lost
// more than the long forwards gained (forward call 2.20-2.25 -> 2.38-2.39 ms).
template <int NN, class Tab>
__global__ void __launch_bounds__(256) k_glv_rows_long_fwd(FwdArgs fa)
{
    load_ctrl_tables();
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= *fa.split || w >= fa.sp.B) return;
    glv_rows_fwd_body<NN, Tab, true, true>(fa, fa.order[w]);
}

// The reverse alone for one seed, as in k_glv_rows_long: one block of three
// waves per instance (the adjoint recursion, the stage values ahead of it, the
// dJ/dr and dJ/dA accumulation), one block per CU.
template <int NN, class Tab>
__global__ void __launch_bounds__(192) k_glv_rows_long_adj(RevArgs ra)
{
    __shared__ StageRing<NN, Tab> ring;
    __shared__ double lds_pad[VA_LONG_LDS_PAD / 8];
    if (ra.sp.B < 0) lds_pad[threadIdx.x] = 0.0;  // never: keeps the reservation
    const int64_t w = blockIdx.x;
    if (w >= *ra.split || w >= ra.sp.B) return;  // the whole block
    const int64_t b = ra.order[w];
    if (threadIdx.x == 0) {
        ring.avail = 0;
        ring.uavail = 0;
        ring.done = 0;
        ring.abort = 0;
    }
    __syncthreads();
    if (threadIdx.x < 64) glv_long_backward<NN, Tab>(ra, b, ring);
    else if (threadIdx.x < 128) glv_long_stage<NN, Tab>(ra, b, ring);
    else glv_long_accumulate<NN, Tab>(ra, b, ring);
}

template <int NN, class Tab>
__global__ void __launch_bounds__(256) k_glv_rows_long_rev(RevArgs ra)
{
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= *ra.split || w >= ra.sp.B) return;
    glv_rows_rev_body<NN, Tab, true>(ra, ra.order[w]);
}

// ---------------------------------------------------------------------------
// Plain form in the same layout: one launch per accepted step, no refill, every
// row (t, h, y) stored in [step][field][instance] order; one reverse launch per
// step, recomputing the step's stages from y_n.
// ---------------------------------------------------------------------------
template <int NN, class Tab, bool Adaptive>
__global__ void __launch_bounds__(256) k_glv_rows_plain_fwd(PlainFwdArgs a)
{
    if constexpr (Adaptive) load_ctrl_tables();
    using L = GlvRows<NN>;
    constexpr int F = 2 + NN;  // plain rows (t, h, y)
    __shared__ double2 slab2[L::WAVES][32];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int grp = lane / NN, i = lane - grp * NN;
    if (grp >= L::GPW) return;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int64_t b = ((int64_t)blockIdx.x * L::WAVES + wv) * L::GPW + grp;
    if (b >= B || a.st.done[b]) return;
    double *slab = reinterpret_cast<double *>(slab2[wv]);
    const int gbase = grp * NN;
    const int32_t step = a.step;
    double v[NN], Arow[NN], K[Tab::SF];
    const double r = growth_rate(a.p[(int64_t)i * B + b]);
#pragma unroll
    for (int j = 0; j < NN; ++j) Arow[j] = a.p[(int64_t)(NN * (i + 1) + j) * B + b];
    double y = a.y[(int64_t)i * B + b];
    double t = step == 0 ? sp.t0 : a.st.t[b];
    double dt = step == 0 ? sp.dt0 : a.st.dt[b];
    int32_t rej = step == 0 ? 0 : a.n_rej[b];
    double *row = a.traj + (size_t)step * F * B + b;
    const bool more = Adaptive ? less_with_sign(t, sp.t1, dt)
                               : less_eq_with_sign(t + sp.dt0, sp.t1, sp.dt0);
    if (!more || step >= sp.cap) {
        int32_t st = LANE_OK;
        if (more) {
            st = LANE_MAX_STEPS;
        } else {
            if (i == 0) {
                row[0] = t;
                row[(size_t)B] = 0.0;
            }
            row[(size_t)(2 + i) * B] = y;
            exchange<NN>(slab, lane, gbase, fabs(y), v);
            bool fin = true;
#pragma unroll
            for (int j = 0; j < NN; ++j) fin = fin && (v[j] <= 1.7976931348623157e308);
            if (!fin) st = LANE_NONFINITE;
        }
        if (i == 0) {
            a.status[b] = st;
            a.n_acc[b] = step;
            a.n_rej[b] = rej;
            a.st.done[b] = 1;
            atomicSub(a.st.active, 1);
            atomicMax(a.st.max_steps_seen, step);
        }
        return;
    }
    if (i == 0) row[0] = t;
    row[(size_t)(2 + i) * B] = y;
    if (Adaptive && less_with_sign(sp.t1, t + dt, dt)) dt = sp.t1 - t;
    if (Tab::FSAL && step > 0) {
        K[0] = a.st.k0[(int64_t)i * B + b];
    } else {
        exchange<NN>(slab, lane, gbase, y, v);
        K[0] = y * growth_row<NN>(Arow, r, v);
    }
    double ynew = y;
    int32_t tries = 0;
    for (;;) {
        const double h = Adaptive ? dt : sp.dt0;
#pragma unroll
        for (int m = 1; m < Tab::SF; ++m) {
            double X = y;
#pragma unroll
            for (int j = 0; j < Tab::SF; ++j)
                if (j < m && Tab::A[m][j] != 0.0) X = X + (Tab::A[m][j] * h) * K[j];
            exchange<NN>(slab, lane, gbase, X, v);
            K[m] = X * growth_row<NN>(Arow, r, v);
        }
        ynew = y;
#pragma unroll
        for (int j = 0; j < Tab::S; ++j)
            if (Tab::B[j] != 0.0) ynew = ynew + (Tab::B[j] * h) * K[j];
        if constexpr (!Adaptive) {
            const double tn = sp.t0 + (double)(step + 1) * sp.dt0;
            if (i == 0) row[(size_t)B] = stored_step(tn, t);
            t = tn;
            break;
        } else {
            constexpr int j0 = ErrFirst<Tab>::value();
            double xerr = (Tab::DB[j0] * h) * K[j0];
#pragma unroll
            for (int j = j0 + 1; j < Tab::SF; ++j)
                if (Tab::DB[j] != 0.0) xerr = xerr + (Tab::DB[j] * h) * K[j];
            const double den = sp.atol + sp.rtol * (fabs(y) + h * fabs(K[0]));
            exchange<NN>(slab, lane, gbase, fabs(xerr) / den, v);
            const double err = group_max<NN>(v);
            double dt_next;
            const bool ok = control_step<Tab>(dt, err, dt_next);
            if (!ok) {
                ++rej;
            } else {
                const double tn = t + dt;
                if (i == 0) row[(size_t)B] = stored_step(tn, t);
                t = tn;
            }
            dt = dt_next;
            ++tries;
            if (tries > sp.max_attempts) {
                if (ok) a.y[(int64_t)i * B + b] = ynew;
                if (i == 0) {
                    a.status[b] = LANE_MAX_FAIL;
                    a.n_acc[b] = step + (ok ? 1 : 0);
                    a.n_rej[b] = rej;
                    a.st.done[b] = 1;
                    atomicSub(a.st.active, 1);
                }
                return;
            }
            if (ok) break;
        }
    }
    a.y[(int64_t)i * B + b] = ynew;
    if constexpr (Tab::FSAL) a.st.k0[(int64_t)i * B + b] = K[Tab::SF - 1];
    if (i == 0) {
        a.st.t[b] = t;
        a.st.dt[b] = dt;
        a.n_rej[b] = rej;
        a.n_acc[b] = step + 1;
        a.status[b] = LANE_OK;
    }
}

template <int NN, class Tab>
__global__ void __launch_bounds__(256) k_glv_rows_plain_rev(PlainRevArgs a)
{
    using L = GlvRows<NN>;
    constexpr int S = Tab::S;
    constexpr int F = 2 + NN;  // plain rows (t, h, y)
    __shared__ double2 xsl2[L::WAVES][S][32];
    __shared__ double2 usl2[L::WAVES][32];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int grp = lane / NN, i = lane - grp * NN;
    if (grp >= L::GPW) return;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int64_t b = ((int64_t)blockIdx.x * L::WAVES + wv) * L::GPW + grp;
    if (b >= B || a.status[b] != LANE_OK) return;
    const int32_t n = a.n_acc[b] - 1 - a.k;
    if (n < 0) return;
    const int gbase = grp * NN;
    double *uslab = reinterpret_cast<double *>(usl2[wv]);
    double v[NN], Arow[NN], Acol[NN];
    const double r = growth_rate(a.p[(int64_t)i * B + b]);
#pragma unroll
    for (int j = 0; j < NN; ++j) {
        Arow[j] = a.p[(int64_t)(NN * (i + 1) + j) * B + b];
        Acol[j] = a.p[(int64_t)(NN * (j + 1) + i) * B + b];
    }
    const double *row = a.traj + (size_t)n * F * B + b;
    const double h = row[(size_t)B];
    const double yn = row[(size_t)(2 + i) * B];
    // recompute the stages from y_n (compute_intermediate_states,
    // detail/backpropagation.hpp:24-64)
    double X[S], g[S], K[S], Kb[S];
#pragma unroll
    for (int m = 0; m < S; ++m) {
        double x = yn;
#pragma unroll
        for (int j = 0; j < S; ++j)
            if (j < m && Tab::A[m][j] != 0.0) x = x + (Tab::A[m][j] * h) * K[j];
        X[m] = x;
        double *xs = reinterpret_cast<double *>(xsl2[wv][m]);
        exchange<NN>(xs, lane, gbase, x, v);
        g[m] = growth_row<NN>(Arow, r, v);
        K[m] = x * g[m];
    }
    double lam = a.lambda[(int64_t)i * B + b];
    double rbar = a.st.pbar[(int64_t)i * B + b];
    double Abar[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) Abar[j] = a.st.pbar[(int64_t)(NN * (i + 1) + j) * B + b];
#pragma unroll
    for (int m = 0; m < S; ++m) Kb[m] = (Tab::B[m] * h) * lam;
    double ybar = lam;
#pragma unroll
    for (int m = S - 1; m >= 0; --m) {
        const double w = Kb[m];
        const double u = w * X[m];
        exchange<NN>(uslab, lane, gbase, u, v);
        const double xi = adjoint_col<NN>(Acol, w * g[m], v);
        rbar = rbar + u;
        gather_vec<NN>(reinterpret_cast<const double *>(xsl2[wv][m]), gbase, v);
#pragma unroll
        for (int j = 0; j < NN; ++j) Abar[j] = Abar[j] + u * v[j];
        ybar = ybar + xi;
#pragma unroll
        for (int j = 0; j < S; ++j)
            if (j < m && Tab::A[m][j] != 0.0) Kb[j] = Kb[j] + (xi * Tab::A[m][j]) * h;
    }
    a.lambda[(int64_t)i * B + b] = ybar;
    a.st.pbar[(int64_t)i * B + b] = rbar;
#pragma unroll
    for (int j = 0; j < NN; ++j) a.st.pbar[(int64_t)(NN * (i + 1) + j) * B + b] = Abar[j];
}

template <int NN, class Tab, bool Ad, bool WithLong = true>
struct LaunchGlvRows {
    // The persistent forward with two rows per lane where N is even and small
    // enough for both rows' parameters to stay in registers, when the instances
    // come longest-first (a.order): a wave then holds instances of similar
    // length.  In index order (a handle's first solve) a wave runs as long as the
    // longest of its instances, and 12 unordered instances per wave made the
    // first solve 18% slower than 6, so that forward keeps one row per lane.
    static constexpr bool kPairs = NN % 2 == 0 && NN <= 16;
    static hipError_t fwd(const FwdArgs &a, int grid, int block, hipStream_t s)
    {
        if constexpr (kPairs) {
            if (a.order) {
                hipLaunchKernelGGL((k_glv_rows2_fwd<NN, Tab, Ad>), dim3(grid), dim3(block), 0, s,
                                   a);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((k_glv_rows_fwd<NN, Tab, Ad>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    static hipError_t rev(const RevArgs &a, int grid, int block, hipStream_t s)
    {
        hipLaunchKernelGGL((k_glv_rows_rev<NN, Tab>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    static hipError_t lng(const FwdArgs &fa, const RevArgs &ra, int grid, int block,
                          hipStream_t s)
    {
        hipLaunchKernelGGL((k_glv_rows_long<NN, Tab>), dim3(grid), dim3(block), 0, s, fa, ra);
        return hipGetLastError();
    }
    static hipError_t lng_fwd(const FwdArgs &a, int grid, int block, hipStream_t s)
    {
        hipLaunchKernelGGL((k_glv_rows_long_fwd<NN, Tab>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    // block 192: one block of three waves per instance (one seed); otherwise one
    // wave per instance (any number of seeds)
    static hipError_t lng_rev(const RevArgs &a, int grid, int block, hipStream_t s)
    {
        if (block == 192)
            hipLaunchKernelGGL((k_glv_rows_long_adj<NN, Tab>), dim3(grid), dim3(block), 0, s, a);
        else
            hipLaunchKernelGGL((k_glv_rows_long_rev<NN, Tab>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    static hipError_t pfwd(const PlainFwdArgs &a, int grid, int block, hipStream_t s)
    {
        hipLaunchKernelGGL((k_glv_rows_plain_fwd<NN, Tab, Ad>), dim3(grid), dim3(block), 0, s,
                           a);
        return hipGetLastError();
    }
    static hipError_t prev(const PlainRevArgs &a, int grid, int block, hipStream_t s)
    {
        hipLaunchKernelGGL((k_glv_rows_plain_rev<NN, Tab>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    static int occ_fwd(int block)
    {
        int nb = 0;
        hipError_t e;
        if constexpr (kPairs)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_glv_rows2_fwd<NN, Tab, Ad>,
                                                             block, 0);
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_glv_rows_fwd<NN, Tab, Ad>,
                                                             block, 0);
        return e == hipSuccess ? nb : 1;
    }
    static int occ_rev(int block)
    {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_glv_rows_rev<NN, Tab>, block, 0) !=
            hipSuccess)
            return 1;
        return nb;
    }
    // Instance slots of one block of `block` threads.
    static int slots(int block) { return (block / 64) * GlvRows<NN>::GPW; }
    static void apply(KernelSet *k)
    {
        k->fwd = &fwd;
        k->rev = &rev;
        k->plain_fwd = &pfwd;
        k->plain_rev = &prev;
        k->occ_fwd = &occ_fwd;
        k->occ_rev = &occ_rev;
        k->slots = &slots;
        k->layout = 1;
        k->s_rev = Tab::S;
        k->s_fwd = Tab::SF;
        if constexpr (Ad && WithLong) {
            k->long_fused = &lng;
            k->long_fwd = &lng_fwd;
            k->long_rev = &lng_rev;
        }
    }
};

// Row-per-lane kernels only (no one-lane-per-instance kernels, whose per-lane
// copy of A does not fit in registers for larger N, and no long-instance kernels:
// solves then run unsplit): the adaptive Cash-Karp 5(4), DOPRI5 and Fehlberg
// 7(8) loops and the fixed-step Euler and RK4 loops.
template <int NN>
inline bool pick_glv_rows_core(int tableau, int adaptive, KernelSet *out)
{
    *out = KernelSet{};
    out->n = NN;
    out->np = LotkaVolterra<NN>::P;
    out->flop_rhs = LotkaVolterra<NN>::FLOP_RHS;
    out->flop_vjp = LotkaVolterra<NN>::FLOP_VJP;
    if (adaptive) {
        switch (tableau) {
        case TAB_CK54: LaunchGlvRows<NN, CashKarp54, true, false>::apply(out); return true;
        case TAB_DOPRI5: LaunchGlvRows<NN, Dopri5, true, false>::apply(out); return true;
        case TAB_F78: LaunchGlvRows<NN, Fehlberg78, true, false>::apply(out); return true;
        default: return false;
        }
    }
    switch (tableau) {
    case TAB_EULER: LaunchGlvRows<NN, Euler, false, false>::apply(out); return true;
    case TAB_RK4: LaunchGlvRows<NN, RK4, false, false>::apply(out); return true;
    default: return false;
    }
}

template <int NN>
inline bool pick_glv_rows(int tableau, int adaptive, KernelSet *out)
{
    if (!pick_tableau<LotkaVolterra<NN>>(tableau, adaptive, out)) return false;
    if (adaptive) {
        switch (tableau) {
        case TAB_CK54: LaunchGlvRows<NN, CashKarp54, true>::apply(out); return true;
        case TAB_DOPRI5: LaunchGlvRows<NN, Dopri5, true>::apply(out); return true;
        case TAB_F78: LaunchGlvRows<NN, Fehlberg78, true>::apply(out); return true;
        default: return false;
        }
    }
    switch (tableau) {
    case TAB_EULER: LaunchGlvRows<NN, Euler, false>::apply(out); return true;
    case TAB_RK4: LaunchGlvRows<NN, RK4, false>::apply(out); return true;
    case TAB_CK54: LaunchGlvRows<NN, CashKarp54, false>::apply(out); return true;
    case TAB_DOPRI5: LaunchGlvRows<NN, Dopri5, false>::apply(out); return true;
    case TAB_F78: LaunchGlvRows<NN, Fehlberg78, false>::apply(out); return true;
    default: return false;
    }
}

}  // namespace va
