// va_kernels_impl.hpp -- generic (one lane per instance) HIP kernels of the batched
// explicit-RK forward sweep and discrete-adjoint reverse sweep, gfx950.
//
//   k_fwd       persistent forward: lanes pull instances from an atomic work
//               counter (lane refill), run odeint's controlled loop per lane and
//               store (t_n, y_n) rows -- lib/include/detail/runge_kutta.hpp:74-118
//               (fixed-step loop :36-72), StateStorage.hpp:28-35.
//   k_rev       persistent reverse: replays the stored rows backwards, recomputes
//               the stages and applies the RK adjoint recursion --
//               lib/include/detail/backpropagation.hpp:83-158,231-254,280-348.
//   k_plain_*   the plain GPU form: one launch per accepted step, no refill,
//               every row (t, h, y) stored [step][field][instance]; one reverse
//               launch per step, which recomputes the step's stages from y_n as
//               the reference's compute_intermediate_states does.
//
// Every lane runs a flattened state machine (one attempt / one reverse step per
// loop trip) so the wave stays converged on the arithmetic; the refill branch is
// the only divergent path.
#pragma once
#include "va_device.hpp"
#include "va_kernels.hpp"

namespace va {

enum { LANE_OK = 0, LANE_MAX_FAIL = 1, LANE_NONFINITE = 2, LANE_MAX_STEPS = 3 };
enum { COST_REFERENCE = 0, COST_SUM = 1, COST_SQUARED_NORM = 2, COST_LINEAR = 3 };

template <int N>
__device__ __forceinline__ bool all_finite(const double *y)
{
    bool ok = true;
#pragma unroll
    for (int i = 0; i < N; ++i) ok = ok && (fabs(y[i]) <= 1.7976931348623157e308);
    return ok;
}

// ---------------------------------------------------------------------------
// Persistent forward
// ---------------------------------------------------------------------------
template <class Sys, class Tab, bool Adaptive>
__global__ void __launch_bounds__(256) k_fwd(FwdArgs a)
{
    if constexpr (Adaptive) load_ctrl_tables();
    constexpr int N = Sys::N;
    constexpr int F = N + 2;  // row (t_n, h_n, y_n)
    using R = RK<Sys, Tab>;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;

    typename Sys::Params q;
    double y[N], K[Tab::SF][N], ynew[N], xerr[N];
    double t = 0.0, dt = 0.0;
    int64_t b = 0;
    int32_t count = 0, rej = 0, tries = 0, st = LANE_OK;
    bool new_step = true, have_k0 = false, need = true;
    double *rows = nullptr;

    for (;;) {
        if (need) {
            b = atomicAdd(a.counter, 1);
            if (b >= B) break;
            if (a.order) b = a.order[b];
            Sys::load(q, a.p, B, b, sp.kconst);
#pragma unroll
            for (int i = 0; i < N; ++i) y[i] = a.y[(int64_t)i * B + b];
            t = sp.t0;
            dt = sp.dt0;
            count = rej = tries = 0;
            st = LANE_OK;
            new_step = true;
            have_k0 = false;
            need = false;
            rows = a.traj + (size_t)b * (size_t)(sp.cap + 1) * F;
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
                double *r = rows + (size_t)count * F;
                r[0] = t;
#pragma unroll
                for (int i = 0; i < N; ++i) r[2 + i] = y[i];
                if (Adaptive && less_with_sign(sp.t1, t + dt, dt)) dt = sp.t1 - t;
                if (!Tab::FSAL || !have_k0) {
                    Sys::rhs(q, y, K[0]);
                    have_k0 = true;
                }
                tries = 0;
                new_step = false;
            }
        }
        if (!finish) {
            const double h = Adaptive ? dt : sp.dt0;
            R::attempt(q, y, K, h, ynew, xerr);
            if constexpr (Adaptive) {
                const double err = R::error_norm(sp.atol, sp.rtol, h, y, K[0], xerr);
                double dt_next;
                const bool accept = control_step<Tab>(dt, err, dt_next);
                if (!accept) {
                    ++rej;
                } else {
                    const double tn = t + dt;
                    rows[(size_t)count * F + 1] = stored_step(tn, t);
                    t = tn;
#pragma unroll
                    for (int i = 0; i < N; ++i) y[i] = ynew[i];
                    if constexpr (Tab::FSAL) {
#pragma unroll
                        for (int i = 0; i < N; ++i) K[0][i] = K[Tab::SF - 1][i];
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
#pragma unroll
                for (int i = 0; i < N; ++i) y[i] = ynew[i];
                if constexpr (Tab::FSAL) {
#pragma unroll
                    for (int i = 0; i < N; ++i) K[0][i] = K[Tab::SF - 1][i];
                }
                ++count;
                const double tn = sp.t0 + (double)count * sp.dt0;
                rows[(size_t)(count - 1) * F + 1] = stored_step(tn, t);
                t = tn;
                new_step = true;
            }
        }
        if (finish) {
            if (st == LANE_OK) {
                double *r = rows + (size_t)count * F;
                r[0] = t;
                r[1] = 0.0;
#pragma unroll
                for (int i = 0; i < N; ++i) r[2 + i] = y[i];
                if (!all_finite<N>(y)) st = LANE_NONFINITE;
            }
#pragma unroll
            for (int i = 0; i < N; ++i) a.y[(int64_t)i * B + b] = y[i];
            a.n_acc[b] = count;
            a.n_rej[b] = rej;
            a.status[b] = st;
            atomicMax(a.max_steps_seen, count);
            need = true;
        }
    }
}

// ---------------------------------------------------------------------------
// Reverse step (detail/backpropagation.hpp:83-158) on one stored row pair.
// ---------------------------------------------------------------------------
template <class Sys, class Tab>
__device__ __forceinline__ void back_prop_step(const typename Sys::Params &q, double h,
                                               const double *yn, double *lam, double *pbar)
{
    // h = t_{n+1} - t_n as stored by the forward (the reference's GetDt).
    constexpr int N = Sys::N, S = Tab::S;
    using R = RK<Sys, Tab>;
    double K[Tab::SF][N], X[S][N], Kb[S][N];
#pragma unroll
    for (int m = 0; m < S; ++m) {
        R::stage_input(m, h, yn, K, X[m]);
        Sys::rhs(q, X[m], K[m]);
    }
#pragma unroll
    for (int m = 0; m < S; ++m) {
        const double bh = Tab::B[m] * h;
#pragma unroll
        for (int i = 0; i < N; ++i) Kb[m][i] = bh * lam[i];
    }
    double ybar[N];
#pragma unroll
    for (int i = 0; i < N; ++i) ybar[i] = lam[i];
#pragma unroll
    for (int m = S - 1; m >= 0; --m) {
        double xi[N];
        Sys::vjp(q, X[m], Kb[m], xi, pbar);
#pragma unroll
        for (int i = 0; i < N; ++i) ybar[i] = ybar[i] + xi[i];
#pragma unroll
        for (int j = 0; j < S; ++j) {
            if (j < m && Tab::A[m][j] != 0.0) {
#pragma unroll
                for (int i = 0; i < N; ++i) Kb[j][i] = Kb[j][i] + (xi[i] * Tab::A[m][j]) * h;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) lam[i] = ybar[i];
}

template <int N>
__device__ __forceinline__ double terminal_cost(int cost, const double *w, const double *y,
                                                double *lam)
{
    double J = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (cost == COST_SUM) {
            J = J + y[i];
            lam[i] = 1.0;
        } else if (cost == COST_SQUARED_NORM) {
            J = fma(y[i], y[i], J);
            lam[i] = 2.0 * y[i];
        } else {
            J = fma(w[i], y[i], J);
            lam[i] = w[i];
        }
    }
    return J;
}

// ---------------------------------------------------------------------------
// Persistent reverse
// ---------------------------------------------------------------------------
template <class Sys, class Tab>
__global__ void __launch_bounds__(256) k_rev(RevArgs a)
{
    constexpr int N = Sys::N, P = Sys::P, F = N + 2;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const double qnan = __builtin_nan("");

    typename Sys::Params q;
    double lam[N], pbar[P];
    int64_t b = 0;
    int32_t o = 0, n = -1;
    bool need = true;
    const double *rows = nullptr;

    for (;;) {
        if (need) {
            b = atomicAdd(a.counter, 1);
            if (b >= B) break;
            if (a.order) b = a.order[b];
            if (a.status[b] != LANE_OK) {
                const int nseed = a.cost ? 1 : a.nout;
                for (int oo = 0; oo < nseed; ++oo) {
                    for (int i = 0; i < N; ++i) a.lambda[((int64_t)oo * N + i) * B + b] = qnan;
                    for (int k = 0; k < P; ++k) a.mu[((int64_t)oo * P + k) * B + b] = qnan;
                }
                if (a.cost) a.J[b] = qnan;
                continue;
            }
            Sys::load(q, a.p, B, b, sp.kconst);
            rows = a.traj + (size_t)b * (size_t)(sp.cap + 1) * F;
            o = 0;
            if (a.cost) {
                double yT[N];
#pragma unroll
                for (int i = 0; i < N; ++i) yT[i] = a.yT[(int64_t)i * B + b];
                a.J[b] = terminal_cost<N>(a.cost, a.cost_w, yT, lam);
            } else {
#pragma unroll
                for (int i = 0; i < N; ++i) lam[i] = a.lambda[(int64_t)i * B + b];
            }
#pragma unroll
            for (int k = 0; k < P; ++k) pbar[k] = 0.0;
            n = a.n_acc[b] - 1;
            need = false;
        }
        if (n >= 0) {
            const double *r = rows + (size_t)n * F;
            back_prop_step<Sys, Tab>(q, r[1], r + 2, lam, pbar);
            --n;
        }
        if (n < 0) {
            const int64_t lo = (int64_t)o * N, po = (int64_t)o * P;
#pragma unroll
            for (int i = 0; i < N; ++i) a.lambda[(lo + i) * B + b] = lam[i];
            if (a.cost) {
#pragma unroll
                for (int k = 0; k < P; ++k) a.mu[(po + k) * B + b] = 0.0 + pbar[k];
            } else {
#pragma unroll
                for (int k = 0; k < P; ++k) a.mu[(po + k) * B + b] = a.mu[(po + k) * B + b] + pbar[k];
            }
            ++o;
            if (a.cost || o >= a.nout) {
                need = true;
            } else {
#pragma unroll
                for (int i = 0; i < N; ++i) lam[i] = a.lambda[((int64_t)o * N + i) * B + b];
#pragma unroll
                for (int k = 0; k < P; ++k) pbar[k] = 0.0;
                n = a.n_acc[b] - 1;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Plain GPU form
// ---------------------------------------------------------------------------
template <class Sys, class Tab, bool Adaptive>
__global__ void __launch_bounds__(256) k_plain_fwd(PlainFwdArgs a)
{
    if constexpr (Adaptive) load_ctrl_tables();
    constexpr int N = Sys::N;
    constexpr int F = 2 + N;  // fields (t_n, h_n, y_n)
    using R = RK<Sys, Tab>;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || a.st.done[b]) return;
    const int32_t step = a.step;

    typename Sys::Params q;
    Sys::load(q, a.p, B, b, sp.kconst);
    double y[N], K[Tab::SF][N], ynew[N], xerr[N];
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = a.y[(int64_t)i * B + b];
    double t = step == 0 ? sp.t0 : a.st.t[b];
    double dt = step == 0 ? sp.dt0 : a.st.dt[b];
    int32_t rej = step == 0 ? 0 : a.n_rej[b];
    double *row = a.traj + (size_t)step * F * B + b;

    const bool more = Adaptive ? less_with_sign(t, sp.t1, dt)
                               : less_eq_with_sign(t + sp.dt0, sp.t1, sp.dt0);
    int32_t st = LANE_OK;
    if (!more || step >= sp.cap) {
        if (more) {
            st = LANE_MAX_STEPS;
        } else {
            row[0] = t;
            row[(size_t)B] = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) row[(size_t)(2 + i) * B] = y[i];
            if (!all_finite<N>(y)) st = LANE_NONFINITE;
        }
        a.status[b] = st;
        a.n_acc[b] = step;
        a.n_rej[b] = rej;
        a.st.done[b] = 1;
        atomicSub(a.st.active, 1);
        atomicMax(a.st.max_steps_seen, step);
        return;
    }
    row[0] = t;
#pragma unroll
    for (int i = 0; i < N; ++i) row[(size_t)(2 + i) * B] = y[i];
    if (Adaptive && less_with_sign(sp.t1, t + dt, dt)) dt = sp.t1 - t;
    if (Tab::FSAL && step > 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) K[0][i] = a.st.k0[(int64_t)i * B + b];
    } else {
        Sys::rhs(q, y, K[0]);
    }
    if constexpr (Adaptive) {
        int32_t tries = 0;
        for (;;) {
            R::attempt(q, y, K, dt, ynew, xerr);
            const double err = R::error_norm(sp.atol, sp.rtol, dt, y, K[0], xerr);
            double dt_next;
            const bool ok = control_step<Tab>(dt, err, dt_next);
            if (!ok) {
                ++rej;
            } else {
                const double tn = t + dt;
                row[(size_t)B] = stored_step(tn, t);
                t = tn;
            }
            dt = dt_next;
            ++tries;
            if (tries > sp.max_attempts) {
                if (ok) {
#pragma unroll
                    for (int i = 0; i < N; ++i) a.y[(int64_t)i * B + b] = ynew[i];
                }
                a.status[b] = LANE_MAX_FAIL;
                a.n_acc[b] = step + (ok ? 1 : 0);
                a.n_rej[b] = rej;
                a.st.done[b] = 1;
                atomicSub(a.st.active, 1);
                return;
            }
            if (ok) break;
        }
    } else {
        R::attempt(q, y, K, sp.dt0, ynew, xerr);
        const double tn = sp.t0 + (double)(step + 1) * sp.dt0;
        row[(size_t)B] = stored_step(tn, t);
        t = tn;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) a.y[(int64_t)i * B + b] = ynew[i];
    if constexpr (Tab::FSAL) {
#pragma unroll
        for (int i = 0; i < N; ++i) a.st.k0[(int64_t)i * B + b] = K[Tab::SF - 1][i];
    }
    a.st.t[b] = t;
    a.st.dt[b] = dt;
    a.n_rej[b] = rej;
    a.n_acc[b] = step + 1;
    a.status[b] = LANE_OK;
}

template <class Sys, class Tab>
__global__ void __launch_bounds__(256) k_plain_rev(PlainRevArgs a)
{
    constexpr int N = Sys::N, P = Sys::P;
    constexpr int F = 2 + N;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || a.status[b] != LANE_OK) return;
    const int32_t n = a.n_acc[b] - 1 - a.k;
    if (n < 0) return;

    typename Sys::Params q;
    Sys::load(q, a.p, B, b, sp.kconst);
    const double *row = a.traj + (size_t)n * F * B + b;
    const double h = row[(size_t)B];
    double yn[N], lam[N], pbar[P];
#pragma unroll
    for (int i = 0; i < N; ++i) yn[i] = row[(size_t)(2 + i) * B];
#pragma unroll
    for (int i = 0; i < N; ++i) lam[i] = a.lambda[(int64_t)i * B + b];
#pragma unroll
    for (int k = 0; k < P; ++k) pbar[k] = a.st.pbar[(int64_t)k * B + b];
    back_prop_step<Sys, Tab>(q, h, yn, lam, pbar);
#pragma unroll
    for (int i = 0; i < N; ++i) a.lambda[(int64_t)i * B + b] = lam[i];
#pragma unroll
    for (int k = 0; k < P; ++k) a.st.pbar[(int64_t)k * B + b] = pbar[k];
}

// ---------------------------------------------------------------------------
// Launchers and the dispatch table
// ---------------------------------------------------------------------------
template <class Sys, class Tab, bool Ad>
struct Launch {
    static hipError_t fwd(const FwdArgs &a, int grid, int block, hipStream_t s)
    {
        hipLaunchKernelGGL((k_fwd<Sys, Tab, Ad>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    static hipError_t rev(const RevArgs &a, int grid, int block, hipStream_t s)
    {
        hipLaunchKernelGGL((k_rev<Sys, Tab>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    static hipError_t pfwd(const PlainFwdArgs &a, int grid, int block, hipStream_t s)
    {
        hipLaunchKernelGGL((k_plain_fwd<Sys, Tab, Ad>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    static hipError_t prev(const PlainRevArgs &a, int grid, int block, hipStream_t s)
    {
        hipLaunchKernelGGL((k_plain_rev<Sys, Tab>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    static int occ_fwd(int block)
    {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fwd<Sys, Tab, Ad>, block, 0) !=
            hipSuccess)
            return 1;
        return nb;
    }
    static int occ_rev(int block)
    {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_rev<Sys, Tab>, block, 0) !=
            hipSuccess)
            return 1;
        return nb;
    }
    static int slots(int block) { return block; }
    static KernelSet make()
    {
        KernelSet k{};
        k.n = Sys::N;
        k.np = Sys::P;
        k.s_rev = Tab::S;
        k.s_fwd = Tab::SF;
        k.flop_rhs = Sys::FLOP_RHS;
        k.flop_vjp = Sys::FLOP_VJP;
        k.fwd = &fwd;
        k.rev = &rev;
        k.plain_fwd = &pfwd;
        k.plain_rev = &prev;
        k.occ_fwd = &occ_fwd;
        k.occ_rev = &occ_rev;
        k.slots = &slots;
        k.layout = 0;
        return k;
    }
};


enum { TAB_EULER = 1, TAB_RK4 = 2, TAB_CK54 = 3, TAB_DOPRI5 = 4, TAB_F78 = 5 };


template <class Sys>
inline bool pick_tableau(int tableau, int adaptive, KernelSet *out)
{
    if (adaptive) {
        switch (tableau) {
        case TAB_CK54: *out = Launch<Sys, CashKarp54, true>::make(); return true;
        case TAB_DOPRI5: *out = Launch<Sys, Dopri5, true>::make(); return true;
        case TAB_F78: *out = Launch<Sys, Fehlberg78, true>::make(); return true;
        default: return false;
        }
    }
    switch (tableau) {
    case TAB_EULER: *out = Launch<Sys, Euler, false>::make(); return true;
    case TAB_RK4: *out = Launch<Sys, RK4, false>::make(); return true;
    case TAB_CK54: *out = Launch<Sys, CashKarp54, false>::make(); return true;
    case TAB_DOPRI5: *out = Launch<Sys, Dopri5, false>::make(); return true;
    case TAB_F78: *out = Launch<Sys, Fehlberg78, false>::make(); return true;
    default: return false;
    }
}

}  // namespace va
