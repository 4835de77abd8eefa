// va_observed.hpp -- observation-cost forward and reverse sweeps (one lane per
// instance), with the trajectory stored in full or kept as checkpoints.
//
// No reference counterpart: the reference's cost is terminal only
// (examples/VanDerPol/main.cpp:98-117).  The forward is the controlled loop of
// lib/include/detail/runge_kutta.hpp:74-118 run segment by segment: a step that
// would pass the next observation time is shortened to land on it, as at t1
// (:98-101); after such a shortened step is accepted, the next step size is the
// larger of the controller's proposal and the size before shortening (what
// Boost.odeint 1.74 integrate_times does for controlled steppers).  At each
// observation time the running cost J += |y - y_obs_k|^2 is accumulated and the
// stored-point index recorded.  The reverse is the discrete adjoint of
// detail/backpropagation.hpp:83-158 with lambda += 2 (y_n - y_obs_k) at every
// observation point.
//
// Checkpoints.  With a trajectory budget too small for every row, an instance
// keeps `slots` checkpoints (t, dt to try, observations taken, y) at every
// `stride`-th accepted step; when the slots fill up, every other checkpoint is
// dropped and the stride doubles, so any step count up to slots * stride_max
// fits.  The reverse rebuilds each segment's rows from its checkpoint by
// replaying the controlled loop (the same decisions, so bit-identical rows, and
// f(y) from the checkpoint equals the FSAL value carried by the forward), then
// sweeps the segment backwards.  Memory: slots rows per instance plus
// stride_max rows per lane; work: one extra forward pass.
#pragma once

#include "va_kernels_impl.hpp"

namespace va {

// The time the controlled loop heads for after observation kobs - 1: observation
// kobs's time, or t1 once every observation is taken.  It waits for its load at
// once (s_waitcnt vmcnt(0)): it runs only where an observation is passed, and
// with the load retired there the steps in between test the target without
// waiting for the wave's earlier row stores (gfx9 counts loads and stores in
// one in-order counter).
__device__ __forceinline__ double obs_time(const double *obs_t, int K, int kobs, double t1)
{
    const double v = kobs < K ? obs_t[kobs] : t1;
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt 0; expcnt and lgkmcnt left at their maxima
    return v;
}

// The stored-point index of observation kobs, or -1 (no point) once kobs < 0.
__device__ __forceinline__ int32_t obs_point(const int32_t *orow, int kobs)
{
    return kobs >= 0 ? orow[kobs] : -1;
}

// Observation jumps at stored point `point` (state yn), newest observation first.
// `next` holds obs_point(orow, kobs) and is reloaded only when kobs moves, so
// the steps between observations test a register instead of waiting on a load.
template <int N>
__device__ __forceinline__ int obs_jumps(const int32_t *orow, const double *obs_y, int kobs,
                                         int32_t &next, int32_t point, const double *yn,
                                         double *lam)
{
    while (kobs >= 0 && next == point) {
#pragma unroll
        for (int i = 0; i < N; ++i) lam[i] = fma(2.0, yn[i] - obs_y[kobs * N + i], lam[i]);
        --kobs;
        next = obs_point(orow, kobs);
    }
    return kobs;
}

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------
template <class Sys, class Tab, bool Ckpt>
__global__ void __launch_bounds__(256) k_obs_fwd(ObsFwdArgs a)
{
    load_ctrl_tables();
    constexpr int N = Sys::N;
    constexpr int F = N + 2, FC = N + 3;
    using R = RK<Sys, Tab>;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int K = a.K;

    typename Sys::Params q;
    double y[N], Kst[Tab::SF][N], ynew[N], xerr[N];
    // tnext: the next target, obs_time(a.obs_t, K, kobs, sp.t1), reloaded only
    // when kobs moves, so a step's clipping test waits on no load
    double t = 0.0, dt = 0.0, dt_keep = 0.0, Jacc = 0.0, tnext = 0.0;
    int64_t b = 0;
    int32_t count = 0, rej = 0, tries = 0, st = LANE_OK, kobs = 0, stride = 1, nck = 0;
    bool new_step = true, have_k0 = false, need = true, clipped = false;
    double *rows = nullptr, *ck = nullptr;
    int32_t *orow = nullptr;

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
            Jacc = 0.0;
            stride = 1;
            if constexpr (Ckpt) {
                // Start from the spacing the previous solve of this instance reached
                // (its step count is still in n_acc until this one finishes): the
                // smallest with room for that many steps, so that a solve taking
                // the same steps again never fills the slots.  Each doubling moves
                // half the checkpoints, one lane at a time, and with one lane per
                // instance the whole wave waits for any lane that moves them (half
                // that spacing, which filled the slots once per solve, cost 10% of
                // the observed Van der Pol solve).  Any spacing gives the same
                // results: the reverse replays the forward from each checkpoint.
                if (a.order) {
                    const int32_t prev = a.n_acc[b];
                    while ((int64_t)stride * a.st.slots <= prev) stride *= 2;
                }
            }
            count = rej = tries = kobs = nck = 0;
            tnext = obs_time(a.obs_t, K, 0, sp.t1);
            st = LANE_OK;
            new_step = true;
            have_k0 = false;
            need = false;
            if constexpr (Ckpt)
                ck = a.st.ck + (size_t)b * (size_t)a.st.slots * FC;
            else
                rows = a.st.traj + (size_t)b * (size_t)(sp.cap + 1) * F;
            orow = a.st.obs_row + (size_t)b * (size_t)K;
        }
        bool finish = false;
        if (new_step) {
            // observations reached at the current point, then the next target
            double target = sp.t1;
            bool is_obs = false;
            for (;;) {
                is_obs = kobs < K;
                target = tnext;
                if (!is_obs || less_with_sign(t, target, dt)) break;
                orow[kobs] = count;
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    const double d = y[i] - a.obs_y[kobs * N + i];
                    Jacc = fma(d, d, Jacc);
                }
                ++kobs;
                tnext = obs_time(a.obs_t, K, kobs, sp.t1);
            }
            if (!less_with_sign(t, target, dt)) {
                finish = true;
            } else if (count >= sp.cap) {
                st = LANE_MAX_STEPS;
                finish = true;
            } else {
                if constexpr (Ckpt) {
                    // stride is a power of two (it starts at 1 and is only doubled
                    // or halved), so this is count % stride == 0 without the
                    // integer division's two dozen instructions on every step
                    if ((count & (stride - 1)) == 0) {
                        if (nck >= a.st.slots) {
                            // drop every other checkpoint, double the spacing
                            for (int m = 1; m < a.st.slots / 2; ++m)
#pragma unroll
                                for (int f = 0; f < FC; ++f)
                                    ck[(size_t)m * FC + f] = ck[(size_t)(2 * m) * FC + f];
                            stride *= 2;
                            nck = a.st.slots / 2;
                        }
                        double *c = ck + (size_t)nck * FC;
                        c[0] = t;
                        c[1] = dt;
                        c[2] = (double)kobs;
#pragma unroll
                        for (int i = 0; i < N; ++i) c[3 + i] = y[i];
                        ++nck;
                    }
                } else {
                    double *r = rows + (size_t)count * F;
                    r[0] = t;
#pragma unroll
                    for (int i = 0; i < N; ++i) r[2 + i] = y[i];
                }
                clipped = false;
                dt_keep = dt;
                if (less_with_sign(target, t + dt, dt)) {
                    dt = target - t;
                    clipped = is_obs;
                }
                if (!Tab::FSAL || !have_k0) {
                    Sys::rhs(q, y, Kst[0]);
                    have_k0 = true;
                }
                tries = 0;
                new_step = false;
            }
        }
        if (!finish) {
            const double h = dt;
            R::attempt(q, y, Kst, h, ynew, xerr);
            const double err = R::error_norm(sp.atol, sp.rtol, h, y, Kst[0], xerr);
            double dt_next;
            if (!control_step<Tab>(dt, err, dt_next)) {
                ++rej;
                clipped = false;
            } else {
                const double tn = t + dt;
                if constexpr (!Ckpt) rows[(size_t)count * F + 1] = stored_step(tn, t);
                t = tn;
#pragma unroll
                for (int i = 0; i < N; ++i) y[i] = ynew[i];
                if constexpr (Tab::FSAL) {
#pragma unroll
                    for (int i = 0; i < N; ++i) Kst[0][i] = Kst[Tab::SF - 1][i];
                }
                if (clipped && fabs(dt_keep) > fabs(dt_next)) dt_next = dt_keep;
                ++count;
                new_step = true;
            }
            dt = dt_next;
            ++tries;
            if (tries > sp.max_attempts) {
                st = LANE_MAX_FAIL;
                finish = true;
            }
        }
        if (finish) {
            if (st == LANE_OK) {
                if constexpr (!Ckpt) {
                    double *r = rows + (size_t)count * F;
                    r[0] = t;
                    r[1] = 0.0;
#pragma unroll
                    for (int i = 0; i < N; ++i) r[2 + i] = y[i];
                }
                if (!all_finite<N>(y)) st = LANE_NONFINITE;
            }
#pragma unroll
            for (int i = 0; i < N; ++i) a.y[(int64_t)i * B + b] = y[i];
            a.J[b] = st == LANE_OK ? Jacc : __builtin_nan("");
            a.n_acc[b] = count;
            a.n_rej[b] = rej;
            a.status[b] = st;
            if constexpr (Ckpt) a.st.stride[b] = stride;
            atomicMax(a.max_steps_seen, count);
            need = true;
        }
    }
}

// ---------------------------------------------------------------------------
// Reverse
// ---------------------------------------------------------------------------

// Replay `len` accepted steps from the checkpoint c (t, dt to try, observations
// taken, y), running the forward's controlled loop with the same observation
// stops, and either (Rows) rebuild their rows (h_n, y_n) into seg, h_n =
// t_{n+1} - t_n, or (!Rows) store a checkpoint of the same form as c at the start
// of every `sub`-th step into seg (the level-2 checkpoints of a long segment).
// Field f of row / checkpoint k is seg[(k W + f) ss] with W = N+1 / N+3: the
// lanes' buffers are interleaved (ss = the grid's thread count), so a wave's
// stores and loads of one field coalesce.  One trip of the loop is one attempt,
// whether the lane starts a step or retries one, so the lanes of a wave do not
// wait for each other's rejected attempts step by step (with 64 lanes some lane
// rejects in most steps); the values formed are those of the forward.
template <class Sys, class Tab, bool Rows = true>
__device__ void recompute_segment(const typename Sys::Params &q, const SolveParams &sp,
                                  const double *obs_t, int K, const double *c, int32_t len,
                                  double *seg, int64_t ss, int32_t sub = 1)
{
    constexpr int N = Sys::N, FS = N + 1, FC = N + 3;
    using R = RK<Sys, Tab>;
    double y[N], Kst[Tab::SF][N], ynew[N], xerr[N];
    double t = c[0], dt = c[1], dt_keep = 0.0;
    int kobs = (int)c[2];
    double tnext = obs_time(obs_t, K, kobs, sp.t1);  // as in the forward
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = c[3 + i];
    // f(y) here equals the FSAL value the forward carried into this step: the
    // last stage input of DOPRI5 is y_{n+1} computed with the same operations.
    Sys::rhs(q, y, Kst[0]);
    int32_t k = 0, tries = 0;
    bool new_step = true, clipped = false;
    double *r = seg;  // field 0 of row k
    while (k < len) {
        if (new_step) {
            double target;
            bool is_obs;
            for (;;) {
                is_obs = kobs < K;
                target = tnext;
                if (!is_obs || less_with_sign(t, target, dt)) break;
                ++kobs;
                tnext = obs_time(obs_t, K, kobs, sp.t1);
            }
            if constexpr (Rows) {
                r = seg + (size_t)k * FS * ss;
#pragma unroll
                for (int i = 0; i < N; ++i) r[(size_t)(1 + i) * ss] = y[i];
            } else if (k % sub == 0) {
                double *c2 = seg + (size_t)(k / sub) * FC * ss;
                c2[0] = t;
                c2[ss] = dt;
                c2[2 * ss] = (double)kobs;
#pragma unroll
                for (int i = 0; i < N; ++i) c2[(size_t)(3 + i) * ss] = y[i];
            }
            clipped = false;
            dt_keep = dt;
            if (less_with_sign(target, t + dt, dt)) {
                dt = target - t;
                clipped = is_obs;
            }
            if (!Tab::FSAL && k > 0) Sys::rhs(q, y, Kst[0]);
            tries = 0;
            new_step = false;
        }
        const double h = dt;
        R::attempt(q, y, Kst, h, ynew, xerr);
        const double err = R::error_norm(sp.atol, sp.rtol, h, y, Kst[0], xerr);
        double dt_next;
        const bool ok = control_step<Tab>(dt, err, dt_next);
        if (ok) {
            const double tn = t + dt;
            if constexpr (Rows) r[0] = stored_step(tn, t);
            t = tn;
#pragma unroll
            for (int i = 0; i < N; ++i) y[i] = ynew[i];
            if constexpr (Tab::FSAL) {
#pragma unroll
                for (int i = 0; i < N; ++i) Kst[0][i] = Kst[Tab::SF - 1][i];
            }
            if (clipped && fabs(dt_keep) > fabs(dt_next)) dt_next = dt_keep;
            dt = dt_next;
            ++k;
            new_step = true;
        } else {
            clipped = false;
            dt = dt_next;
            // (an instance whose forward ended OK never exhausts its tries here;
            // the bound only keeps the loop finite)
            if (++tries > sp.max_attempts) {
                ++k;
                new_step = true;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Checkpointed reverse in pairs of waves (one level of checkpoints, segments
// of at most VA_OBS_PAIR_MAX_ROWS rows)
// ---------------------------------------------------------------------------
// A lane's reverse is two dependent chains one after the other, segment by
// segment: the replay of the controlled loop from the segment's checkpoint and
// the backward sweep over the replayed rows.  Here wave 2p of a block replays
// while wave 2p+1 sweeps backwards, for the same 64 instances (one per lane):
// in iteration i each replay lane rebuilds one segment into the row buffer of
// parity i & 1 and records which (instance, segment) it was; the backward lane
// of the same index reverses iteration i's segment while the replay lane
// rebuilds iteration i+1's into the other buffer.  The two buffers are the row
// buffers the one-wave reverse gives the two waves' threads, so the memory is
// the same.  The waves meet through two counters in LDS (iterations replayed,
// iterations reversed), written and polled by whole waves, so no lane waits
// on another lane of its own wave; every wait gives up after VA_OBS_SPIN_LIMIT
// polls, and then each wave writes NaN outputs for the instances it holds and
// stops.  The replay runs at most two iterations ahead.  Results are those of
// the one-wave reverse: the same rows, reversed in the same order.
// A wait lasts one iteration of the other wave: one segment replayed or
// reversed, at most VA_OBS_PAIR_MAX_ROWS steps (longer segments take the
// one-wave reverse), tens of milliseconds; a wait gives up only after some
// seconds of polls.
constexpr int VA_OBS_SPIN_LIMIT = 1 << 26;
constexpr int VA_OBS_PAIR_MAX_ROWS = 8192;

struct ObsPipe {
    int item_b[2][64];  // instance of each lane in the iterations of each parity (-1: idle)
    int item_j[2][64];  // its segment (-1: none; the instance's outputs only)
    int rep;            // iterations replayed
    int con;            // iterations reversed: a buffer is free again
    int abort_;         // a wait gave up
};

__device__ __forceinline__ int pipe_load(int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void pipe_store(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Wait until *ctr > k (the same word for every lane, so the whole wave leaves
// the loop together).  Returns false when the other wave gave up or, after
// VA_OBS_SPIN_LIMIT polls, gives up itself.
__device__ __forceinline__ bool pipe_wait(ObsPipe &P, int *ctr, int k)
{
    for (int spins = 0; spins < VA_OBS_SPIN_LIMIT; ++spins) {
        if (pipe_load(ctr) > k) {
            // the other wave's buffer and item writes before its count come first
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            return true;
        }
        if (pipe_load(&P.abort_)) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    pipe_store(&P.abort_, 1);
    return false;
}

template <int N, int P>
__device__ __forceinline__ void obs_nan_outputs(const ObsRevArgs &a, int64_t b)
{
    const int64_t B = a.sp.B;
    const double qnan = __builtin_nan("");
#pragma unroll
    for (int i = 0; i < N; ++i) a.dJdy0[(int64_t)i * B + b] = qnan;
    for (int k = 0; k < P; ++k) a.dJdp[(int64_t)k * B + b] = qnan;
}

// Replay wave: takes the instances (longest first) and rebuilds their segments,
// last first.
template <class Sys, class Tab>
__device__ void obs_rev_replay(const ObsRevArgs &a, ObsPipe &Pp, int pair, int lane)
{
    constexpr int N = Sys::N, P = Sys::P, FC = N + 3;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int K = a.K;
    const int64_t ss = (int64_t)gridDim.x * blockDim.x;
    const int64_t tbase = (int64_t)blockIdx.x * blockDim.x + pair * 128 + lane;
    typename Sys::Params q;
    int64_t b = -1;
    int32_t cnt = 0, stride = 1, j = -1;
    bool done = false, have = false;
    for (int it = 0;; ++it) {
        // this lane's item: the next segment of its instance, or a new instance
        bool seg_item = false;
        int64_t ib = -1;
        if (!done && !have) {
            const int64_t pos = atomicAdd(a.counter, 1);
            if (pos >= B) {
                done = true;
            } else {
                b = a.order ? a.order[pos] : pos;
                if (a.status[b] == LANE_OK) {
                    Sys::load(q, a.p, B, b, sp.kconst);
                    cnt = a.n_acc[b];
                    stride = a.st.stride[b];
                    j = (cnt + stride - 1) / stride - 1;
                    have = j >= 0;
                }
                ib = b;  // an instance without segments: its outputs only
            }
        }
        if (have) {
            ib = b;
            seg_item = true;
        }
        if (__builtin_amdgcn_ballot_w64(!done) == 0) {
            // the final iteration, all lanes idle: the backward wave stops at it
            if (!pipe_wait(Pp, &Pp.con, it - 2)) return;
            Pp.item_b[it & 1][lane] = -1;
            Pp.item_j[it & 1][lane] = -1;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            pipe_store(&Pp.rep, it + 1);
            return;
        }
        // the buffer of this parity was read by the backward wave's iteration it - 2
        if (!pipe_wait(Pp, &Pp.con, it - 2)) {
            // NaN outputs for every instance this lane handed over or holds
            for (int d = 1; d <= 2; ++d)
                if (it - d >= 0 && Pp.item_b[(it - d) & 1][lane] >= 0)
                    obs_nan_outputs<N, P>(a, Pp.item_b[(it - d) & 1][lane]);
            if (ib >= 0) obs_nan_outputs<N, P>(a, ib);
            return;
        }
        Pp.item_b[it & 1][lane] = (int)ib;
        Pp.item_j[it & 1][lane] = seg_item ? j : -1;
        if (seg_item) {
            const double *c = a.st.ck + ((size_t)b * (size_t)a.st.slots + (size_t)j) * FC;
            const int32_t n0 = j * stride;
            const int32_t len = (cnt - n0 < stride) ? cnt - n0 : stride;
            recompute_segment<Sys, Tab>(q, sp, a.obs_t, K, c, len,
                                        a.st.seg + tbase + (it & 1) * 64, ss);
            if (--j < 0) have = false;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        pipe_store(&Pp.rep, it + 1);
    }
}

// Backward wave: reverses each replayed segment and writes the outputs.
template <class Sys, class Tab>
__device__ void obs_rev_backward(const ObsRevArgs &a, ObsPipe &Pp, int pair, int lane)
{
    constexpr int N = Sys::N, P = Sys::P, FS = N + 1;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int K = a.K;
    const int64_t ss = (int64_t)gridDim.x * blockDim.x;
    const int64_t tbase = (int64_t)blockIdx.x * blockDim.x + pair * 128 + lane;
    typename Sys::Params q;
    double lam[N], pbar[P];
    int64_t cur = -1;  // the instance in progress
    int32_t stride = 1, cnt = 0, onext = -1;
    int kobs = -1;
    const int32_t *orow = nullptr;
    // state for instance b, reversed from its end
    auto start = [&](int64_t b) {
        Sys::load(q, a.p, B, b, sp.kconst);
        double yT[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            lam[i] = 0.0;
            yT[i] = a.yT[(int64_t)i * B + b];
        }
        if (a.cost) a.J[b] = terminal_cost<N>(a.cost, a.cost_w, yT, lam);
#pragma unroll
        for (int k = 0; k < P; ++k) pbar[k] = 0.0;
        cnt = a.n_acc[b];
        stride = a.st.stride[b];
        orow = a.st.obs_row + (size_t)b * (size_t)K;
        onext = obs_point(orow, K - 1);
        kobs = obs_jumps<N>(orow, a.obs_y, K - 1, onext, cnt, yT, lam);
        cur = b;
    };
    auto finish = [&](int64_t b) {
#pragma unroll
        for (int i = 0; i < N; ++i) a.dJdy0[(int64_t)i * B + b] = lam[i];
        for (int k = 0; k < P; ++k) a.dJdp[(int64_t)k * B + b] = 0.0 + pbar[k];
        cur = -1;
    };
    for (int it = 0;; ++it) {
        if (!pipe_wait(Pp, &Pp.rep, it)) {
            if (cur >= 0) obs_nan_outputs<N, P>(a, cur);
            return;
        }
        const int ib = Pp.item_b[it & 1][lane];
        const int ij = Pp.item_j[it & 1][lane];
        if (__builtin_amdgcn_ballot_w64(ib >= 0) == 0) return;  // the replay wave is done
        if (ib >= 0) {
            if (ij < 0) {
                // no segments: a failed instance, or one with no accepted step
                if (a.status[ib] != LANE_OK) {
                    obs_nan_outputs<N, P>(a, ib);
                } else {
                    start(ib);
                    finish(ib);
                }
            } else {
                if (ib != cur) start(ib);
                const int32_t n0 = ij * stride;
                const int32_t len = (cnt - n0 < stride) ? cnt - n0 : stride;
                const double *seg = a.st.seg + tbase + (it & 1) * 64;
                // row k-1 is loaded while step k is reversed
                double hn = 0.0, yn[N] = {};
                auto load_row = [&](int32_t k) {
                    const double *r = seg + (size_t)k * FS * ss;
                    hn = r[0];
#pragma unroll
                    for (int i = 0; i < N; ++i) yn[i] = r[(size_t)(1 + i) * ss];
                };
                if (len > 0) load_row(len - 1);
                for (int32_t k = len - 1; k >= 0; --k) {
                    const double hk = hn;
                    double yk[N];
#pragma unroll
                    for (int i = 0; i < N; ++i) yk[i] = yn[i];
                    if (k > 0) load_row(k - 1);
                    back_prop_step<Sys, Tab>(q, hk, yk, lam, pbar);
                    kobs = obs_jumps<N>(orow, a.obs_y, kobs, onext, n0 + k, yk, lam);
                }
                if (ij == 0) finish(ib);
            }
        }
        // this iteration's buffer may be rebuilt
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        pipe_store(&Pp.con, it + 1);
    }
}

template <class Sys, class Tab, bool Ckpt>
__global__ void __launch_bounds__(256) k_obs_rev(ObsRevArgs a)
{
    if constexpr (Ckpt) {
        load_ctrl_tables();
        if (a.st.levels == 1 && a.st.seg_rows <= VA_OBS_PAIR_MAX_ROWS &&
            (blockDim.x & 127) == 0) {
            // pairs of waves (the condition is the same for the whole grid)
            __shared__ ObsPipe pipes[2];
            const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
            if (lane == 0 && (w & 1) == 0) {
                pipes[w >> 1].rep = 0;
                pipes[w >> 1].con = 0;
                pipes[w >> 1].abort_ = 0;
            }
            __syncthreads();
            if (w & 1) obs_rev_backward<Sys, Tab>(a, pipes[w >> 1], w >> 1, lane);
            else obs_rev_replay<Sys, Tab>(a, pipes[w >> 1], w >> 1, lane);
            return;
        }
    }
    constexpr int N = Sys::N, P = Sys::P;
    constexpr int F = N + 2, FC = N + 3, FS = N + 1;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int K = a.K;
    const double qnan = __builtin_nan("");
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double *seg = nullptr;
    // lanes' segment buffers interleaved: field f of row k at seg[(k FS + f) ss]
    const int64_t ss = (int64_t)gridDim.x * blockDim.x;
    if constexpr (Ckpt) seg = a.st.seg + gtid;

    for (;;) {
        int64_t b = atomicAdd(a.counter, 1);
        if (b >= B) break;
        if (a.order) b = a.order[b];
        if (a.status[b] != LANE_OK) {
#pragma unroll
            for (int i = 0; i < N; ++i) a.dJdy0[(int64_t)i * B + b] = qnan;
            for (int k = 0; k < P; ++k) a.dJdp[(int64_t)k * B + b] = qnan;
            continue;
        }
        typename Sys::Params q;
        Sys::load(q, a.p, B, b, sp.kconst);
        double lam[N], pbar[P], yT[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            lam[i] = 0.0;
            yT[i] = a.yT[(int64_t)i * B + b];
        }
        if (a.cost) a.J[b] = terminal_cost<N>(a.cost, a.cost_w, yT, lam);
#pragma unroll
        for (int k = 0; k < P; ++k) pbar[k] = 0.0;
        const int32_t cnt = a.n_acc[b];
        const int32_t *orow = a.st.obs_row + (size_t)b * (size_t)K;
        int32_t onext = obs_point(orow, K - 1);
        int kobs = obs_jumps<N>(orow, a.obs_y, K - 1, onext, cnt, yT, lam);
        if constexpr (!Ckpt) {
            const double *rows = a.st.traj + (size_t)b * (size_t)(sp.cap + 1) * F;
            for (int32_t n = cnt - 1; n >= 0; --n) {
                const double *r = rows + (size_t)n * F;
                back_prop_step<Sys, Tab>(q, r[1], r + 2, lam, pbar);
                kobs = obs_jumps<N>(orow, a.obs_y, kobs, onext, n, r + 2, lam);
            }
        } else {
            const int32_t stride = a.st.stride[b];
            const double *ck = a.st.ck + (size_t)b * (size_t)a.st.slots * FC;
            const int32_t nseg = (cnt + stride - 1) / stride;
            // reverse `len` rows from checkpoint c; the first is point n0
            auto reverse_rows = [&](const double *c, int32_t n0, int32_t len) {
                recompute_segment<Sys, Tab>(q, sp, a.obs_t, K, c, len, seg, ss);
                // row k-1 is loaded while step k is reversed, so no step waits
                // on its own row's loads
                double hn = 0.0, yn[N] = {};
                auto load_row = [&](int32_t k) {
                    const double *r = seg + (size_t)k * FS * ss;
                    hn = r[0];
#pragma unroll
                    for (int i = 0; i < N; ++i) yn[i] = r[(size_t)(1 + i) * ss];
                };
                if (len > 0) load_row(len - 1);
                for (int32_t k = len - 1; k >= 0; --k) {
                    const double hk = hn;
                    double yk[N];
#pragma unroll
                    for (int i = 0; i < N; ++i) yk[i] = yn[i];
                    if (k > 0) load_row(k - 1);
                    back_prop_step<Sys, Tab>(q, hk, yk, lam, pbar);
                    kobs = obs_jumps<N>(orow, a.obs_y, kobs, onext, n0 + k, yk, lam);
                }
            };
            for (int32_t j = nseg - 1; j >= 0; --j) {
                const int32_t n0 = j * stride;
                const int32_t len = (cnt - n0 < stride) ? cnt - n0 : stride;
                if (len <= a.st.seg_rows) {
                    reverse_rows(ck + (size_t)j * FC, n0, len);
                    continue;
                }
                // two levels: level-2 checkpoints every `sub` steps of the segment,
                // then each sub-segment, last first (sub <= seg_rows and
                // ceil(len / sub) <= slots2 by the sizing in va_create)
                const int32_t sub = (len + a.st.slots2 - 1) / a.st.slots2;
                double *ck2 = a.st.ck2 + gtid;
                recompute_segment<Sys, Tab, false>(q, sp, a.obs_t, K, ck + (size_t)j * FC, len,
                                                   ck2, ss, sub);
                for (int32_t i = (len - 1) / sub; i >= 0; --i) {
                    double c2[FC];
#pragma unroll
                    for (int f = 0; f < FC; ++f) c2[f] = ck2[(size_t)(i * FC + f) * ss];
                    const int32_t m0 = i * sub;
                    const int32_t leni = (len - m0 < sub) ? len - m0 : sub;
                    reverse_rows(c2, n0 + m0, leni);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < N; ++i) a.dJdy0[(int64_t)i * B + b] = lam[i];
        for (int k = 0; k < P; ++k) a.dJdp[(int64_t)k * B + b] = 0.0 + pbar[k];
    }
}

template <class Sys, class Tab>
struct LaunchObs {
    static hipError_t fwd(const ObsFwdArgs &a, int grid, int block, hipStream_t s)
    {
        if (a.st.ckpt)
            hipLaunchKernelGGL((k_obs_fwd<Sys, Tab, true>), dim3(grid), dim3(block), 0, s, a);
        else
            hipLaunchKernelGGL((k_obs_fwd<Sys, Tab, false>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    static hipError_t rev(const ObsRevArgs &a, int grid, int block, hipStream_t s)
    {
        if (a.st.ckpt)
            hipLaunchKernelGGL((k_obs_rev<Sys, Tab, true>), dim3(grid), dim3(block), 0, s, a);
        else
            hipLaunchKernelGGL((k_obs_rev<Sys, Tab, false>), dim3(grid), dim3(block), 0, s, a);
        return hipGetLastError();
    }
    // the smaller occupancy of the two storage variants, so one sizing serves both
    static int occ_fwd(int block)
    {
        int n0 = 0, n1 = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n0, k_obs_fwd<Sys, Tab, false>, block,
                                                         0) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&n1, k_obs_fwd<Sys, Tab, true>, block,
                                                         0) != hipSuccess)
            return 1;
        return n0 < n1 ? n0 : n1;
    }
    static int occ_rev(int block)
    {
        int n0 = 0, n1 = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n0, k_obs_rev<Sys, Tab, false>, block,
                                                         0) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&n1, k_obs_rev<Sys, Tab, true>, block,
                                                         0) != hipSuccess)
            return 1;
        return n0 < n1 ? n0 : n1;
    }
    static void apply(KernelSet *k)
    {
        k->obs_fwd = &fwd;
        k->obs_rev = &rev;
        k->occ_obs_fwd = &occ_fwd;
        k->occ_obs_rev = &occ_rev;
    }
};

// Adds the observation kernels for the adaptive error steppers.
template <class Sys>
inline void pick_observed(int tableau, int adaptive, KernelSet *out)
{
    if (!adaptive) return;
    switch (tableau) {
    case TAB_CK54: LaunchObs<Sys, CashKarp54>::apply(out); break;
    case TAB_DOPRI5: LaunchObs<Sys, Dopri5>::apply(out); break;
    case TAB_F78: LaunchObs<Sys, Fehlberg78>::apply(out); break;
    default: break;
    }
}

}  // namespace va
