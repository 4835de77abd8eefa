// va_glv_shared.hip -- generalized Lotka-Volterra with an interaction matrix A shared
// by every instance (N = 128): stage right-hand sides and adjoints as FP64 MFMA
// products over tiles of 16 instances, gfx950.
//
// System: examples/GeneralizedLotkaVolterra/main.cpp:105-119 with A common to
// the batch and r per instance: f_i = x_i g_i, g_i = r_i + sum_j A_ij x_j.
// Forward: the controlled loop of lib/include/detail/runge_kutta.hpp:74-118 per
// instance; reverse: detail/backpropagation.hpp:83-158 per instance, with
// dJ/dA = sum over instances, steps and stages of u x^T (u = w * x, w the stage
// adjoint seed) reduced over the batch.
//
// Work decomposition.  A workgroup of 8 waves owns one tile of 16 instances at a
// time; wave w owns state components [16w, 16w + 16).  The 16 instances of a
// tile step in lockstep (one attempt per iteration, each with its own t, h and
// accept/reject decision); instances that finish idle until the tile is done
// (static tiles, grouped longest-first for the reverse, so little is idle).
// Per stage, every wave writes its components of the stage vectors to LDS and
// computes its 16x16 block of
//     G = X A^T        (growth, forward and recompute)   32 MFMAs 16x16x4
//     xi = U A         (adjoint of the growth)           32 MFMAs
//     Abar += U^T X    (shared-parameter gradient)       32 MFMAs, K = instances
// A stays in LDS for the whole launch (128 KB, one workgroup per CU).
//
// MFMA 16x16x4 f64 lane layout (checked on gfx950): operand A lane l holds
// A[l%16][l/16], operand B lane l holds B[l/16][l%16]; result element jj of
// lane l is D[l/16 + 4 jj][l%16].  So lane l = 16 r + c holds, for every state
// vector, instances r, r+4, r+8, r+12 of the tile at component 16 w + c.
//
// LDS rows are 128 doubles with the column index XOR-ed with the row's low
// four bits, which makes both the row-wise and the column-wise fragment reads
// free of bank conflicts.
#include "va_kernels_impl.hpp"

namespace va {

namespace {

constexpr int GS_N = 128;
constexpr int GS_T = 16;      // instances per tile
constexpr int GS_WAVES = 8;
constexpr int GS_THREADS = 64 * GS_WAVES;
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int sw(int row, int col) { return row * GS_N + (col ^ (row & 15)); }

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// G[inst][16w + c] = sum_k X[inst][k] A[16w + c][k] for the tile (X staged in Xs).
// The forward and the reverse's recompute call this same sequence, so they
// produce the same bits (U, the unroll factor, only sets how far operand loads
// run ahead: the reverse, short of registers, uses 4).
template <int U = 8>
__device__ __forceinline__ d4 gemm_xat(const double *Xs, const double *As, int w, int r, int c)
{
    d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll U
    for (int kb = 0; kb < GS_N / 4; ++kb) {
        const int k = 4 * kb + r;
        acc = mfma(Xs[sw(c, k)], As[sw(16 * w + c, k)], acc);
    }
    return acc;
}

// xi[inst][16w + c] = sum_k U[inst][k] A[k][16w + c].
__device__ __forceinline__ d4 gemm_ua(const double *Us, const double *As, int w, int r, int c)
{
    d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
    for (int kb = 0; kb < GS_N / 4; ++kb) {
        const int k = 4 * kb + r;
        acc = mfma(Us[sw(c, k)], As[sw(k, 16 * w + c)], acc);
    }
    return acc;
}

// Abar[16w + r + 4jj][16jb + c] += sum_inst U[inst][16w + .] X[inst][16jb + .]
__device__ __forceinline__ void gemm_utx(const double *Us, const double *Xs, int w, int r, int c,
                                         d4 (&abar)[GS_N / 16])
{
    // one instance row at a time: the 8 accumulators are live across the
    // call, so operands are not loaded further ahead
#pragma unroll 1
    for (int kb = 0; kb < GS_T / 4; ++kb) {
        const int inst = 4 * kb + r;
        const double a = Us[sw(inst, 16 * w + c)];
#pragma unroll
        for (int jb = 0; jb < GS_N / 16; ++jb) abar[jb] = mfma(a, Xs[sw(inst, 16 * jb + c)], abar[jb]);
    }
}

__device__ __forceinline__ void load_A(double *As, const double *A)
{
    for (int idx = threadIdx.x; idx < GS_N * GS_N; idx += GS_THREADS)
        As[sw(idx / GS_N, idx % GS_N)] = A[idx];
}

// max over the 128 components of each instance of the tile: within the wave over
// the 16 lanes of a row group, then over the 8 waves through LDS.
__device__ __forceinline__ void tile_max(double (&v)[4], double *red, int w, int r, int c)
{
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        double m = v[jj];
#pragma unroll
        for (int o = 8; o >= 1; o /= 2) m = fmax(m, __shfl_xor(m, o, 64));
        v[jj] = m;
    }
    if (c == 0) {
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) red[w * GS_T + r + 4 * jj] = v[jj];
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        double m = 0.0;
        for (int ww = 0; ww < GS_WAVES; ++ww) m = fmax(m, red[ww * GS_T + r + 4 * jj]);
        v[jj] = m;
    }
    __syncthreads();
}

// Stage input X_m = y + sum_{j<m} (a_mj h) K_j of instance slot jj, in the order
// the forward uses.
template <class Tab, int S>
__device__ __forceinline__ double stage_x(int m, double h, double y, const double (&K)[S][4],
                                          int jj)
{
    double x = y;
    for (int j = 0; j < m; ++j)
        if (Tab::A[m][j] != 0.0) x = x + (Tab::A[m][j] * h) * K[j][jj];
    return x;
}

}  // namespace

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------
template <class Tab>
__global__ void __launch_bounds__(GS_THREADS) k_glvs_fwd(SharedFwdArgs a)
{
    __shared__ double As[GS_N * GS_N];
    __shared__ double Xs[GS_T * GS_N];
    __shared__ double red[GS_WAVES * GS_T];
    constexpr int F = GS_N + 2;
    constexpr bool kSpec = !Tab::FSAL;  // f(y_{n+1}) computed with the attempt
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane >> 4, c = lane & 15, comp = 16 * w + c;
    load_A(As, a.A);
    __syncthreads();
    const int64_t ntiles = (B + GS_T - 1) / GS_T;

    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int64_t b[4];
        bool act[4], new_step[4];
        double y[4], rr[4], K[Tab::SF][4], t[4], dt[4];
        int32_t count[4], rej[4], tries[4], st[4];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            b[jj] = tile * GS_T + r + 4 * jj;
            act[jj] = b[jj] < B;
            const int64_t bb = act[jj] ? b[jj] : 0;
            y[jj] = act[jj] ? a.y[(int64_t)comp * B + bb] : 0.0;
            rr[jj] = act[jj] ? a.r[(int64_t)comp * B + bb] : 0.0;
            t[jj] = sp.t0;
            dt[jj] = sp.dt0;
            count[jj] = rej[jj] = tries[jj] = 0;
            st[jj] = LANE_OK;
            new_step[jj] = true;
        }
        // f(y0) for the whole tile
        {
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) Xs[sw(r + 4 * jj, comp)] = y[jj];
            __syncthreads();
            const d4 g = gemm_xat(Xs, As, w, r, c);
            __syncthreads();
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) K[0][jj] = y[jj] * (rr[jj] + g[jj]);
        }
        for (;;) {
            // step start of the instances that accepted (or just started)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                if (!act[jj] || !new_step[jj]) continue;
                double *rows = a.traj + (size_t)b[jj] * (size_t)(sp.cap + 1) * F;
                if (!less_with_sign(t[jj], sp.t1, dt[jj]) || count[jj] >= sp.cap) {
                    if (less_with_sign(t[jj], sp.t1, dt[jj])) {
                        st[jj] = LANE_MAX_STEPS;
                    } else {
                        double *row = rows + (size_t)count[jj] * F;
                        if (w == 0 && c == 0) {
                            row[0] = t[jj];
                            row[1] = 0.0;
                        }
                        row[2 + comp] = y[jj];
                    }
                    act[jj] = false;
                    continue;
                }
                double *row = rows + (size_t)count[jj] * F;
                if (w == 0 && c == 0) row[0] = t[jj];
                row[2 + comp] = y[jj];
                if (less_with_sign(sp.t1, t[jj] + dt[jj], dt[jj])) dt[jj] = sp.t1 - t[jj];
                tries[jj] = 0;
                new_step[jj] = false;
            }
            // every wave holds all 16 instances (rows r, jj), so a ballot is tile-uniform
            const bool any = __builtin_amdgcn_ballot_w64(act[0] || act[1] || act[2] || act[3]) != 0;
            if (!any) break;

            double X[4];
            for (int m = 1; m < Tab::SF; ++m) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    double x = y[jj];
                    for (int j = 0; j < m; ++j)
                        if (Tab::A[m][j] != 0.0) x = x + (Tab::A[m][j] * dt[jj]) * K[j][jj];
                    X[jj] = x;
                    Xs[sw(r + 4 * jj, comp)] = x;
                }
                __syncthreads();
                const d4 g = gemm_xat(Xs, As, w, r, c);
                __syncthreads();
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) K[m][jj] = X[jj] * (rr[jj] + g[jj]);
            }
            double ynew[4], ratio[4];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const double h = dt[jj];
                constexpr int j0 = ErrFirst<Tab>::value();
                double yn = y[jj], xe = (Tab::DB[j0] * h) * K[j0][jj];
                for (int j = 0; j < Tab::S; ++j)
                    if (Tab::B[j] != 0.0) yn = yn + (Tab::B[j] * h) * K[j][jj];
                for (int j = j0 + 1; j < Tab::SF; ++j)
                    if (Tab::DB[j] != 0.0) xe = xe + (Tab::DB[j] * h) * K[j][jj];
                ynew[jj] = yn;
                const double den = sp.atol + sp.rtol * (fabs(y[jj]) + h * fabs(K[0][jj]));
                ratio[jj] = fabs(xe) / den;
            }
            tile_max(ratio, red, w, r, c);
            double k0n[4] = {0.0, 0.0, 0.0, 0.0};
            if constexpr (kSpec) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) Xs[sw(r + 4 * jj, comp)] = ynew[jj];
                __syncthreads();
                const d4 g = gemm_xat(Xs, As, w, r, c);
                __syncthreads();
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) k0n[jj] = ynew[jj] * (rr[jj] + g[jj]);
            }
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                if (!act[jj]) continue;
                double dt_next;
                // the tile buffers and A fill the LDS: the pow tables stay in memory
                if (!control_step<Tab, powref::GlobalTables>(dt[jj], ratio[jj], dt_next)) {
                    ++rej[jj];
                } else {
                    const double tn = t[jj] + dt[jj];
                    if (w == 0 && c == 0)
                        a.traj[(size_t)b[jj] * (size_t)(sp.cap + 1) * F + (size_t)count[jj] * F +
                               1] = stored_step(tn, t[jj]);
                    t[jj] = tn;
                    y[jj] = ynew[jj];
                    K[0][jj] = Tab::FSAL ? K[Tab::SF - 1][jj] : k0n[jj];
                    ++count[jj];
                    new_step[jj] = true;
                }
                dt[jj] = dt_next;
                if (++tries[jj] > sp.max_attempts) {
                    st[jj] = LANE_MAX_FAIL;
                    act[jj] = false;
                }
            }
        }
        // finite check of y(T) over all components, then the per-instance outputs
        double bad[4];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) bad[jj] = (fabs(y[jj]) <= 1.7976931348623157e308) ? 0.0 : 1.0;
        tile_max(bad, red, w, r, c);
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            if (b[jj] >= B) continue;
            a.y[(int64_t)comp * B + b[jj]] = y[jj];
            if (st[jj] == LANE_OK && bad[jj] != 0.0) st[jj] = LANE_NONFINITE;
            if (w == 0 && c == 0) {
                a.n_acc[b[jj]] = count[jj];
                a.n_rej[b[jj]] = rej[jj];
                a.status[b[jj]] = st[jj];
                atomicMax(a.max_steps_seen, count[jj]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Reverse
// ---------------------------------------------------------------------------
template <class Tab>
__global__ void __launch_bounds__(GS_THREADS) k_glvs_rev(SharedRevArgs a)
{
    __shared__ double As[GS_N * GS_N];
    __shared__ double Xs[GS_T * GS_N];
    __shared__ double Us[GS_T * GS_N];
    constexpr int F = GS_N + 2, S = Tab::S;
    const SolveParams sp = a.sp;
    const int64_t B = sp.B;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane >> 4, c = lane & 15, comp = 16 * w + c;
    const double qnan = __builtin_nan("");
    load_A(As, a.A);
    __syncthreads();
    d4 abar[GS_N / 16];
#pragma unroll
    for (int jb = 0; jb < GS_N / 16; ++jb) abar[jb] = d4{0.0, 0.0, 0.0, 0.0};
    const int64_t ntiles = (B + GS_T - 1) / GS_T;
    // growth products and stage inputs of the S stages of the step being reversed,
    // [2][S][64 lanes] per wave: the backward reads them back instead of keeping
    // the stage values live in registers, which the dJ/dA accumulators fill
    d4 *gw = reinterpret_cast<d4 *>(a.gscr) + ((size_t)blockIdx.x * (GS_THREADS / 64) + w) * 2 * S * 64;
    d4 *xw = gw + S * 64;

    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int64_t b[4];
        bool ok[4];
        double lam[4], rbar[4], rr[4];
        int32_t n[4];
        int32_t nmax = 0;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const int64_t k = tile * GS_T + r + 4 * jj;
            b[jj] = k < B ? (a.order ? (int64_t)a.order[k] : k) : -1;
            ok[jj] = b[jj] >= 0 && a.status[b[jj]] == LANE_OK;
            const int64_t bb = ok[jj] ? b[jj] : 0;
            lam[jj] = ok[jj] ? a.lambda[(int64_t)comp * B + bb] : 0.0;
            rr[jj] = ok[jj] ? a.r[(int64_t)comp * B + bb] : 0.0;
            rbar[jj] = 0.0;
            n[jj] = ok[jj] ? a.n_acc[bb] : 0;
            nmax = n[jj] > nmax ? n[jj] : nmax;
        }
        // tile-uniform step count (all rows r of the wave hold all 16 instances)
        nmax = max(nmax, __shfl_xor(nmax, 16, 64));
        nmax = max(nmax, __shfl_xor(nmax, 32, 64));

        for (int32_t it = 0; it < nmax; ++it) {
            double h[4], yn[4], K[S][4];
            bool on[4];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int32_t step = n[jj] - 1 - it;
                on[jj] = ok[jj] && step >= 0;
                h[jj] = 0.0;
                yn[jj] = 0.0;
                if (on[jj]) {
                    const double *row = a.traj + (size_t)b[jj] * (size_t)(sp.cap + 1) * F +
                                        (size_t)step * F;
                    h[jj] = row[1];
                    yn[jj] = row[2 + comp];
                }
            }
            // recompute the stages from y_n (backpropagation.hpp:24-64); the stage
            // inputs X_m and growth products A X_m go to this wave's scratch rows,
            // where the adjoint reads them back
            for (int m = 0; m < S; ++m) {
                double X[4];
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    X[jj] = stage_x<Tab>(m, h[jj], yn[jj], K, jj);
                    Xs[sw(r + 4 * jj, comp)] = X[jj];
                }
                __syncthreads();
                const d4 G = gemm_xat<4>(Xs, As, w, r, c);
                __syncthreads();
                gw[m * 64 + lane] = G;
                xw[m * 64 + lane] = d4{X[0], X[1], X[2], X[3]};
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) K[m][jj] = X[jj] * (rr[jj] + G[jj]);
            }
            // adjoint of the step (:104-153)
            double Kb[S][4], ybar[4];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                ybar[jj] = lam[jj];
                for (int m = 0; m < S; ++m) Kb[m][jj] = (Tab::B[m] * h[jj]) * lam[jj];
            }
            for (int m = S - 1; m >= 0; --m) {
                const d4 G = gw[m * 64 + lane];   // this lane's own growth products of stage m
                const d4 Xm = xw[m * 64 + lane];  // and its stage inputs
                double u[4];
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const double X = Xm[jj];
                    u[jj] = on[jj] ? Kb[m][jj] * X : 0.0;
                    Us[sw(r + 4 * jj, comp)] = u[jj];
                    Xs[sw(r + 4 * jj, comp)] = X;
                }
                __syncthreads();
                const d4 xu = gemm_ua(Us, As, w, r, c);
                gemm_utx(Us, Xs, w, r, c, abar);
                __syncthreads();
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    if (!on[jj]) continue;
                    const double xi = Kb[m][jj] * (rr[jj] + G[jj]) + xu[jj];
                    rbar[jj] = rbar[jj] + u[jj];
                    ybar[jj] = ybar[jj] + xi;
                    for (int j = 0; j < m; ++j)
                        if (Tab::A[m][j] != 0.0)
                            Kb[j][jj] = Kb[j][jj] + (xi * Tab::A[m][j]) * h[jj];
                }
            }
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
                if (on[jj]) lam[jj] = ybar[jj];
        }
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            if (b[jj] < 0) continue;
            a.dJdy0[(int64_t)comp * B + b[jj]] = ok[jj] ? lam[jj] : qnan;
            a.dJdr[(int64_t)comp * B + b[jj]] = ok[jj] ? 0.0 + rbar[jj] : qnan;
        }
    }
    // this workgroup's part of dJ/dA
    double *part = a.abar_part + (size_t)blockIdx.x * GS_N * GS_N;
#pragma unroll
    for (int jb = 0; jb < GS_N / 16; ++jb)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
            part[(size_t)(16 * w + r + 4 * jj) * GS_N + 16 * jb + c] = abar[jb][jj];
}

// dJdA[k] = sum over workgroups of the partial sums, in workgroup order.
__global__ void k_glvs_reduce(const double *part, int nparts, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= GS_N * GS_N) return;
    double s = 0.0;
    for (int g = 0; g < nparts; ++g) s = s + part[(size_t)g * GS_N * GS_N + k];
    out[k] = s;
}

template <class Tab>
struct LaunchShared {
    static hipError_t fwd(const SharedFwdArgs &a, int grid, hipStream_t s)
    {
        hipLaunchKernelGGL((k_glvs_fwd<Tab>), dim3(grid), dim3(GS_THREADS), 0, s, a);
        return hipGetLastError();
    }
    static hipError_t rev(const SharedRevArgs &a, int grid, hipStream_t s)
    {
        hipLaunchKernelGGL((k_glvs_rev<Tab>), dim3(grid), dim3(GS_THREADS), 0, s, a);
        return hipGetLastError();
    }
};

hipError_t launch_glvs_reduce(const double *part, int nparts, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_glvs_reduce, dim3((GS_N * GS_N + 255) / 256), dim3(256), 0, s, part,
                       nparts, out);
    return hipGetLastError();
}

bool kernels_lotka_volterra_shared(int n, int tableau, int adaptive, KernelSet *out)
{
    if (n != GS_N || !adaptive) return false;
    KernelSet k{};
    k.n = GS_N;
    k.np = GS_N;  // per-instance parameters: r; A is passed separately
    k.s_rev = 6;
    k.flop_rhs = 2 * GS_N * GS_N + GS_N;
    k.flop_vjp = 4 * GS_N * GS_N + 3 * GS_N;
    k.layout = 2;
    k.shared_tile = GS_T;
    k.shared_waves = GS_WAVES;
    switch (tableau) {
    case TAB_CK54:
        k.shared_fwd = &LaunchShared<CashKarp54>::fwd;
        k.shared_rev = &LaunchShared<CashKarp54>::rev;
        k.s_fwd = CashKarp54::SF;
        k.s_rev = CashKarp54::S;
        break;
    case TAB_DOPRI5:
        k.shared_fwd = &LaunchShared<Dopri5>::fwd;
        k.shared_rev = &LaunchShared<Dopri5>::rev;
        k.s_fwd = Dopri5::SF;
        k.s_rev = Dopri5::S;
        break;
    default: return false;
    }
    *out = k;
    return true;
}

}  // namespace va
