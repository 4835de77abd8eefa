// va_capi.cpp -- implementation of the C ABI declared in include/vectorizedadjoint_amd.h.
//
// A handle mirrors the reference's Driver (lib/include/Driver.hpp:15-79): it
// owns the trajectory store and the stepper/system choice; the caller owns
// every input and output buffer.  All work is stream-ordered on the caller's
// stream; nothing here synchronises except va_max_steps_seen and the plain
// form's completion polling.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <new>
#include <vector>

#include "../../include/vectorizedadjoint_amd.h"
#include "va_kernels.hpp"

namespace {

thread_local char g_err[512] = "";

va_status fail(va_status s, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return s;
}

#define VA_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(VA_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_),    \
                        __FILE__, __LINE__);                                           \
    } while (0)

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double *p, double v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy_f64(double *dst, const double *src, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_iota_i32(int32_t *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

// The per-instance counts and status to the caller's arrays (any may be null).
__global__ void k_copy_counts(int32_t *n_acc, int32_t *n_rej, int32_t *status,
                              const int32_t *s_acc, const int32_t *s_rej, const int32_t *s_st,
                              int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (n_acc) n_acc[i] = s_acc[i];
    if (n_rej) n_rej[i] = s_rej[i];
    if (status) status[i] = s_st[i];
}

__global__ void k_identity_seeds(double *lambda, int n, int64_t B)
{
    // lambda[o][i][b] = (o == i)
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)n * n * B;
    if (idx >= total) return;
    const int64_t oi = idx / B;
    const int o = (int)(oi / n), i = (int)(oi % n);
    lambda[idx] = (o == i) ? 1.0 : 0.0;
}

// Plain-form rows are [row][field][instance]; gather (t, h, y) of one instance.
__global__ void k_gather_rows(const double *ptraj, int64_t b, int64_t B, int fields, int n,
                              int64_t nrows, double *out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = nrows * (n + 2);
    if (idx >= total) return;
    const int64_t row = idx / (n + 2);
    const int f = (int)(idx % (n + 2));
    out[idx] = ptraj[((size_t)row * fields + f) * (size_t)B + b];
}

// Instance-major copies for the Lotka-Volterra row kernels.  The persistent
// kernels take instances in a data-dependent order, so reading [P][B]
// parameters fetches a whole cache line for each 8-byte value; an instance's
// P values read from one contiguous block of 8P bytes instead.  A block of 256
// threads transposes 32 instances through LDS, 128 parameters per pass: on the
// [P][B] side each wave moves two 256-byte runs of 32 instances per
// instruction, on the [B][P] side each wave moves 64 consecutive parameters of
// one instance (an instance's parameters, and the block's 32 instances, are
// contiguous there).  The row stride of 129 doubles keeps the 32 instances of a
// parameter in distinct banks.
constexpr int VA_TILE_B = 32;   // instances per block
constexpr int VA_TILE_K = 128;  // parameters per pass: 32 x 129 doubles of LDS

__device__ __forceinline__ void to_instance_major_tile(const double *src, double *dst, int64_t P,
                                                       int64_t B,
                                                       double (&tile)[VA_TILE_B][VA_TILE_K + 1])
{
    const int64_t b0 = (int64_t)blockIdx.x * VA_TILE_B;
    const int nb = (int)((B - b0 < VA_TILE_B) ? B - b0 : VA_TILE_B);
    const int bb = threadIdx.x & 31, kr = threadIdx.x >> 5;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t k0 = 0; k0 < P; k0 += VA_TILE_K) {
        const int kn = (int)((P - k0 < VA_TILE_K) ? P - k0 : VA_TILE_K);
        if (bb < nb)
            for (int k = kr; k < kn; k += 8) tile[bb][k] = src[(k0 + k) * B + b0 + bb];
        __syncthreads();
        for (int c = wv; c < nb; c += 4)
            for (int k = lane; k < kn; k += 64) dst[(b0 + c) * P + k0 + k] = tile[c][k];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_to_instance_major(const double *src, double *dst,
                                                           int64_t P, int64_t B)
{
    __shared__ double tile[VA_TILE_B][VA_TILE_K + 1];
    to_instance_major_tile(src, dst, P, B, tile);
}

// A split solve's prologue in one launch: the parameter copy above, y(T) := y0
// (when they differ) and cleared marks for the block's instances, and, in block
// 0, the cleared work counters [0, 4).
__global__ void __launch_bounds__(256) k_split_prologue(const double *p, double *pim, int64_t P,
                                                        int64_t B, const double *y0, double *yT,
                                                        int n, int32_t *marks, int32_t *counters)
{
    __shared__ double tile[VA_TILE_B][VA_TILE_K + 1];
    to_instance_major_tile(p, pim, P, B, tile);
    const int64_t b0 = (int64_t)blockIdx.x * VA_TILE_B;
    const int nb = (int)((B - b0 < VA_TILE_B) ? B - b0 : VA_TILE_B);
    const int bb = threadIdx.x & 31;
    if (y0 != yT && bb < nb)
        for (int i = threadIdx.x >> 5; i < n; i += 8) yT[(int64_t)i * B + b0 + bb] = y0[(int64_t)i * B + b0 + bb];
    if ((int)threadIdx.x < nb) marks[b0 + threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x < 4) counters[threadIdx.x] = 0;
}

// dJ/dp sums of the row kernels, instance-major [nout][B][P], into mu
// [nout][P][B]: 0.0 + sum with a cost, mu + sum without one (the updates the
// kernels make when they write mu themselves); with set, only the instances
// marked there (the others' mu is the long-instance kernel's).  blockIdx.y is
// the seed.
__global__ void __launch_bounds__(256) k_from_instance_major(const double *src, double *dst,
                                                             int64_t P, int64_t B, int add,
                                                             const int32_t *set)
{
    __shared__ double tile[VA_TILE_B][VA_TILE_K + 1];
    const int64_t b0 = (int64_t)blockIdx.x * VA_TILE_B;
    const int nb = (int)((B - b0 < VA_TILE_B) ? B - b0 : VA_TILE_B);
    const int bb = threadIdx.x & 31, kr = threadIdx.x >> 5;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double *sp = src + (int64_t)blockIdx.y * B * P;
    double *dp = dst + (int64_t)blockIdx.y * P * B;
    const bool mine = bb < nb && (!set || set[b0 + bb]);
    for (int64_t k0 = 0; k0 < P; k0 += VA_TILE_K) {
        const int kn = (int)((P - k0 < VA_TILE_K) ? P - k0 : VA_TILE_K);
        for (int c = wv; c < nb; c += 4)
            for (int k = lane; k < kn; k += 64) tile[c][k] = sp[(b0 + c) * P + k0 + k];
        __syncthreads();
        if (mine) {
            for (int k = kr; k < kn; k += 8) {
                double *m = dp + (k0 + k) * B + b0 + bb;
                const double x = tile[bb][k];
                *m = add ? *m + x : 0.0 + x;
            }
        }
        __syncthreads();
    }
}

// Longest-first orders.  The reverse of one instance is a chain of n_acc
// dependent steps, its forward a chain of n_acc + n_rej attempts; when the long
// chains start last they set the end of the launch, and with one lane per
// instance a wave runs as long as its longest instance.  Instances are bucketed
// by the chain length (buckets of 8 below 1024, of 64 above, the last
// open-ended) and taken bucket by bucket from the longest; within a bucket the
// order is arbitrary.  Reverse order (mode 0): accepted steps, failed instances
// last (they only write NaN).  Forward order (mode 1): attempts, every status.
// Each instance's result does not depend on the order.
constexpr int VA_ORDER_BUCKETS = 256;  // = the block size of the order kernels

__device__ __forceinline__ int order_bucket(const int32_t *n_acc, const int32_t *n_rej,
                                            const int32_t *status, int64_t b, int mode)
{
    int32_t len = n_acc[b];
    if (mode == 0 && status[b] != 0) return 0;
    if (mode == 1) len += n_rej[b];
    const int k = len < 1024 ? (len >> 3) : 128 + ((len - 1024) >> 6);
    return k < VA_ORDER_BUCKETS - 1 ? k : VA_ORDER_BUCKETS - 1;
}

__global__ void __launch_bounds__(256) k_order_hist(const int32_t *n_acc, const int32_t *n_rej,
                                                    const int32_t *status, int64_t B, int mode,
                                                    int32_t *hist)
{
    __shared__ int32_t cnt[VA_ORDER_BUCKETS];
    if (threadIdx.x < VA_ORDER_BUCKETS) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) atomicAdd(&cnt[order_bucket(n_acc, n_rej, status, b, mode)], 1);
    __syncthreads();
    if (threadIdx.x < VA_ORDER_BUCKETS && cnt[threadIdx.x])
        atomicAdd(&hist[threadIdx.x], cnt[threadIdx.x]);
}

// cursor[k] = number of instances in buckets above k (longest bucket first).
// With n_long: the instances in buckets >= VA_LONG_BUCKET, at most VA_LONG_MAX
// (the leading positions of the order that the long-instance kernel takes).
constexpr int VA_LONG_BUCKET = 32;  // 256 attempts and more
constexpr int VA_LONG_MAX = 256;    // one wave each

__global__ void __launch_bounds__(VA_ORDER_BUCKETS)
k_order_offsets(const int32_t *hist, int32_t *cursor, int32_t *n_long)
{
    // one thread per bucket: suffix sums by doubling steps
    __shared__ int32_t sfx[VA_ORDER_BUCKETS];
    const int k = threadIdx.x;
    const int32_t own = hist[k];
    sfx[k] = own;
    __syncthreads();
    for (int d = 1; d < VA_ORDER_BUCKETS; d *= 2) {
        const int32_t v = k + d < VA_ORDER_BUCKETS ? sfx[k + d] : 0;
        __syncthreads();
        sfx[k] += v;
        __syncthreads();
    }
    cursor[k] = sfx[k] - own;
    if (k == VA_LONG_BUCKET && n_long) *n_long = sfx[k] < VA_LONG_MAX ? sfx[k] : VA_LONG_MAX;
}

__global__ void __launch_bounds__(256) k_order_scatter(const int32_t *n_acc,
                                                       const int32_t *n_rej,
                                                       const int32_t *status, int64_t B,
                                                       int mode, int32_t *cursor, int32_t *order)
{
    __shared__ int32_t cnt[VA_ORDER_BUCKETS], base[VA_ORDER_BUCKETS];
    if (threadIdx.x < VA_ORDER_BUCKETS) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int k = 0, r = 0;
    if (b < B) {
        k = order_bucket(n_acc, n_rej, status, b, mode);
        r = atomicAdd(&cnt[k], 1);
    }
    __syncthreads();
    if (threadIdx.x < VA_ORDER_BUCKETS && cnt[threadIdx.x])
        base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], cnt[threadIdx.x]);
    __syncthreads();
    if (b < B) order[base[k] + r] = (int32_t)b;
}

// Digest of each instance's stored step sequence: d = 0x9E3779B97F4A7C15, then for
// every stored point n = 0..n_acc: d = mix(d ^ bits(t_n)); d = mix(d ^ bits(h_n)),
// mix the SplitMix64 finaliser (the oracle's orc_step_digest).  0 for instances
// whose forward did not end OK.  plain: rows [row][field][instance], else
// [instance][row][field] with `fields` doubles per row.
__device__ __forceinline__ uint64_t digest_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) k_step_digest(const double *traj, int64_t B, int64_t cap,
                                                     int fields, int plain, const int32_t *n_acc,
                                                     const int32_t *status, uint64_t *out)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    uint64_t d = 0;
    if (status[b] == VA_LANE_OK) {
        d = 0x9E3779B97F4A7C15ull;
        const int64_t T = (int64_t)n_acc[b] + 1;  // <= cap + 1 rows for an OK instance
        for (int64_t k = 0; k < T && k <= cap; ++k) {
            double t, hs;
            if (plain) {
                t = traj[((size_t)k * fields) * (size_t)B + b];
                hs = traj[((size_t)k * fields + 1) * (size_t)B + b];
            } else {
                const double *r = traj + ((size_t)b * (size_t)(cap + 1) + (size_t)k) * fields;
                t = r[0];
                hs = r[1];
            }
            d = digest_mix(d ^ __builtin_bit_cast(uint64_t, t));
            d = digest_mix(d ^ __builtin_bit_cast(uint64_t, hs));
        }
    }
    out[b] = d;
}

inline int blocks_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

}  // namespace

struct va_handle {
    va_problem pb;
    va::KernelSet ks;
    int device = 0;
    int cu_count = 0;
    int block = 256;
    int64_t lanes_fwd = 0, lanes_rev = 0;  // instances in flight of the persistent kernels
    // persistent store
    double *traj = nullptr;
    // plain store + state
    double *ptraj = nullptr, *pl_t = nullptr, *pl_dt = nullptr, *pl_k0 = nullptr,
           *pl_pbar = nullptr;
    int32_t *pl_done = nullptr;
    // bookkeeping [B] and counters
    int32_t *n_acc = nullptr, *n_rej = nullptr, *status = nullptr;
    // [0] fwd queue, [1] rev queue, [2] max steps, [3] active, [nlong_idx] long
    // instances at the head of order_fwd (from the previous forward's attempts),
    // [11 - nlong_idx] those of order_fwd_next
    int32_t *counters = nullptr;
    int32_t *order = nullptr;     // [B] reverse order, longest first
    int32_t *order_fwd = nullptr; // [B] next forward's order (previous attempts, longest first)
    // [B] the order a split solve builds for the solve after it while its own
    // reverse still reads order_fwd; the two are swapped when it returns
    int32_t *order_fwd_next = nullptr;
    int nlong_idx = 5;
    int32_t *order_tmp = nullptr; // [2 * VA_ORDER_BUCKETS] histogram, cursors
    // observed solves
    int64_t lanes_obs_fwd = 0, lanes_obs_rev = 0;
    int32_t obs_ckpt = 0, obs_slots = 0, obs_stride_max = 0;
    int32_t obs_levels = 1, obs_slots2 = 0, obs_seg_rows = 0;
    double *obs_ck = nullptr, *obs_seg = nullptr, *obs_ck2 = nullptr;
    int32_t *obs_stride = nullptr, *obs_row = nullptr;
    // shared-parameter solves
    int shared_grid = 0;
    double *abar_part = nullptr;
    // Lotka-Volterra row layout, persistent: instance-major parameters [B][P]
    // and dJ/dp sums [pbar_nout][B][P]
    double *pim = nullptr, *pbar_im = nullptr;
    int pbar_nout = 0;
    int32_t *pbar_set = nullptr;  // [B]: instances whose sums are in pbar_im
    double *shared_gscr = nullptr;  // per-workgroup growth products of the shared-A reverse
    // VA_FLAG_GRAPH: va_solve recorded once per argument set, then replayed
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    const void *graph_key[12] = {};
    int32_t graph_cost = 0;
    int32_t *host_counter = nullptr;
    bool have_forward = false;
    // order_fwd holds the longest-first order of an earlier forward (and
    // counters[5] its count of long instances); until then forwards take index order
    bool ordered = false;
    // va_solve with the long instances on a second stream (fork / join events,
    // and the end of the bulk's forward)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_fwd = nullptr;
    bool capturing = false;  // solve_enqueue is being recorded into a graph
};

namespace {

va_status fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
    VA_HIP(hipGetLastError());
    return VA_OK;
}

va_status fill_f64(double *p, double v, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_f64, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
    VA_HIP(hipGetLastError());
    return VA_OK;
}

va::SolveParams solve_params(const va_handle *h)
{
    va::SolveParams sp;
    sp.B = h->pb.batch;
    sp.cap = h->pb.max_steps;
    sp.t0 = h->pb.t0;
    sp.t1 = h->pb.t1;
    sp.dt0 = h->pb.dt0;
    sp.atol = h->pb.abs_tol;
    sp.rtol = h->pb.rel_tol;
    sp.kconst = h->pb.system_constant;
    sp.max_attempts = h->pb.max_attempts;
    sp.adaptive = h->pb.adaptive;
    return sp;
}

va_status check_handle(const va_handle *h)
{
    if (!h) return fail(VA_ERR_INVALID, "null handle");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != h->device)
        return fail(VA_ERR_INVALID, "handle belongs to device %d, current device is %d", h->device,
                    dev);
    return VA_OK;
}

// Handles that keep checkpoints have no per-step store: only va_solve_observed.
va_status check_rows(const va_handle *h)
{
    if (!h->ks.fwd)
        return fail(VA_ERR_UNSUPPORTED,
                    "this handle's system has shared parameters: use va_solve_shared");
    if (h->obs_ckpt)
        return fail(VA_ERR_STATE,
                    "this handle keeps checkpoints instead of a trajectory store: use "
                    "va_solve_observed");
    return VA_OK;
}

void release(va_handle *h)
{
    if (!h) return;
    hipFree(h->traj);
    hipFree(h->ptraj);
    hipFree(h->pl_t);
    hipFree(h->pl_dt);
    hipFree(h->pl_k0);
    hipFree(h->pl_pbar);
    hipFree(h->pl_done);
    hipFree(h->n_acc);
    hipFree(h->n_rej);
    hipFree(h->status);
    hipFree(h->counters);
    hipFree(h->order);
    hipFree(h->order_fwd);
    hipFree(h->order_fwd_next);
    hipFree(h->order_tmp);
    hipFree(h->obs_ck);
    hipFree(h->obs_seg);
    hipFree(h->obs_ck2);
    hipFree(h->obs_stride);
    hipFree(h->obs_row);
    hipFree(h->abar_part);
    hipFree(h->pim);
    hipFree(h->pbar_im);
    hipFree(h->pbar_set);
    hipFree(h->shared_gscr);
    if (h->graph_exec) hipGraphExecDestroy(h->graph_exec);
    if (h->cap_stream) hipStreamDestroy(h->cap_stream);
    if (h->side) hipStreamDestroy(h->side);
    if (h->ev_fork) hipEventDestroy(h->ev_fork);
    if (h->ev_join) hipEventDestroy(h->ev_join);
    if (h->ev_fwd) hipEventDestroy(h->ev_fwd);
    if (h->host_counter) hipHostFree(h->host_counter);
}

// Longest-first order over the instances of the last forward: mode 0 into
// h->order (the reverse, by accepted steps), mode 1 into h->order_fwd (the next
// forward, by attempts, with its count of long instances), mode 2 as mode 1 into
// h->order_fwd_next.
va_status order_by_length(va_handle *h, int mode, hipStream_t s)
{
    const int64_t B = h->pb.batch;
    va_status st = fill_i32(h->order_tmp, 0, 2 * VA_ORDER_BUCKETS, s);
    if (st) return st;
    const int gb = blocks_for(B, 256);
    int32_t *out = mode == 0 ? h->order : mode == 1 ? h->order_fwd : h->order_fwd_next;
    int32_t *n_long = mode == 0 ? nullptr
                      : h->counters + (mode == 1 ? h->nlong_idx : 11 - h->nlong_idx);
    const int by = mode == 0 ? 0 : 1;
    if (mode != 0) h->ordered = true;
    hipLaunchKernelGGL(k_order_hist, dim3(gb), dim3(256), 0, s, h->n_acc, h->n_rej, h->status,
                       B, by, h->order_tmp);
    hipLaunchKernelGGL(k_order_offsets, dim3(1), dim3(VA_ORDER_BUCKETS), 0, s, h->order_tmp,
                       h->order_tmp + VA_ORDER_BUCKETS, n_long);
    hipLaunchKernelGGL(k_order_scatter, dim3(gb), dim3(256), 0, s, h->n_acc, h->n_rej,
                       h->status, B, by, h->order_tmp + VA_ORDER_BUCKETS, out);
    VA_HIP(hipGetLastError());
    return VA_OK;
}

va_status order_longest_first(va_handle *h, hipStream_t s) { return order_by_length(h, 0, s); }

// Forwards take instances longest-first by an earlier forward's attempts unless
// the caller asked for index order: where instances are refilled (more
// instances than are in flight), so that the long ones start first, and with one
// lane per instance also when all are resident, because the 64 instances that
// share a wave then have similar lengths (a wave runs as long as its longest
// instance).
bool order_wanted(const va_handle *h, int64_t lanes)
{
    if (h->pb.flags & VA_FLAG_COLD_ORDER) return false;
    return h->pb.batch > lanes || (h->ks.layout == 0 && h->pb.batch > 64);
}

// The order is wanted and an earlier forward has left it (a handle's first
// forward runs in index order).
bool reuse_order(const va_handle *h, int64_t lanes)
{
    return h->ordered && order_wanted(h, lanes);
}

// ---------------- persistent form ----------------
// Split schedule: the instances that ran long in the previous forward (the head
// of order_fwd, counted in counters[5]) run one wave each on the caller's
// stream, while the persistent kernels take the rest of the batch on a second
// stream forked from it.  va_solve runs each long instance's forward and then
// its reverse in one kernel; va_forward and va_adjoint run the forward alone
// and the reverse alone.
bool split_solve(const va_handle *h)
{
    return h->side && reuse_order(h, h->lanes_fwd);
}

// Blocks of a persistent kernel beside a long-instance kernel: all of its
// grid but the room of VA_LONG_MAX waves of its own size.  (GLV N=10,
// B=65536: a solve took 2.46 ms so, 2.68 ms with two thirds of the rest; a
// smaller share speeds up the long instances less than it slows down the
// bulk.  Room for both waves of each of the solve's long-instance blocks made
// the bulk's reverse 9% slower.)
int bulk_grid(const va_handle *h, int64_t lanes)
{
    const int per_block = h->ks.slots(h->block);
    const int reserve = blocks_for((int64_t)VA_LONG_MAX * h->ks.slots(64), per_block);
    const int g = blocks_for(std::min<int64_t>(lanes, h->pb.batch), per_block);
    return std::max(1, g - reserve);
}

// Fork the second stream from s.  The caller enqueues the long-instance kernel
// on s before the persistent kernels on the second stream, so that its waves
// are placed first.
va_status fork_side(va_handle *h, hipStream_t s)
{
    VA_HIP(hipEventRecord(h->ev_fork, s));
    VA_HIP(hipStreamWaitEvent(h->side, h->ev_fork, 0));
    return VA_OK;
}

va_status join_side(va_handle *h, hipStream_t s)
{
    VA_HIP(hipEventRecord(h->ev_join, h->side));
    VA_HIP(hipStreamWaitEvent(s, h->ev_join, 0));
    return VA_OK;
}

// The Lotka-Volterra persistent kernels of a handle read instance-major
// parameters [B][P] (copied from the caller's [P][B] per call) and store
// instance-major dJ/dp sums [nout][B][P], applied to the caller's mu afterwards
// (for the instances they marked, beside a long-instance kernel, which reads
// and writes the caller's arrays itself).  Beside a long-instance kernel the
// copies run on the second stream, off the long kernel's path.
bool instance_major(const va_handle *h) { return h->ks.layout == 1 && h->pim != nullptr; }

va_status to_instance_major(const double *p, double *pim, int64_t P, int64_t B, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_instance_major, dim3(blocks_for(B, VA_TILE_B)), dim3(256), 0, s, p,
                       pim, P, B);
    VA_HIP(hipGetLastError());
    return VA_OK;
}

// mu = 0.0 + sum (with a cost) or mu + sum (without one), [nout][P][B], for the
// instances marked in set (all when set is null)
va_status from_instance_major(const double *pbar_im, double *mu, int64_t P, int64_t B, int nout,
                              bool add, hipStream_t s, const int32_t *set = nullptr)
{
    const dim3 grid(blocks_for(B, VA_TILE_B), (unsigned)nout);
    hipLaunchKernelGGL(k_from_instance_major, grid, dim3(256), 0, s, pbar_im, mu, P, B,
                       add ? 1 : 0, set);
    VA_HIP(hipGetLastError());
    return VA_OK;
}

va_status params_instance_major(va_handle *h, const double *p, hipStream_t s)
{
    return to_instance_major(p, h->pim, h->ks.np, h->pb.batch, s);
}

// dJ/dp sums of nout seeds go to h->pbar_im (grown here for more seeds than it
// holds: only va_adjoint with several seeds, which no graph captures).
va_status pbar_instance_major(va_handle *h, va::RevArgs &a, int nout)
{
    const int64_t B = h->pb.batch, P = h->ks.np;
    if (nout > h->pbar_nout) {
        hipFree(h->pbar_im);
        h->pbar_im = nullptr;
        h->pbar_nout = 0;
        VA_HIP(hipMalloc((void **)&h->pbar_im, sizeof(double) * (size_t)nout * B * P));
        h->pbar_nout = nout;
    }
    a.mu = h->pbar_im;
    return VA_OK;
}

va_status pbar_to_mu(va_handle *h, double *mu, int nout, bool add, hipStream_t s,
                     const int32_t *set = nullptr)
{
    return from_instance_major(h->pbar_im, mu, h->ks.np, h->pb.batch, nout, add, s, set);
}

// The bulk's copy-in, instance marks and arguments beside a long-instance
// kernel, on stream s (the second stream).
va_status bulk_instance_major(va_handle *h, const double *p, va::RevArgs *r, int nout,
                              hipStream_t s)
{
    va_status st = params_instance_major(h, p, s);
    if (st || !r) return st;
    r->p = h->pim;
    if ((st = pbar_instance_major(h, *r, nout))) return st;
    if ((st = fill_i32(h->pbar_set, 0, h->pb.batch, s))) return st;
    r->pbar_set = h->pbar_set;
    return VA_OK;
}

va_status forward_persistent(va_handle *h, double *y, const double *p, hipStream_t s)
{
    const int64_t B = h->pb.batch;
    va_status st = fill_i32(h->counters, 0, 4, s);
    if (st) return st;
    va::FwdArgs a;
    a.sp = solve_params(h);
    a.y = y;
    a.p = p;
    a.traj = h->traj;
    a.n_acc = h->n_acc;
    a.n_rej = h->n_rej;
    a.status = h->status;
    a.counter = h->counters + 0;
    a.max_steps_seen = h->counters + 2;
    const bool im = instance_major(h);
    const bool reuse = reuse_order(h, h->lanes_fwd);
    a.order = reuse ? h->order_fwd : nullptr;
    const int64_t lanes = std::min<int64_t>(h->lanes_fwd, B);
    if (split_solve(h) && h->ks.long_fwd) {
        a.split = h->counters + h->nlong_idx;
        a.part = 1;
        if ((st = fork_side(h, s))) return st;
        VA_HIP(h->ks.long_fwd(a, VA_LONG_MAX / 4, 256, s));
        va::FwdArgs ab = a;
        if (im) {
            if ((st = bulk_instance_major(h, p, nullptr, 0, h->side))) return st;
            ab.p = h->pim;
        }
        VA_HIP(h->ks.fwd(ab, bulk_grid(h, h->lanes_fwd), h->block, h->side));
        if ((st = join_side(h, s))) return st;
    } else {
        if (im) {
            if ((st = params_instance_major(h, p, s))) return st;
            a.p = h->pim;
        }
        VA_HIP(h->ks.fwd(a, blocks_for(lanes, h->ks.slots(h->block)), h->block, s));
    }
    if (order_wanted(h, h->lanes_fwd)) return order_by_length(h, 1, s);
    return VA_OK;
}

va_status reverse_persistent(va_handle *h, const double *p, int cost, const double *cost_w,
                             const double *yT, double *J, double *lambda, double *mu, int nout,
                             hipStream_t s)
{
    const int64_t B = h->pb.batch;
    va_status st = fill_i32(h->counters + 1, 0, 1, s);
    if (st) return st;
    // With the split schedule the last forward left its own instances
    // longest-first by attempts in order_fwd, and their long head counted.
    const bool split = split_solve(h) && h->ks.long_rev;
    if (!split && (st = order_longest_first(h, s))) return st;
    va::RevArgs a;
    a.sp = solve_params(h);
    a.order = split ? h->order_fwd : h->order;
    a.p = p;
    a.traj = h->traj;
    a.n_acc = h->n_acc;
    a.status = h->status;
    a.cost = cost;
    a.cost_w = cost_w;
    a.yT = yT;
    a.J = J;
    a.lambda = lambda;
    a.mu = mu;
    a.nout = nout;
    a.counter = h->counters + 1;
    const bool im = instance_major(h);
    const int nseed = cost ? 1 : nout;
    if (split) {
        a.split = h->counters + h->nlong_idx;
        a.part = 1;
        if ((st = fork_side(h, s))) return st;
        // one seed: a block of three waves per long instance, as in a solve
        if (nseed == 1) VA_HIP(h->ks.long_rev(a, VA_LONG_MAX, 192, s));
        else VA_HIP(h->ks.long_rev(a, VA_LONG_MAX / 4, 256, s));
        va::RevArgs ab = a;
        if (im && (st = bulk_instance_major(h, p, &ab, nseed, h->side))) return st;
        VA_HIP(h->ks.rev(ab, bulk_grid(h, h->lanes_rev), h->block, h->side));
        if (im && (st = pbar_to_mu(h, mu, nseed, cost == 0, h->side, h->pbar_set))) return st;
        return join_side(h, s);
    }
    if (im) {
        if ((st = params_instance_major(h, p, s))) return st;
        a.p = h->pim;
        if ((st = pbar_instance_major(h, a, nseed))) return st;
    }
    const int64_t lanes = std::min<int64_t>(h->lanes_rev, B);
    VA_HIP(h->ks.rev(a, blocks_for(lanes, h->ks.slots(h->block)), h->block, s));
    if (im) return pbar_to_mu(h, mu, nseed, cost == 0, s);
    return VA_OK;
}

va_status copy_out(va_handle *h, int32_t *n_acc, int32_t *n_rej, int32_t *lane_status,
                   hipStream_t s);

va_status solve_split(va_handle *h, const double *y0, double *yT, const double *p, int cost,
                      const double *cost_w, double *J, double *dJdy0, double *dJdp,
                      int32_t *n_accepted, int32_t *n_rejected, int32_t *lane_status,
                      hipStream_t s)
{
    va_status st = VA_OK;
    const int64_t B = h->pb.batch;
    const bool im = instance_major(h);
    if (!im) {
        if ((st = fill_i32(h->counters, 0, 4, s))) return st;
        if (yT != y0) {
            hipLaunchKernelGGL(k_copy_f64, dim3(blocks_for((int64_t)h->ks.n * B, 256)), dim3(256),
                               0, s, yT, y0, (int64_t)h->ks.n * B);
            VA_HIP(hipGetLastError());
        }
    }
    va::FwdArgs fa;
    fa.sp = solve_params(h);
    fa.y = yT;
    fa.p = p;
    fa.traj = h->traj;
    fa.n_acc = h->n_acc;
    fa.n_rej = h->n_rej;
    fa.status = h->status;
    fa.counter = h->counters + 0;
    fa.max_steps_seen = h->counters + 2;
    fa.order = h->order_fwd;
    fa.split = h->counters + h->nlong_idx;
    fa.part = 1;
    va::RevArgs ra;
    ra.sp = fa.sp;
    ra.p = p;
    ra.traj = h->traj;
    ra.n_acc = h->n_acc;
    ra.status = h->status;
    ra.cost = cost;
    ra.cost_w = cost_w;
    ra.yT = yT;
    ra.J = J;
    ra.lambda = dJdy0;
    ra.mu = dJdp;
    ra.nout = 1;
    ra.counter = h->counters + 1;
    ra.order = h->order_fwd;
    ra.split = h->counters + h->nlong_idx;
    ra.part = 1;
    va::FwdArgs fb = fa;
    va::RevArgs rb = ra;
    // the bulk's parameter copy first, alone on the device (beside the long
    // kernel's start it ran about half as fast; the long kernel has time to
    // spare), in one launch with y(T) := y0 and the cleared marks and counters
    if (im) {
        if ((st = pbar_instance_major(h, rb, 1))) return st;
        rb.p = h->pim;
        rb.pbar_set = h->pbar_set;
        fb.p = h->pim;
        hipLaunchKernelGGL(k_split_prologue, dim3(blocks_for(B, VA_TILE_B)), dim3(256), 0, s, p,
                           h->pim, (int64_t)h->ks.np, B, y0, yT, h->ks.n, h->pbar_set,
                           h->counters);
        VA_HIP(hipGetLastError());
    }
    if ((st = fork_side(h, s))) return st;
    // one block of three waves per long instance
    VA_HIP(h->ks.long_fused(fa, ra, VA_LONG_MAX, 192, s));
    VA_HIP(h->ks.fwd(fb, bulk_grid(h, h->lanes_fwd), h->block, h->side));
    // The next solve's order and the caller's counts need every forward's counts:
    // on the caller's stream after the long-instance kernel and the bulk's
    // forward, beside the bulk's reverse (which still reads this solve's order),
    // the order into the spare buffer.  A solve recorded into a graph builds it
    // after the join into the same buffer, so that every replay reads the order
    // the previous replay left.
    const bool ahead = !h->capturing;
    if (ahead) {
        VA_HIP(hipEventRecord(h->ev_fwd, h->side));
        VA_HIP(hipStreamWaitEvent(s, h->ev_fwd, 0));
        if ((st = order_by_length(h, 2, s))) return st;
        if ((st = copy_out(h, n_accepted, n_rejected, lane_status, s))) return st;
    }
    VA_HIP(h->ks.rev(rb, bulk_grid(h, h->lanes_rev), h->block, h->side));
    if (im && (st = pbar_to_mu(h, dJdp, 1, cost == 0, h->side, h->pbar_set))) return st;
    if ((st = join_side(h, s))) return st;
    if (!ahead) {
        if ((st = order_by_length(h, 1, s))) return st;
        return copy_out(h, n_accepted, n_rejected, lane_status, s);
    }
    std::swap(h->order_fwd, h->order_fwd_next);
    h->nlong_idx = 11 - h->nlong_idx;
    return VA_OK;
}

// ---------------- plain form ----------------
va::PlainState plain_state(va_handle *h)
{
    va::PlainState ps;
    ps.t = h->pl_t;
    ps.dt = h->pl_dt;
    ps.k0 = h->pl_k0;
    ps.done = h->pl_done;
    ps.active = h->counters + 3;
    ps.max_steps_seen = h->counters + 2;
    ps.pbar = h->pl_pbar;
    return ps;
}

va_status forward_plain(va_handle *h, double *y, const double *p, hipStream_t s)
{
    const int64_t B = h->pb.batch;
    va_status st;
    if ((st = fill_i32(h->pl_done, 0, B, s))) return st;
    if ((st = fill_i32(h->counters + 2, 0, 1, s))) return st;
    if ((st = fill_i32(h->counters + 3, (int32_t)B, 1, s))) return st;
    va::PlainFwdArgs a;
    a.sp = solve_params(h);
    a.y = y;
    a.p = p;
    a.traj = h->ptraj;
    a.n_acc = h->n_acc;
    a.n_rej = h->n_rej;
    a.status = h->status;
    a.st = plain_state(h);
    const int grid = blocks_for(B, h->ks.slots(h->block));
    const int32_t last = h->pb.max_steps + 1;
    const int chunk = 8;
    for (int32_t step = 0; step <= last; step += chunk) {
        for (int32_t k = step; k < std::min(step + chunk, last + 1); ++k) {
            a.step = k;
            VA_HIP(h->ks.plain_fwd(a, grid, h->block, s));
        }
        VA_HIP(hipMemcpyAsync(h->host_counter, h->counters + 3, sizeof(int32_t),
                              hipMemcpyDeviceToHost, s));
        VA_HIP(hipStreamSynchronize(s));
        if (*h->host_counter <= 0) break;
    }
    return VA_OK;
}

va_status read_max_steps(va_handle *h, int32_t *out, hipStream_t s)
{
    VA_HIP(hipMemcpyAsync(h->host_counter, h->counters + 2, sizeof(int32_t),
                          hipMemcpyDeviceToHost, s));
    VA_HIP(hipStreamSynchronize(s));
    *out = *h->host_counter;
    return VA_OK;
}

// One seed: lambda [N][B] in/out, mu [P][B] +=.
va_status reverse_plain_seed(va_handle *h, const double *p, double *lambda, double *mu,
                             int32_t smax, hipStream_t s)
{
    const int64_t B = h->pb.batch;
    va_status st;
    if ((st = fill_f64(h->pl_pbar, 0.0, (int64_t)h->ks.np * B, s))) return st;
    va::PlainRevArgs a;
    a.sp = solve_params(h);
    a.p = p;
    a.traj = h->ptraj;
    a.n_acc = h->n_acc;
    a.status = h->status;
    a.lambda = lambda;
    a.st = plain_state(h);
    const int grid = blocks_for(B, h->ks.slots(h->block));
    for (int32_t k = 0; k < smax; ++k) {
        a.k = k;
        VA_HIP(h->ks.plain_rev(a, grid, h->block, s));
    }
    VA_HIP(va::launch_plain_finish(lambda, lambda, mu, h->pl_pbar, h->status, B, h->ks.n,
                                   h->ks.np, blocks_for(B, 256), 256, s));
    return VA_OK;
}

va_status copy_out(va_handle *h, int32_t *n_acc, int32_t *n_rej, int32_t *lane_status,
                   hipStream_t s)
{
    const int64_t B = h->pb.batch;
    if (!n_acc && !n_rej && !lane_status) return VA_OK;
    hipLaunchKernelGGL(k_copy_counts, dim3(blocks_for(B, 256)), dim3(256), 0, s, n_acc, n_rej,
                       lane_status, h->n_acc, h->n_rej, h->status, B);
    VA_HIP(hipGetLastError());
    return VA_OK;
}

}  // namespace

extern "C" {

void va_problem_init(va_problem *pb)
{
    if (!pb) return;
    memset(pb, 0, sizeof(*pb));
    pb->abi_version = VA_ABI_VERSION;
    pb->adaptive = 1;
    pb->dt0 = 1e-3;
    pb->abs_tol = 1e-6;
    pb->rel_tol = 1e-6;
    pb->max_attempts = 500;
    pb->max_steps = 4096;
    pb->mode = VA_MODE_PERSISTENT;
}

const char *va_last_error(void) { return g_err; }

}  // extern "C"

namespace {

// The checks of va_create that need no device state: argument ranges, the kernel
// set for (system, N, tableau) -- or the caller's own table for VA_SYSTEM_USER --
// and the parameter count.
va_status check_problem(const va_problem *pb, va::KernelSet *ks_out,
                        const va::KernelSet *given = nullptr)
{
    if (pb->abi_version != VA_ABI_VERSION)
        return fail(VA_ERR_INVALID, "abi_version %d, library implements %d", pb->abi_version,
                    VA_ABI_VERSION);
    if (pb->batch <= 0 || pb->batch > (int64_t)0x7fffffff)
        return fail(VA_ERR_INVALID, "batch must be in [1, 2^31), got %lld", (long long)pb->batch);
    if (pb->max_steps < 1 || pb->max_attempts < 1)
        return fail(VA_ERR_INVALID, "max_steps and max_attempts must be >= 1");
    if (!(pb->dt0 > 0.0) || !(pb->t1 >= pb->t0))
        return fail(VA_ERR_INVALID, "need dt0 > 0 and t1 >= t0");
    if (pb->adaptive && !(pb->abs_tol >= 0.0 && pb->rel_tol >= 0.0))
        return fail(VA_ERR_INVALID, "tolerances must be >= 0");
    if (pb->mode != VA_MODE_PERSISTENT && pb->mode != VA_MODE_PLAIN)
        return fail(VA_ERR_INVALID, "unknown mode %d", pb->mode);
    if (pb->adaptive &&
        (pb->tableau == VA_TABLEAU_EULER || pb->tableau == VA_TABLEAU_RK4))
        return fail(VA_ERR_INVALID,
                    "tableau %d has no embedded error estimate: use adaptive = 0", pb->tableau);
    va::KernelSet ks{};
    if (pb->flags & ~(VA_FLAG_GRAPH | VA_FLAG_COLD_ORDER | VA_FLAG_CHECKPOINT_2LEVEL))
        return fail(VA_ERR_INVALID, "unknown va_problem.flags bits");
    if (pb->n_obs < 0) return fail(VA_ERR_INVALID, "n_obs must be >= 0");
    if (pb->checkpoint_slots < 0 || pb->checkpoint_slots == 1 || pb->checkpoint_slots % 2 != 0)
        return fail(VA_ERR_INVALID, "checkpoint_slots must be 0 or an even number >= 2");

    if (pb->instances_per_wave < 0 || pb->instances_per_wave > 64)
        return fail(VA_ERR_INVALID, "instances_per_wave must be in [0, 64]");
    if (given) {
        if (pb->system != VA_SYSTEM_USER)
            return fail(VA_ERR_INVALID, "a kernel table needs system VA_SYSTEM_USER");
        if (given->table_size != (int32_t)sizeof(va::KernelSet) ||
            given->table_version != va::VA_KERNEL_TABLE_VERSION || !given->fwd || !given->rev)
            return fail(VA_ERR_INVALID,
                        "kernel table of size %d, version %d: this library takes size %d, "
                        "version %d (rebuild the system's translation unit against its headers)",
                        given->table_size, given->table_version, (int)sizeof(va::KernelSet),
                        va::VA_KERNEL_TABLE_VERSION);
        if (given->n != pb->n_state)
            return fail(VA_ERR_INVALID, "the kernel table is for N=%d, the problem says %d",
                        given->n, pb->n_state);
        if (pb->instances_per_wave != 0 && pb->instances_per_wave != 64)
            return fail(VA_ERR_INVALID, "user systems run one lane per instance");
        ks = *given;
    } else if (pb->system == VA_SYSTEM_USER) {
        return fail(VA_ERR_INVALID, "VA_SYSTEM_USER needs va_create_with_kernels");
    } else if (!va::get_kernels(pb->system, pb->n_state, pb->tableau, pb->adaptive,
                                pb->instances_per_wave, &ks)) {
        return fail(VA_ERR_UNSUPPORTED,
                    "no kernels for system %d, N=%d, tableau %d, adaptive %d, "
                    "instances_per_wave %d",
                    pb->system, pb->n_state, pb->tableau, pb->adaptive, pb->instances_per_wave);
    }
    if (pb->n_param != ks.np)
        return fail(VA_ERR_INVALID, "system %d with N=%d has %d parameters, problem says %d",
                    pb->system, pb->n_state, ks.np, pb->n_param);
    *ks_out = ks;
    return VA_OK;
}

// Launch shapes picked by va_autotune in this process, per device and problem
// shape (system, N, tableau, adaptive, observations, batch rounded up to a power
// of two).  va_create applies them to problems whose block_size /
// instances_per_wave are 0.
struct TuneKey {
    int device, system, n, tableau, adaptive, n_obs;
    int64_t bucket;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(device, system, n, tableau, adaptive, n_obs, bucket) <
               std::tie(o.device, o.system, o.n, o.tableau, o.adaptive, o.n_obs, o.bucket);
    }
};
struct TuneChoice {
    int32_t block, ipw, slots;
};
std::mutex g_tune_mtx;
std::map<TuneKey, TuneChoice> g_tuned;

TuneKey tune_key(const va_problem *pb, int device)
{
    int64_t bucket = 1;
    while (bucket < pb->batch) bucket *= 2;
    return TuneKey{device, pb->system, pb->n_state, pb->tableau, pb->adaptive, pb->n_obs, bucket};
}

va_status create_handle(const va_problem *pb_in, const va::KernelSet *given, va_handle **out)
{
    g_err[0] = 0;
    if (!pb_in || !out) return fail(VA_ERR_INVALID, "null argument");
    *out = nullptr;
    va_problem tuned = *pb_in;
    if (tuned.block_size == 0 || tuned.instances_per_wave == 0) {
        int dev = -1;
        if (hipGetDevice(&dev) == hipSuccess) {
            std::lock_guard<std::mutex> lock(g_tune_mtx);
            auto it = g_tuned.find(tune_key(&tuned, dev));
            if (it != g_tuned.end()) {
                if (tuned.block_size == 0) tuned.block_size = it->second.block;
                if (tuned.instances_per_wave == 0) tuned.instances_per_wave = it->second.ipw;
            }
        }
    }
    const va_problem *pb = &tuned;
    va::KernelSet ks{};
    if (va_status st = check_problem(pb, &ks, given)) return st;
    if (ks.layout == 2 && (pb->mode != VA_MODE_PERSISTENT || pb->n_obs > 0))
        return fail(VA_ERR_UNSUPPORTED,
                    "shared-parameter systems run in the persistent mode without observations");
    const bool obs_capable = ks.obs_fwd && pb->mode == VA_MODE_PERSISTENT;
    if ((pb->n_obs > 0 || pb->checkpoint_slots > 0) && !obs_capable)
        return fail(VA_ERR_UNSUPPORTED,
                    "observed and checkpointed solves need an adaptive error stepper on a "
                    "one-lane system (Van der Pol, oscillators) in the persistent mode");

    va_handle *h = new (std::nothrow) va_handle();
    if (!h) return fail(VA_ERR_OUT_OF_MEMORY, "host allocation");
    h->pb = *pb;
    h->ks = ks;
    hipDeviceProp_t prop;
    hipError_t de = hipGetDevice(&h->device);
    if (de == hipSuccess) de = hipGetDeviceProperties(&prop, h->device);
    if (de != hipSuccess) {
        delete h;
        return fail(VA_ERR_HIP, "no HIP device: %s", hipGetErrorString(de));
    }
    h->cu_count = prop.multiProcessorCount;
    h->block = pb->block_size > 0 ? pb->block_size : 256;
    if (h->block % 64 != 0 || h->block > 256) {
        delete h;
        return fail(VA_ERR_INVALID, "block_size must be a multiple of 64 and <= 256");
    }
    const int64_t B = pb->batch;
    if (ks.layout == 2) {
        // shared-parameter tiles: one workgroup per CU (the matrix fills the LDS)
        const int64_t ntiles = (B + ks.shared_tile - 1) / ks.shared_tile;
        h->shared_grid = (int)std::min<int64_t>(ntiles, h->cu_count);
    } else if (pb->lanes > 0) {
        h->lanes_fwd = h->lanes_rev = pb->lanes;
    } else {
        // every resident instance slot of the device
        const int64_t per_block = ks.slots(h->block);
        h->lanes_fwd = (int64_t)std::max(1, ks.occ_fwd(h->block)) * h->cu_count * per_block;
        h->lanes_rev = (int64_t)std::max(1, ks.occ_rev(h->block)) * h->cu_count * per_block;
    }
    const double budget = (double)pb->trajectory_budget_bytes;
    bool full_store = true;  // per-step rows in h->traj / h->ptraj
    const size_t rows_b = (size_t)pb->max_steps + 1;
    const double store =
        (double)sizeof(double) * (double)B *
        (double)(pb->mode == VA_MODE_PERSISTENT
                     ? rows_b * (ks.n + 2)
                     : rows_b * (2 + (size_t)ks.n));
    const bool over = budget > 0.0 && store > budget;
    if (pb->n_obs > 0 || (obs_capable && (over || pb->checkpoint_slots > 0))) {
        // observation handles, and terminal-cost handles whose store exceeds the
        // budget: full rows while they fit, else checkpoints (observation kernels
        // with no observations and the terminal cost)
        const int64_t per_block = h->block;
        h->lanes_obs_fwd = (int64_t)std::max(1, ks.occ_obs_fwd(h->block)) * h->cu_count * per_block;
        h->lanes_obs_rev = (int64_t)std::max(1, ks.occ_obs_rev(h->block)) * h->cu_count * per_block;
        if (pb->lanes > 0) h->lanes_obs_fwd = h->lanes_obs_rev = pb->lanes;
        const double fixed = 4.0 * (double)B * ((double)pb->n_obs + 1.0);
        const double full = 8.0 * (double)B * ((double)pb->max_steps + 1.0) * (ks.n + 2) + fixed;
        if (pb->checkpoint_slots == 0 && (budget <= 0.0 || full <= budget)) {
            h->obs_ckpt = 0;
        } else {
            full_store = false;
            h->obs_ckpt = 1;
            const int64_t rev_threads =
                (int64_t)blocks_for(std::min<int64_t>(h->lanes_obs_rev, B), h->block) * h->block;
            auto cost = [&](int32_t slots, int32_t *stride_max) {
                int32_t sm = 1;
                while ((int64_t)sm * slots < (int64_t)pb->max_steps) sm *= 2;
                *stride_max = sm;
                return 8.0 * (double)B * slots * (ks.n + 3) +
                       8.0 * (double)rev_threads * sm * (ks.n + 1) + fixed;
            };
            // two levels: within a segment of up to stride_max steps, slots2 level-2
            // checkpoints per lane (about sqrt(stride_max)) and a row buffer of
            // ceil(stride_max / slots2) rows
            auto cost2 = [&](int32_t slots, int32_t *stride_max, int32_t *slots2,
                             int32_t *seg_rows) {
                int32_t sm = 1;
                while ((int64_t)sm * slots < (int64_t)pb->max_steps) sm *= 2;
                int32_t s2 = 2;
                while ((int64_t)s2 * s2 < sm) s2 *= 2;
                *stride_max = sm;
                *slots2 = s2;
                *seg_rows = (sm + s2 - 1) / s2;
                return 8.0 * (double)B * slots * (ks.n + 3) +
                       8.0 * (double)rev_threads *
                           ((double)s2 * (ks.n + 3) + (double)*seg_rows * (ks.n + 1)) +
                       fixed;
            };
            const bool force2 = (pb->flags & VA_FLAG_CHECKPOINT_2LEVEL) != 0;
            int32_t slots = pb->checkpoint_slots, sm = 0, s2 = 0, rows2 = 0;
            int32_t levels = force2 ? 2 : 1;
            if (slots > 0) {
                if (levels == 1) cost(slots, &sm);
                else cost2(slots, &sm, &s2, &rows2);
            } else {
                // the most slots that fit (shortest segments, smallest recompute buffers):
                // one level if any fits, else two
                slots = 0;
                if (levels == 1)
                    for (int32_t c = 1 << 16; c >= 2 && !slots; c /= 2)
                        if (cost(c, &sm) <= budget) slots = c;
                if (!slots) {
                    levels = 2;
                    for (int32_t c = 1 << 16; c >= 2 && !slots; c /= 2)
                        if (cost2(c, &sm, &s2, &rows2) <= budget) slots = c;
                }
                if (!slots) {
                    delete h;
                    return fail(VA_ERR_OUT_OF_MEMORY,
                                "trajectory_budget_bytes %lld cannot hold checkpoints for %lld "
                                "instances (at two levels either)",
                                (long long)pb->trajectory_budget_bytes, (long long)B);
                }
            }
            h->obs_slots = slots;
            h->obs_stride_max = sm;
            h->obs_levels = levels;
            h->obs_slots2 = levels == 2 ? s2 : 0;
            h->obs_seg_rows = levels == 2 ? rows2 : sm;
        }
    } else if (over) {
        delete h;
        return fail(VA_ERR_OUT_OF_MEMORY,
                    "trajectory store of %.3g bytes exceeds trajectory_budget_bytes %lld",
                    store, (long long)pb->trajectory_budget_bytes);
    }

    const size_t rows = (size_t)pb->max_steps + 1;
    hipError_t e = hipSuccess;
    auto alloc = [&](void **ptr, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(ptr, bytes);
    };
    alloc((void **)&h->n_acc, sizeof(int32_t) * B);
    alloc((void **)&h->n_rej, sizeof(int32_t) * B);
    alloc((void **)&h->status, sizeof(int32_t) * B);
    alloc((void **)&h->counters, sizeof(int32_t) * 8);
    alloc((void **)&h->order, sizeof(int32_t) * (size_t)pb->batch);
    alloc((void **)&h->order_fwd, sizeof(int32_t) * (size_t)pb->batch);
    alloc((void **)&h->order_fwd_next, sizeof(int32_t) * (size_t)pb->batch);
    alloc((void **)&h->order_tmp, sizeof(int32_t) * 2 * VA_ORDER_BUCKETS);
    if (pb->n_obs > 0 || h->obs_ckpt) {
        alloc((void **)&h->obs_row, sizeof(int32_t) * std::max<size_t>(1, (size_t)B * pb->n_obs));
        if (h->obs_ckpt) {
            const int64_t rev_threads =
                (int64_t)blocks_for(std::min<int64_t>(h->lanes_obs_rev, B), h->block) * h->block;
            alloc((void **)&h->obs_ck, sizeof(double) * (size_t)B * h->obs_slots * (ks.n + 3));
            alloc((void **)&h->obs_stride, sizeof(int32_t) * (size_t)B);
            alloc((void **)&h->obs_seg,
                  sizeof(double) * (size_t)rev_threads * h->obs_seg_rows * (ks.n + 1));
            if (h->obs_levels == 2)
                alloc((void **)&h->obs_ck2,
                      sizeof(double) * (size_t)rev_threads * h->obs_slots2 * (ks.n + 3));
        }
    }
    if (ks.layout == 2) {
        alloc((void **)&h->abar_part, sizeof(double) * (size_t)h->shared_grid * ks.n * ks.n);
        // per workgroup: waves x s stages x 64 lanes x 4 growth values
        // the reverse's growth products and stage inputs, per wave and stage
        alloc((void **)&h->shared_gscr, sizeof(double) * (size_t)h->shared_grid *
                                            ks.shared_waves * 2 * ks.s_rev * 64 * 4);
    }
    if (pb->mode == VA_MODE_PERSISTENT) {
        if (full_store) alloc((void **)&h->traj, sizeof(double) * rows * (size_t)B * (ks.n + 2));
        if (ks.layout == 1 && full_store) {
            // instance-major parameters and dJ/dp sums of the row kernels
            alloc((void **)&h->pim, sizeof(double) * (size_t)B * ks.np);
            alloc((void **)&h->pbar_im, sizeof(double) * (size_t)B * ks.np);
            alloc((void **)&h->pbar_set, sizeof(int32_t) * (size_t)B);
            h->pbar_nout = 1;
        }
    } else {
        const size_t F = 2 + (size_t)ks.n;  // plain rows (t, h, y)
        alloc((void **)&h->ptraj, sizeof(double) * rows * F * (size_t)B);
        alloc((void **)&h->pl_t, sizeof(double) * B);
        alloc((void **)&h->pl_dt, sizeof(double) * B);
        alloc((void **)&h->pl_k0, sizeof(double) * B * ks.n);
        alloc((void **)&h->pl_pbar, sizeof(double) * B * ks.np);
        alloc((void **)&h->pl_done, sizeof(int32_t) * B);
    }
    if (e == hipSuccess) e = hipHostMalloc((void **)&h->host_counter, sizeof(int32_t));
    if (e != hipSuccess) {
        release(h);
        delete h;
        (void)hipGetLastError();
        return fail(e == hipErrorOutOfMemory ? VA_ERR_OUT_OF_MEMORY : VA_ERR_HIP,
                    "allocating the trajectory store (%zu rows x %lld instances): %s", rows,
                    (long long)B, hipGetErrorString(e));
    }
    // the second stream and its events of split solves (created here: va_solve may
    // run under stream capture)
    if (e == hipSuccess && ks.long_fused && pb->adaptive && pb->mode == VA_MODE_PERSISTENT) {
        e = hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_fwd, hipEventDisableTiming);
    }
    // no previous forward: zero attempt counts and no long instances
    if (e == hipSuccess) e = hipMemset(h->counters, 0, sizeof(int32_t) * 8);
    if (e == hipSuccess) e = hipMemset(h->n_acc, 0, sizeof(int32_t) * B);
    if (e == hipSuccess) e = hipMemset(h->n_rej, 0, sizeof(int32_t) * B);
    {  // the first forward takes instances in index order
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_iota_i32, dim3(blocks_for(B, 256)), dim3(256), 0, nullptr,
                               h->order_fwd, B);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) {
            release(h);
            delete h;
            (void)hipGetLastError();
            return fail(VA_ERR_HIP, "initialising streams, step counts and instance order: %s",
                        hipGetErrorString(e));
        }
    }
    *out = h;
    return VA_OK;
}

}  // namespace

extern "C" {

va_status va_create(const va_problem *pb, va_handle **out)
{
    return create_handle(pb, nullptr, out);
}

va_status va_create_with_kernels(const va_problem *pb, const va_kernel_table *kernels,
                                 va_handle **out)
{
    if (!kernels) return fail(VA_ERR_INVALID, "null kernel table");
    return create_handle(pb, reinterpret_cast<const va::KernelSet *>(kernels), out);
}

void va_destroy(va_handle *h)
{
    if (!h) return;
    release(h);
    delete h;
}

va_status va_forward(va_handle *h, double *y, const double *p, int32_t *n_accepted,
                     int32_t *n_rejected, int32_t *lane_status, void *stream)
{
    va_status st = check_handle(h);
    if (st) return st;
    if ((st = check_rows(h))) return st;
    if (!y || !p) return fail(VA_ERR_INVALID, "y and p are required");
    hipStream_t s = (hipStream_t)stream;
    st = h->pb.mode == VA_MODE_PLAIN ? forward_plain(h, y, p, s) : forward_persistent(h, y, p, s);
    if (st) return st;
    h->have_forward = true;
    return copy_out(h, n_accepted, n_rejected, lane_status, s);
}

va_status va_adjoint(va_handle *h, const double *p, int32_t nout, double *lambda, double *mu,
                     void *stream)
{
    va_status st = check_handle(h);
    if (st) return st;
    if ((st = check_rows(h))) return st;
    if (!h->have_forward) return fail(VA_ERR_STATE, "va_adjoint needs a va_forward first");
    if (!p || !lambda || !mu || nout < 1)
        return fail(VA_ERR_INVALID, "p, lambda, mu required and nout >= 1");
    hipStream_t s = (hipStream_t)stream;
    if (h->pb.mode == VA_MODE_PLAIN) {
        int32_t smax = 0;
        if ((st = read_max_steps(h, &smax, s))) return st;
        const int64_t B = h->pb.batch;
        for (int o = 0; o < nout; ++o) {
            st = reverse_plain_seed(h, p, lambda + (size_t)o * h->ks.n * B,
                                    mu + (size_t)o * h->ks.np * B, smax, s);
            if (st) return st;
        }
        return VA_OK;
    }
    return reverse_persistent(h, p, 0, nullptr, nullptr, nullptr, lambda, mu, nout, s);
}

va_status va_sensitivity_matrix(va_handle *h, const double *p, double *S, void *stream)
{
    va_status st = check_handle(h);
    if (st) return st;
    if ((st = check_rows(h))) return st;
    if (!p || !S) return fail(VA_ERR_INVALID, "p and S are required");
    hipStream_t s = (hipStream_t)stream;
    const int64_t B = h->pb.batch;
    const int n = h->ks.n, np = h->ks.np;
    double *lam = nullptr;
    VA_HIP(hipMallocAsync((void **)&lam, sizeof(double) * (size_t)n * n * B, s));
    hipLaunchKernelGGL(k_identity_seeds, dim3(blocks_for((int64_t)n * n * B, 256)), dim3(256), 0,
                       s, lam, n, B);
    if ((st = fill_f64(S, 0.0, (int64_t)n * np * B, s))) {
        hipFreeAsync(lam, s);
        return st;
    }
    st = va_adjoint(h, p, n, lam, S, stream);
    hipFreeAsync(lam, s);
    return st;
}

static va_status solve_enqueue(va_handle *h, const double *y0, const double *p, int32_t cost,
                               const double *cost_weights, double *J, double *yT, double *dJdy0,
                               double *dJdp, int32_t *n_accepted, int32_t *n_rejected,
                               int32_t *lane_status, void *stream);
static va_status observed_enqueue(va_handle *h, const double *y0, const double *p,
                                  const double *obs_t, const double *obs_y, int32_t cost,
                                  const double *cost_w, double *J, double *yT, double *dJdy0,
                                  double *dJdp, int32_t *n_accepted, int32_t *n_rejected,
                                  int32_t *lane_status, void *stream);

va_status va_solve(va_handle *h, const double *y0, const double *p, int32_t cost,
                   const double *cost_weights, double *J, double *yT, double *dJdy0,
                   double *dJdp, int32_t *n_accepted, int32_t *n_rejected, int32_t *lane_status,
                   void *stream)
{
    va_status st = check_handle(h);
    if (st) return st;
    if (h->obs_ckpt && h->pb.n_obs == 0) {
        // terminal cost over a checkpointed trajectory
        if (!y0 || !p || !J || !yT || !dJdy0 || !dJdp)
            return fail(VA_ERR_INVALID, "y0, p, J, yT, dJdy0, dJdp are required");
        if (cost < VA_COST_SUM || cost > VA_COST_LINEAR)
            return fail(VA_ERR_INVALID, "unknown cost %d", cost);
        if (cost == VA_COST_LINEAR && !cost_weights)
            return fail(VA_ERR_INVALID, "VA_COST_LINEAR needs cost_weights");
        return observed_enqueue(h, y0, p, nullptr, nullptr, cost, cost_weights, J, yT, dJdy0,
                                dJdp, n_accepted, n_rejected, lane_status, stream);
    }
    if ((st = check_rows(h))) return st;
    // A handle's first solve (index order, no long instances known yet) runs
    // unrecorded; the graph records the solve that reuses the order.
    if (!(h->pb.flags & VA_FLAG_GRAPH) || h->pb.mode != VA_MODE_PERSISTENT ||
        (!h->ordered && order_wanted(h, h->lanes_fwd)))
        return solve_enqueue(h, y0, p, cost, cost_weights, J, yT, dJdy0, dJdp, n_accepted,
                             n_rejected, lane_status, stream);
    const void *key[12] = {y0, p, cost_weights, J, yT, dJdy0, dJdp, n_accepted, n_rejected,
                           lane_status, nullptr, nullptr};
    bool same = h->graph_exec && h->graph_cost == cost;
    for (int i = 0; i < 12 && same; ++i) same = key[i] == h->graph_key[i];
    if (!same) {
        // record the launch sequence on a private stream, then keep the executable graph
        if (!h->cap_stream) VA_HIP(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
        if (h->graph_exec) {
            (void)hipGraphExecDestroy(h->graph_exec);
            h->graph_exec = nullptr;
        }
        VA_HIP(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
        h->capturing = true;
        st = solve_enqueue(h, y0, p, cost, cost_weights, J, yT, dJdy0, dJdp, n_accepted,
                           n_rejected, lane_status, h->cap_stream);
        h->capturing = false;
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(h->cap_stream, &graph);
        if (st) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        VA_HIP(ec);
        const hipError_t ei = hipGraphInstantiate(&h->graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        VA_HIP(ei);
        for (int i = 0; i < 12; ++i) h->graph_key[i] = key[i];
        h->graph_cost = cost;
    }
    VA_HIP(hipGraphLaunch(h->graph_exec, (hipStream_t)stream));
    h->have_forward = true;
    return VA_OK;
}

static va_status solve_enqueue(va_handle *h, const double *y0, const double *p, int32_t cost,
                               const double *cost_weights, double *J, double *yT, double *dJdy0,
                               double *dJdp, int32_t *n_accepted, int32_t *n_rejected,
                               int32_t *lane_status, void *stream)
{
    va_status st = VA_OK;
    if (!y0 || !p || !J || !yT || !dJdy0 || !dJdp)
        return fail(VA_ERR_INVALID, "y0, p, J, yT, dJdy0, dJdp are required");
    if (cost < VA_COST_SUM || cost > VA_COST_LINEAR)
        return fail(VA_ERR_INVALID, "unknown cost %d", cost);
    if (cost == VA_COST_LINEAR && !cost_weights)
        return fail(VA_ERR_INVALID, "VA_COST_LINEAR needs cost_weights");
    hipStream_t s = (hipStream_t)stream;
    const int64_t B = h->pb.batch;
    const int n = h->ks.n, np = h->ks.np;
    const bool split = h->pb.mode != VA_MODE_PLAIN && split_solve(h);
    if (split) {
        // copies y0 and the counts itself
        st = solve_split(h, y0, yT, p, cost, cost_weights, J, dJdy0, dJdp, n_accepted,
                         n_rejected, lane_status, s);
        if (!st) h->have_forward = true;
        return st;
    }
    if (yT != y0) {
        hipLaunchKernelGGL(k_copy_f64, dim3(blocks_for((int64_t)n * B, 256)), dim3(256), 0, s, yT,
                           y0, (int64_t)n * B);
        VA_HIP(hipGetLastError());
    }
    if (h->pb.mode == VA_MODE_PLAIN) {
        if ((st = forward_plain(h, yT, p, s))) return st;
        h->have_forward = true;
        VA_HIP(va::launch_plain_cost(cost, cost_weights, yT, J, dJdy0, h->status, B, n,
                                     blocks_for(B, h->block), h->block, s));
        if ((st = fill_f64(dJdp, 0.0, (int64_t)np * B, s))) return st;
        int32_t smax = 0;
        if ((st = read_max_steps(h, &smax, s))) return st;
        if ((st = reverse_plain_seed(h, p, dJdy0, dJdp, smax, s))) return st;
    } else {
        if ((st = forward_persistent(h, yT, p, s))) return st;
        h->have_forward = true;
        if ((st = reverse_persistent(h, p, cost, cost_weights, yT, J, dJdy0, dJdp, 1, s)))
            return st;
    }
    return copy_out(h, n_accepted, n_rejected, lane_status, s);
}

// Observation-kernel solve: the running cost over obs (K = pb.n_obs), plus the
// terminal cost when cost > 0 (K = 0 serves terminal-cost handles that keep
// checkpoints).
static va_status observed_enqueue(va_handle *h, const double *y0, const double *p,
                                  const double *obs_t, const double *obs_y, int32_t cost,
                                  const double *cost_w, double *J, double *yT, double *dJdy0,
                                  double *dJdp, int32_t *n_accepted, int32_t *n_rejected,
                                  int32_t *lane_status, void *stream)
{
    va_status st = VA_OK;
    hipStream_t s = (hipStream_t)stream;
    const int64_t B = h->pb.batch;
    const int n = h->ks.n;
    if (yT != y0) {
        hipLaunchKernelGGL(k_copy_f64, dim3(blocks_for((int64_t)n * B, 256)), dim3(256), 0, s, yT,
                           y0, (int64_t)n * B);
        VA_HIP(hipGetLastError());
    }
    if ((st = fill_i32(h->counters, 0, 4, s))) return st;
    va::ObsStore os;
    os.ckpt = h->obs_ckpt;
    os.traj = h->traj;
    os.ck = h->obs_ck;
    os.slots = h->obs_slots;
    os.stride_max = h->obs_stride_max;
    os.stride = h->obs_stride;
    os.seg = h->obs_seg;
    os.levels = h->obs_levels;
    os.seg_rows = h->obs_seg_rows;
    os.slots2 = h->obs_slots2;
    os.ck2 = h->obs_ck2;
    os.obs_row = h->obs_row;
    va::ObsFwdArgs f;
    f.sp = solve_params(h);
    f.y = yT;
    f.p = p;
    f.n_acc = h->n_acc;
    f.n_rej = h->n_rej;
    f.status = h->status;
    f.counter = h->counters + 0;
    f.max_steps_seen = h->counters + 2;
    const bool reuse = reuse_order(h, h->lanes_obs_fwd);
    f.order = reuse ? h->order_fwd : nullptr;
    f.obs_t = obs_t;
    f.obs_y = obs_y;
    f.K = h->pb.n_obs;
    f.J = J;
    f.st = os;
    VA_HIP(h->ks.obs_fwd(f, blocks_for(std::min<int64_t>(h->lanes_obs_fwd, B), h->block),
                         h->block, s));
    h->have_forward = !h->obs_ckpt;
    if (order_wanted(h, h->lanes_obs_fwd) && (st = order_by_length(h, 1, s))) return st;
    if ((st = order_longest_first(h, s))) return st;
    va::ObsRevArgs r;
    r.sp = f.sp;
    r.p = p;
    r.yT = yT;
    r.n_acc = h->n_acc;
    r.status = h->status;
    r.obs_y = obs_y;
    r.obs_t = obs_t;
    r.K = h->pb.n_obs;
    r.dJdy0 = dJdy0;
    r.dJdp = dJdp;
    r.counter = h->counters + 1;
    r.order = h->order;
    r.cost = cost;
    r.cost_w = cost_w;
    r.J = J;
    r.st = os;
    VA_HIP(h->ks.obs_rev(r, blocks_for(std::min<int64_t>(h->lanes_obs_rev, B), h->block),
                         h->block, s));
    return copy_out(h, n_accepted, n_rejected, lane_status, s);
}

va_status va_solve_observed(va_handle *h, const double *y0, const double *p, const double *obs_t,
                            const double *obs_y, double *J, double *yT, double *dJdy0,
                            double *dJdp, int32_t *n_accepted, int32_t *n_rejected,
                            int32_t *lane_status, void *stream)
{
    va_status st = check_handle(h);
    if (st) return st;
    if (h->pb.n_obs <= 0)
        return fail(VA_ERR_STATE, "va_solve_observed needs a handle created with n_obs > 0");
    if (!y0 || !p || !obs_t || !obs_y || !J || !yT || !dJdy0 || !dJdp)
        return fail(VA_ERR_INVALID, "y0, p, obs_t, obs_y, J, yT, dJdy0, dJdp are required");
    return observed_enqueue(h, y0, p, obs_t, obs_y, 0, nullptr, J, yT, dJdy0, dJdp, n_accepted,
                            n_rejected, lane_status, stream);
}

va_status va_solve_shared(va_handle *h, const double *y0, const double *p, const double *p_shared,
                          int32_t cost, const double *cost_w, double *J, double *yT,
                          double *dJdy0, double *dJdp, double *dJdp_shared, int32_t *n_accepted,
                          int32_t *n_rejected, int32_t *lane_status, void *stream)
{
    va_status st = check_handle(h);
    if (st) return st;
    if (h->ks.layout != 2)
        return fail(VA_ERR_STATE, "va_solve_shared needs a shared-parameter system");
    if (!y0 || !p || !p_shared || !J || !yT || !dJdy0 || !dJdp || !dJdp_shared)
        return fail(VA_ERR_INVALID,
                    "y0, p, p_shared, J, yT, dJdy0, dJdp, dJdp_shared are required");
    if (cost < VA_COST_SUM || cost > VA_COST_LINEAR)
        return fail(VA_ERR_INVALID, "unknown cost %d", cost);
    if (cost == VA_COST_LINEAR && !cost_w)
        return fail(VA_ERR_INVALID, "VA_COST_LINEAR needs cost_weights");
    hipStream_t s = (hipStream_t)stream;
    const int64_t B = h->pb.batch;
    const int n = h->ks.n;
    if (yT != y0) {
        hipLaunchKernelGGL(k_copy_f64, dim3(blocks_for((int64_t)n * B, 256)), dim3(256), 0, s, yT,
                           y0, (int64_t)n * B);
        VA_HIP(hipGetLastError());
    }
    if ((st = fill_i32(h->counters, 0, 4, s))) return st;
    va::SharedFwdArgs f;
    f.sp = solve_params(h);
    f.y = yT;
    f.r = p;
    f.A = p_shared;
    f.traj = h->traj;
    f.n_acc = h->n_acc;
    f.n_rej = h->n_rej;
    f.status = h->status;
    f.max_steps_seen = h->counters + 2;
    VA_HIP(h->ks.shared_fwd(f, h->shared_grid, s));
    h->have_forward = true;
    // J and the seed dJ/dy(T), written into dJdy0 and swept back in place
    VA_HIP(va::launch_plain_cost(cost, cost_w, yT, J, dJdy0, h->status, B, n, blocks_for(B, 256),
                                 256, s));
    if ((st = order_longest_first(h, s))) return st;
    va::SharedRevArgs r;
    r.sp = f.sp;
    r.r = p;
    r.A = p_shared;
    r.traj = h->traj;
    r.n_acc = h->n_acc;
    r.status = h->status;
    r.order = h->order;
    r.lambda = dJdy0;
    r.dJdy0 = dJdy0;
    r.dJdr = dJdp;
    r.abar_part = h->abar_part;
    r.gscr = h->shared_gscr;
    VA_HIP(h->ks.shared_rev(r, h->shared_grid, s));
    VA_HIP(va::launch_glvs_reduce(h->abar_part, h->shared_grid, dJdp_shared, s));
    return copy_out(h, n_accepted, n_rejected, lane_status, s);
}

va_status va_trajectory(va_handle *h, int64_t instance, double *rows, int64_t max_rows,
                        int64_t *n_rows, void *stream)
{
    va_status st = check_handle(h);
    if (st) return st;
    if (h->obs_ckpt) return check_rows(h);
    if (!h->have_forward) return fail(VA_ERR_STATE, "va_trajectory needs a va_forward first");
    if (!n_rows || instance < 0 || instance >= h->pb.batch || max_rows < 0 ||
        (max_rows > 0 && !rows))
        return fail(VA_ERR_INVALID, "bad instance, rows or n_rows");
    hipStream_t s = (hipStream_t)stream;
    int32_t cnt = 0, lane_st = 0;
    VA_HIP(hipMemcpyAsync(&cnt, h->n_acc + instance, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    VA_HIP(hipMemcpyAsync(&lane_st, h->status + instance, sizeof(int32_t), hipMemcpyDeviceToHost,
                          s));
    VA_HIP(hipStreamSynchronize(s));
    const int64_t stored = lane_st == VA_LANE_OK ? (int64_t)cnt + 1 : (int64_t)cnt;
    *n_rows = stored;
    const int64_t ncopy = std::min<int64_t>(stored, max_rows);
    if (ncopy <= 0) return VA_OK;
    const int n = h->ks.n;
    const size_t bytes = sizeof(double) * (size_t)ncopy * (n + 2);
    if (h->pb.mode == VA_MODE_PERSISTENT) {
        const double *src = h->traj + (size_t)instance * (size_t)(h->pb.max_steps + 1) * (n + 2);
        VA_HIP(hipMemcpyAsync(rows, src, bytes, hipMemcpyDeviceToHost, s));
    } else {
        double *tmp = nullptr;
        VA_HIP(hipMallocAsync((void **)&tmp, bytes, s));
        const int fields = 2 + n;
        hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(ncopy * (n + 2), 256)), dim3(256), 0, s,
                           h->ptraj, instance, h->pb.batch, fields, n, ncopy, tmp);
        hipError_t e = hipMemcpyAsync(rows, tmp, bytes, hipMemcpyDeviceToHost, s);
        hipFreeAsync(tmp, s);
        VA_HIP(e);
    }
    VA_HIP(hipStreamSynchronize(s));
    return VA_OK;
}

}  // extern "C"

// ---------------- va_solve_noalloc ----------------
// Scratch of the entry point that allocates nothing: trajectory rows, work
// counters and per-instance counts.  Each device the library runs on holds its
// own copy until the process ends.  Host code never passes this name to a
// kernel (on the host it is a stand-in, not a device address): the device
// address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[VA_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int VA_NOALLOC_MAX_DEVICES = 64;

struct NoallocDevice {
    std::once_flag once;
    hipError_t err = hipSuccess;
    unsigned char *ptr = nullptr;  // this device's copy of `scratch`
    int cu_count = 0;
};

NoallocDevice g_noalloc[VA_NOALLOC_MAX_DEVICES];

// First call on a device, with that device current: resolve its copy of the
// scratch array and its compute-unit count.
va_status noalloc_device(int dev, NoallocDevice **out)
{
    if (dev < 0 || dev >= VA_NOALLOC_MAX_DEVICES)
        return fail(VA_ERR_UNSUPPORTED, "device %d: va_solve_noalloc serves devices 0..%d", dev,
                    VA_NOALLOC_MAX_DEVICES - 1);
    NoallocDevice &d = g_noalloc[dev];
    std::call_once(d.once, [&d, dev]() {
        void *ptr = nullptr;
        d.err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (d.err == hipSuccess)
            d.err = hipDeviceGetAttribute(&d.cu_count, hipDeviceAttributeMultiprocessorCount, dev);
        if (d.err == hipSuccess) d.ptr = static_cast<unsigned char *>(ptr);
    });
    if (d.err != hipSuccess || !d.ptr)
        return fail(VA_ERR_HIP, "scratch of device %d: %s", dev, hipGetErrorString(d.err));
    *out = &d;
    return VA_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

va_status va_autotune(va_problem *pb, const va_kernel_table *kernels, const double *y0,
                      const double *p, int32_t cost, const double *cost_weights,
                      const double *obs_t, const double *obs_y, int64_t sample, int32_t reps,
                      void *stream)
{
    g_err[0] = 0;
    if (!pb || !y0 || !p || sample < 1 || reps < 1)
        return fail(VA_ERR_INVALID, "va_autotune needs pb, y0, p, sample >= 1 and reps >= 1");
    if (pb->n_obs > 0 && (!obs_t || !obs_y))
        return fail(VA_ERR_INVALID, "an observed problem needs obs_t and obs_y");
    va::KernelSet ks{};
    {
        va_problem probe = *pb;
        if (va_status st = check_problem(&probe, &ks,
                                         reinterpret_cast<const va::KernelSet *>(kernels)))
            return st;
    }
    if (ks.layout == 2) return fail(VA_ERR_UNSUPPORTED, "shared-parameter systems are not tuned");
    int dev = -1;
    VA_HIP(hipGetDevice(&dev));
    hipStream_t s = (hipStream_t)stream;
    const int64_t B = pb->batch, Bs = std::min<int64_t>(B, sample);
    const int n = pb->n_state, np = pb->n_param;
    // the first Bs instances, compacted: [field][Bs]
    double *buf = nullptr;
    const size_t nd = (size_t)Bs * (size_t)(n + np + 1 + n + n + np);
    VA_HIP(hipMalloc((void **)&buf, sizeof(double) * nd));
    double *ys = buf, *ps = ys + (size_t)n * Bs, *J = ps + (size_t)np * Bs, *yT = J + Bs,
           *gy = yT + (size_t)n * Bs, *gp = gy + (size_t)n * Bs;
    hipError_t e = hipMemcpy2DAsync(ys, sizeof(double) * Bs, y0, sizeof(double) * B,
                                    sizeof(double) * Bs, n, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(ps, sizeof(double) * Bs, p, sizeof(double) * B, sizeof(double) * Bs,
                             np, hipMemcpyDeviceToDevice, s);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    auto cleanup = [&]() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        (void)hipFree(buf);
    };
    if (e != hipSuccess) {
        cleanup();
        return fail(VA_ERR_HIP, "va_autotune set-up: %s", hipGetErrorString(e));
    }
    // candidates: threads per block; the instance layouts that exist; checkpoint slots
    std::vector<int32_t> ipws = {0};
    if (!kernels) {
        // one lane per instance as well, where it is a different kernel set
        va::KernelSet k0{}, k64{};
        const bool has0 = va::get_kernels(pb->system, n, pb->tableau, pb->adaptive, 0, &k0);
        const bool has64 = va::get_kernels(pb->system, n, pb->tableau, pb->adaptive, 64, &k64);
        if (has64 && (!has0 || k64.fwd != k0.fwd)) ipws.push_back(64);
    }
    std::vector<int32_t> slots = {pb->checkpoint_slots};
    if (pb->n_obs > 0) slots = {0, 64, 256};
    TuneChoice best{0, 0, pb->checkpoint_slots};
    float best_ms = 1e30f;
    for (int32_t ipw : ipws)
        for (int32_t block : {64, 128, 256})
            for (int32_t sl : slots) {
                va_problem c = *pb;
                c.batch = Bs;
                c.block_size = block;
                c.instances_per_wave = ipw;
                c.checkpoint_slots = sl;
                c.lanes = 0;
                c.flags &= ~VA_FLAG_GRAPH;
                va_handle *h = nullptr;
                if (create_handle(&c, reinterpret_cast<const va::KernelSet *>(kernels), &h) !=
                    VA_OK)
                    continue;  // e.g. checkpoints that do not fit the budget
                auto run = [&]() {
                    return c.n_obs > 0
                               ? va_solve_observed(h, ys, ps, obs_t, obs_y, J, yT, gy, gp, nullptr,
                                                   nullptr, nullptr, stream)
                               : va_solve(h, ys, ps, cost, cost_weights, J, yT, gy, gp, nullptr,
                                          nullptr, nullptr, stream);
                };
                va_status st = run();  // warm-up: also orders the next solve's instances
                if (st == VA_OK && hipEventRecord(ev0, s) != hipSuccess) st = VA_ERR_HIP;
                for (int r = 0; r < reps && st == VA_OK; ++r) st = run();
                float ms = 0.0f;
                if (st == VA_OK && (hipEventRecord(ev1, s) != hipSuccess ||
                                    hipEventSynchronize(ev1) != hipSuccess ||
                                    hipEventElapsedTime(&ms, ev0, ev1) != hipSuccess))
                    st = VA_ERR_HIP;
                va_destroy(h);
                if (st != VA_OK) {
                    cleanup();
                    return st;
                }
                if (ms < best_ms) {
                    best_ms = ms;
                    best = TuneChoice{block, ipw, sl};
                }
            }
    cleanup();
    if (best_ms >= 1e30f) return fail(VA_ERR_UNSUPPORTED, "no launch configuration ran");
    pb->block_size = best.block;
    pb->instances_per_wave = best.ipw;
    pb->checkpoint_slots = best.slots;
    {
        std::lock_guard<std::mutex> lock(g_tune_mtx);
        g_tuned[tune_key(pb, dev)] = best;
    }
    g_err[0] = 0;
    return VA_OK;
}

va_status va_solve_noalloc(const va_problem *pb, const double *y0, const double *p, int32_t cost,
                           const double *cost_weights, double *J, double *yT, double *dJdy0,
                           double *dJdp, int32_t *n_accepted, int32_t *n_rejected,
                           int32_t *lane_status, void *stream)
{
    g_err[0] = 0;
    if (!pb) return fail(VA_ERR_INVALID, "null problem");
    va::KernelSet ks{};
    if (va_status st = check_problem(pb, &ks)) return st;
    if (pb->mode != VA_MODE_PERSISTENT || ks.layout == 2 || !ks.fwd || pb->n_obs > 0 ||
        pb->checkpoint_slots > 0 || (pb->flags & VA_FLAG_GRAPH))
        return fail(VA_ERR_UNSUPPORTED,
                    "va_solve_noalloc runs the persistent terminal-cost solve of a per-instance "
                    "system: no plain mode, shared parameters, observations, checkpoints or "
                    "VA_FLAG_GRAPH");
    if (!y0 || !p || !J || !yT || !dJdy0 || !dJdp)
        return fail(VA_ERR_INVALID, "y0, p, J, yT, dJdy0, dJdp are required");
    if (cost < VA_COST_SUM || cost > VA_COST_LINEAR)
        return fail(VA_ERR_INVALID, "unknown cost %d", cost);
    if (cost == VA_COST_LINEAR && !cost_weights)
        return fail(VA_ERR_INVALID, "VA_COST_LINEAR needs cost_weights");
    const int block = pb->block_size > 0 ? pb->block_size : 256;
    if (block % 64 != 0 || block > 256)
        return fail(VA_ERR_INVALID, "block_size must be a multiple of 64 and <= 256");
    const int64_t B = pb->batch;
    const int n = ks.n;
    // scratch layout: 8 counters | n_acc | n_rej | status | rows [B][max_steps+1][N+2]
    // | for the Lotka-Volterra row kernels, instance-major parameters and dJ/dp
    // sums [B][P] each
    const size_t cnt_bytes = align256(sizeof(int32_t) * (size_t)B);
    const size_t off_nacc = 256, off_nrej = off_nacc + cnt_bytes, off_stat = off_nrej + cnt_bytes;
    const size_t off_traj = off_stat + cnt_bytes;
    const double traj_bytes =
        8.0 * (double)B * ((double)pb->max_steps + 1.0) * (double)(n + 2);
    const bool im = ks.layout == 1;
    const size_t im_bytes = im ? align256(sizeof(double) * (size_t)B * ks.np) : 0;
    const size_t off_pim = off_traj + align256((size_t)traj_bytes), off_pbar = off_pim + im_bytes;
    const double need = (double)off_pbar + (double)im_bytes;
    if (need > (double)VA_NOALLOC_SCRATCH_BYTES)
        return fail(VA_ERR_OUT_OF_MEMORY,
                    "va_solve_noalloc needs %.0f bytes of scratch (B * (max_steps+1) * (N+2) * "
                    "8 bytes of rows, 12 bytes per instance and, for Lotka-Volterra, 16 P bytes "
                    "per instance) and holds %lld: use fewer instances or a smaller max_steps, "
                    "or va_create + va_solve",
                    need, (long long)VA_NOALLOC_SCRATCH_BYTES);
    if (pb->trajectory_budget_bytes > 0 && traj_bytes > (double)pb->trajectory_budget_bytes)
        return fail(VA_ERR_OUT_OF_MEMORY,
                    "trajectory store of %.3g bytes exceeds trajectory_budget_bytes %lld",
                    traj_bytes, (long long)pb->trajectory_budget_bytes);
    int dev = -1;
    if (hipError_t de = hipGetDevice(&dev); de != hipSuccess)
        return fail(VA_ERR_HIP, "no HIP device: %s", hipGetErrorString(de));
    NoallocDevice *d = nullptr;
    if (va_status st = noalloc_device(dev, &d)) return st;
    int32_t *counters = reinterpret_cast<int32_t *>(d->ptr);
    int32_t *nacc = reinterpret_cast<int32_t *>(d->ptr + off_nacc);
    int32_t *nrej = reinterpret_cast<int32_t *>(d->ptr + off_nrej);
    int32_t *stat = reinterpret_cast<int32_t *>(d->ptr + off_stat);
    double *traj = reinterpret_cast<double *>(d->ptr + off_traj);
    double *pim = im ? reinterpret_cast<double *>(d->ptr + off_pim) : nullptr;
    double *pbar_im = im ? reinterpret_cast<double *>(d->ptr + off_pbar) : nullptr;

    hipStream_t s = (hipStream_t)stream;
    const int per_block = ks.slots(block);
    int64_t lanes_f = pb->lanes, lanes_r = pb->lanes;
    if (lanes_f <= 0) {
        lanes_f = (int64_t)std::max(1, ks.occ_fwd(block)) * d->cu_count * per_block;
        lanes_r = (int64_t)std::max(1, ks.occ_rev(block)) * d->cu_count * per_block;
    }
    lanes_f = std::min<int64_t>(lanes_f, B);
    lanes_r = std::min<int64_t>(lanes_r, B);
    // the counters start at zero: a fill kernel (no memset node under graph capture)
    if (va_status st = fill_i32(counters, 0, 8, s)) return st;
    if (yT != y0) {
        hipLaunchKernelGGL(k_copy_f64, dim3(blocks_for((int64_t)n * B, 256)), dim3(256), 0, s, yT,
                           y0, (int64_t)n * B);
        VA_HIP(hipGetLastError());
    }
    va::SolveParams sp;
    sp.B = B;
    sp.cap = pb->max_steps;
    sp.t0 = pb->t0;
    sp.t1 = pb->t1;
    sp.dt0 = pb->dt0;
    sp.atol = pb->abs_tol;
    sp.rtol = pb->rel_tol;
    sp.kconst = pb->system_constant;
    sp.max_attempts = pb->max_attempts;
    sp.adaptive = pb->adaptive;
    if (im) {
        if (va_status st = to_instance_major(p, pim, ks.np, B, s)) return st;
    }
    va::FwdArgs a{};
    a.sp = sp;
    a.y = yT;
    a.p = im ? pim : p;
    a.traj = traj;
    a.n_acc = nacc;
    a.n_rej = nrej;
    a.status = stat;
    a.counter = counters + 0;
    a.max_steps_seen = counters + 2;
    VA_HIP(ks.fwd(a, blocks_for(lanes_f, per_block), block, s));
    va::RevArgs r{};
    r.sp = sp;
    r.order = nullptr;
    r.p = im ? pim : p;
    r.traj = traj;
    r.n_acc = nacc;
    r.status = stat;
    r.cost = cost;
    r.cost_w = cost_weights;
    r.yT = yT;
    r.J = J;
    r.lambda = dJdy0;
    r.mu = im ? pbar_im : dJdp;
    r.nout = 1;
    r.counter = counters + 1;
    VA_HIP(ks.rev(r, blocks_for(lanes_r, per_block), block, s));
    if (im) {
        if (va_status st = from_instance_major(pbar_im, dJdp, ks.np, B, 1, cost == 0, s))
            return st;
    }
    if (n_accepted || n_rejected || lane_status) {
        hipLaunchKernelGGL(k_copy_counts, dim3(blocks_for(B, 256)), dim3(256), 0, s, n_accepted,
                           n_rejected, lane_status, nacc, nrej, stat, B);
        VA_HIP(hipGetLastError());
    }
    return VA_OK;
}

va_status va_step_digest(va_handle *h, uint64_t *digest, void *stream)
{
    va_status st = check_handle(h);
    if (st) return st;
    if (h->obs_ckpt)
        return fail(VA_ERR_STATE,
                    "this handle keeps checkpoints instead of a trajectory store: no digest");
    if (!h->have_forward) return fail(VA_ERR_STATE, "va_step_digest needs a forward first");
    if (!digest) return fail(VA_ERR_INVALID, "null digest");
    hipStream_t s = (hipStream_t)stream;
    const int64_t B = h->pb.batch;
    const bool plain = h->pb.mode == VA_MODE_PLAIN;
    hipLaunchKernelGGL(k_step_digest, dim3(blocks_for(B, 256)), dim3(256), 0, s,
                       plain ? h->ptraj : h->traj, B, (int64_t)h->pb.max_steps, h->ks.n + 2,
                       plain ? 1 : 0, h->n_acc, h->status, digest);
    VA_HIP(hipGetLastError());
    return VA_OK;
}

va_status va_max_steps_seen(va_handle *h, int32_t *out, void *stream)
{
    va_status st = check_handle(h);
    if (st) return st;
    if (!out) return fail(VA_ERR_INVALID, "null out");
    return read_max_steps(h, out, (hipStream_t)stream);
}

}  // extern "C"
