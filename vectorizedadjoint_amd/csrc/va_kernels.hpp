// va_kernels.hpp -- host-side launch interface of the HIP kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace va {

struct SolveParams {
    int64_t B;
    int64_t cap;          // accepted-step capacity per instance
    double t0, t1, dt0, atol, rtol, kconst;
    int32_t max_attempts;
    int32_t adaptive;
};

// Persistent forward: each lane takes instances from *counter until B is reached.
struct FwdArgs {
    SolveParams sp;
    double *y;             // [N][B] in y0, out y(T)
    const double *p;       // [P][B]; Lotka-Volterra persistent kernels: [B][P]
    double *traj;          // [B][cap+1][N+1] rows (t, y)
    int32_t *n_acc, *n_rej, *status;
    int32_t *counter;
    int32_t *max_steps_seen;
    const int32_t *order = nullptr;  // [B] or null: the k-th instance taken is order[k]
    const int32_t *split = nullptr;  // null: queue positions [0, B); else part 0 takes
    int32_t part = 0;                // [0, *split), part 1 takes [*split, B)
};

// Persistent reverse over the stored trajectory.
//   cost > 0: lambda computed from y(T) (read from yT) and written to dJdy0,
//             dJdp = sum p_bar, J written.
//   cost == 0: lambda [nout][N][B] in/out, mu [nout][P][B] += (reference semantics).
// The Lotka-Volterra persistent kernels take p as [B][P], store the dJ/dp sums
// [nout][B][P] in mu as they are, for either cost (the caller applies them),
// and set pbar_set[b] for each instance they wrote (when pbar_set is given);
// the long-instance kernels take p and mu as above.
struct RevArgs {
    SolveParams sp;
    const double *p;
    const double *traj;
    const int32_t *n_acc, *status;
    int32_t cost;
    const double *cost_w;  // [N] for the linear cost
    const double *yT;      // [N][B]
    double *J;
    double *lambda;        // [nout][N][B]
    double *mu;            // [nout][P][B]
    int32_t nout;
    int32_t *counter;
    const int32_t *order;  // [B] or null: the k-th instance taken is order[k]
    const int32_t *split = nullptr;  // as in FwdArgs
    int32_t part = 0;
    int32_t *pbar_set = nullptr;     // [B] or null
};

// Plain form: one launch per accepted step, state in HBM, no refill, full
// trajectory store [step][field][B] with fields (t, y[N], K[S][N]).
struct PlainState {
    double *t, *dt;        // [B]
    double *k0;            // [N][B] f(y_n) carried for FSAL tableaus
    int32_t *done;         // [B]
    int32_t *active;       // 1 counter
    int32_t *max_steps_seen;
    double *pbar;          // [P][B] reverse accumulator
};

struct PlainFwdArgs {
    SolveParams sp;
    double *y;
    const double *p;
    double *traj;
    int32_t *n_acc, *n_rej, *status;
    PlainState st;
    int32_t step;
};

struct PlainRevArgs {
    SolveParams sp;
    const double *p;
    const double *traj;
    const int32_t *n_acc, *status;
    double *lambda;        // [N][B] seed in/out
    PlainState st;
    int32_t k;             // reverse launch index
};

// Observation cost (forced stops at obs_t, running cost, lambda jumps) with the
// trajectory either stored in full or kept as checkpoints.
struct ObsStore {
    int32_t ckpt;          // 0: full rows (t, h, y) in traj; 1: checkpoints in ck
    double *traj;          // [B][cap+1][N+2]
    double *ck;            // [B][slots][N+3] (t, dt to try, observations taken, y)
    int32_t slots;         // checkpoints per instance (even)
    int32_t stride_max;    // largest checkpoint spacing; slots * stride_max >= cap
    int32_t *stride;       // [B] spacing each instance ended with
    double *seg;           // [lanes][seg_rows][N+1] (h, y) rows recomputed by the reverse
    int32_t *obs_row;      // [B][K] stored-point index of each observation
    // two levels: a segment longer than seg_rows is replayed once to store level-2
    // checkpoints every `sub` steps (sub <= seg_rows) in ck2, then recomputed
    // sub-segment by sub-segment into seg
    int32_t levels = 1;
    int32_t seg_rows = 0;  // rows of the per-lane row buffer (stride_max with one level)
    int32_t slots2 = 0;    // level-2 checkpoints per lane
    double *ck2 = nullptr; // [lanes][slots2][N+3], lanes interleaved like seg
};

struct ObsFwdArgs {
    SolveParams sp;
    double *y;             // [N][B] in y0, out y(T)
    const double *p;
    int32_t *n_acc, *n_rej, *status;
    int32_t *counter;
    int32_t *max_steps_seen;
    const int32_t *order = nullptr;  // [B] or null: the k-th instance taken is order[k]
    const double *obs_t;   // [K]
    const double *obs_y;   // [K][N]
    int32_t K;
    double *J;             // [B]
    ObsStore st;
};

struct ObsRevArgs {
    SolveParams sp;
    const double *p;
    const double *yT;      // [N][B]
    const int32_t *n_acc, *status;
    const double *obs_y;   // [K][N]
    const double *obs_t;
    int32_t K;
    double *dJdy0;         // [N][B]
    double *dJdp;          // [P][B]
    int32_t *counter;
    const int32_t *order;
    int32_t cost;          // > 0: terminal cost added to the seed, J written
    const double *cost_w;
    double *J;
    ObsStore st;
};

// Lotka-Volterra with A shared by the batch (tiles of instances on MFMA).
struct SharedFwdArgs {
    SolveParams sp;
    double *y;             // [N][B] in y0, out y(T)
    const double *r;       // [N][B]
    const double *A;       // [N][N] row-major, shared
    double *traj;          // [B][cap+1][N+2]
    int32_t *n_acc, *n_rej, *status;
    int32_t *max_steps_seen;
};

struct SharedRevArgs {
    SolveParams sp;
    const double *r;
    const double *A;
    const double *traj;
    const int32_t *n_acc, *status;
    const int32_t *order;  // [B] longest first
    const double *lambda;  // [N][B] dJ/dy(T) (may alias dJdy0)
    double *dJdy0;         // [N][B]
    double *dJdr;          // [N][B]
    double *abar_part;     // [grid][N][N] per-workgroup partial dJ/dA
    double *gscr;          // [grid][waves][2][S][64][4] growth products and stage inputs of the step being reversed
};

// Layout version of KernelSet: kernel tables made in a caller's translation unit
// (VA_DEVICE_SYSTEM) are accepted only with the library's size and version.
constexpr int32_t VA_KERNEL_TABLE_VERSION = 1;

struct KernelSet {
    int32_t table_size = (int32_t)sizeof(KernelSet);
    int32_t table_version = VA_KERNEL_TABLE_VERSION;
    int n, np, s_rev, s_fwd;
    int flop_rhs, flop_vjp;
    hipError_t (*fwd)(const FwdArgs &, int grid, int block, hipStream_t);
    hipError_t (*rev)(const RevArgs &, int grid, int block, hipStream_t);
    // forward then reverse of each long-running instance, one block of three waves per
    // instance: grid = instances, block = 192 (null where absent)
    hipError_t (*long_fused)(const FwdArgs &, const RevArgs &, int grid, int block, hipStream_t);
    // the same instances' forward alone (one wave per instance) and reverse
    // alone (block 192: three waves per instance, one seed; block 256: one wave
    // per instance, any seeds) (null where absent)
    hipError_t (*long_fwd)(const FwdArgs &, int grid, int block, hipStream_t);
    hipError_t (*long_rev)(const RevArgs &, int grid, int block, hipStream_t);
    hipError_t (*plain_fwd)(const PlainFwdArgs &, int grid, int block, hipStream_t);
    hipError_t (*plain_rev)(const PlainRevArgs &, int grid, int block, hipStream_t);
    int (*occ_fwd)(int block);   // resident blocks per CU of the persistent kernels
    int (*occ_rev)(int block);
    int (*slots)(int block);     // instances one block holds at a time
    int layout;                  // 0: one lane per instance, 1: one group of N lanes per instance
    // observation-cost kernels (null where the system/tableau has none)
    hipError_t (*obs_fwd)(const ObsFwdArgs &, int grid, int block, hipStream_t);
    hipError_t (*obs_rev)(const ObsRevArgs &, int grid, int block, hipStream_t);
    int (*occ_obs_fwd)(int block);
    int (*occ_obs_rev)(int block);
    // shared-parameter kernels (layout 2), one workgroup per CU
    hipError_t (*shared_fwd)(const SharedFwdArgs &, int grid, hipStream_t);
    hipError_t (*shared_rev)(const SharedRevArgs &, int grid, hipStream_t);
    int shared_tile;
    int shared_waves;            // waves per workgroup of the shared-parameter kernels
};

hipError_t launch_glvs_reduce(const double *part, int nparts, double *out, hipStream_t s);

// Returns false when (system, n, tableau, adaptive) has no compiled kernels.
// instances_per_wave: 64 selects one lane per instance; 64/N (or 0 = default)
// selects the row-per-lane layout where the system has one.
bool get_kernels(int system, int n, int tableau, int adaptive, int instances_per_wave,
                 KernelSet *out);

hipError_t launch_plain_finish(double *lambda_out, const double *lambda_in, double *mu,
                               const double *pbar, const int32_t *status, int64_t B, int n, int np,
                               int grid, int block, hipStream_t s);
hipError_t launch_plain_cost(int cost, const double *cost_w, const double *yT, double *J,
                             double *lambda, const int32_t *status, int64_t B, int n, int grid,
                             int block, hipStream_t s);

}  // namespace va
