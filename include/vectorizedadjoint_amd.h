/*
 * vectorizedadjoint_amd.h -- C ABI of the MI355X (gfx950) batched explicit
 * Runge-Kutta forward sweep and discrete-adjoint reverse sweep, FP64.
 *
 * The reference (RuiMartins1996/VectorizedAdjoint) is a header-only C++ library
 * whose entry points are templates in namespace vectorizedadjoint.  This ABI
 * is what those entry points bind to here; the C++ drop-in headers under
 * include/vectorizedadjoint/ keep the reference's names and call through it.
 * Each entry point below cites the reference interface it replaces.
 *
 * Conventions
 *   - One handle = one problem (system, tableau, tolerances, time span) over a
 *     batch of B independent instances (parameter / initial-condition sets).
 *     A handle owns its device trajectory store; it is not thread-safe.
 *   - Every array argument is a DEVICE pointer (hipMalloc'd memory of the
 *     current device), laid out structure-of-arrays with the instance index
 *     fastest: y[i][b] at y[i*B + b], p[k][b] at p[k*B + b],
 *     lambda[o][i][b] at lambda[(o*N + i)*B + b].
 *   - `stream` is a hipStream_t (NULL = the null stream).  Calls are
 *     stream-ordered and return without synchronising, except where noted.
 *   - Functions return va_status and never throw; va_last_error() gives a
 *     message for the calling thread.
 */
#ifndef VECTORIZEDADJOINT_AMD_H
#define VECTORIZEDADJOINT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VA_ABI_VERSION 2

typedef enum va_status {
    VA_OK = 0,
    VA_ERR_INVALID = 1,        /* bad argument / problem description            */
    VA_ERR_UNSUPPORTED = 2,    /* no kernels for this (system, N, tableau)     */
    VA_ERR_HIP = 3,            /* HIP runtime error                             */
    VA_ERR_OUT_OF_MEMORY = 4,
    VA_ERR_STATE = 5,          /* call order violated (e.g. adjoint before forward) */
    VA_ERR_LANE_FAILURE = 6    /* reserved: some instance failed (see lane status) */
} va_status;

/* Built-in systems (device right-hand sides with hand-written VJPs). */
typedef enum va_system {
    VA_SYSTEM_VAN_DER_POL = 1,         /* examples/VanDerPol/main.cpp:39-43; p = (mu)          */
    VA_SYSTEM_OSCILLATOR = 2,          /* damped oscillator, p = (k, c)                         */
    VA_SYSTEM_OSCILLATOR_FIXED_K = 3,  /* examples/HarmonicOscillator/main.cpp:25-29; p = (mu),
                                          k = va_problem.system_constant                      */
    VA_SYSTEM_LOTKA_VOLTERRA = 4,      /* examples/GeneralizedLotkaVolterra/main.cpp:105-119;
                                          p = (r[N], A[N][N] row-major), P = N + N*N          */
    VA_SYSTEM_LOTKA_VOLTERRA_SHARED = 5, /* same system with A shared by the batch: p = r[N],
                                          P = N; A passed to va_solve_shared; N = 128     */
    VA_SYSTEM_USER = 100               /* a system of the caller's: va_create_with_kernels  */
} va_system;

/* Tableaus: ButcherTable.hpp:39-251 (Euler, RK4, Cash-Karp 5(4), Fehlberg 7(8))
 * plus Dormand-Prince 5(4), which the reference rejects (ButcherTable.hpp:247-250). */
typedef enum va_tableau {
    VA_TABLEAU_EULER = 1,
    VA_TABLEAU_RK4 = 2,
    VA_TABLEAU_CASH_KARP54 = 3,
    VA_TABLEAU_DOPRI5 = 4,
    VA_TABLEAU_FEHLBERG78 = 5
} va_tableau;

/* Terminal cost functionals evaluated on the device by va_solve. */
typedef enum va_cost {
    VA_COST_SUM = 1,           /* J = sum_i y_i(T)                                   */
    VA_COST_SQUARED_NORM = 2,  /* J = sum_i y_i(T)^2                                 */
    VA_COST_LINEAR = 3         /* J = sum_i w_i y_i(T), w = cost_weights (device, N) */
} va_cost;

typedef enum va_mode {
    VA_MODE_PERSISTENT = 0,    /* persistent kernels with lane refill (default)     */
    VA_MODE_PLAIN = 1          /* one launch per accepted step, no refill, full
                                  (t, y, stage K) store: the baseline GPU form       */
} va_mode;

/* Per-instance status written by va_forward / va_solve. */
typedef enum va_lane_status {
    VA_LANE_OK = 0,
    VA_LANE_MAX_FAIL = 1,      /* more than max_attempts tries in one step
                                  (odeint failed_step_checker -> step_adjustment_error) */
    VA_LANE_NONFINITE = 2,     /* y(T) not finite                                    */
    VA_LANE_MAX_STEPS = 3      /* more accepted steps than max_steps                 */
} va_lane_status;

typedef struct va_problem {
    int32_t abi_version;       /* VA_ABI_VERSION                                      */
    int32_t system;            /* va_system                                           */
    int32_t tableau;           /* va_tableau                                          */
    int32_t adaptive;          /* 1: controlled loop (runge_kutta.hpp:74-118),
                                  0: fixed step (runge_kutta.hpp:36-72)               */
    int32_t n_state;           /* N (Nin)                                             */
    int32_t n_param;           /* P (Npar), must match the system                     */
    int64_t batch;             /* B instances                                         */
    double t0, t1, dt0;        /* start/end time, initial (or fixed) step             */
    double abs_tol, rel_tol;   /* make_controlled(abs, rel)                           */
    double system_constant;    /* non-differentiated constant (VA_SYSTEM_OSCILLATOR_FIXED_K) */
    int32_t max_attempts;      /* tries per step before VA_LANE_MAX_FAIL (odeint: 500) */
    int32_t max_steps;         /* accepted-step capacity of each stored trajectory    */
    int32_t mode;              /* va_mode                                             */
    int32_t block_size;        /* threads per block (64..256, multiple of 64); 0 = the
                                  va_autotune choice for this problem shape when this
                                  process tuned one, else 256                          */
    int32_t instances_per_wave;  /* 64: one lane per instance; 64/N: one group of N
                                    lanes per instance (Lotka-Volterra); 0 = the
                                    va_autotune choice when tuned, else the group
                                    layout where the system has one                  */
    int32_t checkpoint_slots;  /* observed solves: 0 = store every step while the store
                                  fits trajectory_budget_bytes, else checkpoint; > 0 =
                                  checkpoint with this many slots per instance (even) */
    int64_t lanes;             /* instances in flight in the persistent kernels, 0 = all
                                  resident slots; fewer than B means lanes are refilled */
    int64_t trajectory_budget_bytes; /* 0 = unlimited; va_create fails if the store
                                        (or, for observed solves, the checkpoint store)
                                        would exceed it                               */
    int32_t n_obs;             /* observation times of va_solve_observed (0 = none)   */
    int32_t flags;             /* VA_FLAG_* bits                                      */
} va_problem;

/* va_problem.flags: va_solve (persistent mode) records its launch sequence once in a
 * HIP graph per distinct set of argument pointers and replays it on the caller's
 * stream afterwards (one graph launch instead of about ten kernel launches). */
#define VA_FLAG_GRAPH 1
/* va_problem.flags: by default a persistent forward whose batch exceeds the instances
 * in flight takes its instances longest-first by the attempt counts of the handle's
 * previous forward (index order the first time), so that instances that ran long
 * last time start first; results do not depend on the order.  With this flag every
 * forward takes instances in index order. */
#define VA_FLAG_COLD_ORDER 2
/* va_problem.flags: an observed (or checkpointed) problem keeps two levels of
 * checkpoints even where one fits trajectory_budget_bytes.  Without the flag two
 * levels are used only when one does not fit: the forward keeps level-1 checkpoints
 * (online, doubling their spacing as an instance runs long); the reverse replays a
 * level-1 segment once to keep about sqrt(segment) level-2 checkpoints per lane,
 * then recomputes and reverses the rows sub-segment by sub-segment.  Memory per
 * instance: slots checkpoints; per lane in flight: slots2 checkpoints and
 * ceil(stride_max / slots2) rows instead of stride_max rows; work: two extra forward
 * passes instead of one.  Results are the same bit for bit. */
#define VA_FLAG_CHECKPOINT_2LEVEL 4

typedef struct va_handle va_handle;

/* Fills defaults: odeint limits (500 tries), dt0 = 1e-3, tolerances 1e-6. */
void va_problem_init(va_problem *pb);

/* Driver(Nin, Nout, Npar) + constructDriverButcherTableau + recordDriverRHSFunction
 * (lib/include/Driver.hpp:39-42,87-100): validates the problem, selects the
 * device kernels for (system, N, tableau) and allocates the trajectory store
 * (B * (max_steps+1) rows).  */
va_status va_create(const va_problem *pb, va_handle **out);

/* Kernels compiled for a system of the caller's (opaque).  Made in the caller's own
 * hipcc translation unit by VA_DEVICE_SYSTEM (include/vectorizedadjoint/device_system.hpp)
 * from a device functor with the right-hand side and its vector-Jacobian products. */
typedef struct va_kernel_table va_kernel_table;

/* recordDriverRHSFunction of an arbitrary functor (lib/include/Driver.hpp:95-100,
 * lib/include/AadData.hpp:124-171 -- the reference tapes any templated functor with
 * AADC): va_create for pb->system = VA_SYSTEM_USER with the kernels of the caller's
 * system.  pb->n_state and pb->n_param must be the functor's N and P; one lane per
 * instance; every mode and entry point of a one-lane built-in system (va_forward,
 * va_adjoint, va_sensitivity_matrix, va_solve, va_solve_observed, the plain form) is
 * available.  A table built against other headers is rejected. */
va_status va_create_with_kernels(const va_problem *pb, const va_kernel_table *kernels,
                                 va_handle **out);

/* delete_driver_handle (lib/include/Driver.hpp:81-85). */
void va_destroy(va_handle *h);

/* runge_kutta(stepper, system, x0, p, t0, t1, dt, driver)
 * (lib/include/runge_kutta.hpp:47-59 -> detail/runge_kutta.hpp:36-118):
 * y: in y0 [N][B], out y(T) (in place, as x0 in the reference); p: [P][B].
 * Stores (t_n, y_n) before every accepted step and at the end, with h_n = t_{n+1} - t_n
 * (the reference's GetDt, StateStorage.hpp:22; its reverse sweep uses it).  n_accepted,
 * n_rejected, lane_status: [B] int32, each may be NULL. */
va_status va_forward(va_handle *h, double *y, const double *p, int32_t *n_accepted,
                     int32_t *n_rejected, int32_t *lane_status, void *stream);

/* setCostGradients + adjointSolve (lib/include/Driver.hpp:103-114,
 * lib/include/backpropagation.hpp:18-43 -> detail/backpropagation.hpp:280-348)
 * over the trajectory stored by the last va_forward on this handle:
 * lambda [nout][N][B]: in dJ/dy(T), out dJ/dy0 (overwritten);
 * mu [nout][P][B]: dJ/dp accumulated (+=).  Instances whose forward failed get NaN. */
va_status va_adjoint(va_handle *h, const double *p, int32_t nout, double *lambda, double *mu,
                     void *stream);

/* computeSensitivityMatrix (lib/include/backpropagation.hpp:67-87,
 * detail/backpropagation.hpp:373-434): S [N][P][B], S[i][k][b] = d y_i(T) / d p_k.
 * Allocates a temporary seed buffer of N*N*B doubles. */
va_status va_sensitivity_matrix(va_handle *h, const double *p, double *S, void *stream);

/* Forward + terminal cost + adjoint in one call (no reference counterpart: the
 * reference's drivers chain runge_kutta / setCostGradients / adjointSolve with
 * the cost written by hand, examples/VanDerPol/main.cpp:98-117).  Outputs: J [B], yT [N][B],
 * dJdy0 [N][B], dJdp [P][B] (overwritten), counts and status [B] (may be NULL).
 * cost_weights: [N] device array for VA_COST_LINEAR, else NULL. */
va_status va_solve(va_handle *h, const double *y0, const double *p, int32_t cost,
                   const double *cost_weights, double *J, double *yT, double *dJdy0,
                   double *dJdp, int32_t *n_accepted, int32_t *n_rejected, int32_t *lane_status,
                   void *stream);

/* Forward + observation cost + adjoint (no reference counterpart: the reference's
 * cost is terminal only, examples/VanDerPol/main.cpp:98-117).  The controlled loop stops at each
 * observation time obs_t[k] (device, [n_obs], increasing, inside [t0, t1]) the way
 * it stops at t1 (detail/runge_kutta.hpp:98-101); after such a shortened step the
 * next step size is the larger of the controller's proposal and the step size
 * before shortening (Boost.odeint integrate_times).  J [B] = sum_k |y(t_k) -
 * obs_y[k]|^2 with obs_y [n_obs][N] (device, shared by the batch); dJdy0 [N][B],
 * dJdp [P][B] overwritten; yT [N][B].  When the trajectory store does not fit
 * trajectory_budget_bytes the forward keeps checkpoints instead (a fixed number of
 * checkpoint slots per instance whose spacing doubles as the instance runs long)
 * and the reverse recomputes each segment from its checkpoint, bit-identically.  Adaptive
 * one-lane systems only (Van der Pol, oscillators); needs pb.n_obs > 0. */
va_status va_solve_observed(va_handle *h, const double *y0, const double *p, const double *obs_t,
                            const double *obs_y, double *J, double *yT, double *dJdy0,
                            double *dJdp, int32_t *n_accepted, int32_t *n_rejected,
                            int32_t *lane_status, void *stream);

/* Forward + terminal cost + adjoint for VA_SYSTEM_LOTKA_VOLTERRA_SHARED (no reference
 * counterpart: the reference's parameters are per instance, examples/
 * GeneralizedLotkaVolterra/main.cpp:105-119).  p = r [N][B]; p_shared = A [N][N]
 * row-major (device), common to the batch.  Outputs as va_solve, dJdp = dJ/dr
 * [N][B], and dJdp_shared [N][N] = sum over the instances whose forward ended OK
 * of dJ/dA (summed in a fixed order).  The stage right-hand sides and adjoints run
 * as FP64 MFMA products over tiles of 16 instances. */
va_status va_solve_shared(va_handle *h, const double *y0, const double *p, const double *p_shared,
                          int32_t cost, const double *cost_w, double *J, double *yT,
                          double *dJdy0, double *dJdp, double *dJdp_shared, int32_t *n_accepted,
                          int32_t *n_rejected, int32_t *lane_status, void *stream);

/* Driver::GetT/GetTime/GetDt/GetState (lib/include/Driver.hpp:53-65,
 * StateStorage.hpp:17-42): copies the rows stored for one instance by the last
 * va_forward into HOST memory, row n = (t_n, h_n, y_n[0..N-1]) with h_n = t_{n+1} - t_n
 * (StateStorage.hpp:22; 0 on the final row).  At most max_rows rows
 * are copied; *n_rows receives the number stored (accepted steps + 1).
 * Synchronises the stream. */
va_status va_trajectory(va_handle *h, int64_t instance, double *rows, int64_t max_rows,
                        int64_t *n_rows, void *stream);

/* Launch-shape autotuner (no reference counterpart: the reference has no launch
 * shape).  Times the launch configurations of pb -- threads per block (64, 128, 256),
 * the instance layouts the system has (one lane per instance, one group of N lanes per
 * instance) and, for observed problems (pb->n_obs > 0, obs_t / obs_y given), the
 * checkpoint slots (0, 64, 256) -- on the first min(sample, B) instances of the
 * caller's own y0 [N][B] and p [P][B], each as a warm-up solve and `reps` timed solves,
 * and writes the fastest into pb->block_size, pb->instances_per_wave and
 * pb->checkpoint_slots.  Every configuration computes the same results.  The choice is
 * also kept for this process per device and problem shape (system, N, tableau,
 * adaptive, n_obs, batch rounded up to a power of two), and va_create applies it to
 * problems of that shape whose block_size / instances_per_wave are 0.  kernels: the
 * table of a VA_SYSTEM_USER problem, else NULL.  Allocates and synchronises the
 * stream (a set-up call). */
va_status va_autotune(va_problem *pb, const va_kernel_table *kernels, const double *y0,
                      const double *p, int32_t cost, const double *cost_weights,
                      const double *obs_t, const double *obs_y, int64_t sample, int32_t reps,
                      void *stream);

/* Scratch held by va_solve_noalloc on each device it runs on (until the process ends). */
#define VA_NOALLOC_SCRATCH_BYTES (32ll << 20)

/* va_solve without a handle and without allocating device memory.  No counterpart in
 * the reference: it is for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The problem is
 * described by pb as for va_create (persistent mode, a per-instance system, no
 * observations, checkpoints or VA_FLAG_GRAPH); the other arguments and outputs are
 * those of va_solve.  The trajectory rows, work counters and per-instance counts
 * live in a library-owned device array of VA_NOALLOC_SCRATCH_BYTES per device: a
 * problem needing more than B * (max_steps+1) * (N+2) * 8 + 12 B + 256 bytes of it
 * is rejected with VA_ERR_OUT_OF_MEMORY (GLV N=10 with max_steps 1024: up to 340
 * instances).  Every call on a device shares that device's array and it keeps the
 * last call's contents: calls that overlap on one device (other streams, other
 * threads) race and give wrong results.  Same kernels, and the same results bit for
 * bit, as va_solve. */
va_status va_solve_noalloc(const va_problem *pb, const double *y0, const double *p, int32_t cost,
                           const double *cost_weights, double *J, double *yT, double *dJdy0,
                           double *dJdp, int32_t *n_accepted, int32_t *n_rejected,
                           int32_t *lane_status, void *stream);

/* Per-instance digest of the step sequence stored by the last forward (no reference
 * counterpart: it lets a caller check that two solves -- this library's and the
 * reference's, or two of this library's -- took the same accepted steps without
 * copying trajectories).  digest: [B] device uint64.  For an instance whose forward
 * ended OK: d = 0x9E3779B97F4A7C15, then for each stored point n = 0..n_accepted,
 * d = mix(d ^ bits(t_n)), d = mix(d ^ bits(h_n)), where mix is the SplitMix64
 * finaliser, bits the IEEE-754 bit pattern, h_n = t_{n+1} - t_n (Driver::GetDt,
 * lib/include/StateStorage.hpp:22) and h = 0 on the final point; 0 otherwise.
 * Not available on handles that keep checkpoints instead of a store. */
va_status va_step_digest(va_handle *h, uint64_t *digest, void *stream);

/* Largest accepted-step count of the last forward pass (synchronises the stream). */
va_status va_max_steps_seen(va_handle *h, int32_t *out, void *stream);

/* Message of the last failing call on this thread ("" if none). */
const char *va_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
