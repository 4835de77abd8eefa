/*
 * va_oracle.h -- CPU restatement (plain C99) of VectorizedAdjoint's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only to compare against or to time it.  The product never
 * links it and never falls back to it.
 *
 * What it restates (file:line into the reference repository):
 *   forward, controlled loop ....... lib/include/detail/runge_kutta.hpp:74-118
 *   forward, fixed-step loop ....... lib/include/detail/runge_kutta.hpp:36-72
 *   trajectory store (t_n, y_n) .... lib/include/StateStorage.hpp:4-43, plus the exact
 *                                    accepted h_n (the reference uses t[n+1]-t[n], :22)
 *   Butcher tableaus ............... lib/include/ButcherTable.hpp:39-251
 *   discrete adjoint ............... lib/include/detail/backpropagation.hpp:24-348
 *   seed batching (lambda overwrite, mu +=) ... detail/backpropagation.hpp:280-348
 *   sensitivity matrix ............. detail/backpropagation.hpp:352-434
 *   example systems ................ examples/VanDerPol/main.cpp:39-43,
 *                                    examples/HarmonicOscillator/main.cpp:25-29,
 *                                    examples/GeneralizedLotkaVolterra/main.cpp:105-119
 *   step controller ................ Boost.odeint 1.74.0 controlled_runge_kutta,
 *       default_error_checker, default_step_adjuster, failed_step_checker,
 *       less_with_sign (third party, not vendored; called at
 *       detail/runge_kutta.hpp:55,85-108).  Restated from its published source.
 *
 * Rounding sequence.  Every multiply-add of the step arithmetic is written as
 * an explicit fma() and the file is compiled with -ffp-contract=off, so the
 * exact IEEE operation sequence is fixed and documented; the HIP kernels
 * follow the same sequence, which makes forward trajectories comparable bit
 * for bit.  The controller's pow(err, -1/q) is orc_root(err, q): a polynomial
 * start and two Newton steps for the q-th root (see va_oracle.c), within about
 * one ulp of the exact root and 1.5 ulp of the C library pow on the
 * controller's domain.
 *
 * How each part is pinned (tests/ file: what fixes it independently):
 *   tableaus ........ test_oracle_tableau: rooted-tree order conditions to the
 *                     stated order, row sums, DOPRI5 FSAL row, observed
 *                     convergence order of the fixed-step loop
 *   rhs / vjp ....... test_oracle_math: the reference functors' closed forms,
 *                     complex-step derivatives; the alphasfile fixtures' layout
 *   forward ......... test_oracle_forward: oscillator closed form, SciPy DOP853
 *                     at rtol 1e-13 for Van der Pol and Lotka-Volterra, fixed
 *                     grid t0 + n dt, FSAL evaluation count, status paths
 *   controller ...... test_oracle_forward::test_controller_invariants (accepted
 *                     err <= 1, shrink in [0.2, 1), growth bounds, dead band,
 *                     landing on t1) and the shrink-exponent mutant; its root
 *                     x^(-1/q): test_oracle_math against the exact root (40-digit
 *                     Decimal, <= 1 ulp) and the C library pow (<= 4 ulp)
 *   adjoint ......... test_oracle_adjoint: complex-step derivative of the frozen
 *                     step map, oscillator transfer matrix, central differences,
 *                     the reference's seed semantics, bitwise replay
 *   mutants ......... test_oracle_mutants: six planted mistakes, each caught
 * Parity unpinned: the exact accept/reject decisions of Boost.odeint 1.74
 * itself (Boost is not installed, and the reference holds no recorded
 * outputs); the restated controller is pinned only by the invariants above.
 */
#ifndef VA_ORACLE_H
#define VA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_STAGES 13
#define ORC_MAX_N 256

/* Identifiers shared with the product C-ABI (include/vectorizedadjoint_amd.h). */
enum {
    ORC_SYS_VAN_DER_POL = 1,           /* x0' = x1, x1' = mu((1-x0^2)x1 - x0); P = 1        */
    ORC_SYS_OSCILLATOR = 2,            /* x0' = x1, x1' = -k x0 - c x1; p = (k, c); P = 2      */
    ORC_SYS_OSCILLATOR_FIXED_K = 3,    /* x0' = x1, x1' = -K x0 - mu x1; p = (mu), K constant  */
    ORC_SYS_LOTKA_VOLTERRA = 4,        /* x_i' = x_i (p_i + sum_j p[N(i+1)+j] x_j); P = N+N^2 */
    ORC_SYS_LOTKA_VOLTERRA_SHARED = 5, /* same right-hand side, A shared by the batch          */
    ORC_SYS_USER = 100                 /* the callbacks of orc_set_user_system               */
};

/* A system of the caller's (the device's VA_SYSTEM_USER): f = f(x, p) and
 * wJx = w^T df/dx, wJp = w^T df/dp (overwritten), the signatures of the device
 * functor's rhs and vjp. */
typedef void (*orc_user_rhs)(const double *x, const double *p, double *f);
typedef void (*orc_user_vjp)(const double *x, const double *p, const double *w, double *wJx,
                             double *wJp);

/* Sets the process-wide ORC_SYS_USER system (before solving; not thread-safe). */
void orc_set_user_system(orc_user_rhs rhs, orc_user_vjp vjp);

enum {
    ORC_TAB_EULER = 1,
    ORC_TAB_RK4 = 2,
    ORC_TAB_CASH_KARP54 = 3,
    ORC_TAB_DOPRI5 = 4,
    ORC_TAB_FEHLBERG78 = 5
};

enum {
    ORC_OK = 0,
    ORC_MAX_FAIL = 1,     /* a step needed more tries than max_attempts (odeint step_adjustment_error) */
    ORC_NONFINITE = 2,    /* final state not finite                                                   */
    ORC_MAX_STEPS = 3,    /* more accepted steps than the trajectory capacity                         */
    ORC_BAD_INPUT = 4
};

enum {
    ORC_COST_SUM = 1,          /* J = sum_i y_i(T)        lambda = 1      */
    ORC_COST_SQUARED_NORM = 2, /* J = sum_i y_i(T)^2      lambda = 2 y(T) */
    ORC_COST_LINEAR = 3        /* J = sum_i w_i y_i(T)    lambda = w      */
};

typedef struct orc_tableau {
    int s;          /* stages taking part in the solution (DOPRI5: 6, its FSAL stage excluded)      */
    int s_fwd;      /* stages evaluated by a forward attempt (DOPRI5: 7, stage 7 = f(y_{n+1}))      */
    int order;      /* stepper order (odeint stepper_order)                                          */
    int err_order;  /* error order (odeint error_order), 0 when there is no embedded estimate       */
    int fsal;
    int has_err;
    double a[ORC_MAX_STAGES][ORC_MAX_STAGES];
    double b[ORC_MAX_STAGES];
    double db[ORC_MAX_STAGES];  /* b - b_embedded, rounded as odeint rounds it */
    double c[ORC_MAX_STAGES];
} orc_tableau;

typedef struct orc_problem {
    int system;
    int n;          /* state size N                      */
    int np;         /* parameters per instance P        */
    int tableau;
    int controlled; /* 1: adaptive loop, 0: fixed step  */
    int max_attempts;   /* odeint failed_step_checker limit, 500 */
    int max_steps;      /* accepted-step capacity of one trajectory */
    int pad_;
    double t0, t1, dt0, abs_tol, rel_tol;
    double sys_const;   /* K of ORC_SYS_OSCILLATOR_FIXED_K */
} orc_problem;

typedef struct orc_trajectory {
    int64_t cap;        /* capacity in stored points */
    int64_t T;          /* stored points = accepted steps + 1 */
    double *t;          /* [cap]   t_n                                   */
    double *h;          /* [cap]   h_n = t_{n+1} - t_n (0 at the end)          */
    double *y;          /* [cap*N] y_n                                   */
} orc_trajectory;

typedef struct orc_stats {
    int32_t accepted;
    int32_t rejected;
    int32_t rhs_evals;
    int32_t status;
} orc_stats;

/* Observation schedule shared by a batch: k times t[0..k-1], increasing in the
 * direction of integration and inside [t0, t1]; values y[k][N].  The running
 * cost is J = sum_k sum_i (y_i(t_k) - y[k][i])^2 (no reference counterpart: the
 * reference's cost is terminal only). */
typedef struct orc_obs {
    int k;
    const double *t;
    const double *y;
} orc_obs;

/* Tableau lookup; returns 0 on success. */
int orc_get_tableau(int id, orc_tableau *out);

/* Right-hand side and vector-Jacobian products of the built-in systems.
 * vjp: wJx[j] = sum_i w_i df_i/dx_j, wJp[k] = sum_i w_i df_i/dp_k (overwritten). */
void orc_rhs(int system, int n, const double *x, const double *p, double t, double sys_const,
             double *dxdt);
void orc_vjp(int system, int n, int np, const double *x, const double *p, double t,
             double sys_const, const double *w, double *wJx, double *wJp);

/* x^(-1/q) for the step controller, q = 2..8 (NaN for other q). */
double orc_root(double x, int q);

/* One explicit RK step without error control: out = y + h sum b_m K_m. */
void orc_rk_step(const orc_problem *pb, const double *y, const double *p, double t, double h,
                 double *out, double *xerr);

/* Trajectory storage helpers. */
int orc_traj_alloc(orc_trajectory *tr, int64_t cap, int n);
void orc_traj_free(orc_trajectory *tr);

/* Forward solve of one instance.  y (in: y0, out: y(T)); records (t_n, h_n, y_n)
 * before every accepted step and once at the end.  Returns the accepted step
 * count, fills stats.  If log != NULL, each attempt appends
 * (t, h, err, accepted) to log[4*k..] up to log_cap attempts. */
int64_t orc_forward(const orc_problem *pb, double *y, const double *p, orc_trajectory *tr,
                    orc_stats *st, double *log, int64_t log_cap, int64_t *log_len);

/* Forward solve with observation stops (controlled loop only).  obs_row[k]
 * receives the stored-point index of observation k; *J the running cost. */
int64_t orc_forward_obs(const orc_problem *pb, const orc_obs *ob, double *y, const double *p,
                        orc_trajectory *tr, orc_stats *st, int32_t *obs_row, double *J);

/* Reverse sweep of the observation cost: lambda (in: terminal seed, usually 0;
 * out: dJ/dy0) jumps by 2 (y_n - y_obs_k) at every observation point. */
void orc_backward_obs(const orc_problem *pb, const orc_obs *ob, const double *p,
                      const orc_trajectory *tr, const int32_t *obs_row, double *lambda,
                      double *mu);

/* Reverse sweep over a stored trajectory for one seed.
 * lambda: in dJ/dy(T), out dJ/dy0 (overwritten).  mu: += dJ/dp. */
void orc_backward(const orc_problem *pb, const double *p, const orc_trajectory *tr,
                  double *lambda, double *mu);

/* Reference adjointSolve semantics for Nout seeds: lambda[Nout][N] overwritten,
 * mu[Nout][P] accumulated (detail/backpropagation.hpp:280-348). */
void orc_adjoint_solve(const orc_problem *pb, const double *p, const orc_trajectory *tr,
                       int nout, double *lambda, double *mu);

/* computeSensitivityMatrix: S[N][P] = d y_i(T) / d p (detail/backpropagation.hpp:352-434). */
void orc_sensitivity_matrix(const orc_problem *pb, const double *p, const orc_trajectory *tr,
                            double *S);

/* Replay the stored step sequence (frozen h_n = t_{n+1} - t_n) from y0 without
 * error control: the discrete map whose exact derivative orc_backward computes
 * (the forward itself took the accepted steps, which can differ from h_n by
 * an ulp of t). */
void orc_replay(const orc_problem *pb, const double *y0, const double *p,
                const orc_trajectory *tr, double *yT);

/* Batched solve over SoA arrays: y0[N][B], p[P][B] (instance index fastest).
 * For ORC_SYS_LOTKA_VOLTERRA_SHARED, p holds r[N][B] and p_shared A[N*N];
 * dJdp_shared receives sum over instances of dJ/dA (may be NULL otherwise).
 * Instances [b_begin, b_begin+b_count) are solved with nthreads threads.
 * Outputs are SoA as well; J, n_acc, n_rej, status are [B]. */
int orc_solve_batch(const orc_problem *pb, int64_t B, int64_t b_begin, int64_t b_count,
                    const double *y0, const double *p, const double *p_shared,
                    int cost, const double *cost_w,
                    double *J, double *yT, double *dJdy0, double *dJdp, double *dJdp_shared,
                    int32_t *n_acc, int32_t *n_rej, int32_t *status, int nthreads);

/* orc_solve_batch that also writes digest[B]: orc_step_digest of each
 * instance's stored step sequence (0 where not written). */
int orc_solve_batch_digest(const orc_problem *pb, int64_t B, int64_t b_begin, int64_t b_count,
                           const double *y0, const double *p, const double *p_shared, int cost,
                           const double *cost_w, double *J, double *yT, double *dJdy0,
                           double *dJdp, double *dJdp_shared, int32_t *n_acc, int32_t *n_rej,
                           int32_t *status, uint64_t *digest, int nthreads);

/* 64-bit digest of a stored step sequence: d = 0x9E3779B97F4A7C15, then for
 * n = 0..T-1: d = mix(d ^ bits(t_n)); d = mix(d ^ bits(h_n)), with mix the
 * SplitMix64 finaliser.  Equal (t_n, h_n) histories give equal digests. */
uint64_t orc_step_digest(const orc_trajectory *tr);

/* Batched solve of the observation cost (controlled loop): J[B] = running cost,
 * dJdy0 / dJdp its gradient; the observations are shared by the batch. */
int orc_solve_batch_obs(const orc_problem *pb, const orc_obs *ob, int64_t B, int64_t b_begin,
                        int64_t b_count, const double *y0, const double *p, double *J,
                        double *yT, double *dJdy0, double *dJdp, int32_t *n_acc,
                        int32_t *n_rej, int32_t *status, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
