"""Pins the oracle's discrete adjoint (detail/backpropagation.hpp:24-348 restated).

  - complex-step differentiation of the frozen-step discrete map (an independent
    complex-arithmetic replay of the RK map over the oracle's stored steps) gives
    dJ/dy0 and dJ/dp exactly to rounding, for every tableau and system;
  - for the linear oscillator the replay is linear in y0, so dJ/dy0 also equals
    the transpose of the transfer matrix applied to the seed;
  - central finite differences of the full adaptive solve agree to the
    tolerance level (the step sequence moves with the inputs);
  - reference semantics: lambda overwritten, mu accumulated, one sweep per
    seed, sensitivity matrix = unit seeds, zero steps => dJ/dy0 = lambda.
"""
import os

import numpy as np
import pytest

import oracle as O
from refmath import complex_step_gradients, read_alphas_csv, rk_replay

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _tab(tid):
    t = O.tableau(tid)
    return {"a": t["a"], "b": t["b"], "s": t["s"]}


def _cases():
    a5 = read_alphas_csv(os.path.join(GOLDEN, "glv_N5_alphasfile_cpp.csv"))
    rng = np.random.default_rng(3)
    glv10 = np.concatenate([rng.uniform(0.1, 1, 10), (0.2 * rng.normal(size=(10, 10))
                                                      - np.diag(rng.uniform(0.5, 1.5, 10))).ravel()])
    return [
        # name, system, N, P, tableau, controlled, t1, dt0, tol, y0, p, K
        ("vdp-dopri5", O.SYS_VAN_DER_POL, 2, 1, O.TAB_DOPRI5, True, 3.0, 1e-3, 1e-8,
         [2.0, 0.0], [1.7], 0.0),
        ("vdp-fehlberg78", O.SYS_VAN_DER_POL, 2, 1, O.TAB_FEHLBERG78, True, 3.0, 1e-3, 1e-8,
         [2.0, 0.1], [2.5], 0.0),
        ("osc-cashkarp", O.SYS_OSCILLATOR, 2, 2, O.TAB_CASH_KARP54, True, 5.0, 1e-3, 1e-9,
         [0.3, -0.8], [2.0, 0.2], 0.0),
        ("osck-rk4-fixed", O.SYS_OSCILLATOR_FIXED_K, 2, 1, O.TAB_RK4, False, 10.0, 0.01, 0.0,
         [0.0, 1.0], [0.151], 1.0),
        ("osc-euler-fixed", O.SYS_OSCILLATOR, 2, 2, O.TAB_EULER, False, 1.0, 0.01, 0.0,
         [0.5, 0.5], [1.5, 0.4], 0.0),
        ("glv5-cashkarp-csv", O.SYS_LOTKA_VOLTERRA, 5, 30, O.TAB_CASH_KARP54, True, 10.0, 1e-3,
         1e-8, [0.1] * 5, a5, 0.0),
        ("glv10-dopri5", O.SYS_LOTKA_VOLTERRA, 10, 110, O.TAB_DOPRI5, True, 8.0, 1e-3, 1e-8,
         list(rng.uniform(0.1, 1, 10)), glv10, 0.0),
    ]


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_adjoint_matches_complex_step_of_frozen_map(case):
    name, system, N, P, tid, ctrl, t1, dt0, tol, y0, p, K = case
    pb = O.problem(system, N, P, tid, controlled=ctrl, t0=0.0, t1=t1, dt0=dt0, abs_tol=tol,
                   rel_tol=tol, sys_const=K)
    s = O.Solve(pb, y0, p)
    assert s.status == O.OK and s.steps > 3
    rng = np.random.default_rng(5)
    c = rng.normal(size=N)
    # J = c . y(T) + 0.5 |y(T)|^2, seed lambda = c + y(T)
    lam, mu = s.backward(c + s.y)
    times = np.concatenate([[0.0], np.cumsum(s.steps_h[:-1])])
    # replay over the stored accepted steps h_n (times only serve to carry h)
    tab = _tab(tid)

    def J(yT):
        return c @ yT + 0.5 * np.sum(yT * yT, axis=0)

    gy, gp = complex_step_gradients(system, tab, times, np.array(y0), np.array(p), J, K)
    # the replayed steps are exactly the stored h
    hs = np.diff(times)
    assert np.array_equal(hs, s.steps_h[:-1]) or np.allclose(hs, s.steps_h[:-1], rtol=1e-15)
    sc = max(np.max(np.abs(gy)), 1e-300)
    assert np.max(np.abs(lam - gy)) / sc <= 1e-11, (lam, gy)
    scp = max(np.max(np.abs(gp)), 1e-300)
    assert np.max(np.abs(mu - gp)) / scp <= 1e-11


def test_linear_oscillator_transfer_matrix():
    """The frozen-step RK map of y' = M y is y(T) = Phi y0 with Phi a product of
    step matrices; dJ/dy0 for J = c . y(T) is Phi^T c."""
    k, cdamp = 1.7, 0.25
    pb = O.problem(O.SYS_OSCILLATOR, 2, 2, O.TAB_DOPRI5, t1=6.0, abs_tol=1e-9, rel_tol=1e-9)
    s = O.Solve(pb, [0.4, 0.9], [k, cdamp])
    Phi = np.column_stack([s.replay(e) for e in np.eye(2)])
    c = np.array([0.7, -1.3])
    lam, _ = s.backward(c)
    assert np.allclose(lam, Phi.T @ c, rtol=1e-13, atol=1e-15)
    # and y(T) itself is Phi y0
    assert np.allclose(s.y, Phi @ np.array([0.4, 0.9]), rtol=1e-13)


def test_adaptive_gradient_matches_central_fd():
    """Full adaptive solve: the step sequence moves with the inputs, so finite
    differences agree only to the tolerance level."""
    tol = 1e-10
    pb = O.problem(O.SYS_VAN_DER_POL, 2, 1, O.TAB_DOPRI5, t1=4.0, abs_tol=tol, rel_tol=tol)
    y0, p = np.array([2.0, 0.0]), np.array([1.3])

    def J(y0_, p_):
        s = O.Solve(pb, y0_, p_)
        return float(np.sum(s.y ** 2))

    s = O.Solve(pb, y0, p)
    lam, mu = s.backward(2.0 * s.y)
    eps = 1e-6
    fd_p = (J(y0, p + eps) - J(y0, p - eps)) / (2 * eps)
    fd_y = np.array([(J(y0 + eps * e, p) - J(y0 - eps * e, p)) / (2 * eps) for e in np.eye(2)])
    assert abs(mu[0] - fd_p) <= 1e-4 * max(1.0, abs(fd_p))
    assert np.allclose(lam, fd_y, rtol=1e-4, atol=1e-4)


def test_reference_seed_semantics():
    """adjointSolve: lambda overwritten with dJ/dy0, mu += dJ/dp, each seed its own
    sweep (detail/backpropagation.hpp:299-347)."""
    a5 = read_alphas_csv(os.path.join(GOLDEN, "glv_N5_alphasfile_cpp.csv"))
    pb = O.problem(O.SYS_LOTKA_VOLTERRA, 5, 30, O.TAB_CASH_KARP54, t1=10.0)
    s = O.Solve(pb, [0.1] * 5, a5)
    lam = np.eye(5)
    mu0 = np.full((5, 30), 0.25)
    lam_out, mu_out = s.adjoint_solve(lam, mu0)
    for o in range(5):
        l1, m1 = s.backward(np.eye(5)[o])
        assert np.array_equal(lam_out[o], l1)
        assert np.allclose(mu_out[o], 0.25 + m1, rtol=0, atol=1e-15)
    # computeSensitivityMatrix: rows are the unit-seed parameter adjoints
    S = s.sensitivity_matrix()
    assert np.allclose(S, mu_out - 0.25, rtol=0, atol=1e-15)


def test_zero_steps_gives_identity():
    pb = O.problem(O.SYS_VAN_DER_POL, 2, 1, O.TAB_DOPRI5, t0=1.0, t1=1.0)
    s = O.Solve(pb, [2.0, 0.0], [1.0])
    assert s.steps == 0
    lam, mu = s.backward([0.3, -0.7])
    assert np.array_equal(lam, [0.3, -0.7]) and np.array_equal(mu, [0.0])


def test_stored_steps_are_time_differences_and_replay_matches_forward():
    """The trajectory keeps h_n = t_{n+1} - t_n (the reference's GetDt,
    StateStorage.hpp:22), which the reverse uses; replaying those steps
    reproduces the forward (which took the accepted steps, an ulp of t away)
    to rounding level."""
    pb = O.problem(O.SYS_VAN_DER_POL, 2, 1, O.TAB_CASH_KARP54, t1=5.0)
    s = O.Solve(pb, [2.0, 0.0], [2.0])
    t, h = s.times, s.steps_h
    assert np.array_equal(h[:-1], t[1:] - t[:-1]) and h[-1] == 0.0
    assert np.allclose(s.replay([2.0, 0.0]), s.y, rtol=1e-13, atol=0.0)
    # and the complex replay (independent code, no fma) agrees to rounding
    tab = _tab(O.TAB_CASH_KARP54)
    times = np.concatenate([[0.0], np.cumsum(s.steps_h[:-1])])
    yr = rk_replay(O.SYS_VAN_DER_POL, tab, times, np.array([[2.0], [0.0]]), np.array([[2.0]]))
    assert np.allclose(yr[:, 0].real, s.y, rtol=1e-12)
