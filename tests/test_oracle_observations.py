"""Pins the oracle's observation cost (forced stops at observation times, running
cost J = sum_k |y(t_k) - y_obs_k|^2, lambda jumps in the reverse sweep).

No reference counterpart exists (the reference's cost is terminal only), so the
pins are mathematical:
  - complex-step differentiation of the frozen-step map with the cost taken at
    the stored observation points gives dJ/dy0 and dJ/dp to rounding;
  - one observation at t1 with y_obs = 0 is the terminal cost |y(T)|^2: same
    trajectory and bit-identical gradient as the terminal-cost sweep;
  - every observation time is reached (within the loop's epsilon), the stored
    observation points are ordered, and J equals the sum recomputed from them;
  - central finite differences of the full adaptive solve agree to the
    tolerance level.
"""
import numpy as np
import pytest

import oracle as O
from refmath import rhs as ref_rhs


def _tab(tid):
    t = O.tableau(tid)
    return t["a"], t["b"], t["s"]


def _replay_cost(system, tid, hs, obs_rows, y_obs, y0, p, K=0.0):
    """J of the frozen-step map (steps hs), cost at the given point indices; complex-safe.
    y0: [N, D], p: [P, D]."""
    a, bb, s = _tab(tid)
    y = np.array(y0, dtype=np.complex128)
    J = np.zeros(y.shape[1], dtype=np.complex128)
    rows = list(obs_rows)
    k = 0

    def observe(n, y, J, k):
        while k < len(rows) and rows[k] == n:
            d = y - y_obs[k][:, None]
            J = J + np.sum(d * d, axis=0)
            k += 1
        return J, k

    J, k = observe(0, y, J, k)
    for n, h in enumerate(hs):
        Ks = []
        for m in range(s):
            X = y.copy()
            for j in range(m):
                if a[m, j] != 0.0:
                    X = X + (a[m, j] * h) * Ks[j]
            Ks.append(ref_rhs(system, X, p, K))
        for j in range(s):
            if bb[j] != 0.0:
                y = y + (bb[j] * h) * Ks[j]
        J, k = observe(n + 1, y, J, k)
    assert k == len(rows)
    return J


def _problem(t1=6.0, tol=1e-8, tid=None):
    return O.problem(O.SYS_VAN_DER_POL, 2, 1, tid or O.TAB_DOPRI5, t0=0.0, t1=t1, dt0=1e-3,
                     abs_tol=tol, rel_tol=tol, max_steps=100000)


def _obs(pb, mu_true=2.0, k=12):
    t_obs = pb.t1 * np.arange(1, k + 1) / k
    g = O.Solve(pb, [2.0, 0.0], [mu_true], obs=(t_obs, np.zeros((k, 2))))
    assert g.status == O.OK
    return t_obs, g.states[g.obs_row]


@pytest.mark.parametrize("tid", [O.TAB_DOPRI5, O.TAB_CASH_KARP54, O.TAB_FEHLBERG78])
def test_observation_adjoint_matches_complex_step(tid):
    pb = _problem(tid=tid)
    t_obs, y_obs = _obs(pb)
    y0, p = np.array([1.9, 0.2]), np.array([1.6])
    s = O.Solve(pb, y0, p, obs=(t_obs, y_obs))
    assert s.status == O.OK
    lam, mu = s.backward_obs()
    hs = s.steps_h[:-1]
    N, P = 2, 1
    eps = 1e-30
    Y = np.tile(y0.astype(np.complex128)[:, None], (1, N + P))
    Pm = np.tile(p.astype(np.complex128)[:, None], (1, N + P))
    for k in range(N):
        Y[k, k] += 1j * eps
    Pm[0, N] += 1j * eps
    J = _replay_cost(O.SYS_VAN_DER_POL, tid, hs, s.obs_row, y_obs, Y, Pm)
    g = np.imag(J) / eps
    assert abs(np.real(J[0]) - s.J) <= 1e-12 * max(abs(s.J), 1e-300)
    assert np.max(np.abs(lam - g[:N])) / np.max(np.abs(g[:N])) <= 1e-11
    assert abs(mu[0] - g[N]) / abs(g[N]) <= 1e-11


def test_single_observation_at_t1_is_the_terminal_cost():
    pb = _problem()
    y0, p = [2.0, 0.3], [1.3]
    s1 = O.Solve(pb, y0, p, obs=(np.array([pb.t1]), np.zeros((1, 2))))
    s2 = O.Solve(pb, y0, p)
    assert np.array_equal(s1.y, s2.y) and s1.steps == s2.steps
    lam1, mu1 = s1.backward_obs()
    lam2, mu2 = s2.backward(2.0 * s2.y)
    assert np.array_equal(lam1, lam2) and np.array_equal(mu1, mu2)


def test_observation_points_are_reached_and_cost_is_their_sum():
    pb = _problem()
    t_obs, y_obs = _obs(pb, k=25)
    s = O.Solve(pb, [2.05, -0.1], [2.4], obs=(t_obs, y_obs))
    times, ys = s.times, s.states
    assert np.all(np.diff(s.obs_row) >= 0)
    eps = np.finfo(float).eps
    assert np.all(np.abs(times[s.obs_row] - t_obs) <= 4 * eps * np.maximum(1.0, np.abs(t_obs)))
    d = ys[s.obs_row] - y_obs
    assert abs(np.sum(d * d) - s.J) <= 1e-13 * s.J
    # stopping at the observations adds steps
    assert s.steps >= O.Solve(pb, [2.05, -0.1], [2.4]).steps


def test_observation_gradient_matches_central_fd():
    """Adaptive solve: finite differences agree to the tolerance level (h = 1e-6)."""
    pb = _problem(t1=20.0)
    t_obs, y_obs = _obs(pb, mu_true=3.0, k=50)
    y0 = np.array([2.0, 0.0])

    def J(mu, y=y0):
        return O.Solve(pb, y, [mu], obs=(t_obs, y_obs)).J

    s = O.Solve(pb, y0, [2.5], obs=(t_obs, y_obs))
    lam, mu = s.backward_obs()
    h = 1e-6
    fd = (J(2.5 + h) - J(2.5 - h)) / (2 * h)
    assert abs(fd - mu[0]) <= 1e-5 * abs(mu[0])
    for i in range(2):
        e = np.zeros(2)
        e[i] = h
        fdi = (J(2.5, y0 + e) - J(2.5, y0 - e)) / (2 * h)
        assert abs(fdi - lam[i]) <= 1e-5 * np.max(np.abs(lam))


def test_batch_observations_match_single_solves():
    pb = _problem()
    t_obs, y_obs = _obs(pb)
    B = 16
    rng = np.random.default_rng(2)
    y0 = np.vstack([2.0 + 0.05 * rng.normal(size=B), 0.05 * rng.normal(size=B)])
    p = rng.uniform(0.5, 3.0, size=(1, B))
    out = O.solve_batch(pb, y0, p, 0, obs=(t_obs, y_obs), nthreads=4)
    for b in (0, 7, 15):
        s = O.Solve(pb, y0[:, b], p[:, b], obs=(t_obs, y_obs))
        lam, mu = s.backward_obs()
        assert out["J"][b] == s.J and np.array_equal(out["yT"][:, b], s.y)
        assert np.array_equal(out["dJdy0"][:, b], lam) and np.array_equal(out["dJdp"][:, b], mu)
        assert out["n_acc"][b] == s.steps
