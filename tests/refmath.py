"""Independent NumPy restatements used to pin the oracle (test infrastructure).

Nothing here calls the oracle: these are the textbook objects the oracle's
results are checked against.
  - rooted trees and Butcher order conditions (B-series elementary weights);
  - right-hand sides written from the reference example functors
    (examples/VanDerPol/main.cpp:41-42, examples/HarmonicOscillator/main.cpp:27-28,
    examples/GeneralizedLotkaVolterra/main.cpp:109-118), in complex arithmetic so
    complex-step differentiation gives exact derivatives;
  - a complex-arithmetic replay of the explicit RK map over a given time grid,
    whose complex-step derivative is the exact derivative of the discrete map.
"""
from __future__ import annotations

import numpy as np

SYS_VAN_DER_POL, SYS_OSCILLATOR, SYS_OSCILLATOR_FIXED_K, SYS_LOTKA_VOLTERRA = 1, 2, 3, 4
SYS_LOTKA_VOLTERRA_SHARED = 5
SYS_USER = 100  # in the tests: the Brusselator of tests/cpp/va_user_system.hip


def rooted_trees(max_order: int):
    """All rooted trees up to max_order as (order, children-indices) in a global list."""
    trees = [(1, ())]
    for n in range(2, max_order + 1):
        new = []

        def rec(remaining, start, acc):
            if remaining == 0:
                new.append(tuple(acc))
                return
            for idx in range(start, len(trees)):
                o = trees[idx][0]
                if o > remaining:
                    break
                rec(remaining - o, idx, acc + [idx])

        rec(n - 1, 0, [])
        trees.extend((n, t) for t in new)
    return trees


def order_condition_residuals(A, b, max_order):
    """For every tree t with order <= max_order: b . Psi(t) - 1/gamma(t)."""
    A = np.asarray(A, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = len(b)
    A = A[:s, :s]
    trees = rooted_trees(max_order)
    psi, gamma, out = [], [], []
    for order, ch in trees:
        v = np.ones(s)
        g = float(order)
        for c in ch:
            v = v * (A @ psi[c])
            g *= gamma[c]
        psi.append(v)
        gamma.append(g)
        out.append((order, float(b @ v) - 1.0 / g))
    return out


def rhs(system, x, p, K=0.0):
    """x: [N, D], p: [P, D] (complex allowed) -> f [N, D]."""
    if system == SYS_VAN_DER_POL:
        return np.stack([x[1], p[0] * ((1.0 - x[0] * x[0]) * x[1] - x[0])])
    if system == SYS_OSCILLATOR:
        return np.stack([x[1], -p[0] * x[0] - p[1] * x[1]])
    if system == SYS_OSCILLATOR_FIXED_K:
        return np.stack([x[1], -K * x[0] - p[0] * x[1]])
    if system == SYS_USER:  # Brusselator: x0' = a + x0^2 x1 - (b+1) x0, x1' = b x0 - x0^2 x1
        a, b = p[0], p[1]
        return np.stack([a + x[0] ** 2 * x[1] - (b + 1.0) * x[0], b * x[0] - x[0] ** 2 * x[1]])
    if system in (SYS_LOTKA_VOLTERRA, SYS_LOTKA_VOLTERRA_SHARED):
        N = x.shape[0]
        r = p[:N]
        A = p[N:N + N * N].reshape((N, N) + p.shape[1:])
        g = r + np.einsum("ij...,j...->i...", A, x)
        return x * g
    raise ValueError(system)


def rk_replay(system, tab, times, y0, p, K=0.0):
    """Explicit RK map over the given time grid (h_n = t[n+1]-t[n]); complex-safe.
    y0: [N, D], p: [P, D]."""
    a, bb = tab["a"], tab["b"]
    s = tab["s"]
    y = np.array(y0, dtype=np.complex128)
    for n in range(len(times) - 1):
        h = times[n + 1] - times[n]
        Ks = []
        for m in range(s):
            X = y.copy()
            for j in range(m):
                if a[m, j] != 0.0:
                    X = X + (a[m, j] * h) * Ks[j]
            Ks.append(rhs(system, X, p, K))
        for j in range(s):
            if bb[j] != 0.0:
                y = y + (bb[j] * h) * Ks[j]
    return y


def complex_step_gradients(system, tab, times, y0, p, cost_grad_fn, K=0.0, eps=1e-30):
    """d J / d y0 and d J / d p of the frozen-step map by complex step.
    cost_grad_fn(yT_complex[N, D]) -> J values [D] (complex)."""
    N, P = len(y0), len(p)
    D = N + P
    Y = np.tile(np.asarray(y0, dtype=np.complex128)[:, None], (1, D))
    Pm = np.tile(np.asarray(p, dtype=np.complex128)[:, None], (1, D))
    for k in range(N):
        Y[k, k] += 1j * eps
    for k in range(P):
        Pm[k, N + k] += 1j * eps
    yT = rk_replay(system, tab, times, Y, Pm, K)
    J = cost_grad_fn(yT)
    g = np.imag(J) / eps
    return g[:N], g[N:]


def read_alphas_csv(path):
    """examples/GeneralizedLotkaVolterra/main.cpp:37-89: one comma-separated line."""
    vals = []
    with open(path) as f:
        for line in f:
            for tok in line.strip().split(","):
                if tok:
                    vals.append(float(tok))
    return np.array(vals)
