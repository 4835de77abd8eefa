"""Pins the oracle's building blocks: the controller root x^(-1/q) against the
exact root and the C library pow, the right-hand sides against the reference example functors, and the
vector-Jacobian products against complex-step derivatives (exact to rounding
for these polynomial systems)."""
import math
import os
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle as O
from refmath import read_alphas_csv, rhs

getcontext().prec = 40
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def ulp_diff(a, b):
    return abs(a - b) / math.ulp(b) if b != 0 else abs(a)


def exact_root(x, q):
    """x^(-1/q) to 40 digits, rounded once to a double."""
    return float(Decimal(x) ** (Decimal(-1) / Decimal(q)))


@pytest.mark.parametrize("q", [2, 3, 4, 5, 6, 7, 8])
def test_root_against_exact_and_libm_on_controller_domain(q):
    # the controller raises err in [5^-8, ~1e6] to -1/q (q = ORDER or ERR_ORDER - 1)
    xs = np.concatenate([np.geomspace(5.0 ** -8, 1e6, 3001), np.linspace(0.01, 3.0, 1501),
                         [5.0 ** -5, 5.0 ** -8, 0.5, 1.0, 1.0 + 2 ** -52, 1.0 - 2 ** -53,
                          2.0, 2.0 ** q, 2.0 ** q - 2 ** -40]])
    worst_exact = max(ulp_diff(O.root_(x, q), exact_root(x, q)) for x in xs)
    worst_libm = max(ulp_diff(O.root_(x, q), x ** (-1.0 / q)) for x in xs)
    assert worst_exact <= 1.0, worst_exact
    assert worst_libm <= 4.0, worst_libm


@pytest.mark.parametrize("q", [2, 3, 5, 8])
def test_root_full_range_and_subnormals(q):
    xs = np.concatenate([np.geomspace(1e-307, 1e307, 2001),
                         [5e-324, 1e-320, 2.2250738585072014e-308, 1.7976931348623157e308]])
    worst = max(ulp_diff(O.root_(x, q), exact_root(x, q)) for x in xs)
    assert worst <= 1.0, worst


def test_root_special_values():
    assert O.root_(0.0, 3) == math.inf
    assert O.root_(math.inf, 3) == 0.0
    assert math.isnan(O.root_(math.nan, 5))
    assert math.isnan(O.root_(-1.0, 5))
    assert math.isnan(O.root_(2.0, 9))
    assert O.root_(8.0, 3) == 0.5 and O.root_(2.0 ** -40, 8) == 32.0
    assert abs(O.root_(1.0, 5) - 1.0) <= 2 ** -53


def _systems():
    rng = np.random.default_rng(7)
    a10 = read_alphas_csv(os.path.join(GOLDEN, "glv_N10_alphasfile_cpp.csv"))
    a5 = read_alphas_csv(os.path.join(GOLDEN, "glv_N5_alphasfile_cpp.csv"))
    return [
        ("vdp", O.SYS_VAN_DER_POL, rng.normal(size=2), np.array([3.7]), 0.0),
        ("osc", O.SYS_OSCILLATOR, rng.normal(size=2), np.array([2.2, 0.3]), 0.0),
        ("osck", O.SYS_OSCILLATOR_FIXED_K, rng.normal(size=2), np.array([0.151]), 1.0),
        ("glv5", O.SYS_LOTKA_VOLTERRA, rng.uniform(0.1, 1, 5), a5, 0.0),
        ("glv10", O.SYS_LOTKA_VOLTERRA, rng.uniform(0.1, 1, 10), a10, 0.0),
    ]


def test_alphas_fixture_layout():
    # examples/GeneralizedLotkaVolterra/main.cpp:150: Npar = N*N + N; diagonal A_ii = -10
    for N in (5, 10):
        a = read_alphas_csv(os.path.join(GOLDEN, f"glv_N{N}_alphasfile_cpp.csv"))
        assert a.size == N + N * N
        assert np.allclose(a[:N], 0.1)
        A = a[N:].reshape(N, N)
        assert np.allclose(np.diag(A), -10.0)


@pytest.mark.parametrize("name,system,x,p,K", _systems(), ids=lambda v: v if isinstance(v, str) else "")
def test_rhs_matches_reference_functor(name, system, x, p, K):
    f = O.rhs(system, x, p, sys_const=K)
    ref = rhs(system, x[:, None], p[:, None], K)[:, 0]
    assert np.allclose(f, ref, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("name,system,x,p,K", _systems(), ids=lambda v: v if isinstance(v, str) else "")
def test_vjp_matches_complex_step(name, system, x, p, K):
    rng = np.random.default_rng(11)
    w = rng.normal(size=x.size)
    wjx, wjp = O.vjp(system, x, p, w, sys_const=K)
    eps = 1e-30
    N, P = x.size, p.size
    D = N + P
    X = np.tile(x.astype(np.complex128)[:, None], (1, D))
    Pm = np.tile(p.astype(np.complex128)[:, None], (1, D))
    for k in range(N):
        X[k, k] += 1j * eps
    for k in range(P):
        Pm[k, N + k] += 1j * eps
    J = np.imag(rhs(system, X, Pm, K)) / eps  # [N, D] Jacobian columns
    ref = w @ J
    scale = np.max(np.abs(ref)) + 1e-300
    assert np.max(np.abs(wjx - ref[:N])) / scale <= 1e-14
    assert np.max(np.abs(wjp - ref[N:])) / scale <= 1e-14
