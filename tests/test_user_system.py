"""A system of the caller's (CPU side): the test suite's Brusselator
(tests/cpp/brusselator.hpp), compiled outside the library in
tests/cpp/va_user_system.hip with VA_DEVICE_SYSTEM.

Its host build (tests/cpp/brusselator_host.cpp, no device code) supplies the
oracle's callbacks (orc_set_user_system).  Pinned
here: the right-hand side against its closed form, the VJP against
complex-step derivatives of an independent NumPy restatement (tests/refmath.py),
and the oracle's adjoint for this system against the complex-step derivative
of the frozen step map.  The GPU parity of the same system is
tests/test_gpu_user_system.py and the drop-in binding (VA_BIND_DEVICE_SYSTEM)
tests/test_dropin.py::test_example_user_system_brusselator."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import refmath as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "libva_user_system.so")
HOST_LIB = os.path.join(ROOT, "tests", "cpp", "libva_brusselator_host.so")


@pytest.fixture(scope="module")
def user():
    if not os.path.exists(HOST_LIB):
        pytest.skip("tests/cpp/libva_brusselator_host.so not built")
    lib = C.CDLL(HOST_LIB)
    dp = C.POINTER(C.c_double)
    lib.brusselator_rhs.argtypes = [dp, dp, dp]
    lib.brusselator_vjp.argtypes = [dp, dp, dp, dp, dp]
    O.set_user_system(lib.brusselator_rhs, lib.brusselator_vjp)
    return lib


def _p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def test_kernel_table_is_exported():
    if not os.path.exists(LIB):
        pytest.skip("not built")
    syms = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True).stdout
    assert " T va_brusselator_kernels" in syms


def test_rhs_closed_form(user):
    rng = np.random.default_rng(1)
    for _ in range(200):
        x, p = rng.uniform(-3, 3, 2), rng.uniform(0.2, 4, 2)
        f = np.empty(2)
        user.brusselator_rhs(_p(x), _p(p), _p(f))
        ref = R.rhs(R.SYS_USER, x[:, None], p[:, None])[:, 0]
        assert np.allclose(f, ref, rtol=4e-16, atol=4e-16 * np.abs(x).max() ** 3 * p.max())


def test_vjp_complex_step(user):
    rng = np.random.default_rng(2)
    h = 1e-30
    for _ in range(200):
        x, p, w = rng.uniform(-3, 3, 2), rng.uniform(0.2, 4, 2), rng.normal(size=2)
        Jx = np.empty((2, 2))
        Jp = np.empty((2, 2))
        for k in range(2):
            xc = x.astype(complex)
            xc[k] += 1j * h
            Jx[:, k] = np.imag(R.rhs(R.SYS_USER, xc[:, None], p[:, None])[:, 0]) / h
            pc = p.astype(complex)
            pc[k] += 1j * h
            Jp[:, k] = np.imag(R.rhs(R.SYS_USER, x[:, None], pc[:, None])[:, 0]) / h
        wjx, wjp = np.empty(2), np.empty(2)
        user.brusselator_vjp(_p(x), _p(p), _p(w), _p(wjx), _p(wjp))
        assert np.allclose(wjx, w @ Jx, rtol=1e-14, atol=1e-14)
        assert np.allclose(wjp, w @ Jp, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("tid", [O.TAB_DOPRI5, O.TAB_CASH_KARP54])
def test_oracle_adjoint_is_the_frozen_map_derivative(user, tid):
    pb = O.problem(O.SYS_USER, 2, 2, tid, t1=8.0, abs_tol=1e-9, rel_tol=1e-9)
    y0, p = np.array([1.2, 3.1]), np.array([1.0, 3.0])
    s = O.Solve(pb, y0, p)
    assert s.status == O.OK and s.steps > 20
    lam, mu = s.backward(np.array([0.7, -1.3]))
    tab = O.tableau(tid)
    tab = dict(tab, s=tab["s"])
    gy, gp = R.complex_step_gradients(R.SYS_USER, tab, s.times, y0, p,
                                      lambda yT: 0.7 * yT[0] - 1.3 * yT[1])
    assert np.allclose(lam, gy, rtol=1e-10, atol=1e-12)
    assert np.allclose(mu, gp, rtol=1e-10, atol=1e-12)
