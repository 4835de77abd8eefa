"""GPU parity of a system of the caller's: the Brusselator of
tests/cpp/va_user_system.hip, compiled outside the library with VA_DEVICE_SYSTEM
and passed to va_create_with_kernels, against the oracle running the same
functor's host build as its callbacks.  Same bar as the built-in systems:
identical step sequences and outputs equal bit for bit."""
import ctypes as C
import dataclasses
import os

import numpy as np
import pytest

import oracle as O
from helpers import run_oracle
from test_gpu_parity import _compare, _dev, _torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "libva_user_system.so")
HOST_LIB = os.path.join(ROOT, "tests", "cpp", "libva_brusselator_host.so")


def _lib():
    lib = C.CDLL(LIB)
    lib.va_brusselator_kernels.restype = C.c_void_p
    lib.va_brusselator_kernels.argtypes = [C.c_int32, C.c_int32]
    host = C.CDLL(HOST_LIB)
    O.set_user_system(host.brusselator_rhs, host.brusselator_vjp)
    return lib


@pytest.mark.parametrize("tableau,mode", [("dopri5", "persistent"), ("cash_karp54", "persistent"),
                                          ("dopri5", "plain"), ("rk4", "persistent")])
def test_user_system_matches_oracle(tableau, mode):
    torch = _torch()
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    lib = _lib()
    tid = {"dopri5": va.TAB_DOPRI5, "cash_karp54": va.TAB_CASH_KARP54, "rk4": va.TAB_RK4}[tableau]
    w = dataclasses.replace(I.WORKLOADS["brusselator"], tableau=tid,
                            controlled=tableau != "rk4", dt0=1e-3 if tableau != "rk4" else 1e-2)
    B = 3000
    inp = I.make_inputs(w, B)
    o = run_oracle(w, inp, max_steps=4096)
    assert (o["status"] == 0).all()
    kern = lib.va_brusselator_kernels(tid, 1 if w.controlled else 0)
    assert kern
    m = {"persistent": va.MODE_PERSISTENT, "plain": va.MODE_PLAIN}[mode]
    pb = va.Problem.from_workload(w, batch=B, mode=m, max_steps=4096)
    print(f"[gpu] user system {tableau} {mode} B={B}", flush=True)
    with va.Solver(pb, kernels=kern) as s:
        out = s.solve(_dev(inp["y0"], torch), _dev(inp["p"], torch), w.cost)
        out["digest"] = s.step_digest()
        torch.cuda.synchronize()
        g = {k: v.cpu().numpy() for k, v in out.items()}
    _compare(g, o)


def test_user_system_observed_matches_oracle():
    """The observation cost (va_solve_observed) with the caller's system."""
    torch = _torch()
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    lib = _lib()
    w = I.WORKLOADS["brusselator"]
    B, K = 1000, 20
    inp = I.make_inputs(w, B)
    t_obs = np.linspace(1.0, 20.0, K)
    y_obs = np.stack([1.0 + 0.3 * np.sin(t_obs), 3.0 - 0.2 * np.cos(t_obs)], axis=1)
    o = run_oracle_obs(w, inp, t_obs, y_obs)
    kern = lib.va_brusselator_kernels(va.TAB_DOPRI5, 1)
    pb = va.Problem.from_workload(w, batch=B, max_steps=4096, n_obs=K)
    print(f"[gpu] user system observed B={B}", flush=True)
    with va.Solver(pb, kernels=kern) as s:
        out = s.solve_observed(_dev(inp["y0"], torch), _dev(inp["p"], torch),
                               _dev(t_obs, torch), _dev(y_obs, torch))
        torch.cuda.synchronize()
        g = {k: v.cpu().numpy() for k, v in out.items()}
    ok = o["status"] == 0
    assert ok.all()
    assert np.array_equal(g["status"], o["status"])
    assert np.array_equal(g["n_acc"], o["n_acc"]) and np.array_equal(g["n_rej"], o["n_rej"])
    for k in ("yT", "J", "dJdy0", "dJdp"):
        assert np.array_equal(g[k], o[k]), k


def run_oracle_obs(w, inp, t_obs, y_obs):
    pb = O.problem(w.system, w.n, w.np, w.tableau, t0=w.t0, t1=w.t1, dt0=w.dt0, abs_tol=w.tol,
                   rel_tol=w.tol, max_steps=4096)
    return O.solve_batch(pb, inp["y0"], inp["p"], 0, obs=(t_obs, y_obs),
                         nthreads=min(16, os.cpu_count() or 1))
