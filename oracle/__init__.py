"""ctypes binding of the CPU oracle (liboracle.so built from va_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or as the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SYS_VAN_DER_POL = 1
SYS_OSCILLATOR = 2
SYS_OSCILLATOR_FIXED_K = 3
SYS_LOTKA_VOLTERRA = 4
SYS_LOTKA_VOLTERRA_SHARED = 5
SYS_USER = 100

TAB_EULER = 1
TAB_RK4 = 2
TAB_CASH_KARP54 = 3
TAB_DOPRI5 = 4
TAB_FEHLBERG78 = 5

OK, MAX_FAIL, NONFINITE, MAX_STEPS, BAD_INPUT = 0, 1, 2, 3, 4

COST_SUM = 1
COST_SQUARED_NORM = 2
COST_LINEAR = 3

MAX_STAGES = 13


class Tableau(C.Structure):
    _fields_ = [
        ("s", C.c_int), ("s_fwd", C.c_int), ("order", C.c_int), ("err_order", C.c_int),
        ("fsal", C.c_int), ("has_err", C.c_int),
        ("a", (C.c_double * MAX_STAGES) * MAX_STAGES),
        ("b", C.c_double * MAX_STAGES),
        ("db", C.c_double * MAX_STAGES),
        ("c", C.c_double * MAX_STAGES),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("system", C.c_int), ("n", C.c_int), ("np", C.c_int), ("tableau", C.c_int),
        ("controlled", C.c_int), ("max_attempts", C.c_int), ("max_steps", C.c_int),
        ("pad_", C.c_int),
        ("t0", C.c_double), ("t1", C.c_double), ("dt0", C.c_double),
        ("abs_tol", C.c_double), ("rel_tol", C.c_double), ("sys_const", C.c_double),
    ]


class Trajectory(C.Structure):
    _fields_ = [("cap", C.c_int64), ("T", C.c_int64), ("t", C.POINTER(C.c_double)),
                ("h", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double))]


class Obs(C.Structure):
    _fields_ = [("k", C.c_int), ("t", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double))]


def observations(t_obs, y_obs):
    """An Obs structure over numpy arrays t_obs [K], y_obs [K, N] (kept alive on it)."""
    t_obs = np.ascontiguousarray(t_obs, dtype=np.float64)
    y_obs = np.ascontiguousarray(y_obs, dtype=np.float64)
    ob = Obs()
    ob.k = int(t_obs.size)
    ob.t = _dp(t_obs)
    ob.y = _dp(y_obs)
    ob._keep = (t_obs, y_obs)
    return ob


class Stats(C.Structure):
    _fields_ = [("accepted", C.c_int32), ("rejected", C.c_int32),
                ("rhs_evals", C.c_int32), ("status", C.c_int32)]


_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=HERE, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.orc_get_tableau.argtypes = [C.c_int, C.POINTER(Tableau)]
        L.orc_rhs.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, C.c_double, dp]
        L.orc_vjp.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, C.c_double, C.c_double, dp, dp, dp]
        L.orc_root.argtypes = [C.c_double, C.c_int]
        L.orc_root.restype = C.c_double
        L.orc_rk_step.argtypes = [C.POINTER(Problem), dp, dp, C.c_double, C.c_double, dp, dp]
        L.orc_traj_alloc.argtypes = [C.POINTER(Trajectory), C.c_int64, C.c_int]
        L.orc_traj_free.argtypes = [C.POINTER(Trajectory)]
        L.orc_forward.argtypes = [C.POINTER(Problem), dp, dp, C.POINTER(Trajectory),
                                  C.POINTER(Stats), dp, C.c_int64, C.POINTER(C.c_int64)]
        L.orc_forward.restype = C.c_int64
        L.orc_backward.argtypes = [C.POINTER(Problem), dp, C.POINTER(Trajectory), dp, dp]
        L.orc_adjoint_solve.argtypes = [C.POINTER(Problem), dp, C.POINTER(Trajectory), C.c_int,
                                        dp, dp]
        L.orc_sensitivity_matrix.argtypes = [C.POINTER(Problem), dp, C.POINTER(Trajectory), dp]
        L.orc_replay.argtypes = [C.POINTER(Problem), dp, dp, C.POINTER(Trajectory), dp]
        L.orc_solve_batch.argtypes = [C.POINTER(Problem), C.c_int64, C.c_int64, C.c_int64,
                                      dp, dp, dp, C.c_int, dp, dp, dp, dp, dp, dp,
                                      ip, ip, ip, C.c_int]
        L.orc_solve_batch.restype = C.c_int
        L.orc_solve_batch_digest.argtypes = [C.POINTER(Problem), C.c_int64, C.c_int64,
                                             C.c_int64, dp, dp, dp, C.c_int, dp, dp, dp, dp, dp,
                                             dp, ip, ip, ip, C.POINTER(C.c_uint64), C.c_int]
        L.orc_solve_batch_digest.restype = C.c_int
        L.orc_set_user_system.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_user_system.restype = None
        L.orc_step_digest.argtypes = [C.POINTER(Trajectory)]
        L.orc_step_digest.restype = C.c_uint64
        L.orc_forward_obs.argtypes = [C.POINTER(Problem), C.POINTER(Obs), dp, dp,
                                      C.POINTER(Trajectory), C.POINTER(Stats), ip, dp]
        L.orc_forward_obs.restype = C.c_int64
        L.orc_backward_obs.argtypes = [C.POINTER(Problem), C.POINTER(Obs), dp,
                                       C.POINTER(Trajectory), ip, dp, dp]
        L.orc_solve_batch_obs.argtypes = [C.POINTER(Problem), C.POINTER(Obs), C.c_int64,
                                          C.c_int64, C.c_int64, dp, dp, dp, dp, dp, dp, ip, ip,
                                          ip, C.c_int]
        L.orc_solve_batch_obs.restype = C.c_int
        _lib = L
    return _lib


class use_library:
    """Context manager: temporarily bind this module to another build of
    va_oracle.c (used by the mutation tests)."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        self._saved = _lib
        _lib = None
        saved_path = globals()["LIB_PATH"]
        globals()["LIB_PATH"] = self.path
        try:
            lib()
        finally:
            globals()["LIB_PATH"] = saved_path
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self._saved


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def tableau(tid: int) -> dict:
    t = Tableau()
    if lib().orc_get_tableau(tid, C.byref(t)) != 0:
        raise ValueError(f"unknown tableau {tid}")
    s = t.s_fwd
    return {
        "s": t.s, "s_fwd": t.s_fwd, "order": t.order, "err_order": t.err_order,
        "fsal": bool(t.fsal), "has_err": bool(t.has_err),
        "a": np.array([[t.a[i][j] for j in range(s)] for i in range(s)]),
        "b": np.array([t.b[i] for i in range(s)]),
        "db": np.array([t.db[i] for i in range(s)]),
        "c": np.array([t.c[i] for i in range(s)]),
    }


def problem(system, n, np_, tableau_id, *, controlled=True, t0=0.0, t1=1.0, dt0=1e-3,
            abs_tol=1e-8, rel_tol=1e-8, sys_const=0.0, max_attempts=500, max_steps=200000):
    pb = Problem()
    pb.system, pb.n, pb.np, pb.tableau = system, n, np_, tableau_id
    pb.controlled = 1 if controlled else 0
    pb.max_attempts, pb.max_steps = max_attempts, max_steps
    pb.t0, pb.t1, pb.dt0 = t0, t1, dt0
    pb.abs_tol, pb.rel_tol, pb.sys_const = abs_tol, rel_tol, sys_const
    return pb


def rhs(system, x, p, t=0.0, sys_const=0.0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    f = np.empty_like(x)
    lib().orc_rhs(system, x.size, _dp(x), _dp(p), t, sys_const, _dp(f))
    return f


def vjp(system, x, p, w, t=0.0, sys_const=0.0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    wjx = np.empty_like(x)
    wjp = np.empty_like(p)
    lib().orc_vjp(system, x.size, p.size, _dp(x), _dp(p), t, sys_const, _dp(w), _dp(wjx), _dp(wjp))
    return wjx, wjp


def root_(x, q):
    """x^(-1/q) as the step controller computes it (q = 2..8)."""
    return lib().orc_root(float(x), int(q))


def set_user_system(rhs_fn, vjp_fn):
    """Use the C functions rhs_fn(x, p, f) and vjp_fn(x, p, w, wJx, wJp) (ctypes
    function pointers, e.g. attributes of a loaded shared library) as SYS_USER."""
    lib().orc_set_user_system(C.cast(rhs_fn, C.c_void_p), C.cast(vjp_fn, C.c_void_p))


def rk_step(pb, y, p, t, h):
    y = np.ascontiguousarray(y, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty_like(y)
    err = np.empty_like(y)
    lib().orc_rk_step(C.byref(pb), _dp(y), _dp(p), t, h, _dp(out), _dp(err))
    return out, err


class Solve:
    """One instance: forward (with stored trajectory) and adjoint."""

    def __init__(self, pb, y0, p, log_cap=0, obs=None):
        self.pb = pb
        self.p = np.ascontiguousarray(p, dtype=np.float64)
        self.y = np.array(y0, dtype=np.float64)
        self.tr = Trajectory()
        cap = pb.max_steps + 2
        if lib().orc_traj_alloc(C.byref(self.tr), cap, pb.n) != 0:
            raise MemoryError
        self.st = Stats()
        self.log = np.zeros((max(log_cap, 1), 4))
        nlog = C.c_int64(0)
        self.obs = None
        if obs is not None:
            self.obs = observations(*obs)
            self.obs_row = np.zeros(self.obs.k, np.int32)
            J = C.c_double(0.0)
            self.steps = lib().orc_forward_obs(C.byref(pb), C.byref(self.obs), _dp(self.y),
                                               _dp(self.p), C.byref(self.tr), C.byref(self.st),
                                               _ip(self.obs_row), C.byref(J))
            self.J = J.value
        else:
            self.steps = lib().orc_forward(C.byref(pb), _dp(self.y), _dp(self.p),
                                           C.byref(self.tr), C.byref(self.st), _dp(self.log),
                                           log_cap, C.byref(nlog))
        self.log = self.log[: nlog.value]

    def __del__(self):
        try:
            lib().orc_traj_free(C.byref(self.tr))
        except Exception:
            pass

    @property
    def status(self):
        return self.st.status

    @property
    def times(self):
        return np.ctypeslib.as_array(self.tr.t, shape=(self.tr.T,)).copy()

    @property
    def steps_h(self):
        """Accepted step sizes h_n (length T, last entry 0)."""
        return np.ctypeslib.as_array(self.tr.h, shape=(self.tr.T,)).copy()

    @property
    def states(self):
        return np.ctypeslib.as_array(self.tr.y, shape=(self.tr.T * self.pb.n,)).reshape(
            self.tr.T, self.pb.n).copy()

    def backward(self, lam, mu=None):
        lam = np.array(lam, dtype=np.float64)
        mu = np.zeros(self.pb.np) if mu is None else np.array(mu, dtype=np.float64)
        lib().orc_backward(C.byref(self.pb), _dp(self.p), C.byref(self.tr), _dp(lam), _dp(mu))
        return lam, mu

    def backward_obs(self, lam=None):
        """Gradient of the observation cost: (dJ/dy0, dJ/dp)."""
        lam = np.zeros(self.pb.n) if lam is None else np.array(lam, dtype=np.float64)
        mu = np.zeros(self.pb.np)
        lib().orc_backward_obs(C.byref(self.pb), C.byref(self.obs), _dp(self.p),
                               C.byref(self.tr), _ip(self.obs_row), _dp(lam), _dp(mu))
        return lam, mu

    def adjoint_solve(self, lam, mu):
        lam = np.ascontiguousarray(lam, dtype=np.float64).copy()
        mu = np.ascontiguousarray(mu, dtype=np.float64).copy()
        nout = lam.shape[0]
        lib().orc_adjoint_solve(C.byref(self.pb), _dp(self.p), C.byref(self.tr), nout, _dp(lam),
                                _dp(mu))
        return lam, mu

    def sensitivity_matrix(self):
        S = np.zeros((self.pb.n, self.pb.np))
        lib().orc_sensitivity_matrix(C.byref(self.pb), _dp(self.p), C.byref(self.tr), _dp(S))
        return S

    @property
    def digest(self):
        """orc_step_digest of the stored step sequence."""
        return int(lib().orc_step_digest(C.byref(self.tr)))

    def replay(self, y0, p=None):
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        p = self.p if p is None else np.ascontiguousarray(p, dtype=np.float64)
        out = np.empty_like(y0)
        lib().orc_replay(C.byref(self.pb), _dp(y0), _dp(p), C.byref(self.tr), _dp(out))
        return out


def solve_batch(pb, y0, p, cost, cost_w=None, p_shared=None, b_begin=0, b_count=None,
                nthreads=None, obs=None):
    """SoA batch: y0 [N][B], p [P][B].  Returns a dict of SoA outputs.
    obs=(t_obs [K], y_obs [K, N]): the observation cost replaces `cost`."""
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    N, B = y0.shape
    if b_count is None:
        b_count = B - b_begin
    nthreads = nthreads or os.cpu_count() or 1
    shared = pb.system == SYS_LOTKA_VOLTERRA_SHARED
    P = p.shape[0]
    out = {
        "J": np.full(B, np.nan), "yT": np.full((N, B), np.nan), "dJdy0": np.full((N, B), np.nan),
        "dJdp": np.full((P, B), np.nan), "n_acc": np.zeros(B, np.int32),
        "n_rej": np.zeros(B, np.int32), "status": np.full(B, -1, np.int32),
        "digest": np.zeros(B, np.uint64),
    }
    if shared:
        out["dJdA"] = np.zeros(N * N)
        p_shared = np.ascontiguousarray(p_shared, dtype=np.float64)
    cw = None if cost_w is None else np.ascontiguousarray(cost_w, dtype=np.float64)
    if obs is not None:
        ob = observations(*obs)
        rc = lib().orc_solve_batch_obs(
            C.byref(pb), C.byref(ob), B, b_begin, b_count, _dp(y0), _dp(p), _dp(out["J"]),
            _dp(out["yT"]), _dp(out["dJdy0"]), _dp(out["dJdp"]), _ip(out["n_acc"]),
            _ip(out["n_rej"]), _ip(out["status"]), nthreads)
        if rc != 0:
            raise ValueError("orc_solve_batch_obs rejected its arguments")
        return out
    rc = lib().orc_solve_batch_digest(
        C.byref(pb), B, b_begin, b_count, _dp(y0), _dp(p), _dp(p_shared) if shared else None,
        cost, _dp(cw), _dp(out["J"]), _dp(out["yT"]), _dp(out["dJdy0"]), _dp(out["dJdp"]),
        _dp(out["dJdA"]) if shared else None, _ip(out["n_acc"]), _ip(out["n_rej"]),
        _ip(out["status"]), out["digest"].ctypes.data_as(C.POINTER(C.c_uint64)), nthreads)
    if rc != 0:
        raise ValueError("orc_solve_batch rejected its arguments")
    return out
