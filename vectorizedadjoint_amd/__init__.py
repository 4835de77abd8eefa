"""vectorizedadjoint_amd -- batched adaptive explicit Runge-Kutta forward sweep and
discrete-adjoint reverse sweep on MI355X (HIP, gfx950, FP64).

Python mirror of the C ABI in include/vectorizedadjoint_amd.h (which the C++
drop-in headers under include/vectorizedadjoint/ also bind).  Device memory is
plain torch tensors (float64 / int32 on a ROCm device); the library only sees
their pointers.  There is no CPU fallback: every call goes to the HIP library
and raises if it is missing.

Reference interface mirrored (file:line in RuiMartins1996/VectorizedAdjoint):
    Driver(Nin, Nout, Npar) ........ lib/include/Driver.hpp:39-42        -> Solver(problem)
    runge_kutta(...) ............... lib/include/runge_kutta.hpp:47-59    -> Solver.forward
    setCostGradients + adjointSolve  lib/include/Driver.hpp:103-114,
                                     lib/include/backpropagation.hpp:18-43 -> Solver.adjoint
    computeSensitivityMatrix ....... lib/include/backpropagation.hpp:67-87 -> Solver.sensitivity_matrix
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

from .constants import *  # noqa: F401,F403
from . import constants as _K

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvectorizedadjoint_amd.so")
ABI_VERSION = 2

C_ABI_FUNCTIONS = (
    "va_problem_init", "va_create", "va_destroy", "va_forward", "va_adjoint",
    "va_sensitivity_matrix", "va_solve", "va_solve_observed", "va_solve_shared", "va_trajectory",
    "va_step_digest", "va_max_steps_seen", "va_last_error", "va_solve_noalloc",
    "va_create_with_kernels", "va_autotune",
)


class VaProblem(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("system", C.c_int32), ("tableau", C.c_int32),
        ("adaptive", C.c_int32), ("n_state", C.c_int32), ("n_param", C.c_int32),
        ("batch", C.c_int64),
        ("t0", C.c_double), ("t1", C.c_double), ("dt0", C.c_double),
        ("abs_tol", C.c_double), ("rel_tol", C.c_double), ("system_constant", C.c_double),
        ("max_attempts", C.c_int32), ("max_steps", C.c_int32), ("mode", C.c_int32),
        ("block_size", C.c_int32), ("instances_per_wave", C.c_int32),
        ("checkpoint_slots", C.c_int32), ("lanes", C.c_int64),
        ("trajectory_budget_bytes", C.c_int64), ("n_obs", C.c_int32), ("flags", C.c_int32),
    ]


class VaError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{_K.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_lib_handle = None


def library_path() -> str:
    return LIB_PATH


def lib() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m vectorizedadjoint_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.va_problem_init.argtypes = [C.POINTER(VaProblem)]
        L.va_problem_init.restype = None
        L.va_create.argtypes = [C.POINTER(VaProblem), C.POINTER(vp)]
        L.va_create_with_kernels.argtypes = [C.POINTER(VaProblem), vp, C.POINTER(vp)]
        L.va_autotune.argtypes = [C.POINTER(VaProblem), vp, vp, vp, C.c_int32, vp, vp, vp,
                                  C.c_int64, C.c_int32, vp]
        L.va_destroy.argtypes = [vp]
        L.va_destroy.restype = None
        L.va_forward.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.va_adjoint.argtypes = [vp, vp, C.c_int32, vp, vp, vp]
        L.va_sensitivity_matrix.argtypes = [vp, vp, vp, vp]
        L.va_solve.argtypes = [vp, vp, vp, C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.va_max_steps_seen.argtypes = [vp, C.POINTER(C.c_int32), vp]
        L.va_step_digest.argtypes = [vp, vp, vp]
        L.va_solve_noalloc.argtypes = [C.POINTER(VaProblem), vp, vp, C.c_int32] + [vp] * 9
        L.va_solve_observed.argtypes = [vp] * 13
        L.va_solve_shared.argtypes = [vp] * 4 + [C.c_int32] + [vp] * 10
        L.va_trajectory.argtypes = [vp, C.c_int64, vp, C.c_int64, C.POINTER(C.c_int64), vp]
        L.va_last_error.argtypes = []
        L.va_last_error.restype = C.c_char_p
        for f in ("va_create", "va_create_with_kernels", "va_autotune", "va_forward", "va_adjoint", "va_sensitivity_matrix", "va_solve",
                  "va_solve_observed", "va_solve_shared", "va_trajectory", "va_max_steps_seen",
                  "va_step_digest", "va_solve_noalloc"):
            getattr(L, f).restype = C.c_int
        _lib_handle = L
    return _lib_handle


def _check(status: int) -> None:
    if status != 0:
        raise VaError(status, lib().va_last_error().decode(errors="replace"))


@dataclass
class Problem:
    system: int
    n_state: int
    n_param: int
    tableau: int
    batch: int
    t0: float = 0.0
    t1: float = 1.0
    dt0: float = 1e-3
    abs_tol: float = 1e-6
    rel_tol: float = 1e-6
    adaptive: bool = True
    system_constant: float = 0.0
    max_attempts: int = 500
    max_steps: int = 4096
    mode: int = _K.MODE_PERSISTENT
    block_size: int = 0
    instances_per_wave: int = 0
    lanes: int = 0
    trajectory_budget_bytes: int = 0
    n_obs: int = 0
    checkpoint_slots: int = 0
    flags: int = 0

    def to_c(self) -> VaProblem:
        pb = VaProblem()
        lib().va_problem_init(C.byref(pb))
        pb.abi_version = ABI_VERSION
        pb.system, pb.tableau = self.system, self.tableau
        pb.adaptive = 1 if self.adaptive else 0
        pb.n_state, pb.n_param, pb.batch = self.n_state, self.n_param, self.batch
        pb.t0, pb.t1, pb.dt0 = self.t0, self.t1, self.dt0
        pb.abs_tol, pb.rel_tol = self.abs_tol, self.rel_tol
        pb.system_constant = self.system_constant
        pb.max_attempts, pb.max_steps = self.max_attempts, self.max_steps
        pb.mode, pb.block_size, pb.lanes = self.mode, self.block_size, self.lanes
        pb.instances_per_wave = self.instances_per_wave
        pb.trajectory_budget_bytes = self.trajectory_budget_bytes
        pb.n_obs, pb.checkpoint_slots = self.n_obs, self.checkpoint_slots
        pb.flags = self.flags
        return pb

    @classmethod
    def from_workload(cls, w, batch=None, **kw) -> "Problem":
        args = dict(system=w.system, n_state=w.n, n_param=w.np, tableau=w.tableau,
                    batch=w.batch if batch is None else batch, t0=w.t0, t1=w.t1, dt0=w.dt0,
                    abs_tol=w.tol, rel_tol=w.tol, adaptive=w.controlled,
                    system_constant=w.sys_const)
        args.update(kw)
        return cls(**args)


def _ptr(t) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def _stream(stream) -> int | None:
    if stream is not None:
        return int(stream)
    import torch
    return torch.cuda.current_stream().cuda_stream


class Solver:
    """One problem over a batch of instances (the reference's Driver, batched).

    Arrays are torch tensors on the current ROCm device, SoA with the instance
    index fastest: y [N, B], p [P, B], lambda [Nout, N, B], mu [Nout, P, B].
    """

    def __init__(self, problem: Problem, kernels=None):
        """kernels: for problem.system == SYS_USER, the va_kernel_table pointer that a
        caller's VA_DEVICE_SYSTEM function returned (an int or ctypes pointer)."""
        self.problem = problem
        self._h = C.c_void_p()
        self._pb = problem.to_c()
        self.generation = 0  # forwards run so far; the reverse uses the last one's trajectory
        if kernels is None:
            _check(lib().va_create(C.byref(self._pb), C.byref(self._h)))
        else:
            _check(lib().va_create_with_kernels(C.byref(self._pb), C.c_void_p(kernels),
                                                C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().va_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _shape_check(self, t, shape, dtype_name):
        import torch
        dt = {"f64": torch.float64, "i32": torch.int32}[dtype_name]
        if t is None:
            return
        if t.dtype != dt or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != shape:
            raise ValueError(f"expected contiguous {dtype_name} device tensor of shape {shape}, "
                             f"got {tuple(t.shape)} {t.dtype} on {t.device}")

    def forward(self, y, p, n_accepted=None, n_rejected=None, lane_status=None, stream=None):
        """runge_kutta: y [N,B] in y0 / out y(T) in place."""
        N, P, B = self.problem.n_state, self.problem.n_param, self.problem.batch
        self._shape_check(y, (N, B), "f64")
        self._shape_check(p, (P, B), "f64")
        for t in (n_accepted, n_rejected, lane_status):
            self._shape_check(t, (B,), "i32")
        _check(lib().va_forward(self._h, _ptr(y), _ptr(p), _ptr(n_accepted), _ptr(n_rejected),
                                _ptr(lane_status), _stream(stream)))
        self.generation += 1
        return y

    def adjoint(self, p, lam, mu, stream=None):
        """adjointSolve: lam [Nout,N,B] overwritten with dJ/dy0, mu [Nout,P,B] += dJ/dp."""
        N, P, B = self.problem.n_state, self.problem.n_param, self.problem.batch
        nout = lam.shape[0]
        self._shape_check(p, (P, B), "f64")
        self._shape_check(lam, (nout, N, B), "f64")
        self._shape_check(mu, (nout, P, B), "f64")
        _check(lib().va_adjoint(self._h, _ptr(p), nout, _ptr(lam), _ptr(mu), _stream(stream)))
        return lam, mu

    def sensitivity_matrix(self, p, S=None, stream=None):
        import torch
        N, P, B = self.problem.n_state, self.problem.n_param, self.problem.batch
        if S is None:
            S = torch.empty((N, P, B), dtype=torch.float64, device=p.device)
        self._shape_check(S, (N, P, B), "f64")
        _check(lib().va_sensitivity_matrix(self._h, _ptr(p), _ptr(S), _stream(stream)))
        return S

    def solve(self, y0, p, cost, cost_weights=None, out=None, stream=None):
        """Forward + terminal cost + adjoint.  Returns a dict of device tensors."""
        import torch
        N, P, B = self.problem.n_state, self.problem.n_param, self.problem.batch
        dev = y0.device
        if out is None:
            out = {
                "J": torch.empty(B, dtype=torch.float64, device=dev),
                "yT": torch.empty((N, B), dtype=torch.float64, device=dev),
                "dJdy0": torch.empty((N, B), dtype=torch.float64, device=dev),
                "dJdp": torch.empty((P, B), dtype=torch.float64, device=dev),
                "n_acc": torch.empty(B, dtype=torch.int32, device=dev),
                "n_rej": torch.empty(B, dtype=torch.int32, device=dev),
                "status": torch.empty(B, dtype=torch.int32, device=dev),
            }
        self._shape_check(y0, (N, B), "f64")
        self._shape_check(p, (P, B), "f64")
        _check(lib().va_solve(self._h, _ptr(y0), _ptr(p), int(cost), _ptr(cost_weights),
                              _ptr(out["J"]), _ptr(out["yT"]), _ptr(out["dJdy0"]),
                              _ptr(out["dJdp"]), _ptr(out.get("n_acc")), _ptr(out.get("n_rej")),
                              _ptr(out.get("status")), _stream(stream)))
        self.generation += 1
        return out

    def solve_observed(self, y0, p, obs_t, obs_y, out=None, stream=None):
        """Forward with stops at obs_t [K], running cost J = sum_k |y(t_k) - obs_y[k]|^2
        (obs_y [K, N], shared by the batch) and its gradient.  Returns a dict of device
        tensors J, yT, dJdy0, dJdp, n_acc, n_rej, status."""
        import torch
        N, P, B = self.problem.n_state, self.problem.n_param, self.problem.batch
        K = self.problem.n_obs
        dev = y0.device
        if out is None:
            out = {
                "J": torch.empty(B, dtype=torch.float64, device=dev),
                "yT": torch.empty((N, B), dtype=torch.float64, device=dev),
                "dJdy0": torch.empty((N, B), dtype=torch.float64, device=dev),
                "dJdp": torch.empty((P, B), dtype=torch.float64, device=dev),
                "n_acc": torch.empty(B, dtype=torch.int32, device=dev),
                "n_rej": torch.empty(B, dtype=torch.int32, device=dev),
                "status": torch.empty(B, dtype=torch.int32, device=dev),
            }
        self._shape_check(y0, (N, B), "f64")
        self._shape_check(p, (P, B), "f64")
        self._shape_check(obs_t, (K,), "f64")
        self._shape_check(obs_y, (K, N), "f64")
        _check(lib().va_solve_observed(
            self._h, _ptr(y0), _ptr(p), _ptr(obs_t), _ptr(obs_y), _ptr(out["J"]),
            _ptr(out["yT"]), _ptr(out["dJdy0"]), _ptr(out["dJdp"]), _ptr(out.get("n_acc")),
            _ptr(out.get("n_rej")), _ptr(out.get("status")), _stream(stream)))
        self.generation += 1
        return out

    def solve_shared(self, y0, r, A, cost, cost_weights=None, out=None, stream=None):
        """Lotka-Volterra with A [N, N] shared by the batch: forward, terminal cost,
        adjoint.  Returns device tensors J, yT, dJdy0, dJdr [N, B] and dJdA [N, N] (summed
        over the instances whose forward ended OK)."""
        import torch
        N, B = self.problem.n_state, self.problem.batch
        dev = y0.device
        if out is None:
            out = {
                "J": torch.empty(B, dtype=torch.float64, device=dev),
                "yT": torch.empty((N, B), dtype=torch.float64, device=dev),
                "dJdy0": torch.empty((N, B), dtype=torch.float64, device=dev),
                "dJdr": torch.empty((N, B), dtype=torch.float64, device=dev),
                "dJdA": torch.empty((N, N), dtype=torch.float64, device=dev),
                "n_acc": torch.empty(B, dtype=torch.int32, device=dev),
                "n_rej": torch.empty(B, dtype=torch.int32, device=dev),
                "status": torch.empty(B, dtype=torch.int32, device=dev),
            }
        self._shape_check(y0, (N, B), "f64")
        self._shape_check(r, (N, B), "f64")
        self._shape_check(A, (N, N), "f64")
        _check(lib().va_solve_shared(
            self._h, _ptr(y0), _ptr(r), _ptr(A), int(cost), _ptr(cost_weights), _ptr(out["J"]),
            _ptr(out["yT"]), _ptr(out["dJdy0"]), _ptr(out["dJdr"]), _ptr(out["dJdA"]),
            _ptr(out.get("n_acc")), _ptr(out.get("n_rej")), _ptr(out.get("status")),
            _stream(stream)))
        self.generation += 1
        return out

    def trajectory(self, instance: int, stream=None):
        """Rows (t_n, h_n, y_n) stored for one instance by the last forward, as a
        numpy array [T, N+2] (Driver::GetTime/GetDt/GetState)."""
        import numpy as np
        n = C.c_int64(0)
        _check(lib().va_trajectory(self._h, instance, None, 0, C.byref(n), _stream(stream)))
        rows = np.empty((max(n.value, 1), self.problem.n_state + 2), dtype=np.float64)
        _check(lib().va_trajectory(self._h, instance, rows.ctypes.data, n.value, C.byref(n),
                                   _stream(stream)))
        return rows[: n.value]

    def step_digest(self, stream=None):
        """Per-instance digest of the stored step sequence (t_n, h_n) of the last
        forward, as an int64 tensor [B] holding the uint64 bits (0 where the forward
        did not end OK); equal to the oracle's orc_step_digest."""
        import torch
        out = torch.empty(self.problem.batch, dtype=torch.int64, device="cuda")
        _check(lib().va_step_digest(self._h, _ptr(out), _stream(stream)))
        return out

    def max_steps_seen(self, stream=None) -> int:
        v = C.c_int32(0)
        _check(lib().va_max_steps_seen(self._h, C.byref(v), _stream(stream)))
        return int(v.value)
