"""Seeded synthetic workloads (identical arrays feed the GPU path and the CPU baseline).

Random numbers: SplitMix64 used as a counter-based generator.  Value k of the
stream for field f is mix64(s_f + (k+1) * 0x9E3779B97F4A7C15) with
s_f = mix64(seed + (f+1) * 0x9E3779B97F4A7C15); uniforms take the top 53 bits;
normals are Box-Muller (cosine branch, two uniforms each).  The counter of
instance b, component i is b * width + i, so the first K instances of a batch
do not depend on the batch size.

Configurations follow the reference examples' systems
(examples/VanDerPol/main.cpp:39-43, examples/HarmonicOscillator/main.cpp:25-29,
examples/GeneralizedLotkaVolterra/main.cpp:105-119) with the distributions of
the workload table: seed 20241000 + k for workload k, initial step 1e-3 as in
examples/VanDerPol/main.cpp:71 and examples/GeneralizedLotkaVolterra/main.cpp:165.
Parameter arrays are structure-of-arrays: p[k][b] with the instance fastest.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import constants as K

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _field_seed(seed: int, fld: int) -> np.uint64:
    with np.errstate(over="ignore"):
        s = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + np.uint64(fld + 1) * _GOLDEN
    return np.uint64(_mix64(np.array([s], dtype=np.uint64))[0])


def uniform(seed: int, fld: int, B: int, width: int) -> np.ndarray:
    """[B, width] uniforms in [0, 1)."""
    k = np.arange(B * width, dtype=np.uint64)
    s = _field_seed(seed, fld)
    with np.errstate(over="ignore"):
        st = s + (k + np.uint64(1)) * _GOLDEN
    z = _mix64(st)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return u.reshape(B, width)


def normal(seed: int, fld: int, B: int, width: int) -> np.ndarray:
    u = uniform(seed, fld, B, 2 * width).reshape(B, width, 2)
    r = np.sqrt(-2.0 * np.log1p(-u[..., 0]))
    return r * np.cos(2.0 * math.pi * u[..., 1])


@dataclass
class Workload:
    name: str
    index: int
    system: int
    n: int
    np: int
    tableau: int
    t0: float
    t1: float
    dt0: float
    tol: float
    batch: int
    cost: int
    description: str
    controlled: bool = True
    sys_const: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def seed(self) -> int:
        return 20241000 + self.index


WORKLOADS = {
    "vdp": Workload(
        "vdp", 1, K.SYS_VAN_DER_POL, 2, 1, K.TAB_DOPRI5, 0.0, 20.0, 1e-3, 1e-8, 1024,
        K.COST_SQUARED_NORM,
        "Van der Pol x1'=mu((1-x0^2)x1-x0), mu~U(0.5,5), y0=(2,0)+N(0,0.01^2), t in [0,20], "
        "DOPRI5 tol 1e-8, J=|y(T)|^2, B=1024"),
    "oscillator": Workload(
        "oscillator", 2, K.SYS_OSCILLATOR, 2, 2, K.TAB_DOPRI5, 0.0, 30.0, 1e-3, 1e-10, 1 << 20,
        K.COST_SQUARED_NORM,
        "damped oscillator y''+c y'+k y=0, k~U(0.5,4), c~U(0,0.5), y0~U(-1,1)^2, t in [0,30], "
        "DOPRI5 tol 1e-10, J=|y(T)|^2, B=2^20"),
    "glv10": Workload(
        "glv10", 3, K.SYS_LOTKA_VOLTERRA, 10, 110, K.TAB_CASH_KARP54, 0.0, 50.0, 1e-3, 1e-8,
        65536, K.COST_SUM,
        "generalized Lotka-Volterra N=10, r~U(0.1,1), A_ij~N(0,0.2^2), A_ii=-U(0.5,1.5), "
        "x0~U(0.1,1), t in [0,50], Cash-Karp 5(4) tol 1e-8, J=sum x_i(T), B=65536"),
    "vdp_obs": Workload(
        "vdp_obs", 5, K.SYS_VAN_DER_POL, 2, 1, K.TAB_DOPRI5, 0.0, 20.0, 1e-3, 1e-8, 262144,
        0,
        "Van der Pol with mu log-uniform on [0.1, 100], y0=(2,0)+N(0,0.01^2), t in [0,20], "
        "DOPRI5 tol 1e-8, J = sum_k |y(t_k) - y_obs(t_k)|^2 at t_k = 20k/50 (k=1..50), "
        "observations from a seeded ground-truth mu, B=262144, 4 GB trajectory budget",
        extra={"n_obs": 50, "budget_bytes": 4 << 30}),
    "brusselator": Workload(
        "brusselator", 6, K.SYS_USER, 2, 2, K.TAB_DOPRI5, 0.0, 20.0, 1e-3, 1e-8, 4096,
        K.COST_SQUARED_NORM,
        "Brusselator x0'=a+x0^2 x1-(b+1)x0, x1'=b x0-x0^2 x1 (a system of the caller's, "
        "compiled outside the library), a~U(0.5,1.5), b~U(1.5,3.5), y0~U(0.5,2)^2, "
        "t in [0,20], DOPRI5 tol 1e-8, J=|y(T)|^2, B=4096"),
    "glv128_shared": Workload(
        "glv128_shared", 4, K.SYS_LOTKA_VOLTERRA_SHARED, 128, 128, K.TAB_CASH_KARP54, 0.0, 50.0,
        1e-3, 1e-8, 16384, K.COST_SUM,
        "generalized Lotka-Volterra N=128 with shared A_ij~N(0,0.05^2), A_ii=-1, per-instance "
        "r~U(0.1,1), x0~U(0.05,0.5), t in [0,50], Cash-Karp 5(4) tol 1e-8, J=sum x_i(T), "
        "B=16384"),
}


def make_inputs(w: Workload | str, B: int | None = None) -> dict:
    """Return {'y0': [N][B], 'p': [P][B], 'p_shared': array or None} (float64, C order)."""
    if isinstance(w, str):
        w = WORKLOADS[w]
    B = w.batch if B is None else int(B)
    s = w.seed
    N = w.n
    if w.system == K.SYS_VAN_DER_POL and w.name == "vdp_obs":
        mu = np.exp(math.log(0.1) + (math.log(100.0) - math.log(0.1)) * uniform(s, 0, B, 1))
        nz = normal(s, 1, B, 2) * 0.01
        y0 = np.stack([2.0 + nz[:, 0], nz[:, 1]])
        return {"y0": np.ascontiguousarray(y0), "p": np.ascontiguousarray(mu.T.copy()),
                "p_shared": None}
    if w.system == K.SYS_VAN_DER_POL:
        mu = 0.5 + 4.5 * uniform(s, 0, B, 1)
        nz = normal(s, 1, B, 2) * 0.01
        y0 = np.stack([2.0 + nz[:, 0], nz[:, 1]])
        p = mu.T.copy()
        return {"y0": np.ascontiguousarray(y0), "p": np.ascontiguousarray(p), "p_shared": None}
    if w.system == K.SYS_OSCILLATOR:
        u = uniform(s, 0, B, 2)
        k = 0.5 + 3.5 * u[:, 0]
        c = 0.5 * u[:, 1]
        y0 = -1.0 + 2.0 * uniform(s, 1, B, 2)
        return {"y0": np.ascontiguousarray(y0.T), "p": np.ascontiguousarray(np.stack([k, c])),
                "p_shared": None}
    if w.system == K.SYS_LOTKA_VOLTERRA:
        r = 0.1 + 0.9 * uniform(s, 0, B, N)
        A = 0.2 * normal(s, 1, B, N * N)
        dg = -(0.5 + uniform(s, 2, B, N))
        A = A.reshape(B, N, N)
        A[:, np.arange(N), np.arange(N)] = dg
        x0 = 0.1 + 0.9 * uniform(s, 3, B, N)
        p = np.concatenate([r, A.reshape(B, N * N)], axis=1)
        return {"y0": np.ascontiguousarray(x0.T), "p": np.ascontiguousarray(p.T),
                "p_shared": None}
    if w.system == K.SYS_USER and w.name == "brusselator":
        u = uniform(s, 0, B, 2)
        a = 0.5 + u[:, 0]
        b = 1.5 + 2.0 * u[:, 1]
        y0 = 0.5 + 1.5 * uniform(s, 1, B, 2)
        return {"y0": np.ascontiguousarray(y0.T), "p": np.ascontiguousarray(np.stack([a, b])),
                "p_shared": None}
    if w.system == K.SYS_LOTKA_VOLTERRA_SHARED:
        A = 0.05 * normal(s, 1, 1, N * N).reshape(N, N)
        A[np.arange(N), np.arange(N)] = -1.0
        r = 0.1 + 0.9 * uniform(s, 0, B, N)
        x0 = 0.05 + 0.45 * uniform(s, 3, B, N)
        return {"y0": np.ascontiguousarray(x0.T), "p": np.ascontiguousarray(r.T),
                "p_shared": np.ascontiguousarray(A.reshape(N * N))}
    raise ValueError(f"no generator for system {w.system}")


def observation_schedule(w: Workload | str):
    """(t_obs [K], mu_true, y0_true) of an observed workload: K uniform times in
    (t0, t1] and a ground-truth parameter drawn from the workload's seed."""
    if isinstance(w, str):
        w = WORKLOADS[w]
    K_obs = int(w.extra["n_obs"])
    t_obs = w.t0 + (w.t1 - w.t0) * np.arange(1, K_obs + 1) / K_obs
    u = uniform(w.seed, 7, 1, 1)[0, 0]
    mu_true = math.exp(math.log(0.1) + (math.log(100.0) - math.log(0.1)) * u)
    return t_obs, mu_true, np.array([2.0, 0.0])
