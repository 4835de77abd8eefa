"""Shared helpers of the parity tests (oracle runs, error metrics)."""
import numpy as np


def rel_err(a, b):
    """Per-instance norm-wise relative error max_i|a-b| / max_i|b| over axis 0 of [F, B] arrays
    (for [B] arrays, elementwise |a-b|/|b|).  Returns the array of per-instance errors."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 1:
        return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    num = np.max(np.abs(a - b), axis=0)
    den = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
    return num / den


def run_oracle(w, inputs, nthreads=8, **pb_kw):
    import oracle as O
    kw = dict(t0=w.t0, t1=w.t1, dt0=w.dt0, abs_tol=w.tol, rel_tol=w.tol,
              controlled=w.controlled, sys_const=w.sys_const)
    kw.update(pb_kw)
    pb = O.problem(w.system, w.n, w.np, w.tableau, **kw)
    return O.solve_batch(pb, inputs["y0"], inputs["p"], w.cost, p_shared=inputs.get("p_shared"),
                         nthreads=nthreads)
