"""GPU parity of the shared-interaction Lotka-Volterra solve (N = 128, A common to the
batch, stage products on FP64 MFMA) against the CPU oracle.

The MFMA products sum in a different order from the oracle's growth sums, so the
forward is compared to rounding level rather than bit for bit.  Bar: identical
accepted/rejected counts on at least 99.9% of the instances (all, at these batch
sizes); y(T) and J within 1e-12 relative; dJ/dy0, dJ/dr and the batch-summed
dJ/dA within 1e-10 relative (the stated floating-point tolerance).
"""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _solve(name_tab, B, t1=None):
    import torch
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import constants as K
    from vectorizedadjoint_amd import inputs as I
    assert torch.cuda.is_available(), "no HIP device"
    w = I.WORKLOADS["glv128_shared"]
    inp = I.make_inputs(w, B)
    kw = {"tableau": name_tab, "max_steps": 512}
    if t1 is not None:
        kw["t1"] = t1
    prob = va.Problem.from_workload(w, batch=B, **kw)
    print(f"[gpu] shared GLV N=128 B={B} {kw}", flush=True)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    with va.Solver(prob) as s:
        out = s.solve_shared(dev(inp["y0"]), dev(inp["p"]),
                             dev(inp["p_shared"].reshape(128, 128)), K.COST_SUM)
        torch.cuda.synchronize()
        g = {k: v.cpu().numpy() for k, v in out.items()}
    import oracle as O
    pb = O.problem(w.system, w.n, w.np, name_tab, t0=w.t0, t1=w.t1 if t1 is None else t1,
                   dt0=w.dt0, abs_tol=w.tol, rel_tol=w.tol, max_steps=512)
    o = O.solve_batch(pb, inp["y0"], inp["p"], w.cost, p_shared=inp["p_shared"], nthreads=8)
    return w, g, o


def _check(g, o):
    assert np.array_equal(g["status"], o["status"])
    same = (g["n_acc"] == o["n_acc"]) & (g["n_rej"] == o["n_rej"])
    assert same.mean() >= 0.999, same.mean()
    ok = (o["status"] == 0) & same
    assert ok.any()
    assert rel_err(g["yT"][:, ok], o["yT"][:, ok]).max() <= 1e-12
    assert rel_err(g["J"][ok], o["J"][ok]).max() <= 1e-12
    assert rel_err(g["dJdy0"][:, ok], o["dJdy0"][:, ok]).max() <= TOL
    assert rel_err(g["dJdr"][:, ok], o["dJdp"][:, ok]).max() <= TOL
    dA = o["dJdA"].reshape(128, 128)
    assert np.max(np.abs(g["dJdA"] - dA)) / np.max(np.abs(dA)) <= TOL


@pytest.mark.parametrize("B", [64, 100])
def test_shared_cash_karp_matches_oracle(B):
    from vectorizedadjoint_amd import constants as K
    _, g, o = _solve(K.TAB_CASH_KARP54, B)
    _check(g, o)
    if B == 100:  # a partial last tile of 4 instances
        assert np.all(np.isfinite(g["dJdr"]))


def test_shared_dopri5_matches_oracle():
    from vectorizedadjoint_amd import constants as K
    _, g, o = _solve(K.TAB_DOPRI5, 48, t1=20.0)
    _check(g, o)


def test_shared_handle_rejects_per_instance_entry_points():
    import torch
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    w = I.WORKLOADS["glv128_shared"]
    with va.Solver(va.Problem.from_workload(w, batch=16, max_steps=64)) as s:
        y = torch.zeros((128, 16), dtype=torch.float64, device="cuda")
        with pytest.raises(va.VaError, match="va_solve_shared"):
            s.forward(y, y)


def test_shared_full_batch_slice_matches_oracle():
    """The shared-A workload at its full size (GLV N=128, B=16384, Cash-Karp 5(4),
    tol 1e-8: 1024 tiles over every CU, each workgroup taking several), with the
    per-instance outputs of a 512-instance slice compared to the oracle (the whole
    batch would take the oracle half a minute; the batch-summed dJ/dA is covered
    against the whole-batch oracle by the smaller cases above).  The handle solves
    twice, and every per-instance output of the second solve must equal the
    first's bit for bit (the batch-summed dJ/dA to 1e-12: the reverse takes its
    tiles longest-first, in an order that may differ within a bucket)."""
    import os
    import torch
    import oracle as O
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import constants as K
    from vectorizedadjoint_amd import inputs as I
    assert torch.cuda.is_available(), "no HIP device"
    w = I.WORKLOADS["glv128_shared"]
    B = w.batch
    assert B == 16384
    inp = I.make_inputs(w, B)
    prob = va.Problem.from_workload(w, max_steps=512)
    print(f"[gpu] shared GLV N=128 full batch B={B}", flush=True)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    with va.Solver(prob) as s:
        runs = []
        for _ in range(2):
            out = s.solve_shared(dev(inp["y0"]), dev(inp["p"]),
                                 dev(inp["p_shared"].reshape(128, 128)), K.COST_SUM)
            torch.cuda.synchronize()
            runs.append({k: v.cpu().numpy() for k, v in out.items()})
    g0, g = runs
    for k in ("status", "n_acc", "n_rej", "yT", "J", "dJdy0", "dJdr"):
        assert np.array_equal(g0[k], g[k]), k
    assert np.max(np.abs(g["dJdA"] - g0["dJdA"])) <= 1e-12 * np.max(np.abs(g0["dJdA"]))
    assert np.all(np.isfinite(g["dJdA"]))
    pb = O.problem(w.system, w.n, w.np, w.tableau, t0=w.t0, t1=w.t1, dt0=w.dt0,
                   abs_tol=w.tol, rel_tol=w.tol, max_steps=512)
    b0, cnt = 7777, 512
    o = O.solve_batch(pb, inp["y0"], inp["p"], w.cost, p_shared=inp["p_shared"], b_begin=b0,
                      b_count=cnt, nthreads=min(16, os.cpu_count() or 1))
    sl = slice(b0, b0 + cnt)
    assert np.array_equal(g["status"][sl], o["status"][sl])
    same = (g["n_acc"][sl] == o["n_acc"][sl]) & (g["n_rej"][sl] == o["n_rej"][sl])
    assert same.mean() >= 0.999, same.mean()
    ok = np.zeros(B, bool)
    ok[sl] = (o["status"][sl] == 0) & same
    assert ok.sum() >= 0.99 * cnt
    assert rel_err(g["yT"][:, ok], o["yT"][:, ok]).max() <= 1e-12
    assert rel_err(g["J"][ok], o["J"][ok]).max() <= 1e-12
    e0 = rel_err(g["dJdy0"][:, ok], o["dJdy0"][:, ok]).max()
    er = rel_err(g["dJdr"][:, ok], o["dJdp"][:, ok]).max()
    print(f"    slice [{b0}, {b0 + cnt}): max rel err dJdy0 {e0:.2e} dJdr {er:.2e}", flush=True)
    assert e0 <= TOL and er <= TOL
