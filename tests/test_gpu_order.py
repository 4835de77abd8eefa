"""Schedule reuse.  A persistent forward whose batch exceeds the instances in flight
takes its instances longest-first by the previous forward's attempt counts (index
order the first time).  Repeated solves on one handle -- the first in index order,
the later ones in the reused order -- must give bit-identical outputs, equal to
those of a handle that keeps index order (VA_FLAG_COLD_ORDER), and the oracle's.
"""
import numpy as np
import pytest

from helpers import run_oracle
from test_gpu_parity import _compare, _dev, _torch

pytestmark = pytest.mark.gpu


def _same(outs):
    for g in outs[1:]:
        for k in outs[0]:
            assert np.array_equal(g[k], outs[0][k], equal_nan=True), k


@pytest.mark.parametrize("name,B,lanes", [("glv10", 3000, 120), ("vdp", 3000, 256)])
def test_reused_order_is_bit_identical(name, B, lanes):
    torch = _torch()
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    w = I.WORKLOADS[name]
    inp = I.make_inputs(w, B)
    o = run_oracle(w, inp, max_steps=1024)
    outs = []
    for flags in (0, va.FLAG_COLD_ORDER):
        pb = va.Problem.from_workload(w, batch=B, max_steps=1024, lanes=lanes, flags=flags)
        print(f"[gpu] {name} B={B} lanes={lanes} flags={flags}", flush=True)
        with va.Solver(pb) as s:
            y0, p = _dev(inp["y0"], torch), _dev(inp["p"], torch)
            for _ in range(3):
                out = s.solve(y0, p, w.cost)
                torch.cuda.synchronize()
                outs.append({k: v.cpu().numpy() for k, v in out.items()})
    _compare(outs[0], o)
    _same(outs)


@pytest.mark.parametrize("kw", [{}, {"checkpoint_slots": 16}])
def test_reused_order_observed(kw):
    torch = _torch()
    from test_gpu_observed import _setup
    B = 1500
    torch, O, va, w, inp, t_obs, y_obs, pb = _setup(B)
    prob = va.Problem.from_workload(w, batch=B, t1=20.0, n_obs=t_obs.size, max_steps=16384,
                                    lanes=256, **kw)
    print(f"[gpu] observed B={B} lanes=256 {kw}", flush=True)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    outs = []
    with va.Solver(prob) as s:
        args = (dev(inp["y0"]), dev(inp["p"]), dev(t_obs), dev(y_obs))
        for _ in range(3):
            out = s.solve_observed(*args)
            torch.cuda.synchronize()
            outs.append({k: v.cpu().numpy() for k, v in out.items()})
    o = O.solve_batch(pb, inp["y0"], inp["p"], 0, obs=(t_obs, y_obs), nthreads=8)
    assert np.array_equal(outs[0]["n_acc"], o["n_acc"])
    assert np.array_equal(outs[0]["yT"], o["yT"]) and np.array_equal(outs[0]["J"], o["J"])
    _same(outs)


def test_separate_calls_split_schedule():
    """va_forward and va_adjoint on a handle whose batch exceeds the instances in
    flight: from the second forward on, the instances that ran long run one per
    wave beside the persistent kernels (the split schedule), in the forward alone
    and in the reverse alone.  Every output stays bit-identical to va_solve's and
    to index order; the sensitivity matrix (10 seeds per instance) likewise."""
    torch = _torch()
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    w = I.WORKLOADS["glv10"]
    B, lanes = 3000, 120
    inp = I.make_inputs(w, B)
    o = run_oracle(w, inp, max_steps=1024)
    # the batch holds instances that ran long (>= 256 attempts), so the split is exercised
    assert int(((o["n_acc"] + o["n_rej"]) >= 256).sum()) > 0
    y0, p = _dev(inp["y0"], torch), _dev(inp["p"], torch)
    N, P = y0.shape[0], p.shape[0]
    res = {}
    for flags in (0, va.FLAG_COLD_ORDER):
        pb = va.Problem.from_workload(w, batch=B, max_steps=1024, lanes=lanes, flags=flags)
        print(f"[gpu] separate calls glv10 B={B} lanes={lanes} flags={flags}", flush=True)
        outs = []
        with va.Solver(pb) as s:
            ref = {k: v.cpu().numpy() for k, v in s.solve(y0, p, w.cost).items()}
            for _ in range(3):
                y = y0.clone()
                n_acc = torch.empty(B, dtype=torch.int32, device="cuda")
                n_rej = torch.empty_like(n_acc)
                st = torch.empty_like(n_acc)
                s.forward(y, p, n_accepted=n_acc, n_rejected=n_rej, lane_status=st)
                lam = torch.ones((1, N, B), dtype=torch.float64, device="cuda")
                mu = torch.zeros((1, P, B), dtype=torch.float64, device="cuda")
                s.adjoint(p, lam, mu)
                S = s.sensitivity_matrix(p)
                torch.cuda.synchronize()
                outs.append({"yT": y.cpu().numpy(), "dJdy0": lam[0].cpu().numpy(),
                             "dJdp": mu[0].cpu().numpy(), "n_acc": n_acc.cpu().numpy(),
                             "n_rej": n_rej.cpu().numpy(), "status": st.cpu().numpy(),
                             "S": S.cpu().numpy()})
        _same(outs)
        for k in ("yT", "dJdy0", "dJdp", "n_acc", "n_rej", "status"):
            assert np.array_equal(outs[0][k], ref[k], equal_nan=True), k
        res[flags] = outs[0]
    _same([res[0], res[va.FLAG_COLD_ORDER]])
    ok = o["status"] == 0
    assert np.array_equal(res[0]["status"], o["status"])
    assert np.array_equal(res[0]["n_acc"][ok], o["n_acc"][ok])
    assert np.array_equal(res[0]["yT"][:, ok], o["yT"][:, ok])


def test_interleaved_solves_and_separate_calls():
    """A split solve builds the next solve's order beside its own reverse into a
    second buffer and swaps the two when it returns, and copies the counts to the
    caller there; separate va_forward / va_adjoint calls build their order after
    the forward into the current buffer.  Any interleaving of the two on one handle
    gives the outputs of the handle's first solve, bit for bit, and the oracle's."""
    torch = _torch()
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    w = I.WORKLOADS["glv10"]
    B, lanes = 3000, 120
    inp = I.make_inputs(w, B)
    o = run_oracle(w, inp, max_steps=1024)
    assert int(((o["n_acc"] + o["n_rej"]) >= 256).sum()) > 0
    y0, p = _dev(inp["y0"], torch), _dev(inp["p"], torch)
    N, P = y0.shape[0], p.shape[0]
    pb = va.Problem.from_workload(w, batch=B, max_steps=1024, lanes=lanes)
    keys = ("yT", "dJdy0", "dJdp", "n_acc", "n_rej", "status")
    print(f"[gpu] interleaved solves and calls glv10 B={B} lanes={lanes}", flush=True)
    with va.Solver(pb) as s:
        def solve():
            out = s.solve(y0, p, w.cost)
            torch.cuda.synchronize()
            return {k: v.cpu().numpy() for k, v in out.items()}

        def calls():
            y = y0.clone()
            n_acc = torch.empty(B, dtype=torch.int32, device="cuda")
            n_rej = torch.empty_like(n_acc)
            st = torch.empty_like(n_acc)
            s.forward(y, p, n_accepted=n_acc, n_rejected=n_rej, lane_status=st)
            lam = torch.ones((1, N, B), dtype=torch.float64, device="cuda")
            mu = torch.zeros((1, P, B), dtype=torch.float64, device="cuda")
            s.adjoint(p, lam, mu)
            torch.cuda.synchronize()
            return {"yT": y.cpu().numpy(), "dJdy0": lam[0].cpu().numpy(),
                    "dJdp": mu[0].cpu().numpy(), "n_acc": n_acc.cpu().numpy(),
                    "n_rej": n_rej.cpu().numpy(), "status": st.cpu().numpy()}

        ref = solve()
        for step in (solve, solve, calls, solve, calls, calls, solve, solve):
            got = step()
            for k in (keys + ("J",) if step is solve else keys):
                assert np.array_equal(got[k], ref[k], equal_nan=True), (step.__name__, k)
    _compare(ref, o)
