"""Schedule reuse.  A persistent forward whose batch exceeds the instances in flight
takes its instances longest-first by the previous forward's attempt counts (index
order the first time).  Repeated solves on one handle -- the first in index order,
the later ones in the reused order -- must give bit-identical outputs, equal to
those of a handle that keeps index order (VA_FLAG_COLD_ORDER), and the oracle's.
"""
import numpy as np
import pytest

from helpers import run_oracle
from test_gpu_parity import GRAD_TIGHT, _compare, _dev, _torch

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
    _compare(outs[0], o, grad_tol=GRAD_TIGHT)
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
