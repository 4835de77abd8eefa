"""va_solve_noalloc (the entry point that allocates nothing) against va_solve,
through the C ABI: the same kernels on the same inputs, so every output must be
equal bit for bit -- on two calls in a row, the second with a smaller batch (the
scratch keeps the first call's contents), and for both instance layouts."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(va, torch, problem, y0, p, cost):
    N, P, B = problem.n_state, problem.n_param, problem.batch
    dev = y0.device
    out = {"J": torch.empty(B, dtype=torch.float64, device=dev),
           "yT": torch.empty((N, B), dtype=torch.float64, device=dev),
           "dJdy0": torch.empty((N, B), dtype=torch.float64, device=dev),
           "dJdp": torch.empty((P, B), dtype=torch.float64, device=dev),
           "n_acc": torch.empty(B, dtype=torch.int32, device=dev),
           "n_rej": torch.empty(B, dtype=torch.int32, device=dev),
           "status": torch.empty(B, dtype=torch.int32, device=dev)}
    pb = problem.to_c()
    st = va.lib().va_solve_noalloc(
        C.byref(pb), y0.data_ptr(), p.data_ptr(), int(cost), None, out["J"].data_ptr(),
        out["yT"].data_ptr(), out["dJdy0"].data_ptr(), out["dJdp"].data_ptr(),
        out["n_acc"].data_ptr(), out["n_rej"].data_ptr(), out["status"].data_ptr(), None)
    assert st == 0, va.lib().va_last_error().decode()
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("name,ipw,max_steps,batches",
                         [("glv10", 6, 1024, (300, 77)), ("vdp", 0, 4096, (200, 51))])
def test_noalloc_matches_va_solve(name, ipw, max_steps, batches):
    import torch
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    assert torch.cuda.is_available()
    w = I.WORKLOADS[name]
    for B in batches:
        inp = I.make_inputs(w, B)
        print(f"[gpu] noalloc {name} B={B}", flush=True)
        problem = va.Problem.from_workload(w, batch=B, max_steps=max_steps,
                                           instances_per_wave=ipw)
        y0 = torch.from_numpy(inp["y0"]).cuda()
        p = torch.from_numpy(inp["p"]).cuda()
        got = _noalloc(va, torch, problem, y0, p, w.cost)
        with va.Solver(problem) as s:
            ref = s.solve(y0, p, w.cost)
            torch.cuda.synchronize()
            ref = {k: v.cpu().numpy() for k, v in ref.items()}
        assert (ref["status"] == 0).sum() > 0.9 * B
        for k in ref:
            assert np.array_equal(got[k], ref[k], equal_nan=True), (name, B, k)


def test_noalloc_rejects_oversized_problem():
    import torch
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    w = I.WORKLOADS["glv10"]
    B = 400  # 400 * 1025 * 12 * 8 bytes > 32 MiB
    problem = va.Problem.from_workload(w, batch=B, max_steps=1024)
    y0 = torch.zeros((w.n, B), dtype=torch.float64, device="cuda")
    p = torch.zeros((w.np, B), dtype=torch.float64, device="cuda")
    out = [torch.empty(B * 120, dtype=torch.float64, device="cuda") for _ in range(4)]
    pb = problem.to_c()
    st = va.lib().va_solve_noalloc(C.byref(pb), y0.data_ptr(), p.data_ptr(), int(w.cost), None,
                                   *[o.data_ptr() for o in out], None, None, None, None)
    assert st == 4  # VA_ERR_OUT_OF_MEMORY
    assert b"scratch" in va.lib().va_last_error()
