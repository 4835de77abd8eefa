"""HIP-graph replay of va_solve (VA_FLAG_GRAPH, and capture by torch.cuda.graph) and
the launch-shape autotuner: every replay and every tuned configuration must give
the same results as the plain eager solve, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(name, B, seed_shift=0):
    from vectorizedadjoint_amd import inputs as I
    w = I.WORKLOADS[name]
    inp = I.make_inputs(w, B + seed_shift)
    return w, inp["y0"][:, seed_shift:].copy(), inp["p"][:, seed_shift:].copy()


def _np(out):
    return {k: v.cpu().numpy() for k, v in out.items()}


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), k


# lanes=120: fewer instances in flight than the batch, so repeated solves reuse the
# order and run the long instances on the handle's second stream inside the graph
@pytest.mark.parametrize("name,lanes", [("glv10", 0), ("vdp", 0), ("glv10", 120)])
def test_graph_flag_replays_match_eager(name, lanes):
    import torch
    import vectorizedadjoint_amd as va
    B = 2048
    w, y0h, ph = _inputs(name, B)
    y0 = torch.from_numpy(y0h).cuda()
    p = torch.from_numpy(ph).cuda()
    eager = va.Solver(va.Problem.from_workload(w, batch=B, max_steps=1024, lanes=lanes))
    graph = va.Solver(va.Problem.from_workload(w, batch=B, max_steps=1024, lanes=lanes,
                                               flags=va.FLAG_GRAPH))
    out_g = None
    print(f"[gpu] graph replay {name} B={B} lanes={lanes}", flush=True)
    for it in range(4):
        # new data in the same buffers between replays, with eager work in between
        _, y2, p2 = _inputs(name, B, seed_shift=it * 7)
        y0.copy_(torch.from_numpy(y2))
        p.copy_(torch.from_numpy(p2))
        junk = (y0 * 2.0).sum()
        torch.cuda.synchronize()
        out_g = graph.solve(y0, p, w.cost, out=out_g)
        ref = eager.solve(y0, p, w.cost)
        torch.cuda.synchronize()
        _same(_np(out_g), _np(ref))
        assert torch.isfinite(junk)
    # other output buffers: the graph is recorded again
    out2 = graph.solve(y0, p, w.cost)
    torch.cuda.synchronize()
    _same(_np(out2), _np(eager.solve(y0, p, w.cost)))
    eager.close()
    graph.close()


@pytest.mark.parametrize("lanes", [0, 120])
def test_torch_graph_capture_of_solve(lanes):
    import torch
    import vectorizedadjoint_amd as va
    B = 1024
    w, y0h, ph = _inputs("glv10", B)
    y0 = torch.from_numpy(y0h).cuda()
    p = torch.from_numpy(ph).cuda()
    print(f"[gpu] torch graph capture glv10 B={B} lanes={lanes}", flush=True)
    s = va.Solver(va.Problem.from_workload(w, batch=B, max_steps=1024, lanes=lanes))
    ref = s.solve(y0, p, w.cost)
    torch.cuda.synchronize()
    ref = _np(ref)
    out = s.solve(y0, p, w.cost)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        s.solve(y0, p, w.cost, out=out)
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        s.solve(y0, p, w.cost, out=out)
    for _ in range(3):
        for v in out.values():
            v.zero_()
        g.replay()
        torch.cuda.synchronize()
        _same(_np(out), ref)
    s.close()


def test_autotune_picks_a_configuration_with_identical_results():
    import torch
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd.autotune import autotune, candidates
    B = 4096
    w, y0h, ph = _inputs("glv10", B)
    y0 = torch.from_numpy(y0h).cuda()
    p = torch.from_numpy(ph).cuda()
    pb = va.Problem.from_workload(w, batch=B, max_steps=1024)
    tuned = autotune(pb, y0, p, w.cost, sample=B, reps=2, use_cache=False, verbose=True)
    assert any(all(getattr(tuned, k) == v for k, v in c.items()) for c in candidates(pb))
    with va.Solver(pb) as a, va.Solver(tuned) as b:
        ra = _np(a.solve(y0, p, w.cost))
        rb = _np(b.solve(y0, p, w.cost))
    # the forward is bit-identical across launch shapes; the reverse of the one-lane
    # and row layouts sums the adjoint terms in the same order too
    for k in ("yT", "J", "n_acc", "n_rej", "status"):
        assert np.array_equal(ra[k], rb[k], equal_nan=True), k
    ok = ra["status"] == 0
    for k in ("dJdy0", "dJdp"):
        d = np.max(np.abs(ra[k][:, ok] - rb[k][:, ok])) / np.max(np.abs(ra[k][:, ok]))
        assert d <= 1e-13, (k, d)
