"""GPU parity of the observation-cost solve (va_solve_observed) against the CPU
oracle, with the trajectory stored in full and kept as checkpoints.

Bar: identical accepted/rejected counts and status, y(T) and J bit-identical to
the oracle, dJ/dy0 and dJ/dp within 1e-10 relative; the checkpointed solve
(which recomputes every segment in the reverse sweep) bit-identical to the
full-store solve.
"""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _setup(B, t1=20.0, K=50):
    import torch
    import oracle as O
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    assert torch.cuda.is_available(), "no HIP device"
    w = I.WORKLOADS["vdp_obs"]
    inp = I.make_inputs(w, B)
    t_obs = w.t0 + (t1 - w.t0) * np.arange(1, K + 1) / K
    _, mu_true, y0_true = I.observation_schedule(w)
    pb = O.problem(w.system, w.n, w.np, w.tableau, t0=w.t0, t1=t1, dt0=w.dt0, abs_tol=w.tol,
                   rel_tol=w.tol, max_steps=16384)
    g = O.Solve(pb, y0_true, [mu_true], obs=(t_obs, np.zeros((K, 2))))
    y_obs = g.states[g.obs_row]
    return torch, O, va, w, inp, t_obs, y_obs, pb


def _gpu(torch, va, w, inp, t_obs, y_obs, t1, **kw):
    B = inp["y0"].shape[1]
    K = t_obs.size
    prob = va.Problem.from_workload(w, batch=B, t1=t1, n_obs=K, max_steps=16384, **kw)
    print(f"[gpu] observed solve B={B} K={K} {kw}", flush=True)
    with va.Solver(prob) as s:
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        out = s.solve_observed(dev(inp["y0"]), dev(inp["p"]), dev(t_obs), dev(y_obs))
        torch.cuda.synchronize()
        return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("store", ["full", "checkpoints"])
def test_observed_solve_matches_oracle(store):
    B = 1024
    torch, O, va, w, inp, t_obs, y_obs, pb = _setup(B)
    kw = {} if store == "full" else {"checkpoint_slots": 16}
    g = _gpu(torch, va, w, inp, t_obs, y_obs, 20.0, **kw)
    o = O.solve_batch(pb, inp["y0"], inp["p"], 0, obs=(t_obs, y_obs), nthreads=8)
    assert np.array_equal(g["status"], o["status"])
    ok = o["status"] == 0
    assert ok.all()
    assert np.array_equal(g["n_acc"], o["n_acc"]) and np.array_equal(g["n_rej"], o["n_rej"])
    # step counts span about two orders of magnitude over mu in [0.1, 100]
    assert o["n_acc"].max() > 20 * o["n_acc"].min()
    assert np.array_equal(g["yT"], o["yT"])
    assert np.array_equal(g["J"], o["J"])
    assert rel_err(g["dJdy0"], o["dJdy0"]).max() <= TOL
    assert rel_err(g["dJdp"], o["dJdp"]).max() <= TOL


def test_checkpointed_reverse_is_bit_identical_to_full_store():
    """Few slots force the spacing to double many times; the recomputed segments
    must reproduce the stored rows exactly."""
    B = 512
    torch, O, va, w, inp, t_obs, y_obs, pb = _setup(B, t1=12.0, K=30)
    full = _gpu(torch, va, w, inp, t_obs, y_obs, 12.0)
    for slots in (2, 6, 64):
        ck = _gpu(torch, va, w, inp, t_obs, y_obs, 12.0, checkpoint_slots=slots)
        for k in ("J", "yT", "dJdy0", "dJdp", "n_acc", "n_rej", "status"):
            assert np.array_equal(full[k], ck[k]), (slots, k)


def test_reverse_in_wave_pairs_with_failed_instances_and_few_lanes():
    """The checkpointed reverse in pairs of waves (one replaying segments while
    the other reverses the previous one; blocks of 256 threads) against the
    one-wave reverse (blocks of 64 threads) and the full store: 128 lanes for
    512 instances, so every lane pair takes several instances one after the
    other, and a step cap that ends the longest instances with MAX_STEPS
    (NaN gradients, no segments replayed for them).  Every output identical."""
    B = 512
    torch, O, va, w, inp, t_obs, y_obs, pb = _setup(B, t1=12.0, K=30)

    def run(**kw):
        prob = va.Problem.from_workload(w, batch=B, t1=12.0, n_obs=t_obs.size, max_steps=700,
                                        lanes=128, **kw)
        print(f"[gpu] observed wave pairs B={B} {kw}", flush=True)
        with va.Solver(prob) as s:
            dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
            out = s.solve_observed(dev(inp["y0"]), dev(inp["p"]), dev(t_obs), dev(y_obs))
            torch.cuda.synchronize()
            return {k: v.cpu().numpy() for k, v in out.items()}

    full = run(block_size=256)
    pairs = run(block_size=256, checkpoint_slots=8)
    single = run(block_size=64, checkpoint_slots=8)
    failed = full["status"] != 0
    assert failed.any() and (~failed).any()
    assert np.isnan(full["dJdp"][:, failed]).all()
    for k in ("J", "yT", "dJdy0", "dJdp", "n_acc", "n_rej", "status"):
        assert np.array_equal(full[k], pairs[k], equal_nan=True), k
        assert np.array_equal(full[k], single[k], equal_nan=True), k


@pytest.mark.parametrize("tab", ["cash_karp54", "fehlberg78"])
def test_checkpointed_reverse_other_tableaus(tab):
    """The checkpointed reverse (in wave pairs) with the other embedded tableaus:
    Cash-Karp 5(4) (no FSAL: the replay evaluates f(y) at each step start) and
    Fehlberg 7(8) (13 stages), against the full store, every output identical."""
    from vectorizedadjoint_amd import constants as K
    B = 256
    torch, O, va, w, inp, t_obs, y_obs, pb = _setup(B, t1=8.0, K=20)
    tid = {"cash_karp54": K.TAB_CASH_KARP54, "fehlberg78": K.TAB_FEHLBERG78}[tab]
    full = _gpu(torch, va, w, inp, t_obs, y_obs, 8.0, tableau=tid)
    ck = _gpu(torch, va, w, inp, t_obs, y_obs, 8.0, tableau=tid, checkpoint_slots=6)
    assert (full["status"] == 0).all()
    for k in ("J", "yT", "dJdy0", "dJdp", "n_acc", "n_rej", "status"):
        assert np.array_equal(full[k], ck[k]), k


def test_two_level_checkpoints_are_bit_identical_to_full_store():
    """Two levels of checkpoints (VA_FLAG_CHECKPOINT_2LEVEL): the reverse replays
    each long segment once to keep level-2 checkpoints, then recomputes its rows
    sub-segment by sub-segment; every output equals the full store's."""
    B = 512
    torch, O, va, w, inp, t_obs, y_obs, pb = _setup(B, t1=12.0, K=30)
    full = _gpu(torch, va, w, inp, t_obs, y_obs, 12.0)
    assert full["n_acc"].max() > 1000
    for slots in (2, 6, 0):
        kw = {"flags": va.FLAG_CHECKPOINT_2LEVEL, "checkpoint_slots": slots}
        if slots == 0:
            kw["trajectory_budget_bytes"] = 64 << 20
        ck = _gpu(torch, va, w, inp, t_obs, y_obs, 12.0, **kw)
        for k in ("J", "yT", "dJdy0", "dJdp", "n_acc", "n_rej", "status"):
            assert np.array_equal(full[k], ck[k]), (slots, k)


def test_budget_below_one_level_uses_two_levels():
    """A budget under the smallest one-level checkpoint store (about 4.2 MB here:
    slots x 40 B per instance plus stride_max rows per lane) but above the
    two-level one (about 1.4 MB) is accepted and gives the full store's bits."""
    B = 512
    torch, O, va, w, inp, t_obs, y_obs, pb = _setup(B, t1=12.0, K=30)
    full = _gpu(torch, va, w, inp, t_obs, y_obs, 12.0)
    ck = _gpu(torch, va, w, inp, t_obs, y_obs, 12.0, trajectory_budget_bytes=2 << 20)
    for k in ("J", "yT", "dJdy0", "dJdp", "n_acc", "n_rej", "status"):
        assert np.array_equal(full[k], ck[k]), k


def test_budget_selects_checkpoints_and_reports_state_errors():
    torch, O, va, w, inp, t_obs, y_obs, pb = _setup(256, t1=8.0, K=10)
    B = 256
    # full store would need 256 x 16385 x 32 B = 134 MB; give it 8 MB
    prob = va.Problem.from_workload(w, batch=B, t1=8.0, n_obs=10, max_steps=16384,
                                    trajectory_budget_bytes=8 << 20)
    with va.Solver(prob) as s:
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        out = s.solve_observed(dev(inp["y0"]), dev(inp["p"]), dev(t_obs), dev(y_obs))
        torch.cuda.synchronize()
        o = O.solve_batch(O.problem(w.system, 2, 1, w.tableau, t0=0.0, t1=8.0, dt0=w.dt0,
                                    abs_tol=w.tol, rel_tol=w.tol, max_steps=16384),
                          inp["y0"], inp["p"], 0, obs=(t_obs, y_obs), nthreads=8)
        assert np.array_equal(out["J"].cpu().numpy(), o["J"])
        assert rel_err(out["dJdp"].cpu().numpy(), o["dJdp"]).max() <= TOL
        # a checkpointing handle has no per-step store
        with pytest.raises(va.VaError, match="checkpoints"):
            s.forward(dev(inp["y0"]), dev(inp["p"]))
    # a budget too small even for checkpoints is refused at creation
    with pytest.raises(va.VaError, match="cannot hold checkpoints"):
        va.Solver(va.Problem.from_workload(w, batch=B, n_obs=10, max_steps=16384,
                                           trajectory_budget_bytes=4096))


def test_gradient_against_central_differences():
    """Central differences (h = 1e-6) of the device solve itself.  The adaptive step
    sequence moves with mu, so they agree with the discrete adjoint to the
    tolerance level only: median 5e-7, 90th percentile 2e-5 relative on this
    sample (the stiff end, mu near 100, is the worst).  The adjoint is the exact
    derivative of the frozen-step map, which the oracle tests pin by complex step."""
    B = 256
    torch, O, va, w, inp, t_obs, y_obs, pb = _setup(B)
    h = 1e-6
    base = _gpu(torch, va, w, inp, t_obs, y_obs, 20.0)
    plus = _gpu(torch, va, w, dict(inp, p=inp["p"] + h), t_obs, y_obs, 20.0)
    minus = _gpu(torch, va, w, dict(inp, p=inp["p"] - h), t_obs, y_obs, 20.0)
    fd = (plus["J"] - minus["J"]) / (2 * h)
    ad = base["dJdp"][0]
    rel = np.abs(fd - ad) / np.abs(ad)
    assert np.median(rel) <= 2e-6 and np.percentile(rel, 90) <= 1e-4, (
        np.median(rel), np.percentile(rel, 90))


@pytest.mark.parametrize("kw", [{"trajectory_budget_bytes": 8 << 20}, {"checkpoint_slots": 4}])
def test_terminal_cost_solve_over_budget_uses_checkpoints(kw):
    """va_solve whose per-step store exceeds the budget (or with forced checkpoints)
    runs on checkpoints with a recomputing reverse: bit-identical to the full store."""
    import torch
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    w = I.WORKLOADS["vdp"]
    B = 1024
    inp = I.make_inputs(w, B)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    y0, p = dev(inp["y0"]), dev(inp["p"])
    print(f"[gpu] terminal cost over budget {kw}", flush=True)
    with va.Solver(va.Problem.from_workload(w, batch=B, max_steps=4096)) as full, \
            va.Solver(va.Problem.from_workload(w, batch=B, max_steps=4096, **kw)) as ck:
        a = {k: v.cpu().numpy() for k, v in full.solve(y0, p, w.cost).items()}
        b = {k: v.cpu().numpy() for k, v in ck.solve(y0, p, w.cost).items()}
        with pytest.raises(va.VaError, match="checkpoints"):
            ck.forward(y0.clone(), p)
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def test_full_workload_under_budget_matches_oracle():
    """The observed workload at its full size: Van der Pol, B=262144, mu
    log-uniform over [0.1, 100], 50 observations over [0, 20], under the 4 GB
    trajectory budget (the per-step store would need 15 GB, so the forward keeps
    checkpoints and the reverse recomputes every segment).  Two solves on one
    handle -- the second takes the first's longest-first order and starts each
    instance's checkpoint spacing from its previous step count -- each against
    the oracle on every instance."""
    import os
    B = 262144
    torch, O, va, w, inp, t_obs, y_obs, pb = _setup(B)
    assert w.batch == B
    budget = int(w.extra["budget_bytes"])
    prob = va.Problem.from_workload(w, batch=B, t1=20.0, n_obs=t_obs.size, max_steps=16384,
                                    trajectory_budget_bytes=budget)
    print(f"[gpu] observed full workload B={B} budget={budget}", flush=True)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    outs = []
    with va.Solver(prob) as s:
        args = (dev(inp["y0"]), dev(inp["p"]), dev(t_obs), dev(y_obs))
        for _ in range(2):
            out = s.solve_observed(*args)
            torch.cuda.synchronize()
            outs.append({k: v.cpu().numpy() for k, v in out.items()})
        with pytest.raises(va.VaError, match="checkpoints"):
            s.forward(dev(inp["y0"]), dev(inp["p"]))
    o = O.solve_batch(pb, inp["y0"], inp["p"], 0, obs=(t_obs, y_obs),
                      nthreads=min(16, os.cpu_count() or 1))
    assert (o["status"] == 0).all()
    for g in outs:
        assert np.array_equal(g["status"], o["status"])
        assert np.array_equal(g["n_acc"], o["n_acc"]) and np.array_equal(g["n_rej"], o["n_rej"])
        assert np.array_equal(g["yT"], o["yT"])
        assert np.array_equal(g["J"], o["J"])
        e0, ep = rel_err(g["dJdy0"], o["dJdy0"]).max(), rel_err(g["dJdp"], o["dJdp"]).max()
        print(f"    max rel err dJdy0 {e0:.2e} dJdp {ep:.2e}", flush=True)
        assert e0 <= TOL and ep <= TOL
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k], equal_nan=True), k
