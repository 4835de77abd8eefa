"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the same
seeded inputs.

The oracle follows the reference's own operation sequence (Boost.odeint's
unfused stage sums and error norm, the C library's pow in the step adjuster,
the example functors' loop orders, the reverse sweep of
detail/backpropagation.hpp with h = t[n+1] - t[n]) and the HIP kernels follow it
operation for operation.  Bar, per instance: identical status, identical
accepted step sequence (t_n, h_n) -- compared through va_step_digest against
orc_step_digest -- on >= 99.9% of the instances, and on those y(T), J, dJ/dy0
and dJ/dp equal to the oracle's bit for bit (exact=True; every case here) or
within a norm-wise relative error of 1e-10.
"""
import numpy as np
import pytest

from helpers import rel_err, run_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _torch():
    import torch
    assert torch.cuda.is_available(), "no HIP device"
    return torch


def _dev(a, torch):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _solve_gpu(w, inputs, mode, **kw):
    torch = _torch()
    import vectorizedadjoint_amd as va
    B = inputs["y0"].shape[1]
    pb = va.Problem.from_workload(w, batch=B, mode=mode, **kw)
    print(f"[gpu] solve {w.name} B={B} mode={mode} {kw}", flush=True)
    with va.Solver(pb) as s:
        y0 = _dev(inputs["y0"], torch)
        p = _dev(inputs["p"], torch)
        out = s.solve(y0, p, w.cost)
        out["digest"] = s.step_digest()
        torch.cuda.synchronize()
        return {k: v.cpu().numpy() for k, v in out.items()}


def _compare(g, o, exact=True, grad_tol=TOL):
    """Per-instance parity; prints the fractions and largest errors it finds."""
    ok = o["status"] == 0
    assert np.array_equal(g["status"], o["status"])
    same_counts = (g["n_acc"] == o["n_acc"]) & (g["n_rej"] == o["n_rej"])
    same = same_counts
    if "digest" in g and "digest" in o:
        same = same & (g["digest"].view(np.uint64) == o["digest"].view(np.uint64))
    frac = same[ok].mean()
    m = ok & same
    e_y = rel_err(g["yT"][:, m], o["yT"][:, m]).max(initial=0.0)
    e_J = rel_err(g["J"][m], o["J"][m]).max(initial=0.0)
    e_y0 = rel_err(g["dJdy0"][:, m], o["dJdy0"][:, m]).max(initial=0.0)
    e_p = rel_err(g["dJdp"][:, m], o["dJdp"][:, m]).max(initial=0.0)
    print(f"    same step sequence {frac:.6f} (counts {same_counts[ok].mean():.6f})  max rel err:"
          f" yT {e_y:.2e} J {e_J:.2e} dJdy0 {e_y0:.2e} dJdp {e_p:.2e}", flush=True)
    assert frac >= 0.999, f"step sequences equal on only {frac:.5f} of instances"
    if exact:
        # bit patterns, so that a zero of the other sign counts as a difference
        for k in ("yT", "J", "dJdy0", "dJdp"):
            gk = np.ascontiguousarray(g[k][..., m], dtype=np.float64).view(np.uint64)
            ok_ = np.ascontiguousarray(o[k][..., m], dtype=np.float64).view(np.uint64)
            assert np.array_equal(gk, ok_), f"{k} differs from the oracle"
    assert e_y <= TOL and e_J <= TOL
    assert e_y0 <= grad_tol and e_p <= grad_tol
    assert np.all(np.isnan(g["dJdp"][:, ~ok]))


CASES = [("vdp", 1024), ("oscillator", 4096), ("glv10", 2048)]


@pytest.mark.parametrize("name,B", CASES)
def test_persistent_matches_oracle(name, B):
    from vectorizedadjoint_amd import inputs as I, MODE_PERSISTENT
    w = I.WORKLOADS[name]
    inp = I.make_inputs(w, B)
    o = run_oracle(w, inp, max_steps=4096)
    g = _solve_gpu(w, inp, MODE_PERSISTENT, max_steps=4096)
    _compare(g, o)


@pytest.mark.parametrize("name,B", CASES)
def test_plain_matches_oracle(name, B):
    from vectorizedadjoint_amd import inputs as I, MODE_PLAIN
    w = I.WORKLOADS[name]
    inp = I.make_inputs(w, B)
    o = run_oracle(w, inp, max_steps=2048)
    g = _solve_gpu(w, inp, MODE_PLAIN, max_steps=2048)
    _compare(g, o)


@pytest.mark.parametrize("ipw", [64, 6])
def test_refill_with_few_lanes_matches_oracle(ipw):
    """Fewer instance slots than instances: every slot solves many instances in turn."""
    from vectorizedadjoint_amd import inputs as I, MODE_PERSISTENT
    w = I.WORKLOADS["glv10"]
    inp = I.make_inputs(w, 1500)
    o = run_oracle(w, inp, max_steps=1024)
    g = _solve_gpu(w, inp, MODE_PERSISTENT, max_steps=1024, lanes=120, block_size=64,
                   instances_per_wave=ipw)
    _compare(g, o)


@pytest.mark.parametrize("ipw", [64, 6])
@pytest.mark.parametrize("tableau", ["dopri5", "fehlberg78"])
def test_glv_layouts_and_tableaus_match_oracle(ipw, tableau):
    """Both Lotka-Volterra layouts (one lane per instance, one 10-lane group per
    instance) with the other adaptive tableaus, bit-identical forward."""
    import dataclasses
    from vectorizedadjoint_amd import inputs as I, MODE_PERSISTENT, TAB_DOPRI5, TAB_FEHLBERG78
    tid = {"dopri5": TAB_DOPRI5, "fehlberg78": TAB_FEHLBERG78}[tableau]
    w = dataclasses.replace(I.WORKLOADS["glv10"], tableau=tid, tol=1e-7)
    inp = I.make_inputs(w, 600)
    o = run_oracle(w, inp, max_steps=1024)
    g = _solve_gpu(w, inp, MODE_PERSISTENT, max_steps=1024, instances_per_wave=ipw)
    _compare(g, o)


def test_glv_fixed_step_rk4_rows_matches_oracle():
    import dataclasses
    from vectorizedadjoint_amd import inputs as I, MODE_PERSISTENT, TAB_RK4
    w = dataclasses.replace(I.WORKLOADS["glv10"], tableau=TAB_RK4, controlled=False, dt0=0.05,
                            t1=10.0)
    inp = I.make_inputs(w, 300)
    o = run_oracle(w, inp, max_steps=1024)
    g = _solve_gpu(w, inp, MODE_PERSISTENT, max_steps=1024, instances_per_wave=6)
    _compare(g, o)


def test_headline_batch_matches_oracle():
    """The headline workload at its full size (GLV N=10, B=65536, Cash-Karp 5(4),
    tol 1e-8), every instance against the oracle: the handle's first solve (index
    order, no split) and its second (longest-first order; the instances that ran
    long run one per wave beside the persistent kernels) give the same bits."""
    torch = _torch()
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    w = I.WORKLOADS["glv10"]
    inp = I.make_inputs(w, w.batch)
    o = run_oracle(w, inp, max_steps=1024)
    pb = va.Problem.from_workload(w, max_steps=1024)
    print(f"[gpu] headline glv10 B={w.batch}, two solves", flush=True)
    outs = []
    with va.Solver(pb) as s:
        y0, p = _dev(inp["y0"], torch), _dev(inp["p"], torch)
        for _ in range(2):
            out = s.solve(y0, p, w.cost)
            out["digest"] = s.step_digest()
            torch.cuda.synchronize()
            outs.append({k: v.cpu().numpy() for k, v in out.items()})
    # instances that needed 256 or more attempts exist, so the second solve splits
    assert int(((o["n_acc"] + o["n_rej"]) >= 256).sum()) > 0
    _compare(outs[1], o)
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k], equal_nan=True), k


def test_oscillator_full_batch():
    """The damped oscillator workload at its full size (B = 2^20, DOPRI5 tol 1e-10,
    J = |y(T)|^2; about 740 accepted steps per instance, a 69 GB trajectory store).
    The system is linear and homogeneous, so the frozen-step map is y(T) = Phi y0
    and J = |Phi y0|^2 is homogeneous of degree 2 in y0: <dJ/dy0, y0> = 2 J holds
    for every instance, checked on all 2^20.  A slice of 16384 instances is
    compared with the oracle; the second solve (longest-first order) is
    bit-identical to the first (index order)."""
    import os
    torch = _torch()
    import vectorizedadjoint_amd as va
    from vectorizedadjoint_amd import inputs as I
    w = I.WORKLOADS["oscillator"]
    B = w.batch
    assert B == 1 << 20
    inp = I.make_inputs(w, B)
    pb = va.Problem.from_workload(w, max_steps=2048)
    print(f"[gpu] oscillator full batch B={B}, two solves", flush=True)
    outs = []
    with va.Solver(pb) as s:
        y0, p = _dev(inp["y0"], torch), _dev(inp["p"], torch)
        for _ in range(2):
            out = s.solve(y0, p, w.cost)
            torch.cuda.synchronize()
            outs.append({k: v.cpu().numpy() for k, v in out.items()})
    g = outs[1]
    for k in outs[0]:
        assert np.array_equal(outs[0][k], g[k], equal_nan=True), k
    assert (g["status"] == 0).all()
    euler = np.sum(g["dJdy0"] * inp["y0"], axis=0)
    dev = np.abs(euler - 2.0 * g["J"]) / (2.0 * np.abs(g["J"]))
    print(f"    <dJ/dy0, y0> vs 2J: max rel dev {dev.max():.2e} over {B} instances", flush=True)
    assert dev.max() <= 1e-10
    b0, cnt = 500000, 16384
    import oracle as O
    opb = O.problem(w.system, w.n, w.np, w.tableau, t0=w.t0, t1=w.t1, dt0=w.dt0,
                    abs_tol=w.tol, rel_tol=w.tol, controlled=w.controlled,
                    sys_const=w.sys_const, max_steps=2048)
    o = O.solve_batch(opb, inp["y0"], inp["p"], w.cost, b_begin=b0, b_count=cnt,
                      nthreads=min(16, os.cpu_count() or 1))
    sl = slice(b0, b0 + cnt)
    _compare({k: v[..., sl] for k, v in g.items()}, {k: v[..., sl] for k, v in o.items()},
             )


@pytest.mark.parametrize("name,ipw", [("vdp", 0), ("glv10", 6)])
def test_fixed_step_euler_matches_oracle(name, ipw):
    """The fixed-step loop (detail/runge_kutta.hpp:36-72) with Euler (ButcherTable.hpp:
    50-65; Boost.odeint euler: y + h f(y)) on the one-lane and the 10-lane-group layouts."""
    import dataclasses
    from vectorizedadjoint_amd import inputs as I, MODE_PERSISTENT, TAB_EULER
    t1 = {"vdp": 2.0, "glv10": 5.0}[name]
    dt = {"vdp": 1e-3, "glv10": 1e-2}[name]
    w = dataclasses.replace(I.WORKLOADS[name], tableau=TAB_EULER, controlled=False, dt0=dt, t1=t1)
    inp = I.make_inputs(w, 700)
    o = run_oracle(w, inp, max_steps=4096)
    assert (o["status"] == 0).all() and (o["n_acc"] > 400).all()
    g = _solve_gpu(w, inp, MODE_PERSISTENT, max_steps=4096, instances_per_wave=ipw)
    _compare(g, o)


def test_glv_signed_zeros_match_oracle():
    """Lotka-Volterra with exact zeros: growth rates r = -0.0, species at 0.0 (whose
    products A_ij x_j are zeros of either sign), a row of A of -0.0, and an
    instance that is zero throughout.  The row kernels start each growth sum at
    the first product and take r as r + 0.0; the reference sums from 0.0 and adds
    r.  Outputs are compared bit for bit, signs of zeros included."""
    import dataclasses
    from vectorizedadjoint_amd import inputs as I, MODE_PERSISTENT, MODE_PLAIN
    w = dataclasses.replace(I.WORKLOADS["glv10"], t1=10.0)
    B, N = 256, 10
    inp = I.make_inputs(w, B)
    y0, p = inp["y0"].copy(), inp["p"].copy()
    p[:N, 0:64] = -0.0                          # r = -0
    y0[[0, 3, 7], 32:128] = 0.0                 # extinct species
    y0[:, 160:192] = 0.0                        # everything extinct ...
    p[:N, 160:176] = -0.0                       # ... with r = -0 on half of them
    for k in range(192, 224):                   # a row of A of -0.0
        i = k % N
        p[N * (i + 1):N * (i + 2), k] = -0.0
    p[N:, 224:240] = -np.abs(p[N:, 224:240])    # all interactions negative ...
    y0[[1, 2], 224:240] = 0.0                   # ... with extinct species
    inp = dict(inp, y0=y0, p=p)
    o = run_oracle(w, inp, max_steps=1024)
    assert (o["status"] == 0).sum() > 200
    for mode in (MODE_PERSISTENT, MODE_PLAIN):
        g = _solve_gpu(w, inp, mode, max_steps=1024)
        _compare(g, o)


@pytest.mark.parametrize("n,tableau", [(7, "cash_karp54"), (16, "cash_karp54"),
                                       (32, "cash_karp54"), (5, "dopri5"), (13, "fehlberg78")])
def test_glv_any_n_rows_match_oracle(n, tableau):
    """Generalized Lotka-Volterra at other N (the reference's driver reads N from the
    command line, examples/GeneralizedLotkaVolterra/main.cpp:125-143): one group of
    N lanes per instance, same distributions as the N=10 workload over t in [0, 10]."""
    import dataclasses
    from vectorizedadjoint_amd import (inputs as I, MODE_PERSISTENT, TAB_CASH_KARP54, TAB_DOPRI5,
                                       TAB_FEHLBERG78)
    tid = {"cash_karp54": TAB_CASH_KARP54, "dopri5": TAB_DOPRI5,
           "fehlberg78": TAB_FEHLBERG78}[tableau]
    w = dataclasses.replace(I.WORKLOADS["glv10"], name=f"glv{n}", n=n, np=n + n * n, tableau=tid,
                            t1=10.0)
    inp = I.make_inputs(w, 900)
    o = run_oracle(w, inp, max_steps=1024)
    assert (o["status"] == 0).sum() > 450
    g = _solve_gpu(w, inp, MODE_PERSISTENT, max_steps=1024)
    _compare(g, o)
