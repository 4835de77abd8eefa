"""Launch-shape autotuner for the persistent solves.

For a problem it times a few launch configurations -- threads per block,
instances per wave (one lane per instance or, for Lotka-Volterra, one group of
N lanes per instance), instances in flight, and for observed solves the number
of checkpoint slots -- on a prefix of the caller's own batch, and returns the
problem with the fastest one filled in.  Every configuration computes the same
results (the arithmetic per instance does not depend on the launch shape), so
tuning changes time only.  Results are cached per device and problem shape in
a JSON file (VA_AUTOTUNE_CACHE, default ~/.cache/vectorizedadjoint_amd/autotune.json).
"""
from __future__ import annotations

import dataclasses
import json
import os
import time

from . import Problem, Solver, VaError
from . import constants as K

CACHE = os.environ.get("VA_AUTOTUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "vectorizedadjoint_amd",
                                    "autotune.json"))


def candidates(pb: Problem):
    """Launch configurations worth timing for this problem."""
    out = []
    blocks = (64, 128, 256)
    if pb.system == K.SYS_LOTKA_VOLTERRA and pb.n_state == 10:
        for ipw in (6, 64):
            for blk in blocks:
                out.append({"instances_per_wave": ipw, "block_size": blk})
    elif pb.system == K.SYS_LOTKA_VOLTERRA_SHARED:
        out.append({})
    else:
        for blk in blocks:
            out.append({"block_size": blk})
    if pb.n_obs > 0:
        out = [dict(c, checkpoint_slots=s) for c in out for s in (0, 64, 256)]
    return out


def _key(pb: Problem, device: str) -> str:
    return (f"{device}|sys{pb.system}|n{pb.n_state}|tab{pb.tableau}|ad{int(pb.adaptive)}|"
            f"obs{pb.n_obs}|B{1 << max(0, int(pb.batch - 1).bit_length())}")


def _time(pb: Problem, y0, p, cost, reps, torch, obs=None):
    with Solver(pb) as s:
        run = (lambda: s.solve_observed(y0, p, *obs)) if obs else (lambda: s.solve(y0, p, cost))
        run()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            run()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps


def autotune(pb: Problem, y0, p, cost=K.COST_SUM, obs=None, sample: int = 8192, reps: int = 3,
             use_cache: bool = True, verbose: bool = False) -> Problem:
    """Return pb with the fastest launch configuration.  y0 [N][B], p [P][B] are the
    caller's device tensors; the first `sample` instances are timed.  obs = (t_obs, y_obs)
    device tensors for observed solves."""
    import torch
    device = torch.cuda.get_device_name()
    key = _key(pb, device)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        return dataclasses.replace(pb, **cache[key])
    B = min(pb.batch, sample)
    ys, ps = y0[:, :B].contiguous(), p[:, :B].contiguous()
    best, best_t = {}, float("inf")
    for c in candidates(pb):
        trial = dataclasses.replace(pb, batch=B, **c)
        try:
            dt = _time(trial, ys, ps, cost, reps, torch, obs)
        except VaError as e:  # e.g. a checkpoint budget that does not fit
            if verbose:
                print(f"[autotune] {c}: {e}")
            continue
        if verbose:
            print(f"[autotune] {c}: {dt * 1e3:.3f} ms")
        if dt < best_t:
            best, best_t = c, dt
    if use_cache:
        cache[key] = best
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1)
    return dataclasses.replace(pb, **best)
