"""Launch-shape autotuner for the persistent solves (Python side of va_autotune).

va_autotune (the C ABI) times the launch configurations of a problem -- threads
per block, instances per wave (one lane per instance or, for Lotka-Volterra,
one group of N lanes per instance) and, for observed solves, the number of
checkpoint slots -- on a prefix of the caller's own batch, returns the fastest,
and remembers it for the process so that va_create applies it to problems of
the same shape whose block_size / instances_per_wave are 0.  Every
configuration computes the same results (the arithmetic per instance does not
depend on the launch shape), so tuning changes time only.  This module adds a
JSON cache across processes (VA_AUTOTUNE_CACHE, default
~/.cache/vectorizedadjoint_amd/autotune.json).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import json
import os

from . import Problem, _check, _ptr, _stream, lib
from . import constants as K

CACHE = os.environ.get("VA_AUTOTUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "vectorizedadjoint_amd",
                                    "autotune.json"))


def candidates(pb: Problem):
    """The launch configurations va_autotune times for this problem."""
    ipws = [0]
    if pb.system == K.SYS_LOTKA_VOLTERRA and (pb.n_state <= 5 or pb.n_state == 10):
        ipws.append(64)
    slots = [0, 64, 256] if pb.n_obs > 0 else [pb.checkpoint_slots]
    return [{"instances_per_wave": i, "block_size": b, "checkpoint_slots": s}
            for i in ipws for b in (64, 128, 256) for s in slots]


def _key(pb: Problem, device: str) -> str:
    return (f"{device}|sys{pb.system}|n{pb.n_state}|tab{pb.tableau}|ad{int(pb.adaptive)}|"
            f"obs{pb.n_obs}|B{1 << max(0, int(pb.batch - 1).bit_length())}")


def autotune(pb: Problem, y0, p, cost=K.COST_SUM, obs=None, sample: int = 16384, reps: int = 3,
             use_cache: bool = True, verbose: bool = False, kernels=None,
             cost_weights=None) -> Problem:
    """Return pb with the fastest launch configuration (va_autotune).  y0 [N][B],
    p [P][B] are the caller's device tensors; the first `sample` instances are timed.
    obs = (t_obs, y_obs) device tensors for observed solves; kernels: the kernel
    table of a SYS_USER problem."""
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
    cpb = pb.to_c()
    ot, oy = (obs if obs is not None else (None, None))
    _check(lib().va_autotune(C.byref(cpb), None if kernels is None else C.c_void_p(kernels),
                             _ptr(y0), _ptr(p), int(cost), _ptr(cost_weights), _ptr(ot), _ptr(oy),
                             int(sample), int(reps), _stream(None)))
    best = {"block_size": int(cpb.block_size), "instances_per_wave": int(cpb.instances_per_wave),
            "checkpoint_slots": int(cpb.checkpoint_slots)}
    if verbose:
        print(f"[autotune] {key}: {best}")
    if use_cache:
        cache[key] = best
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1)
    return dataclasses.replace(pb, **best)
