"""Edge cases of the device path against the oracle: an empty time span, a single
instance and ragged batch sizes (not a multiple of the 64 lanes or of the six
10-lane groups of a wave), and every failure status (too many tries in one step,
too many steps), whose instances must report the oracle's status and NaN
gradients while the others keep the parity bar."""
import dataclasses

import numpy as np
import pytest

from helpers import run_oracle
from test_gpu_parity import _compare, _solve_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,ipw", [("vdp", 0), ("glv10", 6), ("glv10", 64)])
def test_empty_time_span(name, ipw):
    from vectorizedadjoint_amd import inputs as I, MODE_PERSISTENT
    w = dataclasses.replace(I.WORKLOADS[name], t1=I.WORKLOADS[name].t0)
    inp = I.make_inputs(w, 100)
    o = run_oracle(w, inp, max_steps=64)
    g = _solve_gpu(w, inp, MODE_PERSISTENT, max_steps=64, instances_per_wave=ipw)
    assert np.all(g["n_acc"] == 0) and np.all(g["status"] == 0)
    assert np.array_equal(g["yT"], inp["y0"])
    assert np.array_equal(g["dJdy0"], o["dJdy0"])
    assert np.all(g["dJdp"] == 0.0)


@pytest.mark.parametrize("B", [1, 7, 61, 1001])
@pytest.mark.parametrize("name", ["vdp", "glv10"])
def test_ragged_batches(name, B):
    from vectorizedadjoint_amd import inputs as I, MODE_PERSISTENT
    w = I.WORKLOADS[name]
    inp = I.make_inputs(w, B)
    o = run_oracle(w, inp, max_steps=1024)
    g = _solve_gpu(w, inp, MODE_PERSISTENT, max_steps=1024)
    _compare(g, o)


@pytest.mark.parametrize("mode_name", ["persistent", "plain"])
def test_failure_statuses_match_oracle(mode_name):
    from vectorizedadjoint_amd import inputs as I, MODE_PERSISTENT, MODE_PLAIN
    mode = {"persistent": MODE_PERSISTENT, "plain": MODE_PLAIN}[mode_name]
    # too many steps: a 70-step cap (about the median) cuts off part of the batch
    w = I.WORKLOADS["glv10"]
    inp = I.make_inputs(w, 512)
    o = run_oracle(w, inp, max_steps=70)
    g = _solve_gpu(w, inp, mode, max_steps=70)
    assert (o["status"] == 3).sum() > 100 and (o["status"] == 0).sum() > 10
    _compare(g, o)
    # too many tries: one try per step fails every instance that needs a rejection
    w = I.WORKLOADS["vdp"]
    inp = I.make_inputs(w, 256)
    o = run_oracle(w, inp, max_steps=4096, max_attempts=1)
    g = _solve_gpu(w, inp, mode, max_steps=4096, max_attempts=1)
    assert (o["status"] == 1).sum() > 10
    assert np.array_equal(g["status"], o["status"])
    assert np.array_equal(g["n_acc"], o["n_acc"]) and np.array_equal(g["n_rej"], o["n_rej"])
    bad = o["status"] != 0
    assert np.all(np.isnan(g["J"][bad])) and np.all(np.isnan(g["dJdy0"][:, bad]))
