"""Seeded workload generator and the FLOP model's tableau counts."""
import numpy as np
import pytest

import oracle as O
from vectorizedadjoint_amd import constants as K
from vectorizedadjoint_amd import flops as FL
from vectorizedadjoint_amd import inputs as I


def test_splitmix64_reference_values():
    # SplitMix64 from state 0: first outputs of the published generator
    z = I._mix64(np.array([0x9E3779B97F4A7C15], dtype=np.uint64))
    assert int(z[0]) == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("name", sorted(I.WORKLOADS))
def test_prefix_stable_and_deterministic(name):
    w = I.WORKLOADS[name]
    a = I.make_inputs(w, 64)
    b = I.make_inputs(w, 256)
    assert np.array_equal(a["y0"], b["y0"][:, :64])
    assert np.array_equal(a["p"], b["p"][:, :64])
    c = I.make_inputs(w, 64)
    assert np.array_equal(a["y0"], c["y0"]) and np.array_equal(a["p"], c["p"])
    assert a["y0"].shape == (w.n, 64) and a["p"].shape == (w.np, 64)


def test_distributions():
    d = I.make_inputs("glv10", 4096)
    N = 10
    r = d["p"][:N]
    A = d["p"][N:].reshape(N, N, -1)
    assert r.min() >= 0.1 and r.max() < 1.0
    diag = A[np.arange(N), np.arange(N)]
    assert diag.max() <= -0.5 and diag.min() > -1.5
    off = A[~np.eye(N, dtype=bool)]
    assert abs(off.mean()) < 0.01 and abs(off.std() - 0.2) < 0.01
    v = I.make_inputs("vdp", 4096)
    assert v["p"].min() >= 0.5 and v["p"].max() < 5.0
    assert abs(v["y0"][0].mean() - 2.0) < 1e-3 and abs(v["y0"][1].std() - 0.01) < 1e-3


@pytest.mark.parametrize("tid", [K.TAB_EULER, K.TAB_RK4, K.TAB_CASH_KARP54, K.TAB_DOPRI5,
                                 K.TAB_FEHLBERG78])
def test_flop_model_tableau_counts(tid):
    t = O.tableau(tid)
    s, e, nnz_a, nnz_as, nnz_b, nnz_db = FL.TABLEAU_COUNTS[tid]
    assert s == t["s"] and e == t["s_fwd"] - 1
    a = t["a"]
    assert nnz_a == int(np.count_nonzero(a[1:t["s_fwd"], :]))
    assert nnz_as == int(np.count_nonzero(a[1:t["s"], :t["s"]]))
    assert nnz_b == int(np.count_nonzero(t["b"][:t["s"]]))
    assert nnz_db == int(np.count_nonzero(t["db"]))
