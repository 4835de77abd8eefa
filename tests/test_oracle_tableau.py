"""Pins the oracle's Butcher tableaus (ButcherTable.hpp:39-251 plus DOPRI5) by the
mathematics: row sums, order conditions of the solution weights up to the
stepper order, of the embedded weights up to the error order, failure of the
next order, and the observed convergence order of the fixed-step loop on the
damped oscillator, whose exact solution is known."""
import numpy as np
import pytest

import oracle as O
from refmath import order_condition_residuals, rooted_trees

TABS = {
    O.TAB_EULER: ("euler", 1, None),
    O.TAB_RK4: ("rk4", 4, None),
    O.TAB_CASH_KARP54: ("cash_karp54", 5, 4),
    O.TAB_DOPRI5: ("dopri5", 5, 4),
    O.TAB_FEHLBERG78: ("fehlberg78", 8, 7),
}


def test_tree_counts():
    # number of rooted trees of order 1..8 (OEIS A000081)
    counts = [0] * 9
    for o, _ in rooted_trees(8):
        counts[o] += 1
    assert counts[1:] == [1, 1, 2, 4, 9, 20, 48, 115]


@pytest.mark.parametrize("tid", sorted(TABS))
def test_row_sums(tid):
    t = O.tableau(tid)
    a, c = t["a"], t["c"]
    for i in range(t["s_fwd"]):
        if tid == O.TAB_FEHLBERG78 and i == 11:
            # Fehlberg's stage 12 has c = 0 with nonzero a entries summing to 0
            pass
        assert abs(a[i, :i].sum() - c[i]) <= 1e-14, (i, a[i, :i].sum(), c[i])
    assert np.allclose(np.triu(a), 0.0)


@pytest.mark.parametrize("tid", sorted(TABS))
def test_order_conditions(tid):
    t = O.tableau(tid)
    name, p, q = TABS[tid]
    assert t["order"] == p
    s = t["s_fwd"]
    A = t["a"][:s, :s]
    b = np.zeros(s)
    b[: t["s"]] = t["b"][: t["s"]]
    res = order_condition_residuals(A, b, p + 1)
    worst = max(abs(r) for o, r in res if o <= p)
    assert worst <= 1e-13, (name, worst)
    assert max(abs(r) for o, r in res if o == p + 1) > 1e-6, f"{name} looks higher than order {p}"
    if q is not None:
        assert t["err_order"] == q
        bhat = b - t["db"][:s]
        res = order_condition_residuals(A, bhat, q + 1)
        assert max(abs(r) for o, r in res if o <= q) <= 1e-13
        assert max(abs(r) for o, r in res if o == q + 1) > 1e-7


def test_dopri5_fsal_row_is_b():
    t = O.tableau(O.TAB_DOPRI5)
    assert t["fsal"] and t["s"] == 6 and t["s_fwd"] == 7
    assert np.array_equal(t["a"][6, :6], t["b"][:6]) and t["c"][6] == 1.0


def _ho_exact(y0, k, c, T):
    # y'' + c y' + k y = 0, underdamped (c^2 < 4k)
    w = np.sqrt(k - c * c / 4.0)
    a = -c / 2.0
    x0, v0 = y0
    A = x0
    Bc = (v0 - a * x0) / w
    x = np.exp(a * T) * (A * np.cos(w * T) + Bc * np.sin(w * T))
    v = np.exp(a * T) * ((a * A + w * Bc) * np.cos(w * T) + (a * Bc - w * A) * np.sin(w * T))
    return np.array([x, v])


@pytest.mark.parametrize("tid", sorted(TABS))
def test_fixed_step_convergence_order(tid):
    name, p, _ = TABS[tid]
    k, c, T = 2.0, 0.3, 2.0
    y0 = np.array([0.7, -0.4])
    exact = _ho_exact(y0, k, c, T)
    h0 = {1: 1e-3, 4: 0.05, 5: 0.1, 8: 0.4}[p]
    errs = []
    for h in (h0, h0 / 2):
        pb = O.problem(O.SYS_OSCILLATOR, 2, 2, tid, controlled=False, t0=0.0, t1=T, dt0=h)
        s = O.Solve(pb, y0, [k, c])
        assert s.status == O.OK
        errs.append(np.max(np.abs(s.y - exact)))
    rate = np.log2(errs[0] / errs[1])
    assert abs(rate - p) < 0.35, (name, errs, rate)
