"""Algorithmic FLOP and byte counts of the forward and reverse sweeps (FMA = 2 FLOP).

These are design-independent yardsticks: they count the arithmetic the
algorithm needs (no redundant recomputation, no padding lanes), so a kernel
that recomputes or idles shows up as a lower fraction of peak.

Per instance, with A attempts, S accepted steps, N states, s solution stages,
e = stages evaluated per attempt beyond the first (f(y_n) is evaluated once per
accepted step and reused by every retry, or carried over for FSAL tableaus):

  forward = A * (e * F_rhs + L_att) + S * F_rhs
      L_att = 2N nnz(a) + 2N nnz(b) + 2N nnz(db) + 6N   (stage inputs, y', xerr, norm)
  reverse = S * (s * (F_rhs + F_vjp) + L_rev)
      L_rev = 2N nnz(a_s) (stage-input recompute) + N s (seed K_bar = (b h) lambda)
            + N s (y_bar accumulation) + 2N nnz(a_s) (K_bar propagation)
  F_rhs, F_vjp per system: Van der Pol 5 / 13, oscillator 3 / 7,
  Lotka-Volterra(N) 2N^2+N / 4N^2+3N (growth g = r + A x reused from the recompute).
"""
from __future__ import annotations

import numpy as np

from . import constants as K

# (s_rev, e = SF - 1, nnz(a) over the forward stages, nnz(a) over the solution stages,
#  nnz(b), nnz(db))
TABLEAU_COUNTS = {
    K.TAB_EULER: (1, 0, 0, 0, 1, 0),
    K.TAB_RK4: (4, 3, 3, 3, 4, 0),
    K.TAB_CASH_KARP54: (6, 5, 15, 15, 4, 5),
    K.TAB_DOPRI5: (6, 6, 20, 15, 5, 6),
    K.TAB_FEHLBERG78: (13, 12, 55, 55, 7, 4),
}

TABLEAU_NAMES = {
    K.TAB_EULER: "euler", K.TAB_RK4: "rk4", K.TAB_CASH_KARP54: "cash_karp54",
    K.TAB_DOPRI5: "dopri5", K.TAB_FEHLBERG78: "fehlberg78",
}


def system_flops(system: int, n: int) -> tuple[int, int]:
    if system == K.SYS_VAN_DER_POL:
        return 5, 13
    if system in (K.SYS_OSCILLATOR, K.SYS_OSCILLATOR_FIXED_K):
        return 3, 7
    if system in (K.SYS_LOTKA_VOLTERRA, K.SYS_LOTKA_VOLTERRA_SHARED):
        return 2 * n * n + n, 4 * n * n + 3 * n
    raise ValueError(system)


def per_attempt_forward(system, n, tableau) -> int:
    s, e, nnz_a, _, nnz_b, nnz_db = TABLEAU_COUNTS[tableau]
    f_rhs, _ = system_flops(system, n)
    return e * f_rhs + 2 * n * (nnz_a + nnz_b + nnz_db) + 6 * n


def per_step_reverse(system, n, tableau) -> int:
    s, _, _, nnz_as, _, _ = TABLEAU_COUNTS[tableau]
    f_rhs, f_vjp = system_flops(system, n)
    return s * (f_rhs + f_vjp) + 4 * n * nnz_as + 2 * n * s


def forward_flops(system, n, tableau, n_acc, n_rej) -> float:
    f_rhs, _ = system_flops(system, n)
    A = np.asarray(n_acc, np.float64) + np.asarray(n_rej, np.float64)
    S = np.asarray(n_acc, np.float64)
    return float(np.sum(A) * per_attempt_forward(system, n, tableau) + np.sum(S) * f_rhs)


def reverse_flops(system, n, tableau, n_acc) -> float:
    return float(np.sum(np.asarray(n_acc, np.float64)) * per_step_reverse(system, n, tableau))


def trajectory_bytes(n, n_acc) -> float:
    """Bytes of (t_n, y_n) rows written by the forward and read back by the reverse."""
    return float(np.sum(np.asarray(n_acc, np.float64) + 1) * (n + 1) * 8)


def io_bytes(n, np_, B) -> float:
    """Algorithmic HBM bytes of one solve: y0, p in; y(T), dJ/dy0, dJ/dp, J out."""
    return float(B * 8 * (n + np_ + n + n + np_ + 1))
