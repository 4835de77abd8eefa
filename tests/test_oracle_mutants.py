"""The oracle's pins are strong enough to catch plausible mistakes: each planted
mistake (compiled from va_oracle.c with -DORC_MUTANT=k into a temporary
directory, CPU only) must make at least one pinning check fail, while the
unmutated build passes the same checks."""
import os
import shutil
import subprocess

import pytest

import oracle as O
import test_oracle_adjoint as TA
import test_oracle_forward as TF
import test_oracle_math as TM
import test_oracle_observations as TO

HERE = os.path.dirname(os.path.abspath(O.__file__))


def _build(tmp_path, k):
    cc = shutil.which("gcc") or shutil.which("cc")
    out = os.path.join(str(tmp_path), f"liboracle_m{k}.so")
    subprocess.run([cc, "-O2", "-std=c99", "-fPIC", "-ffp-contract=off", "-mfma",
                    f"-DORC_MUTANT={k}", os.path.join(HERE, "va_oracle.c"), "-o", out,
                    "-shared", "-pthread", "-lm"], check=True)
    return out


def _glv_case():
    return [c for c in TA.CASES if c[0] == "glv10-dopri5"][0]


def _vdp_case():
    return [c for c in TA.CASES if c[0] == "vdp-dopri5"][0]


def _vjp_glv():
    for c in TM._systems():
        if c[0] == "glv10":
            TM.test_vjp_matches_complex_step(*c)


def _vjp_vdp():
    for c in TM._systems():
        if c[0] == "vdp":
            TM.test_vjp_matches_complex_step(*c)


MUTANTS = {
    1: ("transposed A in the Lotka-Volterra VJP", [_vjp_glv,
        lambda: TA.test_adjoint_matches_complex_step_of_frozen_map(_glv_case())]),
    2: ("dropped K_bar propagation term", [
        lambda: TA.test_adjoint_matches_complex_step_of_frozen_map(_vdp_case())]),
    3: ("sign error in the Van der Pol VJP", [_vjp_vdp,
        lambda: TA.test_adjoint_matches_complex_step_of_frozen_map(_vdp_case())]),
    4: ("off-by-one stage index in the stage sums", [
        lambda: TF.test_oscillator_against_closed_form(O.TAB_DOPRI5, 1e-6)]),
    6: ("wrong sign of the observation jump", [
        lambda: TO.test_observation_adjoint_matches_complex_step(O.TAB_DOPRI5)]),
    5: ("wrong step-shrink exponent", [
        lambda: TF.test_controller_invariants(O.TAB_DOPRI5, 5, 4)]),
}


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_planted_mistake_is_caught(tmp_path, k):
    desc, checks = MUTANTS[k]
    path = _build(tmp_path, k)
    caught = 0
    with O.use_library(path):
        for chk in checks:
            try:
                chk()
            except AssertionError:
                caught += 1
    assert caught >= 1, f"mutant {k} ({desc}) passed every pin"
    # the same checks pass on the real oracle
    for chk in checks:
        chk()
