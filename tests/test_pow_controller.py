"""The step controller's pow (vectorizedadjoint_amd/csrc/va_pow.hpp), CPU only.

Boost.odeint's step adjuster calls std::pow(err, -1/q); the reference's results
therefore carry the C library's rounding of pow, which is not always the
correctly rounded power.  The device restates that pow's algorithm; here the
same header, built for the host, is compared bit for bit with the host libm
pow -- on random arguments, on every controller exponent, and on the error
ratios an actual solve feeds the controller -- and its tables are regenerated
from their definitions.  mpmath (an independent arbitrary-precision library)
confirms that the C library's pow indeed differs from the correctly rounded
power on a fraction of the controller's arguments, which is why a correctly
rounded device pow would not reproduce the reference's step sequences.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "vectorizedadjoint_amd", "csrc")
SRC = os.path.join(ROOT, "tests", "cpp", "pow_check.cpp")


@pytest.fixture(scope="module")
def pow_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("pow") / "pow_check")
    # -ffp-contract=off: the host build keeps the header's operation sequence
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, SRC, "-o", exe,
                    "-lm"], check=True)
    return exe


def _run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    checked, bad = r.stdout.split()[1], r.stdout.split()[3]
    return int(checked), int(bad)


def test_tables_regenerate_from_their_definitions():
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen", os.path.join(CSRC, "gen_va_pow_tables.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    with open(os.path.join(CSRC, "va_pow_tables.hpp")) as f:
        assert f.read() == gen.render()


def test_matches_libm_on_random_arguments(pow_check):
    checked, bad = _run(pow_check, "random", 250000, 20241203)
    assert checked == 250000 * 8 + 1 and bad == 0


def _controller_arguments():
    """(err, exponent) pairs the controller formed in oracle solves of the
    Lotka-Volterra (Cash-Karp 5(4)) and Van der Pol (DOPRI5) workloads."""
    from vectorizedadjoint_amd import inputs as I
    pairs = []
    for name, count in (("glv10", 120), ("vdp", 60)):
        w = I.WORKLOADS[name]
        inp = I.make_inputs(w, count)
        pb = O.problem(w.system, w.n, w.np, w.tableau, t0=w.t0, t1=w.t1, dt0=w.dt0,
                       abs_tol=w.tol, rel_tol=w.tol, max_steps=4096)
        tab = O.tableau(w.tableau)
        lo = 1.0 / 5.0 ** tab["order"]
        for b in range(count):
            s = O.Solve(pb, inp["y0"][:, b], inp["p"][:, b], log_cap=20000)
            for _, _, err, ok in s.log:
                if ok and err < 0.5:
                    pairs.append((max(err, lo), -1.0 / tab["order"]))
                elif not ok:
                    pairs.append((err, -1.0 / (tab["err_order"] - 1)))
    return pairs


def test_matches_libm_on_controller_arguments(pow_check, tmp_path):
    pairs = _controller_arguments()
    assert len(pairs) > 10000
    path = tmp_path / "pairs.txt"
    path.write_text("".join(f"{x.hex()} {y.hex()}\n" for x, y in pairs))
    checked, bad = _run(pow_check, "file", path)
    assert checked == len(pairs) and bad == 0


def test_libm_pow_is_not_correctly_rounded_on_controller_arguments():
    """The premise: libm's pow differs from the correctly rounded power on some
    of the controller's arguments (about 8 in 10^4 on this C library), so only a
    restatement of that pow reproduces the reference's step sizes."""
    import mpmath
    mpmath.mp.prec = 200
    rng = np.random.default_rng(7)
    xs = np.exp(rng.uniform(np.log(3.2e-4), np.log(0.5), 20000))
    y = -1.0 / 5.0
    differ = 0
    for x in xs:
        x = float(x)
        cr = float(mpmath.power(mpmath.mpf(x), mpmath.mpf(y)))
        differ += (x ** y) != cr  # CPython's float ** calls the C library's pow
    print(f"libm pow != correctly rounded on {differ} of {xs.size}")
    assert 1 <= differ <= 200
