"""The drop-in C++ API (include/vectorizedadjoint/lib.hpp).

CPU: the headers compile as plain C++17 and bind the example functors to the
right device systems (identify_system, run on the host).
GPU: tests/cpp/va_examples runs the reference's three example drivers
(examples/VanDerPol, examples/HarmonicOscillator, examples/GeneralizedLotkaVolterra)
through Driver / runge_kutta / setCostGradients / constructDriverButcherTableau /
recordDriverRHSFunction / adjointSolve / computeSensitivityMatrix, and every
output is compared with the CPU oracle run on the same problem: step count,
stored times and steps, observer calls and y(T) bit for bit; lambda (dJ/dy0),
mu (dJ/dp) and the sensitivity matrix within 1e-10 relative.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from vectorizedadjoint_amd import constants as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
EXAMPLES = os.path.join(ROOT, "tests", "cpp", "va_examples")
GOLDEN = os.path.join(ROOT, "tests", "golden")
TOL = 1e-10

IDENTIFY_SRC = r"""
#include <cmath>
#include <cstdio>
#include <vector>
#include "vectorizedadjoint/lib.hpp"
using namespace vectorizedadjoint;
struct Vdp { template <class T> void operator()(const std::vector<T>&x, std::vector<T>&f,
    const std::vector<T>&p, T) const { f[0]=x[1]; f[1]=p[0]*((1.0-x[0]*x[0])*x[1]-x[0]); } };
struct Ho { double k = 2.5; template <class T> void operator()(const std::vector<T>&r,
    std::vector<T>&f, const std::vector<T>&m, T) const { f[0]=r[1]; f[1]=-k*r[0]-m[0]*r[1]; } };
struct Osc { template <class T> void operator()(const std::vector<T>&r, std::vector<T>&f,
    const std::vector<T>&p, T) const { f[0]=r[1]; f[1]=-p[0]*r[0]-p[1]*r[1]; } };
struct Glv { template <class T> void operator()(const std::vector<T>&x, std::vector<T>&f,
    const std::vector<T>&p, T) { int n=x.size(); for (int i=0;i<n;i++) { T s=0.0;
      for (int j=0;j<n;j++) { s+=p[n*(i+1)+j]*x[j]; }
      f[i]=x[i]*(p[i]+s); } } };
struct Other { template <class T> void operator()(const std::vector<T>&x, std::vector<T>&f,
    const std::vector<T>&p, T) const { f[0]=x[0]*x[1]; f[1]=p[0]; } };
struct Bound { template <class T> void operator()(const std::vector<T>&, std::vector<T>&f,
    const std::vector<T>&, T) const { f[0]=0; f[1]=0; } };
VA_BIND_SYSTEM(Bound, VA_SYSTEM_VAN_DER_POL, 0.0)
// forced oscillator: the damped oscillator plus a forcing term in t
struct Forced { template <class T> void operator()(const std::vector<T>&r, std::vector<T>&f,
    const std::vector<T>&p, T t) const { f[0]=r[1]; f[1]=-p[0]*r[0]-p[1]*r[1]+std::sin(t); } };
// piecewise: Van der Pol below x0 = 2.5, something else above
struct Piecewise { template <class T> void operator()(const std::vector<T>&x, std::vector<T>&f,
    const std::vector<T>&p, T) const { f[0]=x[1];
      f[1] = x[0] < 2.5 ? p[0]*((1.0-x[0]*x[0])*x[1]-x[0]) : -x[0]; } };
int main() {
  Vdp a; Ho b; Osc c; Glv d; Other e; Bound g; Forced fo; Piecewise pw;
  auto s1 = identify_system(a, 2, 1), s2 = identify_system(b, 2, 1), s3 = identify_system(c, 2, 2);
  auto s4 = identify_system(d, 10, 110), s6 = identify_system(g, 2, 1);
  int other = 0, forced = 0, piecewise = 0;
  try { identify_system(e, 2, 1); } catch (std::invalid_argument &) { other = 1; }
  try { identify_system(fo, 2, 2); } catch (std::invalid_argument &) { forced = 1; }
  try { identify_system(pw, 2, 1); } catch (std::invalid_argument &) { piecewise = 1; }
  std::printf("%d %d %.17g %d %d %d %d %d %d\n", s1.id, s2.id, s2.constant, s3.id, s4.id, s6.id,
              other, forced, piecewise);
  return 0;
}
"""


def _compile(src, out, tmp_path):
    f = tmp_path / "t.cpp"
    f.write_text(src)
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I", INCLUDE,
           "-I", "/opt/rocm/include", str(f), "-o", str(out),
           "-L", os.path.join(ROOT, "vectorizedadjoint_amd"), "-lvectorizedadjoint_amd",
           "-L", os.path.join(ROOT, "tests", "cpp"), "-lva_user_system",
           "-L", "/opt/rocm/lib", "-lamdhip64",
           f"-Wl,-rpath,{os.path.join(ROOT, 'vectorizedadjoint_amd')}",
           f"-Wl,-rpath,{os.path.join(ROOT, 'tests', 'cpp')}", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_headers_compile_and_identify_systems(tmp_path):
    if not os.path.exists(os.path.join(ROOT, "vectorizedadjoint_amd",
                                       "libvectorizedadjoint_amd.so")):
        pytest.skip("library not built")
    exe = tmp_path / "identify"
    _compile(IDENTIFY_SRC, exe, tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    ids = r.stdout.split()
    assert int(ids[0]) == K.SYS_VAN_DER_POL
    assert int(ids[1]) == K.SYS_OSCILLATOR_FIXED_K and float(ids[2]) == 2.5
    assert int(ids[3]) == K.SYS_OSCILLATOR
    assert int(ids[4]) == K.SYS_LOTKA_VOLTERRA
    assert int(ids[5]) == K.SYS_VAN_DER_POL      # explicit VA_BIND_SYSTEM
    assert int(ids[6]) == 1                      # unknown functor -> std::invalid_argument
    assert int(ids[7]) == 1                      # time-dependent (forced) functor refused
    assert int(ids[8]) == 1                      # piecewise functor refused


def test_example_source_builds_against_headers(tmp_path):
    """The example drivers compile with -Werror against the drop-in headers."""
    if not os.path.exists(os.path.join(ROOT, "vectorizedadjoint_amd",
                                       "libvectorizedadjoint_amd.so")):
        pytest.skip("library not built")
    src = open(os.path.join(ROOT, "tests", "cpp", "va_examples.cpp")).read()
    _compile(src, tmp_path / "ex", tmp_path)


# ---------------------------------------------------------------- GPU ---------------

def _run_example(*args):
    assert os.path.exists(EXAMPLES), "tests/cpp/va_examples is not built"
    print(f"[gpu] va_examples {' '.join(args)}", flush=True)
    r = subprocess.run([EXAMPLES, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _check_against_oracle(res, pb, nout_identity):
    y0 = np.array(res["y0"])
    p = np.array(res["p"])
    o = O.Solve(pb, y0, p)
    assert o.status == 0
    T = len(o.times)
    assert res["steps"] == o.steps
    assert res["stored"] == T and res["observer_calls"] == T
    assert np.array_equal(np.array(res["t"]), o.times)
    assert np.array_equal(np.array(res["h"]), o.steps_h)
    assert np.array_equal(np.array(res["yT"]), o.y)
    assert res["observer_first"] == [o.times[0], *y0.tolist()]
    assert res["observer_last"] == [o.times[-1], *o.y.tolist()]
    # Driver::Rhs runs the caller's functor on the host (its own summation order)
    assert _rel(res["rhs_y0"], O.rhs(pb.system, y0, p, 0.0, pb.sys_const)) <= 1e-14
    n, npar = len(y0), len(p)
    if nout_identity:
        lam0 = np.eye(n)
    else:
        lam0 = o.y.reshape(1, n).copy()
    lam, mu = o.adjoint_solve(lam0, np.zeros((lam0.shape[0], npar)))
    assert _rel(res["lambda"], lam) <= TOL
    assert _rel(res["mu"], mu) <= TOL
    S = o.sensitivity_matrix()
    assert _rel(res["sensitivity"], S) <= TOL
    return o


@pytest.mark.gpu
def test_example_van_der_pol_fehlberg78():
    res = json.loads(_run_example("vdp", "1e-5"))
    pb = O.problem(K.SYS_VAN_DER_POL, 2, 1, K.TAB_FEHLBERG78, t0=0.0, t1=0.5, dt0=1e-3,
                   abs_tol=1e-5, rel_tol=1e-5)
    _check_against_oracle(res, pb, True)


@pytest.mark.gpu
def test_example_harmonic_oscillator_rk4_fixed():
    res = json.loads(_run_example("ho"))
    pb = O.problem(K.SYS_OSCILLATOR_FIXED_K, 2, 1, K.TAB_RK4, controlled=False, t0=0.0,
                   t1=10.0, dt0=0.01, sys_const=1.0)
    o = _check_against_oracle(res, pb, False)
    assert res["steps"] == 1000 and len(o.times) == 1001


@pytest.mark.gpu
@pytest.mark.parametrize("n", [5, 10])
def test_example_lotka_volterra_cash_karp(n):
    csv = os.path.join(GOLDEN, f"glv_N{n}_alphasfile_cpp.csv")
    res = json.loads(_run_example("glv", "1e-5", str(n), csv))
    pb = O.problem(K.SYS_LOTKA_VOLTERRA, n, n + n * n, K.TAB_CASH_KARP54, t0=0.0, t1=10.0,
                   dt0=1e-3, abs_tol=1e-5, rel_tol=1e-5)
    _check_against_oracle(res, pb, True)


@pytest.mark.gpu
def test_example_user_system_brusselator():
    """A functor the library does not know, through the unchanged reference API:
    recordDriverRHSFunction / runge_kutta bind it (VA_BIND_DEVICE_SYSTEM) to the
    device functor compiled outside the library; the oracle runs the same functor's
    host build as its callbacks."""
    import ctypes
    res = json.loads(_run_example("user", "1e-8"))
    # the callbacks come from the functor's host build, a library without device
    # code: the device library is the example's
    lib = ctypes.CDLL(os.path.join(ROOT, "tests", "cpp", "libva_brusselator_host.so"))
    O.set_user_system(lib.brusselator_rhs, lib.brusselator_vjp)
    pb = O.problem(O.SYS_USER, 2, 2, K.TAB_DOPRI5, t0=0.0, t1=10.0, dt0=1e-3,
                   abs_tol=1e-8, rel_tol=1e-8)
    _check_against_oracle(res, pb, True)


@pytest.mark.gpu
def test_adjoint_preconditions_print_and_return():
    out = _run_example("pre").splitlines()
    assert out[0] == "Must call setCostGradients() first!"
    assert out[1] == "Must call constructDriverButcherTableau() to set Butcher Tableau!"
    assert out[2].startswith("Must call recordDriverRHSFunction()")
    assert out[3].startswith("The Butcher tableau differs")
    res = json.loads(out[4])
    pb = O.problem(K.SYS_VAN_DER_POL, 2, 1, K.TAB_DOPRI5, t0=0.0, t1=1.0, dt0=1e-3,
                   abs_tol=1e-8, rel_tol=1e-8)
    o = O.Solve(pb, np.array([2.0, 0.0]), np.array([1.0]))
    lam, mu = o.backward(np.ones(2))
    assert _rel(res["lambda"], lam) <= TOL and _rel([res["mu"]], mu) <= TOL


@pytest.mark.gpu
def test_batched_adjoint_function():
    """AdjointFunction (forward / backward over 64 Van der Pol instances) against the
    oracle's batch solve with J = |y(T)|^2."""
    res = json.loads(_run_example("fn"))
    B = res["batch"]
    y0 = np.array(res["y0"]).reshape(2, B)
    p = np.array(res["p"]).reshape(1, B)
    pb = O.problem(K.SYS_VAN_DER_POL, 2, 1, K.TAB_DOPRI5, t0=0.0, t1=20.0, dt0=1e-3,
                   abs_tol=1e-8, rel_tol=1e-8, max_steps=4096)
    o = O.solve_batch(pb, y0, p, K.COST_SQUARED_NORM, nthreads=4)
    assert (o["status"] == 0).all()
    assert np.array_equal(np.array(res["yT"]).reshape(2, B), o["yT"])
    assert _rel(np.array(res["dJdy0"]).reshape(2, B), o["dJdy0"]) <= TOL
    assert _rel(np.array(res["dJdp"]).reshape(1, B), o["dJdp"]) <= TOL
