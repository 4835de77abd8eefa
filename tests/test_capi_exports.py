"""The C-ABI library loads and exports every function include/vectorizedadjoint_amd.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "vectorizedadjoint_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(va_[a-z_0-9]+)\s*\(", src)))


def test_header_lists_the_abi():
    import vectorizedadjoint_amd as va
    assert sorted(va.C_ABI_FUNCTIONS) == declared_functions()


def test_library_exports_every_declared_symbol():
    import vectorizedadjoint_amd as va
    lib = ctypes.CDLL(va.library_path())
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_problem_struct_layout_matches_header():
    import vectorizedadjoint_amd as va
    # int32 x6, int64, double x6, int32 x6, int64 x2, int32 x2 -> 128 bytes
    assert ctypes.sizeof(va.VaProblem) == 128
    assert va.VaProblem.batch.offset == 24 and va.VaProblem.max_attempts.offset == 80
    assert va.VaProblem.instances_per_wave.offset == 96
    assert va.VaProblem.lanes.offset == 104
    assert va.VaProblem.trajectory_budget_bytes.offset == 112
    assert va.VaProblem.n_obs.offset == 120 and va.VaProblem.checkpoint_slots.offset == 100


def test_header_constants_match_the_python_mirror():
    # every #define VA_<NAME> <int> of the header has the same value in constants.py
    from vectorizedadjoint_amd import constants as K
    src = open(os.path.join(ROOT, "include", "vectorizedadjoint_amd.h")).read()
    pairs = re.findall(r"#define VA_(FLAG_[A-Z_]+)\s+(\d+)", src)
    assert {name for name, _ in pairs} >= {"FLAG_GRAPH", "FLAG_COLD_ORDER"}
    for name, value in pairs:
        assert getattr(K, name) == int(value), name
