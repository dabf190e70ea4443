"""CPU-only checks of the C ABI library: it loads, exports every function
include/stwo_hip.h declares, and its argument checks fail cleanly without a
device.  No kernel is launched here."""
import ctypes
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "stwo-metal_amd", "libstwo_hip.so")
HEADERS = [os.path.join(REPO, "include", f) for f in os.listdir(os.path.join(REPO, "include")) if f.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(stwo_hip_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(REPO, "stwo-metal_amd"), "-j8", "libstwo_hip.so"], check=True)
    return ctypes.CDLL(LIB)


def test_header_declares_the_hot_path():
    names = declared_functions()
    for required in ["stwo_hip_interpolate_columns", "stwo_hip_evaluate_polynomials", "stwo_hip_bit_reverse_column",
                     "stwo_hip_merkle_commit_on_layer", "stwo_hip_merkle_lifted_build_leaves",
                     "stwo_hip_fold_circle_into_line", "stwo_hip_fold_line", "stwo_hip_fri_commit"]:
        assert required in names


def test_every_declared_symbol_is_exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        getattr(lib, n)


def test_every_exported_symbol_is_declared(lib):
    """The reverse direction: no C-ABI entry point the library exports is
    missing from include/."""
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    undeclared = sorted(n for n in exported if n.startswith("stwo_hip_") and n not in declared_functions())
    assert not undeclared, undeclared


def test_null_arguments_are_rejected_without_a_device(lib):
    vp = ctypes.c_void_p
    lib.stwo_hip_last_error.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    assert lib.stwo_hip_last_error(None, None, 0) == -1
    lib.stwo_hip_sync.argtypes = [vp]
    assert lib.stwo_hip_sync(None) == -1
    lib.stwo_hip_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    assert lib.stwo_hip_ctx_create(0, None) == -1
    lib.stwo_hip_interpolate_columns.argtypes = [vp, vp, vp, ctypes.c_size_t, vp]
    assert lib.stwo_hip_interpolate_columns(None, None, None, 0, None) == -1
    lib.stwo_hip_merkle_commit.argtypes = [vp, vp, vp, ctypes.c_size_t, vp, ctypes.c_int]
    assert lib.stwo_hip_merkle_commit(None, None, None, 0, None, 0) == -1
