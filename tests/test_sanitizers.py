"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

* The oracle and the SimdBackend-equivalent CPU port, rebuilt with
  `-fsanitize=address,undefined` (`make -C oracle san`), run the oracle and
  port test files in a child process with the sanitizer runtimes preloaded.
* The library's host code that runs without a GPU (the host Blake2s and
  channel, the Merkle and FRI decommitment planners and the gather
  bookkeeping; `tests/host_san`) is built with the sanitizers on its host
  side and run over many shapes; its Blake2s digests are checked against
  hashlib and its channel transcript against the oracle's channel.
* A canary with a planted heap overflow shows that the preloaded runtime does
  report one (so a clean run means something).
Each sanitizer report aborts its process; the tests check the exit status and
that no report was printed."""
import hashlib
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(REPO, "oracle")
HOST_SAN = os.path.join(REPO, "tests", "host_san")
sys.path.insert(0, ORACLE)
import oracle as O  # noqa: E402


def _runtime(name):
    p = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


ASAN, UBSAN = _runtime("libasan.so"), _runtime("libubsan.so")
pytestmark = pytest.mark.skipif(not (ASAN and UBSAN), reason="gcc sanitizer runtimes not installed")


def _san_env(**extra):
    env = dict(os.environ)
    env.update(LD_PRELOAD=ASAN + ":" + UBSAN, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    env.update(extra)
    return env


@pytest.fixture(scope="module")
def san_libs():
    subprocess.run(["make", "-s", "-C", ORACLE, "san"], check=True)
    return os.path.join(ORACLE, "_san", "liboracle.so"), os.path.join(ORACLE, "_san", "libsimdport_v3.so")


def test_oracle_and_cpu_port_tests_under_sanitizers(san_libs):
    lib_oracle, lib_port = san_libs
    files = ["test_oracle.py", "test_decommit_oracle.py", "test_quotients_oracle.py", "test_widefib_oracle.py",
             "test_poly_ops_oracle.py", "test_layout_oracle.py", "test_simd_port.py", "test_prove_oracle.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] +
                       [os.path.join(REPO, "tests", f) for f in files],
                       env=_san_env(STWO_ORACLE_LIB=lib_oracle, STWO_SIMDPORT_LIB=lib_port), cwd=REPO,
                       capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "Sanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert " passed" in out


def test_planted_overflow_is_reported(san_libs, tmp_path):
    src = tmp_path / "canary.c"
    src.write_text("#include <stdlib.h>\n"
                   "int canary(int i) { int *a = malloc(4 * sizeof(int)); int v = a[i]; free(a); return v; }\n")
    so = tmp_path / "libcanary.so"
    subprocess.run(["gcc", "-O1", "-g", "-fsanitize=address", "-fPIC", "-shared", "-o", str(so), str(src)],
                   check=True)
    code = "import ctypes; print(ctypes.CDLL(%r).canary(4))" % str(so)
    r = subprocess.run([sys.executable, "-c", code], env=_san_env(), capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr


def test_library_host_code_under_sanitizers():
    subprocess.run(["make", "-s", "-C", HOST_SAN], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([os.path.join(HOST_SAN, "_build", "host_checks")], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    lines = r.stdout.splitlines()
    # Blake2s of every length 0..200 of the pattern i * 7 + 3
    msg = bytes((i * 7 + 3) & 0xFF for i in range(200))
    b2 = [ln.split() for ln in lines if ln.startswith("b2 ")]
    assert len(b2) == 201
    for _, n, digest in b2:
        assert digest == hashlib.blake2s(msg[: int(n)]).hexdigest(), n
    # the channel transcript, replayed on the oracle's channel
    ch = O.Channel()
    ch.mix_root([0x01020304 * (i + 1) & 0xFFFFFFFF for i in range(8)])
    ch.draw_u32s()
    ch.mix_felts([[(i * 2654435761) & 0xFFFFFFFF & 0x7FFFFFFF for i in range(4 * k, 4 * k + 4)] for k in range(37)])
    ch.draw_u32s()
    ch.draw_u32s()
    ch.mix_u64(0x123456789ABCDEF)
    felt = ch.draw_secure_felt()
    chan = [ln.split() for ln in lines if ln.startswith("chan ")]
    assert chan == [["chan", ch.digest_bytes().hex(), str(ch.n_draws), str(felt[0])]]
    ok = [ln.split() for ln in lines if ln.startswith("ok ")]
    assert len(ok) == 1 and int(ok[0][1]) > 0 and int(ok[0][2]) > 0
