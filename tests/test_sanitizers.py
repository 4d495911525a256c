"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only): the oracle's C
restatement (oracle/v3dm_oracle.c) and the library's robust-orientation host code
(eth-v3d-lgpl_amd/csrc/v3dm_ransac.cpp), driven by tests/cpp/sanitize_host.cpp on ragged and
degenerate inputs.  Built with gcc / g++ (no GPU code is involved); any sanitizer report
aborts the run (-fno-sanitize-recover=all)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="gcc / g++ not available")

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def test_host_code_under_sanitizers(tmp_path):
    oracle_o = tmp_path / "oracle.o"
    exe = tmp_path / "sanitize_host"
    r = subprocess.run(["gcc", "-std=gnu99", "-fopenmp", "-ffp-contract=off", "-c", *SAN,
                        os.path.join(ROOT, "oracle", "v3dm_oracle.c"), "-o", str(oracle_o)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run(["g++", "-std=c++17", "-fopenmp", *SAN, os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp"),
                        os.path.join(ROOT, "eth-v3d-lgpl_amd", "csrc", "v3dm_ransac.cpp"), str(oracle_o),
                        "-o", str(exe), "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # verify_asan_link_order=0: a library preloaded by the environment may come before the
    # ASan runtime; the run is still instrumented
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    assert "sanitized host run: ok" in r.stdout
