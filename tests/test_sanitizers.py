"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

Builds tests/cpp/sanitize_driver.cpp with g++ -fsanitize=address,undefined
together with the CPU oracle, the synthetic-input toolkit and the CSV
measurement reader / writer (no device code on the line), then runs it: the
oracle's CKF in FP64 and FP32, its Kalman fit and measurement votes on the
C1-, C2-, C5-like and telescope geometries (solenoid map included), its
greedy ambiguity resolution on random track sets, the CSV
round trip of the reference's mock measurement file and a set of malformed
CSV files.  Any sanitizer finding aborts the driver with its report.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = ["tests/cpp/sanitize_driver.cpp", "oracle/ckf_oracle.cpp", "oracle/ambiguity_oracle.cpp",
           "traccc_amd/csrc/sim/simulation.cpp", "traccc_amd/csrc/common/config.cpp",
           "traccc_amd/csrc/common/last_error.cpp", "traccc_amd/csrc/io/csv.cpp"]


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the HIP headers")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "sanitize_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-Iinclude", "-Ioracle",
           "-Itraccc_amd/csrc", *SOURCES, "-lpthread", "-o", str(exe)]
    b = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), os.path.join(ROOT, "tests", "golden",
                                               "mock_event000000000-measurements.csv"),
                        str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    report = r.stdout[-2000:] + r.stderr[-6000:]
    assert r.returncode == 0, report
    assert "sanitized run: 0 failure(s)" in r.stdout, report
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
