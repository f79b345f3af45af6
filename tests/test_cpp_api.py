"""The C++ host interface (include/traccc_amd/combinatorial_kalman_filter_algorithm.hpp),
driven from a C++ program built with g++ against the in-tree libraries.

CPU: the constructor rejects the configurations traccc's constructors reject
(core/src/finding/combinatorial_kalman_filter_algorithm.cpp:21-32,
device/common/src/finding/combinatorial_kalman_filter_algorithm.cpp:33-46)
with std::invalid_argument.
GPU: both operator() forms (host views, HBM-resident inputs) return the FP64
oracle's track candidates on a 200-muon barrel event.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "traccc_amd", "_lib")


@pytest.fixture(scope="module")
def cpp_api(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "ckf_cpp_api")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror",
                    f"-I{os.path.join(ROOT, 'include')}",
                    os.path.join(ROOT, "tests", "cpp", "ckf_cpp_api.cpp"),
                    f"-L{LIB}", "-ltraccc_amd", "-ltraccc_amd_sim", f"-Wl,-rpath,{LIB}",
                    "-o", exe], check=True)
    return exe


def test_constructor_rejects_invalid_configs(cpp_api):
    r = subprocess.run([cpp_api, "config"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok invalid") == 3


def test_cpp_event_matches_python_generator(cpp_api):
    """The C++ program's 200-muon barrel event is the Python C1 event."""
    from traccc_amd import simulation as S
    r = subprocess.run([cpp_api, "event", "200"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"seeds (\d+) measurements (\d+) hash ([0-9a-f]+)", r.stdout)
    assert m, r.stdout
    det, fld, gun = S.workload("C1", seed=42, n_particles=200)
    ev = S.generate_event(det, fld, gun)
    assert int(m.group(1)) == ev.n_seeds
    assert int(m.group(2)) == ev.n_meas
    bits = np.ascontiguousarray(ev.local_position, dtype=np.float32).view(np.uint32).ravel()
    assert int(m.group(3), 16) == _fnv(bits)


def _fnv(values):
    h = 1469598103934665603
    for v in values:
        h = ((h ^ int(v)) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.gpu
def test_cpp_operator_matches_oracle(cpp_api, gpu_available):
    from traccc_amd import simulation as S
    from tests.helpers import run_finder

    print("[cpp_api] C1 barrel, 200 muons, seed 42", flush=True)
    r = subprocess.run([cpp_api, "run", "200"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"seeds (\d+) tracks (\d+) constituents (\d+) hash ([0-9a-f]+)", r.stdout)
    assert m, r.stdout
    det, fld, gun = S.workload("C1", seed=42, n_particles=200)
    ev = S.generate_event(det, fld, gun)
    ref = run_finder("oracle", S.finding_config_for("C1"), det, fld, ev)
    assert int(m.group(1)) == ev.n_seeds
    assert int(m.group(2)) == len(ref["offsets"]) - 1
    assert int(m.group(3)) == len(ref["constituents"])
    expect = _fnv(np.concatenate([ref["constituents"].astype(np.uint64),
                                  ref["offsets"].astype(np.uint64)]))
    assert int(m.group(4), 16) == expect
