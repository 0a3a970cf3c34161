"""AddressSanitizer and UndefinedBehaviorSanitizer over the host-only C++:
the library's time FE algebra, spec parser and mesh hierarchy
(dealii-stfem_amd/csrc/{time_fe,spec,mesh}.cpp, driven by
tests/sanitize/host_driver.cpp) and the oracle (oracle/stfem_oracle.cpp,
driven by tests/sanitize/oracle_driver.cpp), built with g++ and run on the
CPU.  Any memory error, leak-free exit aside, or undefined behaviour fails
the test (UBSan without recovery)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dealii-stfem_amd", "csrc")
SAN = os.path.join(ROOT, "tests", "sanitize")
FLAGS = ["-std=c++17", "-O1", "-g", "-fopenmp", "-fsanitize=address,undefined",
         "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _build_and_run(tmp_path, name, sources, includes):
    exe = str(tmp_path / name)
    cmd = ["g++"] + FLAGS + [f"-I{i}" for i in includes] + sources + ["-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert "0 failures" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]


def test_library_host_code_under_sanitizers(tmp_path):
    _build_and_run(tmp_path, "host_driver",
                   [os.path.join(SAN, "host_driver.cpp")] +
                   [os.path.join(CSRC, f) for f in ("time_fe.cpp", "spec.cpp", "mesh.cpp")],
                   [CSRC])


def test_oracle_under_sanitizers(tmp_path):
    _build_and_run(tmp_path, "oracle_driver",
                   [os.path.join(SAN, "oracle_driver.cpp"),
                    os.path.join(ROOT, "oracle", "stfem_oracle.cpp")], [])
