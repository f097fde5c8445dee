"""The host-side C++ under AddressSanitizer (no GPU): the library's host-only translation units (file formats,
generators) and the FP64 oracle, compiled with g++ -fsanitize=address together with tests/asan_driver.cpp, which
exercises them on good and malformed inputs.  The run must exit 0 with no sanitizer report."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT

CSRC = os.path.join(ROOT, "application_classification_amd", "csrc")
REF_DATA = os.path.join(ROOT, "tests", "golden", "reference_data")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_and_oracle_are_clean_under_asan(tmp_path):
    exe = str(tmp_path / "asan_driver")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fopenmp",
           f"-I{os.path.join(ROOT, 'include')}", os.path.join(ROOT, "tests", "asan_driver.cpp"),
           os.path.join(CSRC, "ac_io.cpp"), os.path.join(CSRC, "ac_gen.cpp"),
           os.path.join(ROOT, "oracle", "ac_oracle.cpp"), "-o", exe, "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([exe, REF_DATA, str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and "Sanitizer" not in r.stderr, (r.stdout + r.stderr)[-4000:]
    assert "0 failed checks" in r.stdout
