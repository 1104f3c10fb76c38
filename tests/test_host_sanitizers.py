"""Host AddressSanitizer + UndefinedBehaviorSanitizer runs (CPU only).

* the CPU oracle (oracle/umi_oracle.c) and the descriptor validator
  (umi_amd/csrc/validate.hip, host code compiled as C++) driven by
  tests/cpp_host/sanitize_driver.cpp on ragged inputs in exact-size buffers;
* the Execute interpreter (oracle/umi_interp.h) through its known-answer
  program tests/cpp_host/test_interp.cpp.
Any out-of-bounds access, leak or undefined operation fails the run.
"""
import os
import shutil
import subprocess
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compiler")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, f"{' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr[-4000:]}"
    return r


def test_oracle_and_validator_under_sanitizers():
    with tempfile.TemporaryDirectory() as d:
        obj = Path(d) / "oracle.o"
        exe = Path(d) / "sanitize_driver"
        _run(["gcc", "-std=c11", "-ffp-contract=off", "-fopenmp", *SAN, "-c", str(ROOT / "oracle" / "umi_oracle.c"),
              "-o", str(obj)])
        _run(["g++", "-std=c++17", *SAN, "-fopenmp", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
              f"-I{ROOT / 'include'}", f"-I{ROOT / 'umi_amd' / 'csrc'}", "-x", "c++",
              str(ROOT / "umi_amd" / "csrc" / "validate.hip"), "-x", "none",
              str(ROOT / "tests" / "cpp_host" / "sanitize_driver.cpp"), str(obj), "-lm", "-o", str(exe)])
        r = _run([str(exe)], env=ENV, timeout=300)
        assert "sanitized run complete" in r.stdout
        assert "validator: " in r.stdout


def test_interpreter_under_sanitizers():
    with tempfile.TemporaryDirectory() as d:
        exe = Path(d) / "test_interp"
        _run(["g++", "-std=c++17", *SAN, f"-I{ROOT / 'umi_amd' / 'src'}", f"-I{ROOT / 'oracle'}",
              str(ROOT / "tests" / "cpp_host" / "test_interp.cpp"), "-o", str(exe)])
        r = _run([str(exe)], env=ENV, timeout=300)
        assert "all known answers match" in r.stdout
