"""Pins the CPU interpreter of UMI::Execute (oracle/umi_interp.h) on the host.

Compiles tests/cpp_host/test_interp.cpp with g++ (no GPU) and runs it: the
interpreter must reproduce the reference examples' known answers.
"""
import subprocess
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_interpreter_known_answers():
    with tempfile.TemporaryDirectory() as d:
        exe = Path(d) / "test_interp"
        cmd = ["g++", "-std=c++17", "-O1", f"-I{ROOT / 'umi_amd' / 'src'}", f"-I{ROOT / 'oracle'}",
               str(ROOT / "tests" / "cpp_host" / "test_interp.cpp"), "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout
        assert "all known answers match" in r.stdout
