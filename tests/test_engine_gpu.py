"""The generic UMI::Execute<> engine on the GPU.

* tests/cpp/bin/test_engine: strategy structs run by the engine and by the CPU
  interpreter (oracle/umi_interp.h), every output compared;
* tests/dropin/bin/<name>_gpu: the reference's own example programs (main/*.cu)
  compiled unchanged against this project's UMI.h / UMI_gpu.h /
  SyncedMemory.h; each asserts its results with its own CPU loop.
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
EXAMPLES = ["filter3x3", "wfilter3x3", "bilateral", "iimg", "full_me", "cost_volume", "fprop", "gemm"]
pytestmark = pytest.mark.gpu


def run(exe: Path, timeout=180):
    print(f"[run] {exe.name}", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    return r


def test_engine_vs_interpreter(gpu):
    exe = ROOT / "tests" / "cpp" / "bin" / "test_engine"
    assert exe.exists(), "tests/cpp/bin/test_engine not built (run __graft_entry__.build())"
    r = run(exe)
    assert r.returncode == 0, r.stdout
    assert "all cases passed" in r.stdout


@pytest.mark.parametrize("name", EXAMPLES)
def test_reference_example_drop_in(gpu, name):
    exe = ROOT / "tests" / "dropin" / "bin" / f"{name}_gpu"
    if not exe.exists():
        pytest.skip("drop-in build needs the reference sources at build time")
    r = run(exe)
    assert r.returncode == 0, f"{name}_gpu exited {r.returncode}"
