"""Recipe for tests/dropin/bin/: the reference's own example programs
(main/*.cu, read where they lie under /root/reference) compiled unchanged
against this project's UMI headers (umi_amd/src) and library -- the drop-in
test of the C++ layer.

TEST INFRASTRUCTURE ONLY.  Each program checks its results with its own CPU
loop; tests/test_engine_gpu.py runs them on the GPU.  main/bilateral.cu calls
two CUDA runtime functions outside the UMI API (bilateral.cu:101); the recipe
maps them to their HIP equivalents on the command line.  Outputs go only to
tests/dropin/bin/ (kept out of version control; they travel to the GPU box
with the tree).  Nothing is built when /root/reference is absent.

    python tests/dropin/build_dropin.py
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REF_MAIN = Path("/root/reference/main")
OUT = HERE / "bin"
LIB = ROOT / "umi_amd" / "lib" / "libumi_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-x", "hip", f"--offload-arch={os.environ.get('UMI_OFFLOAD_ARCH', 'gfx950')}", "-O2", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-DcudaGetLastError=hipGetLastError", "-DcudaGetErrorString=hipGetErrorString", "-UNDEBUG", "-w"]


def _one(src: Path) -> Path:
    out = OUT / f"{src.stem}_gpu"
    rpath = os.path.relpath(LIB.parent, OUT)
    cmd = [HIPCC, *FLAGS, f"-I{ROOT / 'umi_amd' / 'src'}", f"-I{ROOT / 'include'}", str(src), "-o", str(out),
           f"-L{LIB.parent}", "-lumi_hip", f"-Wl,-rpath,$ORIGIN/{rpath}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build() -> list[Path]:
    if not REF_MAIN.is_dir():
        return []
    if not LIB.exists():
        raise RuntimeError(f"{LIB} is missing: build the library first")
    OUT.mkdir(parents=True, exist_ok=True)
    srcs = sorted(REF_MAIN.glob("*.cu"))
    with cf.ThreadPoolExecutor(max_workers=max(1, len(srcs))) as ex:
        return list(ex.map(_one, srcs))


if __name__ == "__main__":
    for p in build():
        print(p)
    sys.exit(0)
