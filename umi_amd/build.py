"""Build recipe for libumi_hip.so (hipcc, gfx950) and the CPU oracle.

Compiles every ``umi_amd/csrc/*.hip`` to an object with hipcc for gfx950 and
links them into ``umi_amd/lib/libumi_hip.so``.  Objects go to
``umi_amd/_build/``; a source is recompiled when it or a header changed,
unless ``force`` is set.  ``python -m umi_amd.build`` runs it.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "lib" / "libumi_hip.so"
ARCH = os.environ.get("UMI_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    # packed-fp32 VALU forms are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


def _headers_mtime() -> float:
    hs = list(CSRC.glob("*.h")) + list((ROOT / "include").glob("*.h"))
    return max((h.stat().st_mtime for h in hs), default=0.0)


def _compile(src: Path, force: bool) -> Path:
    obj = OBJ / (src.stem + ".o")
    if not force and obj.exists() and obj.stat().st_mtime >= max(src.stat().st_mtime, _headers_mtime()):
        return obj
    cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(force: bool = True, jobs: int | None = None) -> Path:
    """Compile all HIP sources and link libumi_hip.so; returns its path."""
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


TEST_CPP = ROOT / "tests" / "cpp"


def _hipcc_program(src: Path, out: Path, extra=()) -> None:
    out.parent.mkdir(parents=True, exist_ok=True)
    rpath = os.path.relpath(LIB.parent, out.parent)
    cmd = [HIPCC, "-x", "hip", *HIP_FLAGS[:2], "-O2", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", f"-I{PKG / 'src'}", f"-I{ROOT / 'include'}", f"-I{ROOT / 'oracle'}", *extra,
           str(src), "-o", str(out), f"-L{LIB.parent}", "-lumi_hip", f"-Wl,-rpath,$ORIGIN/{rpath}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")


def build_tests() -> list[Path]:
    """C++ test programs of the generic engine (tests/cpp/*.cpp -> tests/cpp/bin/)."""
    outs = []
    srcs = sorted(TEST_CPP.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=max(1, len(srcs))) as ex:
        for src, _ in zip(srcs, ex.map(lambda s: _hipcc_program(s, TEST_CPP / "bin" / s.stem), srcs)):
            outs.append(TEST_CPP / "bin" / src.stem)
    return outs


def build_dropin() -> list[Path]:
    """The reference's example programs against this library (tests/dropin/build_dropin.py)."""
    sys.path.insert(0, str(ROOT / "tests" / "dropin"))
    try:
        import build_dropin as recipe
    finally:
        sys.path.pop(0)
    return recipe.build()


def build_oracle() -> Path:
    """Build the CPU oracle (test infrastructure) with its Makefile."""
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return ROOT / "oracle" / "lib" / "liboracle.so"


if __name__ == "__main__":
    lib = build_library(force="--force" in sys.argv)
    print(lib)
    print(build_oracle())
    for p in build_tests() + build_dropin():
        print(p)
