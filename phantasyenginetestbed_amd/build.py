"""Builds libphx.so (host C++ + HIP kernels for gfx950) in-tree.

Usage: python -m phantasyenginetestbed_amd.build   (or build() from __graft_entry__)

Every HIP translation unit is compiled with
  -ffp-contract=off                          no a*b+c fusion: the traversal's float results must
                                             equal the CPU reference's operation by operation
  -fhip-fp32-correctly-rounded-divide-sqrt   1/x and sqrt are IEEE, as on the CPU
  -fno-gpu-flush-denormals-to-zero           denormals kept, as on the CPU
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libphx.so"
CHECK = PKG / "libhiptracer_check.so"   # engine-side check of the C++ renderer interface
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-pthread",
    "-Wall", "-Wno-unused-function", f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [str(ROCM / "bin" / "hipcc"), *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = True) -> Path:
    """Compiles every source under csrc/ and links libphx.so. Always rebuilds."""
    if not (ROCM / "bin" / "hipcc").exists():
        raise RuntimeError(f"hipcc not found under {ROCM}")
    BUILD.mkdir(exist_ok=True)
    sources = sorted(p for p in CSRC.iterdir() if p.suffix in (".cpp", ".hip"))
    jobs = min(len(sources), max(1, min(16, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([str(ROCM / "bin" / "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC",
          *map(str, objs), "-o", str(tmp), "-lpthread"])
    os.replace(tmp, LIB)
    # The C++ renderer interface over the C ABI and its engine-side check program.
    check_tmp = CHECK.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", "-I", str(ROOT / "include"),
          str(ROOT / "tests" / "cpp" / "hiptracer_check.cpp"),
          str(PKG / "cpp" / "HipTracerRenderer.cpp"), "-L", str(PKG), "-lphx",
          "-Wl,-rpath,$ORIGIN", "-o", str(check_tmp)])
    os.replace(check_tmp, CHECK)
    if verbose:
        print(f"built {LIB.relative_to(ROOT)} from {len(sources)} sources and "
              f"{CHECK.relative_to(ROOT)}", flush=True)
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    LIB.unlink(missing_ok=True)
    CHECK.unlink(missing_ok=True)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "clean":
        clean()
    else:
        build()
