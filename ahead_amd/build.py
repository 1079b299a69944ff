"""Builds libahead_hip.so (the gfx950 AN operators behind include/ahead_hip.h) in-tree.

Usage: python -m ahead_amd.build [--jobs N]

Every .hip file under ahead_amd/csrc/ is compiled for gfx950 only and linked into
ahead_amd/lib/libahead_hip.so.  The CPU oracle (oracle/) is built separately by its own Makefile
and is not part of this library.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libahead_hip.so"
ROOT = PKG.parent

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # no packed fp32 instructions anywhere (none of these kernels use floating point)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


CXX = os.environ.get("CXX", "g++")
CPP_TESTS = ROOT / "tests" / "cpp"


def build_cpp_tests(verbose: bool = True) -> list[Path]:
    """Compiles the C++ programs under tests/cpp/ (clients of include/ahead/ and the C ABI, checked
    against the oracle) into tests/cpp/_build/.  Needs libahead_hip.so and the oracle built."""
    out_dir = CPP_TESTS / "_build"
    out_dir.mkdir(exist_ok=True)
    outs = []
    for src in sorted(CPP_TESTS.glob("*.cpp")):
        exe = out_dir / src.stem
        cmd = [CXX, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
               "-I/opt/rocm/include", f"-I{ROOT / 'include'}", f"-I{ROOT / 'include' / 'ahead'}", str(src),
               "-o", str(exe), f"-L{LIB_DIR}", "-lahead_hip", f"-L{ROOT / 'oracle' / '_build'}", "-loracle",
               "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../../ahead_amd/lib",
               "-Wl,-rpath,$ORIGIN/../../../oracle/_build", "-Wl,-rpath,/opt/rocm/lib"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"g++ failed for {src.name}:\n{res.stdout}\n{res.stderr}")
        outs.append(exe)
    if verbose:
        print(f"built {len(outs)} C++ test program(s) in {out_dir}", file=sys.stderr)
    return outs


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args()
    build(args.jobs)


if __name__ == "__main__":
    main()
