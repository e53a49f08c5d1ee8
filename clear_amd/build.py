"""Build recipe for libclear_amd.so (gfx950) and the CPU oracle library.

`build_all()` always recompiles: every HIP source under clear_amd/csrc is
compiled for gfx950 with hipcc and linked into clear_amd/libclear_amd.so; the
oracle's C++ sources are compiled with g++ into oracle/liboracle.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "clear_amd" / "csrc"
LIB = ROOT / "clear_amd" / "libclear_amd.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
OBJ_DIR = ROOT / "clear_amd" / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off keeps every a*b+c as two rounded operations, as the CPU
# oracle computes them; IEEE division/sqrt are requested explicitly.
# Packed FP32 instructions (v_pk_*_f32) are disabled: some of their forms
# have given wrong results on this hardware without any error.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]
ORACLE_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                "-fno-fast-math", "-Wall"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        l for l in proc.stdout.splitlines()
        if "packed-fp32-ops" not in l and "not a recognized feature" not in l
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip() and os.environ.get("CLR_BUILD_VERBOSE"):
        sys.stderr.write(out + "\n")


def build_hip(jobs: int = 8) -> Path:
    """Compiles every .hip/.cpp under csrc for gfx950 and links the shared library."""
    if not Path(HIPCC).exists() and shutil.which(HIPCC) is None:
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    sources = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    objs = []
    cmds = []
    for src in sources:
        obj = OBJ_DIR / (src.name + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.suffix == ".hip" else []
        cmds.append([HIPCC, *HIP_FLAGS, *lang, "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lz"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    """Compiles the CPU oracle (test infrastructure) with g++."""
    sources = sorted(ORACLE_DIR.glob("*.cpp"))
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["g++", *ORACLE_FLAGS, "-o", str(tmp), *map(str, sources), "-lpthread"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


CPP_DRIVER_SRC = ROOT / "tests" / "cpp" / "wrapper_driver.cpp"
CPP_DRIVER = ROOT / "tests" / "cpp" / "wrapper_driver"


def build_cpp_driver() -> Path:
    """Host-only C++ driver of the include/clear_amd.hpp wrappers (run by a
    GPU test), linked against libclear_amd.so by a relative rpath."""
    rocm = Path(HIPCC).resolve().parent.parent
    tmp = CPP_DRIVER.with_suffix(".tmp")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{rocm / 'include'}",
          f"-I{ROOT / 'include'}", str(CPP_DRIVER_SRC), "-o", str(tmp), f"-L{LIB.parent}", "-lclear_amd",
          f"-L{rocm / 'lib'}", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../clear_amd", f"-Wl,-rpath,{rocm / 'lib'}"])
    os.replace(tmp, CPP_DRIVER)
    return CPP_DRIVER


def build_all(jobs: int = 8) -> None:
    build_hip(jobs)
    build_oracle()
    build_cpp_driver()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB} and {ORACLE_LIB}")
