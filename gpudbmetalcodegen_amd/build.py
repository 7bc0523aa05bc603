"""Build the native libraries in-tree.

libgdbhip.so    — the C-ABI over HIP/hipRTC/hipGraph plus the fixed kernels
                  (hipcc, gfx950, packed-FP32 instructions disabled).
libgdbengine.so — the C++ engine (plan builder, emitters, runner, collector,
                  generators); plain g++, it reaches the GPU only through
                  libgdbhip.so.

Run ``python -m gpudbmetalcodegen_amd.build`` (or ``__graft_entry__.build()``).
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
BUILD = ROOT / "_build"
INCLUDE = ROOT / "include"

ARCH = os.environ.get("GDB_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}"]

HIP_SOURCES = [
    CSRC / "hipabi" / "gdb_hip.cpp",
    CSRC / "hipabi" / "gdb_kernels.cpp",
]
ENGINE_DIRS = ["core", "operators", "planning", "execution", "infra", "runtime"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _obj(src: Path) -> Path:
    rel = src.relative_to(CSRC)
    return BUILD / (str(rel).replace(os.sep, "__") + ".o")


def _stale(src: Path, obj: Path, deps: list[Path]) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(d.stat().st_mtime > t for d in [src, *deps])


def _headers() -> list[Path]:
    hs = list(CSRC.rglob("*.h")) + list(INCLUDE.rglob("*.h"))
    return hs


def build(verbose: bool = False, jobs: int | None = None, force: bool = False) -> dict[str, Path]:
    BUILD.mkdir(exist_ok=True)
    headers = _headers()
    hip_srcs = [s for s in HIP_SOURCES if s.exists()]
    eng_srcs = sorted(
        p for d in ENGINE_DIRS if (CSRC / d).exists() for p in (CSRC / d).glob("*.cpp")
    ) + [CSRC / "engine_capi.cpp"]

    jobs_list = []
    for s in hip_srcs:
        o = _obj(s)
        if force or _stale(s, o, headers):
            jobs_list.append([HIPCC, "-x", "hip", "-c", str(s), "-o", str(o), *HIP_FLAGS])
    for s in eng_srcs:
        o = _obj(s)
        if force or _stale(s, o, headers):
            jobs_list.append([CXX, "-c", str(s), "-o", str(o), *CXX_FLAGS])
    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 4)) as pool:
        for cmd, fut in [(c, pool.submit(_run, c)) for c in jobs_list]:
            fut.result()
            if verbose:
                print("built", cmd[cmd.index("-o") + 1], flush=True)

    hip_lib = PKG / "libgdbhip.so"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(hip_lib),
          *[str(_obj(s)) for s in hip_srcs], "-lhiprtc"])
    eng_lib = PKG / "libgdbengine.so"
    _run([CXX, "-shared", "-fPIC", "-o", str(eng_lib), *[str(_obj(s)) for s in eng_srcs],
          f"-L{PKG}", "-lgdbhip", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    return {"hip": hip_lib, "engine": eng_lib}


if __name__ == "__main__":
    out = build(verbose=True, force="--force" in sys.argv)
    for k, v in out.items():
        print(k, v)
