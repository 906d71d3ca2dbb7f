"""Build libcxg.so (the HIP library behind include/cxg.h) in-tree for gfx950.

    python -m calculixpp_amd.build          # build calculixpp_amd/libcxg.so

Every source is compiled with hipcc for gfx950 only; FP64 throughout. Packed FP32
instructions are disabled (the library computes in FP64; the flag keeps hipcc from
forming packed FP32 arithmetic anywhere in the library).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libcxg.so")
OBJ = os.path.join(ROOT, "_build", "cxg")

SOURCES = ["cxg_tables.cpp", "cxg_api.cpp", "cxg_graph.cpp", "cxg_pattern.hip", "cxg_elem.hip",
           "cxg_solve.hip"]
ARCH = os.environ.get("CXG_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return [
        "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", INCLUDE, "-I", CSRC, "-Wall",
    ]


def _compile(src: str) -> str:
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [hipcc()] + lang + flags() + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        sys.stderr.write(_filter_warnings(r.stderr))
    return out


def _filter_warnings(text: str) -> str:
    # hipcc's host pass does not know the packed-fp32 target feature; harmless.
    keep = [l for l in text.splitlines(True) if "packed-fp32-ops" not in l
            and "not a recognized feature" not in l]
    return "".join(keep)


def build(verbose: bool = False) -> str:
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
