"""Build libams.so (HIP for gfx950) in-tree with hipcc.

The library is plain C-ABI (include/ams.h); Python reaches it with ctypes.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libams.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["ams_capi.cpp", "ams_bvh.cpp", "ams_trace.hip", "ams_contract.hip", "ams_objective.hip", "ams_phasemap.hip",
           "ams_surrogate.hip", "ams_probe.hip"]

# -packed-fp32-ops: keep the compiler from emitting v_pk_{add,mul,fma}_f32
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False, jobs: int = 4) -> Path:
    objdir = ROOT / "_build"
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = objdir / (src + ".o")
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", str(path), "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-x", "hip", *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", str(path), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs)
            procs = []
    _wait(procs)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if out.returncode != 0:
        raise RuntimeError("link of libams.so failed:\n" + out.stdout)
    os.replace(tmp, LIB)
    return LIB


def _wait(procs):
    failed = []
    for name, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(f"--- {name} ---\n{text}")
        elif "error" in text.lower():
            print(text, file=sys.stderr)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(failed))


if __name__ == "__main__":
    print(build(verbose=True))
