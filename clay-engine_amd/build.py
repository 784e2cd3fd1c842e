"""Builds libclay_hip.so (the C ABI in include/clay_hip.h) for gfx950 with hipcc.

Every kernel source is compiled with contraction off (-ffp-contract=off): the
plain path evaluates the reference's f32 expressions with one rounding per
operation, which is what makes it bit-exact against an IEEE-f32 evaluation.
Packed FP32 instructions are disabled (-packed-fp32-ops).

Usage: python clay-engine_amd/build.py [-v]
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
OBJ = PKG / "_build"
LIB = PKG / "libclay_hip.so"
SOURCES = ["api.hip", "peaks.hip", "edit.hip", "fluid.hip", "pbd.hip", "query.hip", "sort.hip", "splat.hip", "volume.hip", "wire.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (looked at $HIPCC, /opt/rocm/bin/hipcc, PATH)")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
        f"-I{ROOT / 'include'}",
        f"-I{CSRC}",
    ]


def _compile(src: str, verbose: bool) -> Path:
    out = OBJ / (Path(src).stem + ".o")
    cmd = [hipcc(), *flags(), "-c", str(CSRC / src), "-o", str(out)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stdout}\n{res.stderr}")
    noise = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if verbose and noise:
        print("\n".join(noise), flush=True)
    return out


def build(verbose: bool = False) -> Path:
    """Compiles every HIP source for gfx950 and links libclay_hip.so in-tree."""
    OBJ.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
