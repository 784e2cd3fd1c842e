"""Builds the N-API addon (napi/build/Release/clay_hip.node) with node-gyp against
the system Node headers (no download).  It links clay-engine_amd/libclay_hip.so,
which must be built first.

Usage: python napi/build.py
"""
from __future__ import annotations

import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ADDON = HERE / "build" / "Release" / "clay_hip.node"


def available() -> bool:
    return shutil.which("node") is not None and shutil.which("node-gyp") is not None and \
        Path("/usr/include/node/node_api.h").exists()


def build() -> Path:
    res = subprocess.run(["node-gyp", "rebuild", "--nodedir=/usr"], cwd=HERE, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"node-gyp failed:\n{res.stdout[-3000:]}\n{res.stderr[-3000:]}")
    return ADDON


if __name__ == "__main__":
    print(build() if available() else "node / node-gyp / node_api.h not found", file=sys.stderr)
