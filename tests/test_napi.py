"""The N-API binding (napi/) on the CPU: its host-only entry points called from
Node, and the JS BrushSet.make against the Python one, byte for byte."""
import importlib.util
import json
import subprocess
from pathlib import Path

import pytest

from clayhip import make_brush

ROOT = Path(__file__).resolve().parent.parent
_spec = importlib.util.spec_from_file_location("clay_napi_build", ROOT / "napi" / "build.py")
napi_build = importlib.util.module_from_spec(_spec)  # napi/build.py
_spec.loader.exec_module(napi_build)

pytestmark = pytest.mark.skipif(not napi_build.available(), reason="node, node-gyp or node_api.h missing")

BRUSHES = [
    {"kind": "roundBox", "size": [11, 1, 11], "radius": 0.25, "pos": [0, -1.4, 0]},
    {"kind": "box", "size": [9.5, 1.4, 0.9], "pos": [0, 0.4, 0]},
    {"kind": "sphere", "size": 2.0, "pos": [2.5, 0.8, 0], "op": "cut"},
    {"kind": "capsule", "size": [0.3, 0.8, 0.3], "radius": 0.2, "euler": [0.3, -1.1, 2.0], "material": 3,
     "smooth": 0.05, "onlyMaterial": 2, "slot": 4},
    {"kind": "plane", "radius": -0.5, "op": "paint"},
]


@pytest.fixture(scope="module")
def addon():
    if not napi_build.ADDON.exists():
        napi_build.build()
    return napi_build.ADDON


def test_host_entry_points_and_brush_records(addon, tmp_path):
    spec = tmp_path / "brushes.json"
    spec.write_text(json.dumps(BRUSHES))
    res = subprocess.run(["node", str(ROOT / "napi" / "test_host.js"), str(spec)], capture_output=True, text=True,
                         timeout=60)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert out["ok"]
    for desc, hexrec in zip(BRUSHES, out["records"]):
        want = bytes(make_brush(desc)).hex()
        assert hexrec == want, desc["kind"]
