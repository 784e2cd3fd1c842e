"""One FluidSim.step driven from Node through the N-API binding, checked bit for
bit against the oracle: positions, contacts and every mip of the surface codes."""
import importlib.util
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from clayhip import make_brush, scenes

ROOT = Path(__file__).resolve().parent.parent
_spec = importlib.util.spec_from_file_location("clay_napi_build", ROOT / "napi" / "build.py")
napi_build = importlib.util.module_from_spec(_spec)  # napi/build.py
_spec.loader.exec_module(napi_build)

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not napi_build.ADDON.exists(), reason="N-API addon not built")]


def test_fluidsim_step_from_node_matches_the_oracle(tmp_path):
    pos = scenes.lattice(16, 32, 16, 0.05, (-0.4, 0.0, -0.4), seed=2)
    n = pos.shape[0]
    opts = scenes.dam_break_options(n, 64, 4)
    brushes = [scenes.floor_brush()]
    (tmp_path / "pos.f32").write_bytes(np.ascontiguousarray(pos.T).astype(np.float32).tobytes())
    case = {"options": dict(opts, origin=list(opts["origin"]), gravity=list(opts["gravity"]), splatMode="tiled"),
            "brushes": [dict(b, size=list(b["size"]), pos=list(b["pos"])) for b in brushes],
            "n": n, "pos": str(tmp_path / "pos.f32")}
    (tmp_path / "case.json").write_text(json.dumps(case))
    print(f"node FluidSim step: {n} particles", flush=True)
    res = subprocess.run(["node", str(ROOT / "napi" / "test_gpu.js"), str(tmp_path / "case.json"), str(tmp_path)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    got = np.frombuffer((tmp_path / "pos.f32").read_bytes(), dtype=np.float32).reshape(3, n).T
    orc = O.Fluid(opts, pos, collider=O.Collider().add_analytic([make_brush(b) for b in brushes]))
    for _ in range(opts["substeps"]):
        orc.substep()
    want = orc.get("pos").astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    result = json.loads((tmp_path / "result.json").read_text())
    assert result["particleCount"] == n
    raw, pts = orc.read_contacts()
    assert sorted(tuple(p) for p in result["contacts"]) == sorted(pts)
    grids = O.splat_bake(O.splat_opts(64, 4, opts["worldSize"], opts["origin"], opts["bakeBand"],
                                      opts["surfaceRadius"]), want)
    for m in range(4):
        codes = np.frombuffer((tmp_path / f"codes{m}.u32").read_bytes(), dtype=np.uint32)
        assert np.array_equal(codes, grids[m].reshape(-1)), f"mip {m}"
