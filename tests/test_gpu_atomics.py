"""The splat's atomic counts and the measured atomic peaks.

The tallying re-bake must leave the codes exactly as the plain bake of the same
state wrote them, and count at least one LDS min per voxel the walk lowered (every
voxel below the saturated code got its value from some particle's ds_min_u32).
"""
import math

import numpy as np
import pytest

import clayhip
from clayhip import scenes

pytestmark = pytest.mark.gpu


def test_tallying_bake_keeps_codes_and_counts_every_lowered_voxel(root):
    sc = scenes.dam_break_small()
    world = clayhip.analyticField(root, sc["collider"]["brushes"])
    sim = clayhip.FluidSim(root, world, dict(sc["options"], splatMode="tiled"))
    sim.set_state(sc["pos"])
    for _ in range(3):
        sim.step()
    mips = sc["options"]["bakeMipLevels"]
    before = [sim.surface.codes(m) for m in range(mips)]
    lds, glob = sim.splat_atomics()
    after = [sim.surface.codes(m) for m in range(mips)]
    for m in range(mips):
        assert np.array_equal(before[m], after[m]), f"mip {m}"
    lowered = sum(int((c < 65535).sum()) for c in after)
    assert lowered > 0
    assert lds >= lowered and glob >= lowered
    assert lds < 2000 * sc["pos"].shape[0]  # about 1 100 per particle per bake in the reference's walk
    sim.destroy()
    world.destroy()


def test_splat_atomics_need_a_step(root):
    sc = scenes.dam_break_small()
    world = clayhip.analyticField(root, sc["collider"]["brushes"])
    sim = clayhip.FluidSim(root, world, dict(sc["options"], splatMode="tiled"))
    sim.set_state(sc["pos"])
    with pytest.raises(clayhip.ClayError, match="after a step"):
        sim.splat_atomics()
    sim.destroy()
    world.destroy()


@pytest.mark.parametrize("kind", ["lds_min", "global_min"])
def test_atomic_peaks_are_measured(root, kind):
    rate = root.atomic_peak(kind, rounds=1024)
    assert math.isfinite(rate) and 1e9 < rate < 1e15


def test_valu_peak_is_measured(root):
    rates = {f: root.valu_peak(f, rounds=4096) for f in root.VALU_FORMS}
    for f, rate in rates.items():
        # 256 CUs at 1-3 GHz retire between ~1e11 and a few 1e12 wave64 instructions per second.
        assert math.isfinite(rate) and 1e11 < rate < 1e13, (f, rate)
    print({f: f"{r:.3e}" for f, r in rates.items()})
    # gfx950 issues the two-operand all-VGPR forms at up to about twice the rate of an
    # operation with an SGPR operand or a compare/select/max form, and the
    # transcendental unit at a quarter of it.
    full = max(rates["fmac"], rates["add_u32"], rates["mul_f32"])
    for half in ("fmac_sgpr", "max", "cndmask", "fma_mix"):
        assert full > 1.25 * rates[half], (half, rates)
    assert full > 2.5 * rates["rsq"]
    with pytest.raises(clayhip.ClayError, match="rounds"):
        root.valu_peak(rounds=0)
    with pytest.raises(ValueError, match="unknown VALU form"):
        root.valu_peak("v_pk_fma")
