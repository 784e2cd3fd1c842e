"""The cell sort at 9.2M particles, where the radix sort switches to 8192-key tiles.

The replacement of the reference's insert pass (fluid.ts:308-331) must give the
oracle's stable (key, id) order at any size.  The 16k parity test covers the
2048-key tiles; here a 480 x 40 x 480 lattice at spacing 0.05 (9 216 000
particles, above the 2^23 switch) is sorted once by a substep, and the sorted keys
and the permutation are compared with the oracle's cell keys of the same positions
and its stable sort, bit for bit.
"""
import numpy as np
import pytest

import clayhip
import oracle as O
from clayhip import scenes

pytestmark = pytest.mark.gpu


def test_sort_above_the_big_tile_switch(root):
    pos = scenes.lattice(480, 40, 480, 0.05, (-12.0, 0.0, -12.0), seed=3)
    n = pos.shape[0]
    assert n == 9_216_000 and n >= 1 << 23
    opts = scenes.dam_break_options(n, 64, 1)
    world = clayhip.analyticField(root, [scenes.floor_brush()])
    sim = clayhip.FluidSim(root, world, opts)
    sim.set_state(pos)
    sim.substep()
    perm = sim.debug("perm")
    skeys = sim.debug("keys")
    raw = O.keys(opts, pos)
    g = O.fluid_consts(opts).grid_res ** 3
    want_keys = np.where(raw == O.INVALID_KEY, g, raw).astype(np.uint32)
    want_perm = O.stable_sort(want_keys)
    assert np.array_equal(skeys, want_keys[want_perm]), "sorted keys differ"
    assert np.array_equal(perm, want_perm), "stable (key, id) permutation differs"
    # Many particles share a cell, so stability is exercised: equal keys, ascending ids.
    same = skeys[1:] == skeys[:-1]
    assert same.sum() > n // 2 and np.all(perm[1:][same] > perm[:-1][same])
    sim.destroy()
    world.destroy()
