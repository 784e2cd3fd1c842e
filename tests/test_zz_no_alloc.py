"""clay_splat_bake_static (no allocation) against clay_splat_bake (plain schedule).

The static bake keeps its code grid in a device array that every call on the GPU
shares, so two calls run back to back here, the second with fewer particles and a
smaller volume: each must write exactly the f16 distance planes that a SplatField
of the same cloud and options writes, bit for bit, at every mip.
"""
import ctypes as C

import numpy as np
import pytest

import clayhip
from clayhip import scenes

pytestmark = pytest.mark.gpu


def _planes_of(sf, mips):
    return np.concatenate([sf.volume.download(m)[..., 0].reshape(-1) for m in range(mips)])


def _case(root, pos, opts, prefix=None):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(pos.T)).cuda()
    live = ("prefix", prefix) if prefix is not None else None
    cloud = clayhip.ParticleCloud(pos.shape[0], (t[0], t[1], t[2]), live=live)
    res, mips = opts["resolution"], opts["mipLevels"]
    texels = sum((res >> m) ** 3 for m in range(mips))
    # filled with a pattern no bake writes, so an unwritten texel shows
    out = torch.full((texels,), 0x7E00, dtype=torch.int16, device="cuda")
    clayhip.bake_static(root, cloud, opts, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint16)
    ref = clayhip.SplatField(root, cloud, opts, mode="plain")
    ref.bake()
    want = _planes_of(ref, mips)
    ref.destroy()
    return got, want, t


def test_static_bake_equals_the_allocating_bake_twice(root):
    big = scenes.lattice(16, 64, 16, 0.05, (-0.4, 0.0, -0.4), seed=3)
    opts_big = dict(radius=0.05, resolution=64, mipLevels=4, band=4, worldSize=24, origin=(-12, -4, -12))
    got, want, _ = _case(root, big, opts_big)
    assert got.shape == want.shape
    assert np.array_equal(got, want), f"first call: {(got != want).sum()} texels differ"
    assert (got != np.uint16(0x3C00)).any(), "first call wrote no surface"
    rng = np.random.default_rng(5)
    small = rng.uniform(-1.5, 1.5, size=(3000, 3)).astype(np.float32)
    opts_small = dict(radius=0.12, resolution=32, mipLevels=3, band=2, worldSize=8, origin=(-4, -4, -4))
    got, want, _ = _case(root, small, opts_small, prefix=2500)
    assert got.shape == want.shape
    assert np.array_equal(got, want), f"second call: {(got != want).sum()} texels differ"


def test_static_bake_refuses_what_does_not_fit(root):
    import torch

    t = torch.zeros((3, 16), dtype=torch.float32, device="cuda")
    cloud = clayhip.ParticleCloud(16, (t[0], t[1], t[2]))
    out = torch.zeros((256 ** 3 + 128 ** 3 + 64 ** 3,), dtype=torch.int16, device="cuda")
    with pytest.raises(clayhip.ClayError, match="static scratch holds"):
        clayhip.bake_static(root, cloud, dict(radius=0.05, resolution=256, mipLevels=3), out)
    with pytest.raises(clayhip.ClayError, match="plain schedule"):
        clayhip._lib.check(clayhip.lib().clay_splat_bake_static(
            root.handle, C.byref(cloud.desc), C.byref(_tiled_desc()), C.c_void_p(int(out.data_ptr())), None))


def _tiled_desc():
    from clayhip.splat import splat_desc

    return splat_desc(dict(radius=0.05, resolution=64, mipLevels=4, mode="tiled"))
