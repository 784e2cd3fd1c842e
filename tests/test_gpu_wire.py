"""The baked volume's stream on the GPU against the numpy restatement, byte for byte.

pack of a baked world equals oracle.wire_pack of the same texels (every byte:
header, table, bitmap, payload) at 64^3 x 3 and at the full 288^3 x 3 world;
unpack restores every mip exactly (into a fresh volume, and into one instance of a
multi-instance volume); mismatched, truncated and corrupted streams are refused.
"""
import numpy as np
import pytest

import clayhip
import oracle as O
from clayhip import _lib as L
from clayhip import scenes

pytestmark = pytest.mark.gpu


def texels(vol, mips, instance=None):
    dl = [vol.download(m, instance=instance) for m in range(mips)]
    return [t[..., 0] for t in dl], [t[..., 1] for t in dl]


def baked(root, res, mips, world, origin, instances=1):
    v = clayhip.SdfVolume(root, resolution=res, mipLevels=mips, worldSize=world, origin=origin, band=4,
                          instances=instances)
    v.bakeBrushes([dict(b, material=float(i + 1)) for i, b in enumerate(scenes.ridge_world_brushes())])
    return v


@pytest.mark.parametrize("res,world,origin", [(64, 8.0, (-4.0, -2.0, -4.0)),
                                              (288, scenes.WORLD, scenes.ORIGIN)])
def test_pack_equals_the_oracle_and_unpack_restores(root, res, world, origin):
    v = baked(root, res, 3, world, origin)
    dist, mat = texels(v, 3)
    want = O.wire_pack(dist, mat, 4.0, world, origin)
    got = v.pack()
    assert len(got) == len(want), f"{len(got)} vs {len(want)} bytes"
    if got != want:
        a, b = np.frombuffer(got, np.uint8), np.frombuffer(want, np.uint8)
        first = int(np.flatnonzero(a != b)[0])
        pytest.fail(f"streams differ first at byte {first} of {len(want)}")
    h = np.frombuffer(got[:96], dtype=O.WIRE_HEADER)[0]
    assert 0 < h["stored_tiles"] < h["tiles"]
    fresh = clayhip.SdfVolume(root, resolution=res, mipLevels=3, worldSize=world, origin=origin, band=4)
    fresh.unpack(got)
    d2, m2 = texels(fresh, 3)
    for m in range(3):
        assert np.array_equal(d2[m], dist[m]) and np.array_equal(m2[m], mat[m]), f"mip {m}"
    fresh.destroy()
    v.destroy()


def test_instances_round_trip(root):
    src = baked(root, 64, 3, 8.0, (-4.0, -2.0, -4.0))
    data = src.pack()
    multi = clayhip.SdfVolume(root, resolution=64, mipLevels=3, worldSize=8.0, origin=(-4.0, -2.0, -4.0), band=4,
                              instances=3)
    multi.unpack(data, instance=1)
    dist, mat = texels(src, 3)
    d1, m1 = texels(multi, 3, instance=1)
    for m in range(3):
        assert np.array_equal(d1[m], dist[m]) and np.array_equal(m1[m], mat[m])
    d0, _ = texels(multi, 3, instance=0)
    assert not np.array_equal(d0[0], dist[0]), "instance 0 was overwritten"
    assert multi.pack(instance=1) == data
    multi.destroy()
    src.destroy()


def test_bad_streams_are_refused(root):
    src = baked(root, 64, 3, 8.0, (-4.0, -2.0, -4.0))
    data = src.pack()
    other = clayhip.SdfVolume(root, resolution=32, mipLevels=2, worldSize=8.0, origin=(-4.0, -2.0, -4.0), band=4)
    with pytest.raises(L.ClayError, match="does not fit"):
        other.unpack(data)
    with pytest.raises(L.ClayError, match="truncated"):
        src.unpack(data[:-1])
    bad = bytearray(data)
    bad[0] ^= 1
    with pytest.raises(L.ClayError, match="bad magic"):
        src.unpack(bytes(bad))
    h = np.frombuffer(data[:96], dtype=O.WIRE_HEADER)[0]
    words = np.frombuffer(data, "<u4", (int(h["tiles"]) + 31) // 32, int(h["bitmap_offset"]))
    stored = int(np.flatnonzero(((words[:, None] >> np.arange(32, dtype=np.uint32)) & 1).ravel())[0])
    bad = bytearray(data)
    at = int(h["table_offset"]) + 4 * stored
    bad[at:at + 4] = np.uint32(h["stored_tiles"]).tobytes()
    with pytest.raises(L.ClayError, match="outside the payload"):
        src.unpack(bytes(bad))
    # Asynchronous form: the caller's header, no synchronisation; the flag reports later.
    import torch

    hdr = L.WireHeader.from_buffer_copy(data[:96])
    dev = torch.frombuffer(bytearray(bad), dtype=torch.uint8).cuda()
    src.unpack_device(dev, len(bad), header=hdr)
    with pytest.raises(L.ClayError, match="outside the payload"):
        src.wireStatus()
    assert src.wireStatus() == 0                      # cleared once reported
    good = torch.frombuffer(bytearray(data), dtype=torch.uint8).cuda()
    src.unpack_device(good, len(data), header=hdr)
    assert src.wireStatus() == 0
    other.destroy()
    src.destroy()
