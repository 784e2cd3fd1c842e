"""The baked volume's stream format, restated in numpy (oracle.wire_pack / wire_unpack).

The format is this library's (the reference keeps its bake in a GPU texture), so it
is pinned by its definition: byte offsets of every section, the table word of a
constant and of a stored tile, the bitmap bit, and the payload texel order, on a
hand-built volume; then round trips over random mixes of constant and stored tiles,
the bitwise tile test (-0 is not +0, a NaN tile is still constant), and the errors.
"""
import numpy as np
import pytest

import oracle as O

ONE, NEG, TWO = 0x3C00, 0xBC00, 0x4000


def planes(res=32, mips=2, fill=ONE, mat=0):
    return ([np.full((res >> m,) * 3, fill, np.uint16) for m in range(mips)],
            [np.full((res >> m,) * 3, mat, np.uint16) for m in range(mips)])


def test_layout_table_bitmap_and_payload_of_a_hand_built_volume():
    dist, mat = planes(32, 2)
    ramp = np.arange(512, dtype=np.uint16).reshape(8, 8, 8)      # [z, y, x] within the tile
    dist[0][8:16, 0:8, 16:24] = ramp                              # tile (x 2, y 0, z 1) of mip 0
    mat[0][8:16, 0:8, 16:24] = 7
    mat[1][0:8, 8:16, 0:8] = TWO                                  # mip 1 tile (0, 1, 0): constant
    data = O.wire_pack(dist, mat, 4.0, 8.0, (-4.0, -2.0, -4.0), 1.5)
    tiles = 4 ** 3 + 2 ** 3
    h = np.frombuffer(data[:96], dtype=O.WIRE_HEADER)[0]
    assert (h["magic"], h["version"], h["resolution"], h["mip_levels"]) == (0x42564C43, 1, 32, 2)
    assert (h["band"], h["world_size"], h["gradient_voxels"]) == (4.0, 8.0, 1.5)
    assert tuple(h["origin"]) == (-4.0, -2.0, -4.0)
    assert (h["tiles"], h["stored_tiles"]) == (tiles, 1)
    assert (h["table_offset"], h["bitmap_offset"]) == (96, 96 + 4 * tiles)
    assert h["payload_offset"] == 512 and h["payload_offset"] % 256 == 0   # 96 + 288 + 12 -> 512
    assert h["total_bytes"] == len(data) == 512 + 4096
    table = np.frombuffer(data, "<u4", tiles, 96)
    stored_tile = 2 + 4 * (0 + 4 * 1)
    assert table[stored_tile] == 0
    assert table[0] == ONE                                         # constant (dist 1, material 0)
    assert table[64 + 0 + 2 * (1 + 2 * 0)] == (TWO << 16) | ONE    # mip 1, tile (0, 1, 0)
    words = np.frombuffer(data, "<u4", (tiles + 31) // 32, 96 + 4 * tiles)
    assert words[stored_tile // 32] == 1 << (stored_tile % 32) and words.sum() == words[stored_tile // 32]
    tex = np.frombuffer(data, "<u2", 2048, 512).reshape(512, 4)
    assert np.array_equal(tex[:, 0], np.arange(512))              # x fastest, then y, then z
    assert (tex[:, 1] == 7).all() and (tex[:, 2] == 0).all() and (tex[:, 3] == ONE).all()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_round_trip_over_mixed_tiles(seed):
    rng = np.random.default_rng(seed)
    res, mips = 64, 3
    dist, mat = planes(res, mips)
    for m in range(mips):
        r = res >> m
        t = r // 8
        for _ in range(3 * t):
            z, y, x = rng.integers(0, t, 3) * 8
            if rng.random() < 0.5:
                dist[m][z:z + 8, y:y + 8, x:x + 8] = rng.integers(0, 65536, (8, 8, 8))
            else:
                mat[m][z:z + 8, y:y + 8, x:x + 8] = rng.integers(0, 65536)   # constant, other material
        dist[m][rng.integers(0, r), rng.integers(0, r), rng.integers(0, r)] = NEG
    data = O.wire_pack(dist, mat, 4.0, 24.0, (-12.0, -4.0, -12.0))
    h, d2, m2 = O.wire_unpack(data)
    for m in range(mips):
        assert np.array_equal(d2[m], dist[m]) and np.array_equal(m2[m], mat[m])
    assert len(data) <= O.wire_layout(int(h["tiles"]))[2] + 4096 * int(h["tiles"])


def test_tiles_are_compared_bit_for_bit():
    dist, mat = planes(16, 1, fill=0x0000)
    dist[0][3, 2, 1] = 0x8000                 # -0 beside +0: stored
    dist[0][8:16, 8:16, 8:16] = 0xFFFF        # a NaN pattern, uniform: constant
    data = O.wire_pack(dist, mat, 4.0, 4.0, (0.0, 0.0, 0.0))
    h = np.frombuffer(data[:96], dtype=O.WIRE_HEADER)[0]
    table = np.frombuffer(data, "<u4", 8, 96)
    assert h["stored_tiles"] == 1 and table[0] == 0
    assert table[7] == 0xFFFF
    _, d2, _ = O.wire_unpack(data)
    assert np.array_equal(d2[0], dist[0])


def test_unpack_rejects_bad_streams():
    dist, mat = planes(16, 1)
    dist[0][0, 0, 0] = 0
    data = bytearray(O.wire_pack(dist, mat, 4.0, 4.0, (0.0, 0.0, 0.0)))
    with pytest.raises(ValueError, match="truncated"):
        O.wire_unpack(bytes(data[:-1]))
    bad = bytearray(data)
    bad[0] ^= 1
    with pytest.raises(ValueError, match="magic"):
        O.wire_unpack(bytes(bad))
    bad = bytearray(data)
    bad[96:100] = np.uint32(5).tobytes()       # tile 0 is stored; only one payload tile exists
    with pytest.raises(ValueError, match="outside the payload"):
        O.wire_unpack(bytes(bad))
