"""Wire codec bit packing (the reference's response format,
src/server.cpp:950-1009, src/client.cpp:315-361), CPU only.

Pinned by what the format means: the byte string is the little-endian
encoding of the integer sum(v_i << (i * width)), zero-padded to whole bytes.
"""
import numpy as np
import pytest


def _reference_bytes(vals, width):
    acc = 0
    for i, v in enumerate(vals):
        acc |= (int(v) & ((1 << width) - 1)) << (i * width)
    return acc.to_bytes((len(vals) * width + 7) // 8, "little")


@pytest.mark.parametrize("width", [1, 7, 8, 13, 22, 29, 50, 57, 58, 59, 60, 61, 63, 64])
def test_pack_bits_is_little_endian_concatenation(width):
    import onionpirv2_amd.host as H
    rng = np.random.default_rng(width)
    hi = (1 << width) - 1
    vals = rng.integers(0, 1 << 63, size=37, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    vals = vals & np.uint64(hi)
    vals[:3] = [0, hi, 1]
    got = H.pack_bits(vals, width)
    assert got == _reference_bytes(vals, width)
    back, used = H.unpack_bits(got, len(vals), width)
    assert used == len(got) and np.array_equal(back, vals)


def test_pack_bits_drops_bits_above_width():
    import onionpirv2_amd.host as H
    vals = np.array([0xFFFF, 0x1234], np.uint64)
    assert H.pack_bits(vals, 4) == bytes([0x4F])


def test_unpack_bits_truncated_raises():
    import onionpirv2_amd.host as H
    data = H.pack_bits(np.arange(10, dtype=np.uint64), 22)
    with pytest.raises(RuntimeError):
        H.unpack_bits(data[:-1], 10, 22)
