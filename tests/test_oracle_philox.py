"""Philox4x32-10 oracle vs the Random123 known-answer vectors."""
import numpy as np

from oracle import philox


def test_random123_kat_vectors():
    ctr = np.array([[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]],
                   dtype=np.uint64)
    key = np.array([[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]], dtype=np.uint64)
    want = [[0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8],
            [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD],
            [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]]
    assert philox.philox4x32_10(ctr, key).tolist() == want


def test_uniforms_range_and_stream_split():
    u = philox.uniforms(12345, 0, 4096)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.03
    assert np.array_equal(philox.uniforms(12345, 100, 10), u[100:110])
    assert not np.array_equal(philox.uniforms(12346, 0, 10), u[:10])
