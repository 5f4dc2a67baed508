"""The serving loop's request checks and bit unpacking, on the CPU."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "aes-walsh-ckks_amd"))
import aes_serve  # noqa: E402


def test_request_shape_checks():
    ok = np.zeros((3, 16), np.uint8)
    c, x = aes_serve.check_request(ok, ok)
    assert c.shape == (3, 16)
    with pytest.raises(ValueError):
        aes_serve.check_request(np.zeros((0, 16), np.uint8), np.zeros((0, 16), np.uint8))
    with pytest.raises(ValueError):
        aes_serve.check_request(np.zeros((1025, 16), np.uint8), np.zeros((1025, 16), np.uint8))
    with pytest.raises(ValueError):
        aes_serve.check_request(np.zeros((2, 16), np.uint8), np.zeros((3, 16), np.uint8))


def test_values_to_bytes_reads_bit_b_of_byte_k_at_8k_plus_b():
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, (5, 16), dtype=np.uint8)
    bits = np.unpackbits(data[:, :, None], axis=2, bitorder="little").reshape(5, 128).astype(float)
    noisy = bits + rng.uniform(-0.3, 0.3, bits.shape)
    assert np.array_equal(aes_serve.values_to_bytes(noisy), data)


def test_batches_are_fifo_whole_requests_within_one_evaluation():
    """The native server's packing rule (ckks_aes_serve_take): requests are
    taken first come first served, whole, while they fit 1024 blocks."""
    H = aes_serve._host()
    take = H.ckks_aes_serve_take([32, 500, 400, 100, 7, 32], 1024)
    assert list(take) == [0, 32, 532]          # the 100-block request would overflow: it waits
    assert list(H.ckks_aes_serve_take([1024, 1], 1024)) == [0]
    assert list(H.ckks_aes_serve_take([], 1024)) == []
