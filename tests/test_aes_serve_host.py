"""The serving loop's request checks and bit unpacking, on the CPU."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "aes-walsh-ckks_amd"))
import aes_serve  # noqa: E402


def test_request_shape_checks():
    ok = np.zeros((3, 16), np.uint8)
    assert aes_serve.Request(ok, ok).blocks == 3
    with pytest.raises(ValueError):
        aes_serve.Request(np.zeros((0, 16), np.uint8), np.zeros((0, 16), np.uint8))
    with pytest.raises(ValueError):
        aes_serve.Request(np.zeros((33, 16), np.uint8), np.zeros((33, 16), np.uint8))
    with pytest.raises(ValueError):
        aes_serve.Request(np.zeros((2, 16), np.uint8), np.zeros((3, 16), np.uint8))


def test_values_to_bytes_reads_bit_b_of_byte_k_at_8k_plus_b():
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, (5, 16), dtype=np.uint8)
    bits = np.unpackbits(data[:, :, None], axis=2, bitorder="little").reshape(5, 128).astype(float)
    noisy = bits + rng.uniform(-0.3, 0.3, bits.shape)
    assert np.array_equal(aes_serve.values_to_bytes(noisy), data)


def test_batches_are_fifo_whole_requests_within_one_evaluation():
    from collections import deque
    srv = aes_serve.AesCtrServer.__new__(aes_serve.AesCtrServer)   # packing only: no lanes, no device
    sizes = [32, 500, 400, 100, 7, 32]
    srv.queue = deque(aes_serve.Request(np.zeros((min(n, 32), 16), np.uint8), np.zeros((min(n, 32), 16), np.uint8))
                      for n in sizes)
    # requests hold at most 32 blocks: emulate larger ones by their block counts
    for r, n in zip(srv.queue, sizes):
        r.counters = np.zeros((n, 16), np.uint8)
    batch = srv._take_batch()
    assert [r.blocks for r, _ in batch] == [32, 500, 400]          # the 100-block request would overflow 1024
    assert [off for _, off in batch] == [0, 32, 532]
    assert [r.blocks for r in srv.queue] == [100, 7, 32]           # FIFO: nothing jumps the queue
