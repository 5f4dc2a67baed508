"""Served walsh15 AES-CTR with continuous batching (aes_serve.AesCtrServer):
a request's result is the same alone and in a mix of requests (plaintext
bytes equal and exact; decrypted bit values within 1e-3 of each other), and
a burst larger than one evaluation spreads over both lanes."""
import os
import sys

import numpy as np
import pytest

from helpers import host

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "aes-walsh-ckks_amd"))
import aes_serve  # noqa: E402

KEY = bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c")


def make_request(H, rng, n, first):
    base = int.from_bytes(bytes.fromhex("f0f1f2f3f4f5f6f7f8f9fafbfcfdfeff"), "big")
    counters = np.array([list((base + first + i).to_bytes(16, "big")) for i in range(n)], np.uint8)
    plains = rng.integers(0, 256, (n, 16), dtype=np.uint8)
    cts = np.array([H.aes128_ctr_crypt_block(list(KEY), list(c), list(p)) for c, p in zip(counters, plains)],
                   np.uint8)
    return counters, cts, plains


def test_request_is_the_same_alone_and_in_the_mix():
    H = host()
    srv = aes_serve.AesCtrServer(KEY, lanes=2, contexts_per_lane=2)
    try:
        assert len(srv.lanes) == 2
        rng = np.random.default_rng(31)
        c, x, p = make_request(H, rng, 32, 0)
        alone = srv.submit(aes_serve.Request(c, x))
        assert alone.done.wait(300) and alone.error is None
        assert np.array_equal(alone.plaintexts, p)
        # a mix of 60 requests (more than 1024 blocks) with the same request inside it
        mix, want, first = [], [], 32
        for k in range(60):
            if k == 17:
                mix.append(srv.submit(aes_serve.Request(c, x)))
                want.append(p)
                continue
            n = int(rng.integers(16, 33))
            cc, xx, pp = make_request(H, rng, n, first)
            first += n
            mix.append(srv.submit(aes_serve.Request(cc, xx)))
            want.append(pp)
        for r, w in zip(mix, want):
            assert r.done.wait(300) and r.error is None
            assert np.array_equal(r.plaintexts, w)
        again = mix[17]
        assert np.array_equal(again.plaintexts, alone.plaintexts)
        diff = float(np.max(np.abs(again.values - alone.values)))
        print(f"alone vs in the mix: max value difference {diff:.2e}; evaluations {srv.evaluations}", flush=True)
        assert diff < 1e-3
        assert srv.evaluations >= 3   # the warm-ups are not counted; the burst needs two evaluations
    finally:
        srv.close()


def test_lanes_are_admitted_by_hbm_footprint():
    """A second lane opens only if its measured pool footprint fits the budget."""
    srv = aes_serve.AesCtrServer(KEY, lanes=2, contexts_per_lane=1, memory_budget=1)
    try:
        assert len(srv.lanes) == 1 and srv.footprints[0] > 0
    finally:
        srv.close()
