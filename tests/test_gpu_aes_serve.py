"""Served walsh15 AES-CTR with continuous batching (the native server behind
aes_serve.AesCtrServer: lanes of one key-owning context).

* a request's result is the same alone and in a mix of requests that spans
  several evaluations on both lanes (plaintext bytes exact; decrypted bit
  values within 1e-3 of each other);
* admission: a request beyond one evaluation, and requests beyond the lanes'
  admitted work, are refused with their own error while every other request
  of the burst completes correctly; a second lane opens only if its measured
  footprint fits the memory budget."""
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


def test_request_is_the_same_alone_and_in_the_mix_and_refusals_are_per_request():
    H = host()
    srv = aes_serve.AesCtrServer(KEY, lanes=2, max_queued_evaluations=2)
    try:
        assert srv.lanes == 2 and srv.evaluation_bytes > 0
        print(f"lanes {srv.lanes}, evaluation footprint {srv.evaluation_bytes / 2**30:.2f} GiB", flush=True)
        rng = np.random.default_rng(31)
        c, x, p = make_request(H, rng, 32, 0)
        alone = srv.result(srv.submit(c, x))
        assert alone["error"] == "" and not alone["refused"]
        assert np.array_equal(alone["plaintexts"], p)
        # a burst: 60 requests (more than 1024 blocks, so several evaluations on
        # both lanes) with the same request inside it, one oversize request
        # (1025 blocks) and then a flood past the admitted work
        ids, want, first = [], [], 32
        for k in range(60):
            if k == 17:
                ids.append(srv.submit(c, x))
                want.append(p)
                continue
            if k == 30:
                big_c, big_x, _ = make_request(H, rng, 1025, 10_000)
                oversize = srv.submit(big_c, big_x)
            n = int(rng.integers(16, 33))
            cc, xx, pp = make_request(H, rng, n, first)
            first += n
            ids.append(srv.submit(cc, xx))
            want.append(pp)
        flood_c, flood_x, _ = make_request(H, rng, 1024, 20_000)
        flood = [srv.submit(flood_c, flood_x) for _ in range(12)]
        results = [srv.result(i) for i in ids]  # each result is handed out once
        for r, w in zip(results, want):
            assert r["error"] == "", r["error"]
            assert np.array_equal(r["plaintexts"], w)
        r_over = srv.result(oversize)
        assert r_over["refused"] and "1..1024 blocks" in r_over["error"]
        outcomes = [srv.result(i) for i in flood]
        refused = sum(r["refused"] for r in outcomes)
        assert refused > 0 and all(r["error"] == "" for r in outcomes if not r["refused"])
        assert len({r["lane"] for r in results}) == 2
        print(f"flood: {refused} of {len(flood)} refused at admission; evaluations {srv.evaluations}", flush=True)
        r17 = results[17]
        assert np.array_equal(r17["plaintexts"], alone["plaintexts"])
        diff = float(np.max(np.abs(r17["values"] - alone["values"])))
        print(f"alone vs in the mix: max value difference {diff:.2e}", flush=True)
        assert diff < 1e-3
    finally:
        srv.close()


def test_lanes_are_admitted_by_hbm_footprint():
    """A second lane opens only if its measured pool footprint fits the budget."""
    srv = aes_serve.AesCtrServer(KEY, lanes=2, memory_budget=1)
    try:
        assert srv.lanes == 1 and srv.evaluation_bytes > 0
    finally:
        srv.close()
