"""AES-128 CTR with the XBOOT S-box (ckks_aes_ctr.cpp:280-779) on the
reference's AESXBoot14SecureSparse1024 preset: the reference bench's first
1024 blocks (the xboot14-sparse-1024 row, aes_bench.cpp:378-384 and
29-101) decrypt with 0 mismatched blocks and every slot's error (the
periodic expectation, aes_bench.cpp:104-156) below 0.5, and the
batched evaluation (16 S-boxes per launch sequence, the 64 pair bootstraps
as one batch) equals the one-at-a-time evaluation limb for limb."""
import numpy as np
import pytest

from helpers import host
from test_gpu_aes_ctr import KEY, make_blocks

pytestmark = pytest.mark.gpu


def test_xboot14_sparse1024_decrypts_and_batched_is_bit_exact():
    H = host()
    P = H.CKKSParams.AESXBoot14SecureSparse1024()
    ctx = H.CKKSContext(P)
    active = 1024
    counters, plains, cts = make_blocks(H, active)
    keys = H.ckks_aes128_encrypt_expanded_round_key_bits(ctx, H.aes128_expand_key(KEY), 4)
    assert len(keys) == 1408
    print("xboot14-sparse-1024: batched", flush=True)
    out = H.ckks_aes_ctr(ctx, keys, counters, cts)
    logical = P.getBootstrapActiveSlotCount()
    slots = P.getSlots()
    decoded = np.zeros((slots, 16), np.uint8)
    worst = 0.0
    for bit in range(128):
        v = np.asarray(H.ckks_decrypt(ctx, out[bit]))
        want = np.array([(plains[s % logical][bit // 8] >> (bit % 8)) & 1 if s % logical < active else 0
                         for s in range(slots)], np.float64)
        worst = max(worst, float(np.max(np.abs(v.real - want))), float(np.max(np.abs(v.imag))))
        decoded[:, bit // 8] |= (v.real >= 0.5).astype(np.uint8) << (bit % 8)
    mismatches = int(np.sum(np.any(decoded[:active] != plains, axis=1)))
    print(f"mismatches {mismatches} max error {worst:.3e} level {out[0].getLevel()}", flush=True)
    assert mismatches == 0 and worst < 0.5
    print("one at a time", flush=True)
    seq = H.ckks_aes_ctr(ctx, keys, counters, cts, batch_independent_work=False)
    for bit in range(128):
        assert seq[bit].getLevel() == out[bit].getLevel() and seq[bit].getScale() == out[bit].getScale()
        for k in range(2):
            assert np.array_equal(seq[bit].getPoly(k), out[bit].getPoly(k)), (bit, k)
