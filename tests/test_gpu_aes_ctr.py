"""Homomorphic AES-128 CTR decryption with the mask-packed Walsh S-box
(ckks_aes_ctr_walsh_sbox.cpp:936-1051) at the reference's walsh15 parameters
(ckks_aes_params.hpp:35-82), on the blocks of the reference bench
(aes_bench.cpp:29-101): every decrypted output bit rounds to the plaintext
bit of its block in both lanes (0 mismatched blocks, as aes_bench.cpp:156-196
counts them).  The plaintexts come from the host AES-128, which the CPU tests
pin to FIPS-197 and SP 800-38A known answers."""
import numpy as np
import pytest

from helpers import host

pytestmark = pytest.mark.gpu

KEY = list(bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c"))
COUNTER = bytes.fromhex("f0f1f2f3f4f5f6f7f8f9fafbfcfdfeff")
PLAINTEXT = bytes.fromhex("6bc1bee22e409f96e93d7e117393172a")


def counter_for_block(i):
    """counter_for_block (aes_bench.cpp:62-74): the 128-bit big-endian counter + i."""
    return list((int.from_bytes(COUNTER, "big") + i).to_bytes(16, "big"))


def plaintext_for_block(i):
    """plaintext_for_block (aes_bench.cpp:76-88)"""
    if i == 0:
        return list(PLAINTEXT)
    return [(0x5A + 17 * i + 29 * b) & 0xFF for b in range(16)]


def make_blocks(H, count):
    counters = np.array([counter_for_block(i) for i in range(count)], np.uint8)
    plains = np.array([plaintext_for_block(i) for i in range(count)], np.uint8)
    cts = np.array([H.aes128_ctr_crypt_block(KEY, list(c), list(p)) for c, p in zip(counters, plains)], np.uint8)
    return counters, plains, cts


def validate(H, ctx, bits, expected, complex_blocks):
    """validate_walsh_outputs (aes_bench.cpp:156-196)"""
    assert len(bits) == 128
    decoded = np.zeros((2 * complex_blocks, 16), np.uint8)
    errors = []
    for bit in range(128):
        v = np.asarray(H.ckks_decrypt(ctx, bits[bit]))[:complex_blocks]
        for lane, vals in enumerate((v.real, v.imag)):
            rows = np.arange(complex_blocks) + lane * complex_blocks
            want = (expected[rows, bit // 8] >> (bit % 8)) & 1
            decoded[rows, bit // 8] |= ((vals >= 0.5).astype(np.uint8) << (bit % 8))
            errors.append(np.abs(vals - want))
    errors = np.concatenate(errors)
    mismatches = int(np.sum(np.any(decoded != expected, axis=1)))
    return mismatches, float(errors.max()), float(errors.mean())


def test_aes_ctr_walsh15_decrypts_every_block():
    H = host()
    ctx = H.CKKSContext(H.CKKSParams.AESWalsh15())
    cb = ctx.getParams().getAesWalshBlockCount()
    counters, plains, cts = make_blocks(H, 2 * cb)
    keys = H.ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(ctx, H.aes128_expand_key(KEY), cb)
    assert len(keys) == 1408
    print(f"walsh15 AES-CTR: {2 * cb} blocks", flush=True)
    out = H.ckks_aes_ctr_walsh_sbox_complex_blocks(ctx, keys, counters, cts)
    mismatches, max_err, avg_err = validate(H, ctx, out, plains, cb)
    print(f"mismatches {mismatches} max error {max_err:.3e} average error {avg_err:.3e} "
          f"level {out[0].getLevel()}", flush=True)
    assert mismatches == 0
    assert max_err < 0.05, max_err
    # the batched decryption equals the one-at-a-time decryption, value for value
    batch = H.ckks_decrypt_batch(ctx, out)
    for bit in (0, 57, 127):
        assert np.array_equal(np.asarray(batch[bit]), np.asarray(H.ckks_decrypt(ctx, out[bit])))


def test_aes_ctr_rejects_bad_block_spans():
    H = host()
    ctx = H.CKKSContext(H.CKKSParams.AESWalsh15())
    keys = H.ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(ctx, H.aes128_expand_key(KEY), 512)
    c, _, x = make_blocks(H, 4)
    with pytest.raises(Exception, match="non-empty, equal, and even"):
        H.ckks_aes_ctr_walsh_sbox_complex_blocks(ctx, keys, c[:3], x[:3])
    with pytest.raises(Exception, match="non-empty, equal, and even"):
        H.ckks_aes_ctr_walsh_sbox_complex_blocks(ctx, keys, c[:4], x[:2])
    with pytest.raises(Exception, match="packing capacity"):
        H.ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(ctx, H.aes128_expand_key(KEY), 513)


def test_aes_ctr_batched_and_concurrent_sboxes_are_bit_exact():
    """Three forms of the same decryption, every limb of the 128 output bits
    equal: the plain one (S-boxes one after another on one stream, the
    reference's per-operation order for the linear layer), the batched one
    (the 16 S-boxes of each round and the final-refresh packs as one
    ciphertext batch, MixColumns + AddRoundKey as one linear combination)
    and the concurrent one (S-boxes over the main context and two more built
    from the same parameters, one host thread and stream each)."""
    H = host()
    params = H.CKKSParams.AESWalsh15()
    ctxs = [H.CKKSContext(params) for _ in range(3)]
    ctx = ctxs[0]
    cb = ctx.getParams().getAesWalshBlockCount()
    counters, plains, cts = make_blocks(H, 2 * cb)
    keys = H.ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(ctx, H.aes128_expand_key(KEY), cb)
    print("plain", flush=True)
    seq = H.ckks_aes_ctr_walsh_sbox_complex_blocks(ctx, keys, counters, cts, fuse_linear_layer=False,
                                                   batch_sboxes=False)
    print("batched", flush=True)
    bat = H.ckks_aes_ctr_walsh_sbox_complex_blocks(ctx, keys, counters, cts)
    print("concurrent", flush=True)
    par = H.ckks_aes_ctr_walsh_sbox_complex_blocks(ctx, keys, counters, cts, sbox_contexts=ctxs[1:],
                                                   batch_sboxes=False)
    for other in (bat, par):
        for bit in range(128):
            assert other[bit].getLevel() == seq[bit].getLevel() and other[bit].getScale() == seq[bit].getScale()
            for k in range(2):
                assert np.array_equal(np.asarray(other[bit].getPoly(k)), np.asarray(seq[bit].getPoly(k))), (bit, k)
    mismatches, max_err, _ = validate(H, ctx, bat, plains, cb)
    print(f"batched S-boxes: mismatches {mismatches} max error {max_err:.3e}", flush=True)
    assert mismatches == 0


def test_aes_ctr_rejects_a_context_with_other_keys():
    H = host()
    ctx = H.CKKSContext(H.CKKSParams.AESWalsh15())
    c = H.aes_walsh15_config()
    c.seed = 12345
    other = H.CKKSContext(H.CKKSParams(c))
    keys = H.ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(ctx, H.aes128_expand_key(KEY), 512)
    counters, _, cts = make_blocks(H, 1024)
    with pytest.raises(Exception, match="share the main context"):
        H.ckks_aes_ctr_walsh_sbox_complex_blocks(ctx, keys, counters, cts, sbox_contexts=[other],
                                                 batch_sboxes=False)
    # contexts built from separately drawn parameter sets hold separate keys
    with pytest.raises(Exception, match="share the main context"):
        H.ckks_aes_ctr_walsh_sbox_complex_blocks(ctx, keys, counters, cts,
                                                 sbox_contexts=[H.CKKSContext(H.CKKSParams.AESWalsh15())],
                                                 batch_sboxes=False)
