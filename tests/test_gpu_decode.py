"""Decoding on the device (mi_decode_batch: CRT, the extended-precision
scale division, the inverse special FFT; ckks_encoding.cpp:254-329,
fft.cpp:37-117) equals the host decode of the same coefficients bit for bit,
for ciphertexts at several levels and scales, in both domains, one at a time
and batched (mixed levels)."""
import numpy as np
import pytest

from helpers import host

pytestmark = pytest.mark.gpu


def _ctx(H, N=1 << 12):
    c = H.CKKSParamsConfig()
    c.N = N
    c.mult_depth = 4
    c.q0_bit = 50
    c.q_bit = 40
    c.p_bits = [60, 60]
    c.dnum = 3
    c.secure = False
    c.seed = 5
    return H.CKKSContext(H.CKKSParams(c))


@pytest.mark.parametrize("N", [1 << 12, 1 << 15])
def test_device_decode_equals_host_decode(N):
    H = host()
    ctx = _ctx(H, N)
    slots = ctx.getParams().getSlots()
    rng = np.random.default_rng(N)
    cts = []
    for level in (4, 2, 0):
        v = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)
        ct = H.ckks_encrypt(ctx, v, level)
        cts.append(ct)
    # a non-power-of-two scale: a product and a rescale
    sq = H.ckks_encrypt(ctx, rng.uniform(-1, 1, slots).astype(np.complex128), 4)
    H.ckks_mul_inplace(ctx, sq, sq.clone())
    H.ckks_relin_hybrid_inplace(ctx, sq)
    H.ckks_rescale(ctx, sq)
    cts.append(sq)
    for dom in (H.CKKSDecryptDomain.Slots, H.CKKSDecryptDomain.Coefficients):
        batch = H.ckks_decrypt_batch(ctx, cts, dom)
        for ct, got in zip(cts, batch):
            want = H.ckks_decrypt_host_decode(ctx, ct, dom)
            assert np.array_equal(np.asarray(got).view(np.uint64), np.asarray(want).view(np.uint64))
            one = H.ckks_decrypt(ctx, ct, dom)
            assert np.array_equal(np.asarray(one).view(np.uint64), np.asarray(want).view(np.uint64))
