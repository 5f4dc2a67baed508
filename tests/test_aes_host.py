"""Host AES-128 (the plain reference the homomorphic CTR path is checked
against), the walsh15 parameter set and the default EvalMod polynomial, on
the CPU."""
import os

import numpy as np

from helpers import host

KEY = bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c")


def test_aes128_fips197_known_answers():
    H = host()
    rk = bytes(H.aes128_expand_key(list(KEY)))
    # FIPS-197 Appendix A.1: w[40..43]
    assert rk[160:].hex() == "d014f9a8c9ee2589e13f0cc8b6630ca6"
    assert rk[16:20].hex() == "a0fafe17"
    # FIPS-197 Appendix B
    pt = list(bytes.fromhex("3243f6a8885a308d313198a2e0370734"))
    assert bytes(H.aes128_encrypt_block(list(KEY), pt)).hex() == "3925841d02dc09fbdc118597196a0b32"
    assert bytes(H.aes128_encrypt_block_expanded(list(rk), pt)).hex() == "3925841d02dc09fbdc118597196a0b32"
    # FIPS-197 Appendix C.1 (another key)
    k2 = list(range(16))
    p2 = list(bytes.fromhex("00112233445566778899aabbccddeeff"))
    assert bytes(H.aes128_encrypt_block(k2, p2)).hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"


def test_aes128_ctr_sp800_38a_first_block():
    """NIST SP 800-38A F.5.1 CTR-AES128.Encrypt, block 1: the counter and
    plaintext the reference bench starts from (aes_bench.cpp:29-39)."""
    H = host()
    ctr = list(bytes.fromhex("f0f1f2f3f4f5f6f7f8f9fafbfcfdfeff"))
    pt = list(bytes.fromhex("6bc1bee22e409f96e93d7e117393172a"))
    ct = bytes(H.aes128_ctr_crypt_block(list(KEY), ctr, pt))
    assert ct.hex() == "874d6191b620e3261bef6864990db6ce"
    assert bytes(H.aes128_ctr_crypt_block(list(KEY), ctr, list(ct))) == bytes(pt)


def test_aes_sbox_table_from_its_definition():
    H = host()
    s = H.aes_sbox_table()
    assert s[0x00] == 0x63 and s[0x01] == 0x7C and s[0x53] == 0xED and s[0xFF] == 0x16
    assert sorted(s) == list(range(256))


def test_walsh15_parameters():
    """AESWalsh15 (ckks_aes_params.hpp:35-82): the chain, P, K and packing."""
    H = host()
    P = H.CKKSParams.AESWalsh15()
    assert P.getN() == 1 << 15 and P.getSlots() == 1 << 14
    assert P.getMaxLevel() == 12 and P.getDnum() == 3
    assert [int(q).bit_length() for q in P.getModuli()] == [36, 61, 42, 42] + [38] * 7 + [50, 50]
    assert [int(p).bit_length() for p in P.getPModuli()] == [61] * 4
    assert P.getScale() == 2.0 ** 35
    assert P.usesAesWalsh() and P.getAesWalshBlockCount() == 512 and P.getAesWalshPackedSegmentCount() == 32
    assert P.getBootstrapS2CDepth() == 1 and P.getBootstrapEvalModK() == 11


def test_default_evalmod_polynomial_fit_and_fallback():
    """walsh15's default EvalMod polynomial is the reference's discrete fit
    (reproduces the mod-2 map near every point k/(2K) the bootstrap meets);
    parameters outside the presets fall back to the degree-d Chebyshev
    interpolant of the same target on [-1, 1]."""
    H = host()
    P = H.CKKSParams.AESWalsh15()
    K = P.getBootstrapEvalModK()
    c = np.array(H.default_evalmod_polynomial(P))
    assert H.evalmod_polynomial_is_reference_fit(P) and c.shape == (100,)
    u = np.concatenate([(np.arange(-2 * K + 1, 2 * K) / 2 + d) / K for d in (-0.2, 0.0, 0.2)])
    np.testing.assert_allclose(np.polynomial.chebyshev.chebval(u, c), 0.5 - 0.5 * np.cos(2 * np.pi * K * u), atol=1e-8)
    cfg = H.aes_walsh15_config()
    bp = cfg.bootstrap_params
    bp.evalmod_polynomial_degree = 90
    cfg.bootstrap_params = bp
    P2 = H.CKKSParams(cfg)
    assert not H.evalmod_polynomial_is_reference_fit(P2)
    c2 = np.array(H.default_evalmod_polynomial(P2))
    want = np.polynomial.chebyshev.chebinterpolate(lambda x: 0.5 - 0.5 * np.cos(2 * np.pi * K * x), 90)
    want[1::2] = 0.0
    assert c2.shape == (91,)
    np.testing.assert_allclose(c2, want, atol=1e-12)


def test_xboot_presets_reproduce_reference_chains():
    """The XBOOT presets (ckks_aes_params.hpp:12-211) give the reference's
    own prime chains (the harness dumped them from CKKSParams::XBoot*), and
    their EvalMod settings select the reference's fitted polynomial."""
    import json
    from helpers import GOLDEN
    H = host()
    with open(os.path.join(GOLDEN, "primes.json")) as f:
        primes = json.load(f)
    for tag, make in [("xboot13", H.CKKSParams.XBootParamAes13Reference), ("xboot14", H.CKKSParams.AESXBoot14Secure),
                      ("xboot15", H.CKKSParams.AESXBoot15)]:
        P = make()
        assert [int(v) for v in P.getModuli()] == primes[tag]["q"], tag
        assert [int(v) for v in P.getPModuli()] == primes[tag]["p"], tag
        assert H.evalmod_polynomial_is_reference_fit(P), tag
    assert H.CKKSParams.AESXBoot14SecureSparse1024().getBootstrapActiveSlotCount() == 1024
    assert H.CKKSParams.AESXBoot15Sparse1024().getBootstrapActiveSlotCount() == 1024
    assert H.CKKSParams.AESXBoot15().getBootstrapActiveSlotCount() == 1 << 14


def test_aes_scale_up_factor_uses_the_reference_tolerance():
    """add_term's integer scale-up (ckks_aes_ctr_walsh_sbox.cpp:80-124): a
    ratio within 1e-3 of an integer rounds to it (4 (1 + 1e-5) -> 4), one
    further out is refused as non-integral, scales equal to 2^-45 need no
    factor, and a scale-down is refused."""
    H = host()
    s = 2.0 ** 40
    assert H.aes_scale_up_factor(s, 4 * s * (1 + 1e-5)) == (4, None)
    assert H.aes_scale_up_factor(s, 3 * s) == (3, None)
    assert H.aes_scale_up_factor(s, s) == (1, None)
    f, err = H.aes_scale_up_factor(s, 4 * s * (1 + 2e-3))
    assert err == "ckks_aes_walsh: non-integral scalar scale-up"
    f, err = H.aes_scale_up_factor(4 * s, s)
    assert err == "ckks_aes_walsh: cannot scale down ciphertext for addition"
