"""CPU checks of the C++ host API (aes-walsh-ckks_amd/host): the prime search,
the host FFT tables and the floating-point encode/decode, against the
reference's own outputs (golden fixtures) and closed forms.  Nothing here
touches the GPU."""
import cmath
import json
import os

import numpy as np
import pytest

import oracle as O
from helpers import GOLDEN, host, load_npz


def test_host_prime_search_reproduces_reference_chains():
    """Bit sizes read off the reference's chains give back the same primes."""
    H = host()
    with open(os.path.join(GOLDEN, "primes.json")) as f:
        primes = json.load(f)
    Ns = {"walsh15": 1 << 15, "xboot13": 1 << 14, "xboot14": 1 << 14, "xboot15": 1 << 15,
          "chain14": 1 << 14, "ks16": 1 << 16, "ntt16": 1 << 16}
    for tag, ent in primes.items():
        N = Ns[tag]
        qb = [int(v).bit_length() for v in ent["q"]]
        pb = [int(v).bit_length() for v in ent["p"]]
        p, q = H.find_p_and_q_primes(pb, qb, N)
        assert [int(v) for v in q] == ent["q"], tag
        assert [int(v) for v in p] == ent["p"], tag


def test_host_params_chain_and_defaults():
    H = host()
    c = H.CKKSParamsConfig()
    c.N = 1 << 14
    c.mult_depth = 7
    c.q0_bit = 50
    c.q_bit = 50
    c.p_bits = [50, 50]
    c.dnum = 4
    P = H.CKKSParams(c)
    p, q = O.find_primes([50, 50], [50] * 8, 1 << 14)
    assert list(P.getModuli()) == list(q) and list(P.getPModuli()) == list(p)
    assert P.getMaxLevel() == 7 and P.getAlpha() == 2 and P.getScale() == 2.0 ** 50
    c.q_mult_bits = [40] * 3
    with pytest.raises(ValueError, match="q_mult_bits size must equal mult_depth"):
        H.CKKSParams(c)


def test_host_perm_table_and_fft_twiddles_closed_form():
    H = host()
    N = 1 << 8
    perm = H.perm_table_generation(N)
    logN = 8
    for i in range(N // 2):
        e = pow(5, i, 2 * N)
        assert perm[i] == int(format((e - 1) >> 1, "0%db" % logN)[::-1], 2)
    tw = H.twiddle_ifft_generation(N)
    # stage m, entry i: omega^(brev_m(i) * N/2m) * psi^(N/2m), omega = e^(2 pi i/N), psi = e^(pi i/N)
    for m in (1, 2, 8, 64, 128):
        bits = m.bit_length() - 1
        for i in range(m):
            r = int(format(i, "0%db" % bits)[::-1], 2) if bits else 0
            k = N // (2 * m)
            want = cmath.exp(2j * cmath.pi * r * k / N) * cmath.exp(1j * cmath.pi * k / N)
            assert abs(tw[m + i] - want) < 1e-12


@pytest.mark.parametrize("tag", ["ev10", "ev10b"])
def test_host_encode_matches_reference(tag):
    """Slot encoding (special FFT, llrint, residues) is bit-exact with the
    reference's ckks_encode output (evaluation domain via the oracle NTT)."""
    H = host()
    d = load_npz(tag + ".npz")
    N = len(d["m1"]) * 2
    q = d["q"]
    enc = H.encode_slots_custom_basis_tables(N, d["m1"], 2.0 ** 40, list(q))
    ctx = O.Ctx(N, q, d["p"], 2)
    assert np.array_equal(ctx.ntt_fwd(enc, len(q)), d["encode_m1"])


@pytest.mark.parametrize("tag", ["ev10", "ev10b"])
def test_host_decode_matches_reference(tag):
    """Decrypt's decode (centred CRT of the leading limbs, long-double scale
    division, inverse special FFT) equals the reference's decrypted slots."""
    H = host()
    d = load_npz(tag + ".npz")
    N = len(d["m1"]) * 2
    q = d["q"]
    L = len(q)
    W = L * N
    b, a = d["ct1"][:W], d["ct1"][W:2 * W]
    sk = d["sk"]
    m = np.zeros(W, np.uint64)
    for l in range(L):
        sl = slice(l * N, (l + 1) * N)
        ql = int(q[l])
        m[sl] = np.array([(int(x) * int(y) + int(z)) % ql for x, y, z in zip(a[sl], sk[sl], b[sl])], np.uint64)
    coeff = O.Ctx(N, q, d["p"], 2).ntt_inv(m, L)
    scale = float(d["ct1_meta"][2:3].view(np.float64)[0])
    dec = H.decode_plaintext_custom_basis(N, coeff, list(q), scale)
    assert np.array_equal(dec, d["dec_ct1"])
    # the decoded slots are the encrypted message up to the encryption noise
    assert np.max(np.abs(dec - d["m1"])) < 1e-6


def test_host_decode_coefficient_domain_closed_form():
    """Coefficient-domain decode: slot i holds (c[brev(i)], c[N/2 + brev(i)]) / scale."""
    H = host()
    N = 1 << 6
    q = [int(v) for v in O.find_primes([], [50, 50], N)[1]]
    rng = np.random.default_rng(4)
    vals = rng.integers(-(1 << 40), 1 << 40, N)
    coeffs = np.concatenate([np.array([int(v) % m for v in vals], np.uint64) for m in q])
    out = H.decode_plaintext_custom_basis(N, coeffs, q, 2.0 ** 20, H.CKKSDecryptDomain.Coefficients)
    for i in range(N // 2):
        j = int(format(i, "05b")[::-1], 2)
        assert out[i] == complex(vals[j] / 2.0 ** 20, vals[N // 2 + j] / 2.0 ** 20)


def test_host_params_seed_is_drawn_unless_set():
    """An unset seed is drawn from std::random_device per CKKSParams (so two
    default contexts hold different secret keys, as the reference's
    getrandom-drawn keys do); a set seed is kept for replay."""
    H = host()
    seeds = {H.CKKSParams.AESWalsh15().getSeed() for _ in range(4)}
    assert len(seeds) == 4
    c = H.aes_walsh15_config()
    assert c.seed is None
    c.seed = 77
    assert H.CKKSParams(c).getSeed() == 77 and H.CKKSParams(c).getSeed() == 77
