"""GPU parity: batched NTT through the C-ABI against the oracle and the
reference's golden vectors (bit-exact)."""
import numpy as np
import pytest

import oracle as O
from helpers import digest, load_json, product, rand_residues, splitmix64_stream

pytestmark = pytest.mark.gpu


def _ctx(N, qbits, pbits=(61,)):
    p, q = O.find_primes(list(pbits), list(qbits), N)
    M = product()
    return M, M.Context(N, q, p, 1), q, p


@pytest.mark.parametrize("logN", [10, 11, 12, 13, 14, 15, 16])
def test_ntt_matches_oracle(logN):
    N = 1 << logN
    M, ctx, q, p = _ctx(N, [60, 50, 36, 61, 45])
    octx = O.Ctx(N, q, p, 1)
    rng = np.random.default_rng(logN)
    npolys, limbs = 3, len(q)
    x = np.concatenate([rand_residues(rng, q, N) for _ in range(npolys)])
    d = M.to_dev(x)
    ctx.ntt_forward(d, npolys, limbs)
    got = M.from_dev(d).reshape(npolys, limbs * N)
    for k in range(npolys):
        assert np.array_equal(got[k], octx.ntt_fwd(x[k * limbs * N:(k + 1) * limbs * N], limbs)), k
    ctx.ntt_inverse(d, npolys, limbs)
    assert np.array_equal(M.from_dev(d), x)
    # inverse alone against the oracle
    d2 = M.to_dev(x)
    ctx.ntt_inverse(d2, npolys, limbs)
    got = M.from_dev(d2).reshape(npolys, limbs * N)
    for k in range(npolys):
        assert np.array_equal(got[k], octx.ntt_inv(x[k * limbs * N:(k + 1) * limbs * N], limbs)), k


def test_ntt_strided_batch_and_p_basis():
    N = 1 << 12
    M, ctx, q, p = _ctx(N, [50, 50, 50], pbits=(55, 56))
    octx = O.Ctx(N, q, p, 1)
    rng = np.random.default_rng(5)
    stride = 6 * N  # padding between polys is left untouched
    buf = rng.integers(0, 1 << 62, 4 * stride, dtype=np.uint64)
    for k in range(4):
        buf[k * stride: k * stride + 2 * N] = rand_residues(rng, p, N)
    d = M.to_dev(buf)
    ctx.ntt_forward(d, 4, 2, stride=stride, basis0=3)  # the two P moduli
    got = M.from_dev(d)
    for k in range(4):
        sl = slice(k * stride, k * stride + 2 * N)
        basis = np.array([3, 4], np.int32)
        ref = buf[sl].copy()
        O.lib().oc_ntt_fwd_basis(octx.h, ref.ctypes.data, basis.ctypes.data, 2, 0)
        assert np.array_equal(got[sl], ref)
        assert np.array_equal(got[k * stride + 2 * N:(k + 1) * stride], buf[k * stride + 2 * N:(k + 1) * stride])


@pytest.mark.parametrize("tag", ["ntt10", "ntt12", "ntt16one"])
def test_ntt_matches_reference_golden(tag):
    ent = load_json("ntt.json")[tag]
    N = ent["N"]
    M = product()
    q = ent["q"]
    ctx = M.Context(N, q, [O.find_primes([58], [], N)[0][0]], 1)
    for i, e in enumerate(ent["limbs"]):
        x = splitmix64_stream(ent["seed_base"] + i, N) % np.uint64(q[i])
        d = M.to_dev(x)
        ctx.ntt_forward(d, 1, 1, basis0=i)
        assert digest(M.from_dev(d)) == e["fwd_out"]
        d = M.to_dev(x)
        ctx.ntt_inverse(d, 1, 1, basis0=i)
        assert digest(M.from_dev(d)) == e["inv_out"]


def test_ntt_large_batch_config2_shape_sample():
    """Config-2 primes (40 x 60-bit, N=2^16) on a 2-ciphertext sample."""
    N = 1 << 16
    pj = load_json("primes.json")["ntt16"]["q"]
    M = product()
    ctx = M.Context(N, pj, [O.find_primes([61], [], N)[0][0]], 1)
    rng = np.random.default_rng(9)
    x = np.concatenate([rand_residues(rng, pj, N) for _ in range(4)])  # 2 ct x 2 polys
    d = M.to_dev(x)
    ctx.ntt_forward(d, 4, 40)
    ctx.ntt_inverse(d, 4, 40)
    assert np.array_equal(M.from_dev(d), x)
    octx = O.Ctx(N, pj[:3], [O.find_primes([61], [], N)[0][0]], 1)
    d = M.to_dev(x[: 40 * N])
    ctx.ntt_forward(d, 1, 40)
    got = M.from_dev(d)[: 3 * N]
    assert np.array_equal(got, octx.ntt_fwd(x[: 3 * N], 3))
