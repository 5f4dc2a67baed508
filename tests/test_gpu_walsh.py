"""GPU parity of the fused Walsh S-box product sum (mi_walsh_product_sum)
against the oracle's composition of pinned primitives, bit for bit."""
import numpy as np
import pytest

import oracle as O
from helpers import product, rand_residues

pytestmark = pytest.mark.gpu


def _case(N, qbits, nq, level, seed, coeff_max, pt_mont, big_consts):
    M = product()
    p, q = O.find_primes([qbits + 1], [qbits] * nq, N)
    ctx, octx = M.Context(N, q, p, 1), O.Ctx(N, q, p, 1)
    rng = np.random.default_rng(seed)
    qs = q[: level + 1]
    rnd = lambda: np.stack([rand_residues(rng, qs, N), rand_residues(rng, qs, N)])
    low = [rnd() for _ in range(15)]
    high = [rnd() for _ in range(15)]
    pt = rand_residues(rng, qs, N)
    cr = rng.integers(-coeff_max, coeff_max + 1, (16, 16))
    ci = rng.integers(-coeff_max, coeff_max + 1, (16, 16))
    cr[3, :] = 0          # an all-zero row
    ci[7, 2] = 0
    consts = []
    for h in range(16):
        if big_consts:
            vr = int(rng.integers(-2**62, 2**62)) * (2**60) + int(rng.integers(0, 2**60))
            vi = -int(rng.integers(0, 2**40)) * (2**70)
        else:
            vr, vi = int(rng.integers(-64, 65)) * (2**40), int(rng.integers(-64, 65)) * (2**40)
        consts.append((vr, vi))
    consts[4] = (0, 0)
    acc_o, raw_o = octx.walsh_product_sum(level, low, high, cr, ci, consts, pt, pt_mont)
    dev = lambda x: [M.to_dev(x[0]), M.to_dev(x[1])]
    W = (level + 1) * N
    acc = [M.to_dev(np.zeros(W, np.uint64)) for _ in range(2)]
    raw = [M.to_dev(np.zeros(W, np.uint64)) for _ in range(3)]
    ctx.walsh_product_sum(level, [dev(x) for x in low], [dev(x) for x in high], cr, ci, consts,
                          M.to_dev(pt), acc, raw, pt_montgomery=pt_mont)
    for k in range(2):
        assert np.array_equal(M.from_dev(acc[k]), acc_o[k]), ("acc", k)
    for k in range(3):
        assert np.array_equal(M.from_dev(raw[k]), raw_o[k]), ("raw", k)


@pytest.mark.parametrize("N,qbits,nq,level,coeff_max,pt_mont,big", [
    (1 << 12, 50, 4, 3, 64, True, False),      # Walsh-table sized coefficients
    (1 << 11, 61, 3, 2, 32767, False, True),   # extreme coefficients, 61-bit moduli, 128-bit constants
    (1 << 13, 40, 6, 5, 8, True, True),
])
def test_walsh_product_sum_matches_oracle(N, qbits, nq, level, coeff_max, pt_mont, big):
    print(f"walsh N={N} qbits={qbits} level={level} coeff_max={coeff_max} pt_mont={pt_mont}", flush=True)
    _case(N, qbits, nq, level, 100 + level, coeff_max, pt_mont, big)


def test_walsh_product_sum_rejects_bad_arguments():
    import torch
    M = product()
    N = 1 << 11
    p, q = O.find_primes([51], [50, 50], N)
    ctx = M.Context(N, q, p, 1)
    z = lambda: M.to_dev(np.zeros(2 * N, np.uint64))
    low = [(z(), z()) for _ in range(15)]
    high = [(z(), z()) for _ in range(15)]
    acc, raw, pt = [z(), z()], [z(), z(), z()], z()
    cr = np.zeros((16, 16), np.int16)
    consts = [(0, 0)] * 16
    with pytest.raises(M.MiError, match="alias"):
        ctx.walsh_product_sum(1, low, high, cr, cr, consts, pt, [low[0][0], acc[1]], raw)
    bad = cr.copy()
    bad[1, 1] = -32768
    with pytest.raises(M.MiError, match="coefficient"):
        ctx.walsh_product_sum(1, low, high, bad, cr, consts, pt, acc, raw)
    with pytest.raises(M.MiError, match="level"):
        ctx.walsh_product_sum(2, low, high, cr, cr, consts, pt, acc, raw)
    torch.cuda.synchronize()
