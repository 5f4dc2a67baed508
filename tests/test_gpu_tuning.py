"""The autotuned launch plan: every configuration the tuner can choose (NTT
variants 1 = one limb per workgroup, 2 = two limbs per workgroup, 3 = one
limb with 32-bit LDS exchanges, in every launch-size bucket; key-switch
groups of 8, 16 and 32) gives the same residues as the oracle, and the
context reports the plan it runs with."""
import numpy as np
import pytest

import oracle as O
from helpers import product, rand_residues

pytestmark = pytest.mark.gpu


def test_tuned_plan_is_reported():
    M = product()
    p, q = O.find_primes([60] * 2, [50] * 6, 1 << 13)
    ctx = M.Context(1 << 13, q, p, 3)
    r = ctx.tune_report()
    assert len(r["ntt_variant"]) == 6 and set(r["ntt_variant"]) <= {1, 2, 3}
    assert r["keyswitch_group"] in (8, 16, 32)
    assert r["device"].split()[-1] == "13"


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_every_ntt_plan_matches_the_oracle(variant):
    M = product()
    N = 1 << 15
    p, q = O.find_primes([61] * 4, [36, 61, 42, 42] + [38] * 7 + [50, 50], N)
    ctx = M.Context(N, q, p, 3)
    ctx.tune_set([variant] * 6, 32)
    rng = np.random.default_rng(90 + variant)
    basis = list(q) + list(p)
    # 3 polys x 17 limbs (odd count: the last workgroup of a 2-limb plan has one limb)
    host = np.concatenate([rand_residues(rng, basis, N) for _ in range(3)])
    d = M.to_dev(host)
    ctx.ntt_forward(d, 3, 17)
    fwd = M.from_dev(d)
    octx = O.Ctx(N, q, p, 3)
    for k in (0, 16, 33, 50):  # spot limbs against the oracle (basis index = limb)
        poly, limb = divmod(k, 17)
        want = octx.ntt_fwd_basis(host[k * N:(k + 1) * N], [limb])
        assert np.array_equal(fwd[k * N:(k + 1) * N], want), k
    ctx.ntt_inverse(d, 3, 17)
    assert np.array_equal(M.from_dev(d), host)


@pytest.mark.parametrize("variant", [2, 3])
@pytest.mark.parametrize("logn", [10, 11, 12, 13, 14, 16])
def test_ntt_variants_every_ring_dimension(variant, logn):
    """Each ring dimension has its own column and row group sizes: variants
    2 and 3 against variant 1 (itself pinned by the NTT parity tests)."""
    M = product()
    N = 1 << logn
    p, q = O.find_primes([60] * 2, [50] * 5, N)
    ctx = M.Context(N, q, p, 2)
    rng = np.random.default_rng(logn * 10 + variant)
    basis = list(q) + list(p)
    host = np.concatenate([rand_residues(rng, basis, N) for _ in range(3)])
    outs = []
    for v in (1, variant):
        ctx.tune_set([v] * 6, 32)
        d = M.to_dev(host)
        ctx.ntt_forward(d, 3, 7)
        outs.append(M.from_dev(d))
        ctx.ntt_inverse(d, 3, 7)
        assert np.array_equal(M.from_dev(d), host), v
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("group", [8, 16, 32])
def test_every_keyswitch_group_matches_the_oracle(group):
    M = product()
    N = 1 << 12
    p, q = O.find_primes([55] * 3, [45] * 9, N)
    ctx = M.Context(N, q, p, 3)
    octx = O.Ctx(N, q, p, 3)
    ctx.tune_set([1] * 6, group)
    rng = np.random.default_rng(group)
    basis = list(q) + list(p)
    key = np.concatenate([rand_residues(rng, basis, N) for _ in range(2 * octx.parts(len(q) - 1))])
    dkey = M.to_dev(key)
    level = 6
    items, want = [], []
    for i in range(20):  # 20 switches: groups of 8 leave a partial group
        kt, b, a = (rand_residues(rng, q[:level + 1], N) for _ in range(3))
        ob, oa = octx.keyswitch(level, kt, key, b, a, 0, 1)
        items.append((level, M.to_dev(kt), dkey, M.to_dev(b), M.to_dev(a), 0, 1))
        want.append((ob, oa))
    ctx.keyswitch_batch(items)
    for (_, _, _, db, da, _, _), (ob, oa) in zip(items, want):
        assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa)


@pytest.mark.parametrize("fuse", ["1", "0"])
@pytest.mark.parametrize("mont", [True, False])
def test_folded_key_product_matches_the_oracle(fuse, mont, monkeypatch):
    """Key switches with the key inner product folded into ModDown's NTT
    passes (the default) and as the separate k_mac pass (MICKKS_FUSE_MAC=0),
    at 1, 2 and 3 digits, with add and assign modes, against the oracle
    (the b term the rotations add is covered by the batched rotation
    tests)."""
    monkeypatch.setenv("MICKKS_FUSE_MAC", fuse)
    M = product()
    N = 1 << 13
    p, q = O.find_primes([58] * 3, [45] * 9, N)
    ctx = M.Context(N, q, p, 3)
    octx = O.Ctx(N, q, p, 3)
    rng = np.random.default_rng(7 + int(fuse) + 2 * mont)
    basis = list(q) + list(p)
    key = np.concatenate([rand_residues(rng, basis, N) for _ in range(2 * octx.parts(len(q) - 1))])
    dkey = M.to_dev(key)
    for level, bm, am in ((8, 0, 1), (5, 1, 1), (2, 0, 0)):   # 3, 2 and 1 digits
        kt, b, a = (rand_residues(rng, q[:level + 1], N) for _ in range(3))
        ob, oa = octx.keyswitch(level, kt, key, b, a, bm, am, montgomery=mont)
        db, da = M.to_dev(b), M.to_dev(a)
        ctx.keyswitch(level, M.to_dev(kt), dkey, db, da, bm, am, montgomery=mont)
        assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa), (level, bm, am)
