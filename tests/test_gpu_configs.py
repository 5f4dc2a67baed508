"""The batched kernels at the full shapes of the benchmark configurations,
checked bit for bit against the oracle.

* Batched forward + inverse NTT (N = 2^16, 64 ciphertexts x 2 polys x 40
  limbs of 60-bit primes, 5120 limb-NTTs, one launch pair per direction):
  the round trip gives back every residue, and the forward outputs of
  limbs of the first and the last poly equal the oracle's NTT
  (ntt.cpp:235-297).
* Batched hybrid key switching (N = 2^16, Q = 30 x 50-bit, P = 10 x 50-bit,
  dnum 3 so alpha = 10 and K = 10; 128 items at seeded levels 7..29, 16
  keys, item i on key i mod 16; ckks_keyswitch.cpp:1330-1641): every item
  of a seeded subset that covers each level present (including 29, 19 and
  7) equals the oracle's key switch limb for limb, and a second run of the
  whole batch from the same inputs gives the same limbs for every item."""
import numpy as np
import pytest

import oracle as O
from helpers import load_json, product

pytestmark = pytest.mark.gpu


def _rand_limbs(torch, q, N, count, seed):
    """count polys of len(q) limbs, residues uniform in [0, q_i), on the GPU."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    out = torch.empty((count, len(q), N), dtype=torch.int64, device="cuda")
    for i, qi in enumerate(q):
        out[:, i, :] = torch.randint(0, int(qi), (count, N), generator=g, device="cuda", dtype=torch.int64)
    return out


def test_ntt_config2_full_batch_round_trip_and_oracle_limbs():
    import torch
    M = product()
    N, L, polys = 1 << 16, 40, 128
    q = load_json("primes.json")["ntt16"]["q"]
    assert len(q) == L and all(int(v).bit_length() == 60 for v in q)
    p = O.find_primes([61], [], N)[0]
    ctx = M.Context(N, q, [p[0]], 1)
    x = _rand_limbs(torch, q, N, polys, seed=2024)
    d = x.clone()
    print(f"config 2: {polys * L} limb-NTTs in one launch pair per direction", flush=True)
    ctx.ntt_forward(d.reshape(-1), polys, L)
    torch.cuda.synchronize()
    # oracle limbs of the first and the last poly (limbs 0, 1, 38, 39 of each)
    octx = O.Ctx(N, q, [p[0]], 1)
    for poly in (0, polys - 1):
        src = M.from_dev(x[poly].reshape(-1))
        want = octx.ntt_fwd(src, L)
        got = M.from_dev(d[poly].reshape(-1))
        for limb in (0, 1, L - 2, L - 1):
            sl = slice(limb * N, (limb + 1) * N)
            assert np.array_equal(got[sl], want[sl]), (poly, limb)
    ctx.ntt_inverse(d.reshape(-1), polys, L)
    torch.cuda.synchronize()
    assert torch.equal(d, x)


def test_keyswitch_config3_full_batch_matches_oracle():
    import torch
    M = product()
    N, dnum, nkeys, nct = 1 << 16, 3, 16, 128
    p, q = O.find_primes([50] * 10, [50] * 30, N)
    ctx = M.Context(N, q, p, dnum)
    octx = O.Ctx(N, q, p, dnum)
    assert octx.parts(29) == 3
    basis = list(q) + list(p)
    rng = np.random.default_rng(3)
    levels = rng.integers(7, 30, nct)
    levels[:3] = [29, 19, 7]
    keys_dev = [_rand_limbs(torch, basis, N, 2 * dnum, seed=100 + k).reshape(-1) for k in range(nkeys)]
    items, bufs = [], []
    for i, lv in enumerate(levels):
        Lc = int(lv) + 1
        kt = _rand_limbs(torch, q[:Lc], N, 1, seed=1000 + i).reshape(-1)
        b = _rand_limbs(torch, q[:Lc], N, 1, seed=2000 + i).reshape(-1)
        a = _rand_limbs(torch, q[:Lc], N, 1, seed=3000 + i).reshape(-1)
        bufs.append((kt, b.clone(), a.clone(), b, a))
        items.append((int(lv), kt, keys_dev[i % nkeys], b, a, 0, 1))
    print(f"config 3: {nct} key switches, levels {sorted(set(int(v) for v in levels))}", flush=True)
    ctx.keyswitch_batch(items)
    torch.cuda.synchronize()
    # one item per level present, the first one drawn at that level
    check = {}
    for i, lv in enumerate(levels):
        check.setdefault(int(lv), i)
    assert {29, 19, 7} <= set(check)
    key_host = {}
    for lv, i in sorted(check.items()):
        k = i % nkeys
        if k not in key_host:
            key_host[k] = M.from_dev(keys_dev[k])
        kt, b0, a0, b, a = bufs[i]
        ob, oa = octx.keyswitch(lv, M.from_dev(kt), key_host[k], M.from_dev(b0), M.from_dev(a0), 0, 1)
        assert np.array_equal(M.from_dev(b), ob), (i, lv)
        assert np.array_equal(M.from_dev(a), oa), (i, lv)
    print(f"checked {len(check)} items against the oracle", flush=True)
    # the whole batch again from the same inputs: identical outputs (no item
    # depends on launch grouping or stream order)
    first = [(b.clone(), a.clone()) for (_, _, _, b, a) in bufs]
    for (kt, b0, a0, b, a) in bufs:
        b.copy_(b0)
        a.copy_(a0)
    ctx.keyswitch_batch(items)
    torch.cuda.synchronize()
    for (b1, a1), (_, _, _, b, a) in zip(first, bufs):
        assert torch.equal(b1, b) and torch.equal(a1, a)
