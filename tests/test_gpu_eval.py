"""GPU parity of the evaluator ops through the C-ABI: bit-exact against the
reference's own outputs (golden fixtures with captured keys) and against the
oracle on seeded random inputs at larger ring sizes."""
import numpy as np
import pytest

import oracle as O
from helpers import digest, load_json, load_npz, product, rand_residues

pytestmark = pytest.mark.gpu


def _fixture(tag):
    d = load_npz(tag + ".npz")
    outs = load_json(tag + "_outputs.json")
    N = len(d["m1"]) * 2
    L = len(d["q"])
    dnum = len(d["key_relin"]) // (2 * (L + len(d["p"])) * N)
    return d, outs, N, L, dnum


@pytest.mark.parametrize("tag", ["ev10", "ev10b"])
def test_evaluator_matches_reference_outputs(tag):
    d, outs, N, L, dnum = _fixture(tag)
    M = product()
    ctx = M.Context(N, d["q"], d["p"], dnum)
    lv = L - 1
    W = L * N
    dev = M.to_dev
    got = {}
    ct1 = d["ct1"]
    ct2 = d["ct2"]
    keys = {k[4:]: dev(v) for k, v in d.items() if k.startswith("key_")}

    # tensor -> relin -> rescale -> rotate(1)
    b1, a1 = dev(ct1[:W]), dev(ct1[W:])
    b2, a2 = dev(ct2[:W]), dev(ct2[W:])
    c2 = dev(np.zeros(W, np.uint64))
    ctx.tensor(lv, b1, a1, b2, a2, c2)
    got["mul"] = np.concatenate([M.from_dev(b1), M.from_dev(a1), M.from_dev(c2)])
    ctx.relin(lv, b1, a1, c2, keys["relin"])
    got["mul_relin"] = np.concatenate([M.from_dev(b1), M.from_dev(a1)])
    pair = dev(np.concatenate([M.from_dev(b1), M.from_dev(a1)]))
    ctx.rescale(lv, pair, 2)
    r = M.from_dev(pair).reshape(2, W)[:, : (L - 1) * N]
    got["mul_relin_rescale"] = r.ravel()
    rb, ra = dev(r[0]), dev(r[1])
    ctx.rotate(lv - 1, rb, ra, O.galois_rotation(N, 1), keys["rot1"])
    got["chain_rot1"] = np.concatenate([M.from_dev(rb), M.from_dev(ra)])

    b, a, c2 = dev(ct1[:W]), dev(ct1[W:]), dev(np.zeros(W, np.uint64))
    ctx.square(lv, b, a, c2)
    got["square"] = np.concatenate([M.from_dev(b), M.from_dev(a), M.from_dev(c2)])

    for name, r_ in [("rot1_out", 1), ("rot5_out", 5)]:
        if "rot%d" % r_ not in keys:
            continue
        b, a = dev(ct1[:W]), dev(ct1[W:])
        ctx.rotate(lv, b, a, O.galois_rotation(N, r_), keys["rot%d" % r_])
        got[name] = np.concatenate([M.from_dev(b), M.from_dev(a)])
    if "conj" in keys:
        b, a = dev(ct1[:W]), dev(ct1[W:])
        ctx.conjugate(lv, b, a, keys["conj"])
        got["conj_out"] = np.concatenate([M.from_dev(b), M.from_dev(a)])

    x = dev(ct1)
    ctx.rescale(lv, x, 2)
    r1 = M.from_dev(x).reshape(2, W)[:, : (L - 1) * N]
    got["rescale_ct1"] = r1.ravel()
    x = dev(r1.ravel())
    ctx.rescale(lv - 1, x, 2)
    got["rescale2_ct1"] = M.from_dev(x).reshape(2, (L - 1) * N)[:, : (L - 2) * N].ravel()

    x = dev(ct1)
    ctx.add(lv, x, dev(ct2), 2)
    got["add"] = M.from_dev(x)
    x = dev(ct1)
    ctx.sub(lv, x, dev(ct2), 2)
    got["sub"] = M.from_dev(x)
    x = dev(ct1)
    ctx.mul_scalar(lv, x, 2, [(-7) % int(q) for q in d["q"]])
    got["mulint"] = M.from_dev(x)
    scale = float(d["ct1_meta"][2:3].view(np.float64)[0])
    b = dev(ct1[:W])
    ctx.add_scalar(lv, b, [int(round(3 * scale)) % int(q) for q in d["q"]])
    got["addconst"] = np.concatenate([M.from_dev(b), ct1[W:]])
    x = dev(ct1)
    ctx.mul_plain(lv, x, 2, dev(d["pt_m2"]))
    got["mulplain"] = M.from_dev(x)

    b, a = dev(ct1[: 2 * N]), dev(ct1[W: W + 2 * N])
    ctx.rotate(1, b, a, O.galois_rotation(N, 1), keys["rot1"])
    got["rot1_level1"] = np.concatenate([M.from_dev(b), M.from_dev(a)])
    for name, arr in got.items():
        assert digest(arr) == outs[name]["digest"], name


def _random_setup(N, nq, np_, dnum, qbits=50, pbits=55, seed=0):
    p, q = O.find_primes([pbits] * np_, [qbits] * nq, N)
    M = product()
    return M, M.Context(N, q, p, dnum), O.Ctx(N, q, p, dnum), q, p, np.random.default_rng(seed)


def _rand_key(rng, q, p, N, parts):
    basis = list(q) + list(p)
    return np.concatenate([rand_residues(rng, basis, N) for _ in range(2 * parts)])


@pytest.mark.parametrize("N,nq,np_,dnum", [(1 << 12, 7, 3, 3), (1 << 13, 9, 2, 4), (1 << 14, 13, 4, 3),
                                            (1 << 15, 13, 4, 3), (1 << 16, 12, 4, 3)])
def test_keyswitch_matches_oracle(N, nq, np_, dnum):
    M, ctx, octx, q, p, rng = _random_setup(N, nq, np_, dnum, qbits=45 if N < 1 << 15 else 50,
                                            pbits=60, seed=N + nq)
    parts_max = octx.parts(nq - 1)
    key = _rand_key(rng, q, p, N, parts_max)
    dkey = M.to_dev(key)
    for level in sorted({nq - 1, nq // 2, 0}):
        L = level + 1
        kt = rand_residues(rng, q[:L], N)
        b = rand_residues(rng, q[:L], N)
        a = rand_residues(rng, q[:L], N)
        for bm, am in [(0, 1), (0, 0), (1, 1)]:
            ob, oa = octx.keyswitch(level, kt, key, b, a, bm, am)
            db, da, dkt = M.to_dev(b), M.to_dev(a), M.to_dev(kt)
            ctx.keyswitch(level, dkt, dkey, db, da, bm, am)
            assert np.array_equal(M.from_dev(db), ob), (level, bm, am)
            assert np.array_equal(M.from_dev(da), oa), (level, bm, am)
            assert np.array_equal(M.from_dev(dkt), kt)
        # key term aliasing a_out (rotation form)
        ob, oa = octx.keyswitch(level, a, key, b, a, 0, 1)
        db, da = M.to_dev(b), M.to_dev(a)
        ctx.keyswitch(level, da, dkey, db, da, 0, 1)
        assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa)


def test_keyswitch_normal_form_keys():
    N = 1 << 12
    M, ctx, octx, q, p, rng = _random_setup(N, 6, 2, 3, seed=3)
    key = _rand_key(rng, q, p, N, 3)
    kt, b, a = (rand_residues(rng, q, N) for _ in range(3))
    ob, oa = octx.keyswitch(5, kt, key, b, a, 0, 0, montgomery=False)
    db, da = M.to_dev(b), M.to_dev(a)
    ctx.keyswitch(5, M.to_dev(kt), M.to_dev(key), db, da, 0, 0, montgomery=False)
    assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa)


def test_keyswitch_batch_mixed_levels():
    N = 1 << 13
    M, ctx, octx, q, p, rng = _random_setup(N, 10, 3, 3, seed=11)
    keys = [_rand_key(rng, q, p, N, octx.parts(9)) for _ in range(3)]
    dkeys = [M.to_dev(k) for k in keys]
    items, expect, bufs = [], [], []
    for i, level in enumerate([9, 2, 5, 0, 7, 9]):
        kt, b, a = (rand_residues(rng, q[: level + 1], N) for _ in range(3))
        ob, oa = octx.keyswitch(level, kt, keys[i % 3], b, a, 0, 1)
        db, da, dkt = M.to_dev(b), M.to_dev(a), M.to_dev(kt)
        items.append((level, dkt, dkeys[i % 3], db, da, 0, 1))
        expect.append((ob, oa))
        bufs.append((db, da))
    ctx.keyswitch_batch(items)
    for (db, da), (ob, oa) in zip(bufs, expect):
        assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa)


@pytest.mark.parametrize("count,streams", [(37, "1"), (80, "1"), (80, "0")])
def test_keyswitch_batch_groups_and_dependencies(count, streams, monkeypatch):
    """Items at one level are switched together (up to 32 per launch; with 80
    items, two full groups at level 4 that wait for the flush and run on the
    side streams, or on one stream); an item that reads another's output, or
    adds into the same buffers, still sees the result of switching the items
    one by one in order."""
    monkeypatch.setenv("MICKKS_MULTI_STREAM", streams)
    N = 1 << 12
    M, ctx, octx, q, p, rng = _random_setup(N, 6, 2, 2, seed=13)
    keys = [_rand_key(rng, q, p, N, octx.parts(5)) for _ in range(2)]
    dkeys = [M.to_dev(k) for k in keys]
    host, specs = {}, []
    print(f"keyswitch batch: {count} items, streams {streams}", flush=True)
    for i in range(count):
        level = 4 if i % 5 else 2
        for name in ("kt", "b", "a"):
            host[(name, i)] = rand_residues(rng, q[: level + 1], N)
        bm, am = (0, 1) if i % 3 else (1, 0)
        specs.append([level, ("kt", i), i % 2, ("b", i), ("a", i), bm, am])
    # item 20 switches item 10's a output (same level); item 21 adds into item
    # 11's outputs
    specs[20][0], specs[20][1] = specs[10][0], ("a", 10)
    specs[21] = [specs[11][0], ("kt", 21), 1, ("b", 11), ("a", 11), 0, 0]
    host[("kt", 21)] = rand_residues(rng, q[: specs[11][0] + 1], N)
    dev = {k: M.to_dev(v) for k, v in host.items()}
    items = [(lv, dev[kt], dkeys[k], dev[b], dev[a], bm, am) for lv, kt, k, b, a, bm, am in specs]
    for lv, kt, k, b, a, bm, am in specs:
        host[b], host[a] = octx.keyswitch(lv, host[kt], keys[k], host[b], host[a], bm, am)
    ctx.keyswitch_batch(items)
    for name, v in host.items():
        assert np.array_equal(M.from_dev(dev[name]), v), name


@pytest.mark.parametrize("N,nq", [(1 << 12, 6), (1 << 15, 13), (1 << 16, 8)])
def test_rescale_and_elementwise_match_oracle(N, nq):
    M, ctx, octx, q, p, rng = _random_setup(N, nq, 2, 2, qbits=40 + nq % 7, seed=nq)
    lv = nq - 1
    W = nq * N
    x = np.stack([rand_residues(rng, q, N) for _ in range(3)])
    d = M.to_dev(x.ravel())
    ctx.rescale(lv, d, 3)
    assert np.array_equal(M.from_dev(d).reshape(3, W)[:, : lv * N], octx.rescale(lv, x))
    b1, a1, b2, a2 = (rand_residues(rng, q, N) for _ in range(4))
    db1, da1, dc = M.to_dev(b1), M.to_dev(a1), M.to_dev(np.zeros(W, np.uint64))
    ctx.tensor(lv, db1, da1, M.to_dev(b2), M.to_dev(a2), dc)
    ob, oa, oc = octx.tensor(lv, b1, a1, b2, a2)
    assert np.array_equal(M.from_dev(db1), ob) and np.array_equal(M.from_dev(da1), oa)
    assert np.array_equal(M.from_dev(dc), oc)
    for g in [O.galois_rotation(N, 3), O.galois_rotation(N, -17), 2 * N - 1]:
        src = M.to_dev(b1)
        dst = M.to_dev(np.zeros(W, np.uint64))
        ctx.automorphism(nq, g, src, dst)
        assert np.array_equal(M.from_dev(dst), O.apply_automorphism(b1, N, nq, g))


def test_relin_rotate_conjugate_random_n15():
    """Walsh15-shaped ring (N=2^15, 13 + 4 limbs, dnum 3) with random keys."""
    N = 1 << 15
    M, ctx, octx, q, p, rng = _random_setup(N, 13, 4, 3, qbits=42, pbits=61, seed=15)
    key = _rand_key(rng, q, p, N, 3)
    dkey = M.to_dev(key)
    for level in (12, 6):
        b, a, c2 = (rand_residues(rng, q[: level + 1], N) for _ in range(3))
        ob, oa = octx.relin(level, b, a, c2, key)
        db, da = M.to_dev(b), M.to_dev(a)
        ctx.relin(level, db, da, M.to_dev(c2), dkey)
        assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa)
        ob, oa = octx.rotate(level, b, a, 37, key)
        db, da = M.to_dev(b), M.to_dev(a)
        ctx.rotate(level, db, da, O.galois_rotation(N, 37), dkey)
        assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa)
        ob, oa = octx.conjugate(level, b, a, key)
        db, da = M.to_dev(b), M.to_dev(a)
        ctx.conjugate(level, db, da, dkey)
        assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa)


def test_invalid_arguments_raise():
    N = 1 << 12
    M, ctx, octx, q, p, rng = _random_setup(N, 4, 1, 2, seed=1)
    x = M.to_dev(np.zeros(4 * N, np.uint64))
    with pytest.raises(M.MiError, match="ckks_rescale: invalid level"):
        ctx.rescale(0, x, 1)
    with pytest.raises(M.MiError, match="invalid level"):
        ctx.keyswitch(4, x, x, x, x, 0, 1)
    with pytest.raises(M.MiError, match="must differ"):
        ctx.automorphism(4, 5, x, x)


def test_device_samplers_match_philox_and_their_distributions():
    """mi_sample_uniform / mi_sample_gaussian equal the Philox4x32-10 draws
    recomputed on the host (uniform: floor(r q / 2^128) of the 128-bit draw;
    Gaussian: inversion of the cumulative table), are reproducible for one
    (seed, counter) and differ across counters; at N = 2^16 the Gaussian has
    mean ~0 and variance ~sigma^2 and the uniform limbs mean ~q/2."""
    import math
    from helpers import sampler_draw
    N = 1 << 12
    M, ctx, octx, q, p, rng = _random_setup(N, 3, 1, 1, seed=17)
    out = M.to_dev(np.zeros(3 * N, np.uint64))
    ctx.sample_uniform(out, 3, seed=0x1234567890ABCDEF, counter=5)
    got = M.from_dev(out)
    for l in range(3):
        for i in range(0, N, 97):
            lo, hi = sampler_draw(0x1234567890ABCDEF, 5, l, i)
            assert int(got[l * N + i]) == ((hi << 64) + lo) * int(q[l]) >> 128, (l, i)
    again = M.to_dev(np.zeros(3 * N, np.uint64))
    ctx.sample_uniform(again, 3, seed=0x1234567890ABCDEF, counter=5)
    assert np.array_equal(M.from_dev(again), got)
    ctx.sample_uniform(again, 3, seed=0x1234567890ABCDEF, counter=6)
    assert not np.array_equal(M.from_dev(again), got)
    sigma = 3.2
    bound = math.ceil(12.00610553538285 * sigma)
    w = np.cumsum([math.exp(-(x * x) / (2 * sigma * sigma)) for x in range(-bound, bound + 1)])
    cdf = w / w[-1]
    ctx.sample_gaussian(out, 3, sigma, seed=99, counter=1)
    g = M.from_dev(out)
    for i in range(0, N, 31):
        lo, _ = sampler_draw(99, 1, 0xFFFFFFFF, i)
        u = (lo >> 11) * 2.0 ** -53
        e = int(np.searchsorted(cdf, u, side="left")) - bound
        for l in range(3):
            assert int(g[l * N + i]) == e % int(q[l]), (i, l)
    # distributions at N = 2^16
    N2 = 1 << 16
    M2, ctx2, _, q2, _, _ = _random_setup(N2, 2, 1, 1, seed=19)
    big = M2.to_dev(np.zeros(2 * N2, np.uint64))
    ctx2.sample_gaussian(big, 2, sigma, seed=7, counter=3)
    v = M2.from_dev(big)[:N2].astype(np.int64)
    v = np.where(v > int(q2[0]) // 2, v - int(q2[0]), v)
    assert abs(v.mean()) < 0.1 and abs(v.var() / sigma ** 2 - 1) < 0.05
    ctx2.sample_uniform(big, 2, seed=7, counter=4)
    u = M2.from_dev(big)
    for l in range(2):
        m = u[l * N2:(l + 1) * N2].astype(np.float64).mean() / float(q2[l])
        assert abs(m - 0.5) < 0.01


@pytest.mark.parametrize("streams", ["1", "0"])
def test_rotate_batch_matches_oracle(streams, monkeypatch):
    """mi_rotate_batch: 35 independent ciphertexts (two chunks of key
    switches, on the library's side streams and on one stream) rotated in
    place, each equal to the oracle's rotation; batches whose ciphertexts
    overlap are rejected."""
    monkeypatch.setenv("MICKKS_MULTI_STREAM", streams)
    N = 1 << 12
    M, ctx, octx, q, p, rng = _random_setup(N, 6, 2, 3, seed=21)
    level = 4
    L = level + 1
    keys = [_rand_key(rng, q, p, N, octx.parts(5)) for _ in range(5)]
    dkeys = [M.to_dev(k) for k in keys]
    rots = [1, 3, N // 2 - 7, 100, N // 2 - 1] * 7
    host_cts = [(rand_residues(rng, q[:L], N), rand_residues(rng, q[:L], N)) for _ in rots]
    dev = [(M.to_dev(b), M.to_dev(a)) for b, a in host_cts]
    ctx.rotate_batch(level, dev, [O.galois_rotation(N, r) for r in rots], [dkeys[i % 5] for i in range(len(rots))])
    for i, ((b, a), (db, da), r) in enumerate(zip(host_cts, dev, rots)):
        ob, oa = octx.rotate(level, b, a, r, keys[i % 5])
        assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa), i
    with pytest.raises(M.MiError, match="overlap"):
        ctx.rotate_batch(level, [dev[0], (dev[0][0], dev[1][1])], [5, 5], [dkeys[0], dkeys[0]])
