"""GPU parity of hoisting, the QP-basis key switches, the BSGS linear-transform
stage, ModRaise and the level-0 q0p0 switch, through the C-ABI.

Bit-exact against the reference's own outputs (lt10 / lq10 / bt10 golden
fixtures, captured keys) and against the oracle on seeded random inputs at larger ring
sizes, with random residues standing in for keys and diagonals."""
import numpy as np
import pytest

import oracle as O
from helpers import digest, load_json, load_npz, lt_diags, product, rand_residues

pytestmark = pytest.mark.gpu


def _zeros(M, words):
    return M.to_dev(np.zeros(words, np.uint64))


def test_qp_path_matches_reference_lt10():
    d = load_npz("lt10.npz")
    o = load_json("lt10_outputs.json")
    M = product()
    N = 1 << 10
    q, p = d["q"], d["p"]
    L, K = len(q), len(p)
    T = L + K
    ctx = M.Context(N, q, p, 2)
    lv = L - 1
    b, a = d["ct"].reshape(2, L * N)
    da, db = M.to_dev(a), M.to_dev(b)
    keys = {1: M.to_dev(d["key_rot1"]), 2: M.to_dev(d["key_rot2"]), 4: M.to_dev(d["key_rot4"])}

    dig = _zeros(M, ctx.parts(lv) * T * N)
    ctx.hoist(lv, da, dig)
    assert digest(M.from_dev(dig)) == o["hoist_digits"]

    hb, ha = M.to_dev(b), _zeros(M, L * N)
    ctx.keyswitch_hoisted(lv, dig, keys[1], hb, ha, 0, 1, galois=O.galois_rotation(N, 1))
    assert digest(np.concatenate([M.from_dev(hb), M.from_dev(ha)])) == o["hoisted_rot1"]

    qb, qa = _zeros(M, T * N), _zeros(M, T * N)
    ctx.keyswitch_hoisted_qp(lv, dig, keys[2], db, qb, qa, O.galois_rotation(N, 2))
    assert digest(np.concatenate([M.from_dev(qb), M.from_dev(qa)])) == o["hoisted_qp_rot2"]

    sb, sa = _zeros(M, T * N), _zeros(M, T * N)
    ctx.keyswitch_qp(lv, da, keys[4], sb, sa)
    assert digest(np.concatenate([M.from_dev(sb), M.from_dev(sa)])) == o["qp_from_key_term_rot4"]

    diags = lt_diags(N, np.concatenate([q, p]), 6).reshape(6, T * N)
    dd = [M.to_dev(x) for x in diags]
    ao, bo = _zeros(M, L * N), _zeros(M, L * N)
    ctx.linear_transform_qp(lv, da, db, o["diag_indices"], dd, o["baby_steps"], keys, ao, bo)
    out = np.concatenate([M.from_dev(bo), M.from_dev(ao)])
    assert digest(out) == o["lt"]
    pair = M.to_dev(out)
    ctx.rescale(lv, pair, 2)
    r = M.from_dev(pair).reshape(2, L * N)[:, : (L - 1) * N]
    assert digest(r.ravel()) == o["lt_rescaled"]
    # inputs untouched
    assert np.array_equal(M.from_dev(da), a) and np.array_equal(M.from_dev(db), b)


def test_q_evaluator_matches_reference_lq10():
    """The Q-evaluator BSGS stage (owned stage through a plan with
    Montgomery-form diagonals; the stage-view API with normal-form diagonals
    and a rescale) reproduces the reference's outputs; a batch of the same
    ciphertext three times (once in place) gives the same limbs."""
    d = load_npz("lq10.npz")
    o = load_json("lq10_outputs.json")
    M = product()
    N = 1 << 10
    q, p = d["q"], d["p"]
    L = len(q)
    lv = L - 1
    ctx = M.Context(N, q, p, 2)
    b, a = d["ct"].reshape(2, L * N)
    keys = {r: M.to_dev(d[f"key_rot{r}"]) for r in (1, 2, 3, 6)}
    diags = lt_diags(N, q, 14, seed_base=o["ltq"]["diag_seed_base"]).reshape(14, L * N)
    for name in ("ltq", "bsgs"):
        e = o[name]
        dd = [M.to_dev(diags[t]) for t in e["terms"]]
        da, db = M.to_dev(a), M.to_dev(b)
        ao, bo = _zeros(M, L * N), _zeros(M, L * N)
        ctx.linear_transform_q(lv, da, db, e["diag_indices"], dd, e["baby_steps"], keys, ao, bo,
                               pt_montgomery=e["montgomery"])
        out = np.concatenate([M.from_dev(bo), M.from_dev(ao)])
        if e["rescale"]:
            pair = M.to_dev(out)
            ctx.rescale(lv, pair, 2)
            out = M.from_dev(pair).reshape(2, L * N)[:, : (L - 1) * N].ravel()
        assert digest(out) == e["digest"], name
        if name == "ltq":
            want = out
            cts = [(M.to_dev(a), M.to_dev(b)) for _ in range(3)]
            outs = [(_zeros(M, L * N), _zeros(M, L * N)), cts[1], (_zeros(M, L * N), _zeros(M, L * N))]
            ctx.linear_transform_qp_batch(lv, cts, e["diag_indices"], dd, e["baby_steps"], keys, outs,
                                          pt_montgomery=True, q_evaluator=True)
            for ao, bo in outs:
                assert np.array_equal(np.concatenate([M.from_dev(bo), M.from_dev(ao)]), want)
        assert np.array_equal(M.from_dev(da), a) and np.array_equal(M.from_dev(db), b)
    with pytest.raises(M.MiError, match="duplicate diagonal"):
        ctx.linear_transform_q(lv, da, db, [1, 1], [dd[0], dd[1]], 3, keys, ao, bo)
    with pytest.raises(M.MiError, match="exceeds slot count"):
        ctx.linear_transform_q(lv, da, db, [0, 9], [dd[0], dd[1]], 3, keys, ao, bo, slots=8)


def test_bootstrap_pieces_match_reference_bt10():
    d = load_npz("bt10.npz")
    o = load_json("bt10_outputs.json")
    M = product()
    N = 1 << 10
    ctx = M.Context(N, d["q"], d["p"], 2)
    b, a = d["ct0"].reshape(2, N)
    db, da = M.to_dev(b), M.to_dev(a)
    ctx.q0p0_keyswitch(db, da, M.to_dev(d["key_d2s"]))
    sb, sa = M.from_dev(db), M.from_dev(da)
    assert digest(np.concatenate([sb, sa])) == o["q0p0"]
    lv = o["modraise_level"]
    rb, ra = _zeros(M, (lv + 1) * N), _zeros(M, (lv + 1) * N)
    ctx.modraise(lv, db, rb)
    ctx.modraise(lv, da, ra)
    assert digest(np.concatenate([M.from_dev(rb), M.from_dev(ra)])) == o["modraise"]


def _setup(N, nq, np_, dnum, seed, qbits=45, pbits=60):
    p, q = O.find_primes([pbits] * np_, [qbits] * nq, N)
    M = product()
    return M, M.Context(N, q, p, dnum), O.Ctx(N, q, p, dnum), q, p, np.random.default_rng(seed)


def _key(rng, q, p, N, parts):
    basis = list(q) + list(p)
    return np.concatenate([rand_residues(rng, basis, N) for _ in range(2 * parts)])


@pytest.mark.parametrize("N,nq,np_,dnum,mont", [(1 << 13, 9, 3, 3, True), (1 << 14, 12, 4, 4, False),
                                                 (1 << 16, 10, 3, 2, True)])
def test_hoisted_and_qp_switches_match_oracle(N, nq, np_, dnum, mont):
    M, ctx, octx, q, p, rng = _setup(N, nq, np_, dnum, seed=N + nq)
    key = _key(rng, q, p, N, octx.parts(nq - 1))
    dkey = M.to_dev(key)
    for level in sorted({nq - 1, nq // 2, 0}):
        L, T = level + 1, level + 1 + np_
        qp = list(q[:L]) + list(p)
        a = rand_residues(rng, q[:L], N)
        b = rand_residues(rng, q[:L], N)
        da, db = M.to_dev(a), M.to_dev(b)
        dig_o = octx.hoist(level, a)
        dig = _zeros(M, ctx.parts(level) * T * N)
        ctx.hoist(level, da, dig)
        assert np.array_equal(M.from_dev(dig), dig_o), level
        g3 = O.galois_rotation(N, 3)
        for bm, am, g in [(0, 1, g3), (0, 0, 0), (1, 1, g3)]:
            ob, oa = octx.keyswitch_hoisted(level, dig_o, key, b, a, bm, am, galois=g, montgomery=mont)
            hb, ha = M.to_dev(b), M.to_dev(a)
            ctx.keyswitch_hoisted(level, dig, dkey, hb, ha, bm, am, galois=g, montgomery=mont)
            assert np.array_equal(M.from_dev(hb), ob), (level, bm, am, g)
            assert np.array_equal(M.from_dev(ha), oa), (level, bm, am, g)
        ob, oa = octx.keyswitch_hoisted_qp_rotated(level, dig_o, key, b, g3, montgomery=mont)
        qb, qa = _zeros(M, T * N), _zeros(M, T * N)
        ctx.keyswitch_hoisted_qp(level, dig, dkey, db, qb, qa, g3, montgomery=mont)
        assert np.array_equal(M.from_dev(qb), ob) and np.array_equal(M.from_dev(qa), oa), level
        ob, oa = octx.keyswitch_qp_from_key_term(level, a, key, montgomery=mont)
        ctx.keyswitch_qp(level, da, dkey, qb, qa, montgomery=mont)
        assert np.array_equal(M.from_dev(qb), ob) and np.array_equal(M.from_dev(qa), oa), level
        x = rand_residues(rng, qp, N)
        dx = M.to_dev(x)
        ctx.moddown_qp(level, dx)
        assert np.array_equal(M.from_dev(dx)[: L * N], octx.moddown_qp_to_q(level, x)), level


def test_rotate_hoisted_batch_matches_single_rotations_and_reference():
    """mi_rotate_hoisted_batch equals one mi_keyswitch_hoisted (b Add, a
    Assign, galois) per rotation, bit for bit, for 37 rotations (more than
    one launch sequence takes); on lt10 it reproduces the reference's hoisted
    rotation by 1."""
    d = load_npz("lt10.npz")
    o = load_json("lt10_outputs.json")
    M = product()
    N = 1 << 10
    q, p = d["q"], d["p"]
    L, K = len(q), len(p)
    ctx = M.Context(N, q, p, 2)
    lv = L - 1
    b, a = d["ct"].reshape(2, L * N)
    dig = _zeros(M, ctx.parts(lv) * (L + K) * N)
    ctx.hoist(lv, M.to_dev(a), dig)
    db = M.to_dev(b)
    key1 = M.to_dev(d["key_rot1"])
    outs = [(_zeros(M, L * N), _zeros(M, L * N))]
    ctx.rotate_hoisted_batch(lv, dig, db, [O.galois_rotation(N, 1)], [key1], outs)
    assert digest(np.concatenate([M.from_dev(outs[0][0]), M.from_dev(outs[0][1])])) == o["hoisted_rot1"]
    # many rotations with random keys at a larger ring
    M, ctx, octx, q, p, rng = _setup(1 << 13, 9, 3, 3, seed=29)
    N, lv = 1 << 13, 7
    L, T = lv + 1, lv + 1 + len(p)
    a = rand_residues(rng, q[:L], N)
    b = rand_residues(rng, q[:L], N)
    dig = _zeros(M, ctx.parts(lv) * T * N)
    ctx.hoist(lv, M.to_dev(a), dig)
    keys = [M.to_dev(_key(rng, q, p, N, octx.parts(len(q) - 1))) for _ in range(5)]
    rots = [1 + 3 * i for i in range(37)]
    gal = [O.galois_rotation(N, r) for r in rots]
    ks = [keys[i % 5] for i in range(37)]
    outs = [(_zeros(M, L * N), _zeros(M, L * N)) for _ in rots]
    db = M.to_dev(b)
    ctx.rotate_hoisted_batch(lv, dig, db, gal, ks, outs)
    for i, (g, k) in enumerate(zip(gal, ks)):
        hb, ha = M.to_dev(b), _zeros(M, L * N)
        ctx.keyswitch_hoisted(lv, dig, k, hb, ha, 0, 1, galois=g)
        assert np.array_equal(M.from_dev(outs[i][0]), M.from_dev(hb)), i
        assert np.array_equal(M.from_dev(outs[i][1]), M.from_dev(ha)), i


def _lt_case(M, ctx, octx, q, p, rng, N, level, diag_idx, bs, mont=True, pt_mont=True, in_place=False):
    L, T = level + 1, level + 1 + len(p)
    qp = list(q[:L]) + list(p)
    parts = octx.parts(len(q) - 1)
    a = rand_residues(rng, q[:L], N)
    b = rand_residues(rng, q[:L], N)
    diags = [rand_residues(rng, qp, N) for _ in diag_idx]
    rots = set()
    for dgi in diag_idx:
        if dgi % bs:
            rots.add(dgi % bs)
        if dgi // bs:
            rots.add((dgi // bs) * bs)
    keys = {r: _key(rng, q, p, N, parts) for r in sorted(rots)}
    oa, ob = octx.lt_stage_qp(level, a, b, diag_idx, np.concatenate(diags), bs, keys, pt_mont=pt_mont,
                              montgomery=mont)
    da, db = M.to_dev(a), M.to_dev(b)
    dkeys = {r: M.to_dev(k) for r, k in keys.items()}
    dd = [M.to_dev(x) for x in diags]
    if in_place:
        ao, bo = da, db
    else:
        ao, bo = _zeros(M, L * N), _zeros(M, L * N)
    ctx.linear_transform_qp(level, da, db, diag_idx, dd, bs, dkeys, ao, bo, pt_montgomery=pt_mont,
                            montgomery=mont)
    assert np.array_equal(M.from_dev(ao), oa), (level, diag_idx, bs)
    assert np.array_equal(M.from_dev(bo), ob), (level, diag_idx, bs)


@pytest.mark.parametrize("diag_idx,bs", [
    ([0, 1, 3, 5, 6, 9, 12, 14], 4),   # every giant step, diagonal 0 present
    ([2, 7, 13], 4),                    # no direct term
    ([5, 9], 4),                        # giant 0 empty: first output from a key switch
    ([0], 1),                           # single plaintext product
    ([1, 2, 3, 4, 5, 6, 7, 8], 8),      # babies only
])
def test_linear_transform_matches_oracle(diag_idx, bs):
    M, ctx, octx, q, p, rng = _setup(1 << 13, 9, 3, 3, seed=len(diag_idx) * 7 + bs)
    _lt_case(M, ctx, octx, q, p, rng, 1 << 13, 8, diag_idx, bs)


def test_linear_transform_modes_levels_and_aliasing():
    M, ctx, octx, q, p, rng = _setup(1 << 14, 12, 4, 4, seed=11)
    _lt_case(M, ctx, octx, q, p, rng, 1 << 14, 6, [0, 3, 4, 10, 11], 4, mont=False, pt_mont=False)
    _lt_case(M, ctx, octx, q, p, rng, 1 << 14, 11, [0, 1, 6, 7], 3, in_place=True)
    _lt_case(M, ctx, octx, q, p, rng, 1 << 14, 0, [0, 2, 5], 2)


def test_linear_transform_many_terms_per_giant():
    """More terms in one giant step than one inner-product launch takes."""
    M, ctx, octx, q, p, rng = _setup(1 << 12, 6, 2, 2, seed=5)
    _lt_case(M, ctx, octx, q, p, rng, 1 << 12, 5, list(range(40)) + [64, 70], 64)


def _lt_batch_case(M, ctx, octx, q, p, rng, N, level, diag_idx, bs, n, in_place=()):
    L = level + 1
    qp = list(q[:L]) + list(p)
    parts = octx.parts(len(q) - 1)
    diags = [rand_residues(rng, qp, N) for _ in diag_idx]
    rots = {d % bs for d in diag_idx if d % bs} | {(d // bs) * bs for d in diag_idx if d // bs}
    keys = {r: _key(rng, q, p, N, parts) for r in sorted(rots)}
    dkeys = {r: M.to_dev(k) for r, k in keys.items()}
    dd = [M.to_dev(x) for x in diags]
    cts, outs, want = [], [], []
    for i in range(n):
        a = rand_residues(rng, q[:L], N)
        b = rand_residues(rng, q[:L], N)
        want.append(octx.lt_stage_qp(level, a, b, diag_idx, np.concatenate(diags), bs, keys))
        da, db = M.to_dev(a), M.to_dev(b)
        cts.append((da, db))
        outs.append((da, db) if i in in_place else (_zeros(M, L * N), _zeros(M, L * N)))
    ctx.linear_transform_qp_batch(level, cts, diag_idx, dd, bs, dkeys, outs)
    for i, ((ao, bo), (oa, ob)) in enumerate(zip(outs, want)):
        assert np.array_equal(M.from_dev(ao), oa), (level, i)
        assert np.array_equal(M.from_dev(bo), ob), (level, i)


def test_linear_transform_batch_matches_oracle():
    """Several ciphertexts through one stage in one call, some in place."""
    M, ctx, octx, q, p, rng = _setup(1 << 13, 9, 3, 3, seed=41)
    _lt_batch_case(M, ctx, octx, q, p, rng, 1 << 13, 8, [0, 1, 3, 5, 6, 9, 12, 14], 4, 5, in_place={1, 4})
    _lt_batch_case(M, ctx, octx, q, p, rng, 1 << 13, 4, [5, 9, 10], 4, 3)


def test_linear_transform_batch_beyond_one_launch():
    """More ciphertexts than one launch sequence takes (32), and more baby
    steps than the inner-product kernel holds in registers (16)."""
    M, ctx, octx, q, p, rng = _setup(1 << 12, 6, 2, 2, seed=43)
    _lt_batch_case(M, ctx, octx, q, p, rng, 1 << 12, 3, [0, 2, 7, 9], 4, 35, in_place={0, 33})
    _lt_batch_case(M, ctx, octx, q, p, rng, 1 << 12, 5, list(range(20)) + [32, 41], 32, 3)


@pytest.mark.parametrize("N,nq,mont", [(1 << 14, 10, True), (1 << 16, 24, False)])
def test_modraise_and_q0p0_match_oracle(N, nq, mont):
    M, ctx, octx, q, p, rng = _setup(N, nq, 2, 3, seed=N + 1, qbits=50, pbits=58)
    for target in (1, nq - 1):
        x = rng.integers(0, int(q[0]), N, dtype=np.uint64)
        out = _zeros(M, (target + 1) * N)
        ctx.modraise(target, M.to_dev(x), out)
        assert np.array_equal(M.from_dev(out), octx.modraise(target, x)), target
    # in place: out aliases the q0 limb
    x = rng.integers(0, int(q[0]), N, dtype=np.uint64)
    buf = M.to_dev(np.concatenate([x, np.zeros((nq - 1) * N, np.uint64)]))
    ctx.modraise(nq - 1, buf, buf)
    assert np.array_equal(M.from_dev(buf), octx.modraise(nq - 1, x))
    b = rng.integers(0, int(q[0]), N, dtype=np.uint64)
    a = rng.integers(0, int(q[0]), N, dtype=np.uint64)
    key = np.concatenate([rng.integers(0, int(m), N, dtype=np.uint64) for m in (q[0], p[0], q[0], p[0])])
    ob, oa = octx.q0p0_keyswitch(b, a, key, montgomery=mont)
    db, da = M.to_dev(b), M.to_dev(a)
    ctx.q0p0_keyswitch(db, da, M.to_dev(key), montgomery=mont)
    assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa)


def test_qp_entry_points_reject_bad_arguments():
    M, ctx, octx, q, p, rng = _setup(1 << 12, 4, 2, 2, seed=3)
    N, L, T = 1 << 12, 4, 6
    a, b = M.to_dev(rand_residues(rng, q, N)), M.to_dev(rand_residues(rng, q, N))
    key = M.to_dev(_key(rng, q, p, N, 2))
    dig = _zeros(M, 2 * T * N)
    qb, qa = _zeros(M, T * N), _zeros(M, T * N)
    pt = M.to_dev(rand_residues(rng, list(q) + list(p), N))
    with pytest.raises(M.MiError, match="invalid level"):
        ctx.hoist(4, a, dig)
    with pytest.raises(M.MiError, match="odd"):
        ctx.keyswitch_hoisted_qp(3, dig, key, b, qb, qa, 4)
    with pytest.raises(M.MiError, match="missing rotation key"):
        ctx.linear_transform_qp(3, a, b, [0, 1], [pt, pt], 2, {}, qa, qb)
    with pytest.raises(M.MiError, match="non-negative"):
        ctx.linear_transform_qp(3, a, b, [-1], [pt], 2, {}, qa, qb)
    with pytest.raises(M.MiError, match="slot range"):
        ctx.linear_transform_qp(3, a, b, [8], [pt], 2, {8: key}, qa, qb, slots=4)
    with pytest.raises(M.MiError, match="target level"):
        ctx.modraise(0, a, qa)
    with pytest.raises(M.MiError, match="combine mode"):
        ctx.keyswitch_hoisted(3, dig, key, qb, qa, 2, 1)


def _parr(ts):
    import ctypes
    return (ctypes.c_void_p * len(ts))(*[ptr for ptr in ts])


def test_batched_bootstrap_pieces_match_oracle():
    """The batched forms the bootstrap and the mask-packed unpack use, each
    against the oracle's single-ciphertext form: mi_modraise_batch over 34
    polys (two launch sequences; one output aliases its input),
    mi_q0p0_keyswitch_batch over 35 level-0 ciphertexts, and mi_hoist_batch
    over 33 ciphertexts with a padded digit stride."""
    import torch
    N, nq = 1 << 12, 6
    M, ctx, octx, q, p, rng = _setup(N, nq, 2, 3, seed=91, qbits=50, pbits=58)
    lib = M.lib()
    s = M._stream(torch.cuda.current_stream())
    target = nq - 1
    xs = [rng.integers(0, int(q[0]), N, dtype=np.uint64) for _ in range(34)]
    ins = [M.to_dev(np.concatenate([x, np.zeros(target * N, np.uint64)])) for x in xs]
    outs = [ins[i] if i == 5 else _zeros(M, (target + 1) * N) for i in range(34)]
    print("modraise batch of 34", flush=True)
    M.check(lib.mi_modraise_batch(ctx.h, target, 34, _parr([M.ptr(t) for t in ins]),
                                  _parr([M.ptr(t) for t in outs]), s))
    for i, x in enumerate(xs):
        assert np.array_equal(M.from_dev(outs[i]), octx.modraise(target, x)), i
    for mont in (True, False):
        key = np.concatenate([rng.integers(0, int(m), N, dtype=np.uint64) for m in (q[0], p[0], q[0], p[0])])
        bs = [rng.integers(0, int(q[0]), N, dtype=np.uint64) for _ in range(35)]
        as_ = [rng.integers(0, int(q[0]), N, dtype=np.uint64) for _ in range(35)]
        db, da = [M.to_dev(b) for b in bs], [M.to_dev(a) for a in as_]
        print("q0p0 batch of 35, montgomery", mont, flush=True)
        M.check(lib.mi_q0p0_keyswitch_batch(ctx.h, 35, _parr([M.ptr(t) for t in db]), _parr([M.ptr(t) for t in da]),
                                            M.ptr(M.to_dev(key)), int(mont), s))
        for i in range(35):
            ob, oa = octx.q0p0_keyswitch(bs[i], as_[i], key, montgomery=mont)
            assert np.array_equal(M.from_dev(db[i]), ob) and np.array_equal(M.from_dev(da[i]), oa), (mont, i)
    level = 4
    L = level + 1
    parts = octx.parts(level)
    dw = parts * (L + len(p)) * N
    stride = dw + 3 * N
    a_in = [rand_residues(rng, q[:L], N) for _ in range(33)]
    da = [M.to_dev(a) for a in a_in]
    dig = _zeros(M, 33 * stride)
    print("hoist batch of 33", flush=True)
    M.check(lib.mi_hoist_batch(ctx.h, level, 33, _parr([M.ptr(t) for t in da]), M.ptr(dig), stride, s))
    got = M.from_dev(dig)
    for i, a in enumerate(a_in):
        assert np.array_equal(got[i * stride:i * stride + dw], octx.hoist(level, a)), i
    # a digit stride below one ciphertext's digits is refused
    assert lib.mi_hoist_batch(ctx.h, level, 2, _parr([M.ptr(t) for t in da[:2]]), M.ptr(dig), dw - 1, s) < 0


@pytest.mark.parametrize("fuse", ["1", "0"])
def test_rotate_hoisted_items_per_member_digits(fuse, monkeypatch):
    """mi_rotate_hoisted_items over 35 members' own hoists (mi_hoist_batch,
    one digit stride apart) at one rotation, and over one shared hoist
    (stride 0) at several rotations, each equal to mi_keyswitch_hoisted of
    that member (b Add, a Assign, galois), with the key product folded into
    ModDown and as its own pass."""
    import torch
    monkeypatch.setenv("MICKKS_FUSE_MAC", fuse)
    M, ctx, octx, q, p, rng = _setup(1 << 12, 9, 3, 3, seed=61)
    lib = M.lib()
    s = M._stream(torch.cuda.current_stream())
    N, lv = 1 << 12, 7
    L, T = lv + 1, lv + 1 + len(p)
    dw = ctx.parts(lv) * T * N
    B = 35
    a_in = [M.to_dev(rand_residues(rng, q[:L], N)) for _ in range(B)]
    b_in = [M.to_dev(rand_residues(rng, q[:L], N)) for _ in range(B)]
    dig = _zeros(M, B * dw)
    M.check(lib.mi_hoist_batch(ctx.h, lv, B, _parr([M.ptr(t) for t in a_in]), M.ptr(dig), dw, s))
    key = M.to_dev(_key(rng, q, p, N, octx.parts(len(q) - 1)))
    g = O.galois_rotation(N, 5)
    outs = [(_zeros(M, L * N), _zeros(M, L * N)) for _ in range(B)]
    gal = (ctypes_u64(B))(*([g] * B))
    print("rotate_hoisted_items: 35 members, fuse", fuse, flush=True)
    M.check(lib.mi_rotate_hoisted_items(ctx.h, lv, B, M.ptr(dig), dw, 0, _parr([M.ptr(t) for t in b_in]), gal,
                                        _parr([M.ptr(key)] * B), _parr([M.ptr(o[0]) for o in outs]),
                                        _parr([M.ptr(o[1]) for o in outs]), 1, s))
    for m in range(B):
        one = _zeros(M, dw)
        ctx.hoist(lv, a_in[m], one)
        hb, ha = M.to_dev(M.from_dev(b_in[m])), _zeros(M, L * N)
        ctx.keyswitch_hoisted(lv, one, key, hb, ha, 0, 1, galois=g)
        assert np.array_equal(M.from_dev(outs[m][0]), M.from_dev(hb)), m
        assert np.array_equal(M.from_dev(outs[m][1]), M.from_dev(ha)), m


def ctypes_u64(n):
    import ctypes
    return ctypes.c_uint64 * n


@pytest.mark.parametrize("streams", ["1", "0"])
def test_rotate_hoisted_items_rotation_major_over_members(streams, monkeypatch):
    """mi_rotate_hoisted_items over 3 rotations x 35 members (105 items,
    digit period 35: chunks of 32 and 3 inside each rotation's run of
    members), the chunks on the library's side streams and on one stream;
    every output equal to mi_keyswitch_hoisted of its member and rotation,
    and again under the NTT probe (which keeps every chunk on one stream)."""
    import torch
    monkeypatch.setenv("MICKKS_MULTI_STREAM", streams)
    M, ctx, octx, q, p, rng = _setup(1 << 12, 9, 3, 3, seed=71)
    lib = M.lib()
    s = M._stream(torch.cuda.current_stream())
    N, lv = 1 << 12, 6
    L, T = lv + 1, lv + 1 + len(p)
    dw = ctx.parts(lv) * T * N
    B, R = 35, 3
    a_in = [M.to_dev(rand_residues(rng, q[:L], N)) for _ in range(B)]
    b_in = [M.to_dev(rand_residues(rng, q[:L], N)) for _ in range(B)]
    dig = _zeros(M, B * dw)
    M.check(lib.mi_hoist_batch(ctx.h, lv, B, _parr([M.ptr(t) for t in a_in]), M.ptr(dig), dw, s))
    keys = [M.to_dev(_key(rng, q, p, N, octx.parts(len(q) - 1))) for _ in range(R)]
    gals = [O.galois_rotation(N, r) for r in (3, 17, 40)]
    outs = [(_zeros(M, L * N), _zeros(M, L * N)) for _ in range(R * B)]
    gal = (ctypes_u64(R * B))(*[gals[i // B] for i in range(R * B)])
    print("rotate_hoisted_items: 3 rotations x 35 members, streams", streams, flush=True)
    M.check(lib.mi_rotate_hoisted_items(ctx.h, lv, R * B, M.ptr(dig), dw, B,
                                        _parr([M.ptr(b_in[i % B]) for i in range(R * B)]), gal,
                                        _parr([M.ptr(keys[i // B]) for i in range(R * B)]),
                                        _parr([M.ptr(o[0]) for o in outs]), _parr([M.ptr(o[1]) for o in outs]), 1, s))
    for i in range(R * B):
        r, m = divmod(i, B)
        one = _zeros(M, dw)
        ctx.hoist(lv, a_in[m], one)
        hb, ha = M.to_dev(M.from_dev(b_in[m])), _zeros(M, L * N)
        ctx.keyswitch_hoisted(lv, one, keys[r], hb, ha, 0, 1, galois=gals[r])
        assert np.array_equal(M.from_dev(outs[i][0]), M.from_dev(hb)), i
        assert np.array_equal(M.from_dev(outs[i][1]), M.from_dev(ha)), i
    # the NTT probe keeps the side-stream chunks on the caller's stream while
    # it is on: the same outputs, and every pair counted at 16 N bytes per limb
    import ctypes
    M.check(lib.mi_probe_begin(ctx.h))
    again = [(_zeros(M, L * N), _zeros(M, L * N)) for _ in range(R * B)]
    M.check(lib.mi_rotate_hoisted_items(ctx.h, lv, R * B, M.ptr(dig), dw, B,
                                        _parr([M.ptr(b_in[i % B]) for i in range(R * B)]), gal,
                                        _parr([M.ptr(keys[i // B]) for i in range(R * B)]),
                                        _parr([M.ptr(o[0]) for o in again]), _parr([M.ptr(o[1]) for o in again]), 1, s))
    stats = (ctypes.c_double * 4)()
    M.check(lib.mi_probe_end(ctx.h, stats))
    assert stats[0] > 0 and stats[1] >= 1 and stats[3] >= stats[1]
    assert stats[2] == 16.0 * N * stats[3]
    for i in range(R * B):
        assert np.array_equal(M.from_dev(again[i][0]), M.from_dev(outs[i][0])), i
        assert np.array_equal(M.from_dev(again[i][1]), M.from_dev(outs[i][1])), i
    # overlapping outputs are refused
    assert lib.mi_rotate_hoisted_items(ctx.h, lv, 2, M.ptr(dig), dw, B, _parr([M.ptr(b_in[0])] * 2), gal,
                                       _parr([M.ptr(keys[0])] * 2), _parr([M.ptr(outs[0][0])] * 2),
                                       _parr([M.ptr(outs[0][1]), M.ptr(outs[1][1])]), 1, s) < 0
