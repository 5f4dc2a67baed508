"""Pins the CPU oracle (oracle/ckks_oracle.c) before it is trusted as a checker.

Each restated function is checked against at least one of:
  * golden vectors produced by the reference's own C++ code (tests/golden/),
  * closed forms computed with Python big integers (CRT, brute-force DFT,
    exact rounding division, negacyclic convolution),
  * invariants (inverse of forward, decryption of a key-switched ciphertext).
"""
import numpy as np
import pytest

import oracle as O
from helpers import GOLDEN, brev, crt, digest, load_json, load_npz, lt_diags, rand_residues, splitmix64_stream

RNG = np.random.default_rng(1234)


# ----------------------------------------------------------------- modarith
def _edge_operands(q):
    return [0, 1, 2, q - 1, q - 2, q // 2, q // 2 + 1] + [int(v) for v in RNG.integers(0, q, 40, dtype=np.uint64)]


@pytest.mark.parametrize("q", [68719403009, 1152921504606584833, 2305843009213616129, 4611686018427322369])
def test_modarith_closed_form(q):
    L = O.lib()
    ninv = L.oc_montgomery_neg_inverse(q)
    assert (q * ninv + 1) % (1 << 64) == 0
    assert L.oc_montgomery_radix_mod_q(q) == (1 << 64) % q
    for a in _edge_operands(q):
        for b in _edge_operands(q)[:12]:
            assert L.oc_barrett_mul(a, b, q) == a * b % q
            assert L.oc_shoup_mul(a, b, q) == a * b % q
            # REDC(a * b) = a*b*2^-64 mod q
            assert L.oc_montgomery_mul(a, b, q) == a * b * pow(1 << 64, -1, q) % q
        assert L.oc_compute_shoup(a, q) == (a << 64) // q
        if a:
            assert L.oc_mod_inverse(a, q) == pow(a, -1, q)
        assert L.oc_barrett_reduce_u64(a, q) == a % q
    for a in [q, 2 * q - 1, (1 << 64) - 1, 3 * q + 5]:
        if a < (1 << 64):
            assert L.oc_barrett_reduce_u64(a, q) == a % q
    assert L.oc_powmod(3, q - 1, q) == 1


# ------------------------------------------------------------ prime search
PRESET_BITS = {
    # (q_bits, p_bits, N) from the reference presets (ckks_aes_params.hpp) and
    # the benchmark configurations.
    "walsh15": ([36, 61, 42, 42] + [38] * 7 + [50, 50], [61] * 4, 1 << 15),
    "xboot13": ([32, 28, 31, 31, 31, 32, 32, 32, 32, 32, 32, 28, 28], [32], 1 << 14),
    "xboot14": ([31, 27, 30, 30, 30, 31, 31, 31, 31, 31, 31, 27, 27], [32], 1 << 14),
    "xboot15": ([38, 36, 36, 37, 37, 37] + [38] * 7 + [38] * 4, [38] * 5, 1 << 15),
    "chain14": ([50] * 8, [50, 50], 1 << 14),
    "ks16": ([50] * 30, [50] * 10, 1 << 16),
    "ntt16": ([60] * 40, [], 1 << 16),
}


@pytest.mark.parametrize("tag", sorted(PRESET_BITS))
def test_prime_search_matches_reference(tag):
    ref = load_json("primes.json")[tag]
    qb, pb, N = PRESET_BITS[tag]
    p, q = O.find_primes(pb, qb, N)
    assert [int(v) for v in q] == ref["q"]
    assert [int(v) for v in p] == ref["p"]
    for v, b in zip(list(q) + list(p), qb + pb):
        v = int(v)
        assert v % (2 * N) == 1 and v.bit_length() == b and O.lib().oc_is_prime(v)
    assert len(set(map(int, list(q) + list(p)))) == len(q) + len(p)


def test_is_prime_small_brute_force():
    L = O.lib()
    for n in range(2, 3000):
        isp = all(n % d for d in range(2, int(n ** 0.5) + 1))
        assert bool(L.oc_is_prime(n)) == isp, n


# --------------------------------------------------------------------- NTT
def test_ntt_tables_match_reference():
    ent = load_json("ntt.json")["ntt10"]
    N = ent["N"]
    for q, e in zip(ent["q"], ent["limbs"]):
        fwd, fs, inv, is_, ni = O.ntt_tables(N, q)
        assert digest(fwd) == e["tab_fwd"]
        assert digest(fs) == e["tab_fwds"]
        assert digest(inv) == e["tab_inv"]
        assert digest(is_) == e["tab_invs"]
        assert ni * N % q == 1


@pytest.mark.parametrize("tag", ["ntt10", "ntt12", "ntt16one"])
def test_ntt_outputs_match_reference(tag):
    ent = load_json("ntt.json")[tag]
    N = ent["N"]
    for i, (q, e) in enumerate(zip(ent["q"], ent["limbs"])):
        x = splitmix64_stream(ent["seed_base"] + i, N) % np.uint64(q)
        t = O.ntt_tables(N, q)
        assert digest(O.ntt_forward(x, q, t)) == e["fwd_out"]
        assert digest(O.ntt_forward(x, q, t, lazy=True)) == e["fwd_lazy"]
        assert digest(O.ntt_inverse(x, q, t)) == e["inv_out"]


@pytest.mark.parametrize("N,bits", [(16, 30), (64, 40), (128, 61)])
def test_ntt_brute_force_evaluation(N, bits):
    _, (q,) = O.find_primes([], [bits], N)
    q = int(q)
    psi = O.lib().oc_find_root_of_unity(N, q)
    # psi is the first a >= 2 whose ((q-1)/2N)-th power has order exactly 2N
    e = (q - 1) // (2 * N)
    a = 2
    while pow(pow(a, e, q), N, q) != q - 1:
        a += 1
    assert psi == pow(a, e, q)
    t = O.ntt_tables(N, q)
    x = [int(v) for v in RNG.integers(0, q, N, dtype=np.uint64)]
    y = O.ntt_forward(np.array(x, np.uint64), q, t)
    logn = N.bit_length() - 1
    for i in range(N):
        root = pow(psi, 2 * brev(i, logn) + 1, q)
        assert int(y[i]) == sum(c * pow(root, j, q) for j, c in enumerate(x)) % q
    assert list(O.ntt_inverse(y, q, t)) == x


def test_ntt_negacyclic_convolution():
    N = 256
    _, (q,) = O.find_primes([], [55], N)
    q = int(q)
    t = O.ntt_tables(N, q)
    a = [int(v) for v in RNG.integers(0, q, N, dtype=np.uint64)]
    b = [int(v) for v in RNG.integers(0, q, N, dtype=np.uint64)]
    c = [0] * N
    for i in range(N):
        for j in range(N):
            k = i + j
            if k < N:
                c[k] += a[i] * b[j]
            else:
                c[k - N] -= a[i] * b[j]
    c = [v % q for v in c]
    fa = O.ntt_forward(np.array(a, np.uint64), q, t)
    fb = O.ntt_forward(np.array(b, np.uint64), q, t)
    prod = np.array([int(u) * int(v) % q for u, v in zip(fa, fb)], np.uint64)
    assert [int(v) for v in O.ntt_inverse(prod, q, t)] == c


def test_lazy_forward_range_and_value():
    N = 1024
    _, (q,) = O.find_primes([], [61], N)
    q = int(q)
    t = O.ntt_tables(N, q)
    x = RNG.integers(0, q, N, dtype=np.uint64)
    lazy = O.ntt_forward(x, q, t, lazy=True)
    assert int(lazy.max()) < 4 * q
    assert np.array_equal(lazy % np.uint64(q), O.ntt_forward(x, q, t))


# ------------------------------------------------------------ small context
def small_ctx(N=64, nq=5, np_=2, dnum=2, qbits=40, pbits=50):
    p, q = O.find_primes([pbits] * np_, [qbits] * nq, N)
    return O.Ctx(N, q, p, dnum)


def test_modup_closed_form():
    ctx = small_ctx(N=32, nq=7, np_=3, dnum=3)
    N = ctx.N
    for level in (6, 4, 1):
        L = level + 1
        for part in range(ctx.parts(level)):
            s0 = part * ctx.alpha
            s1 = min(s0 + ctx.alpha, L)
            src_mod = [int(v) for v in ctx.q[s0:s1]]
            dst_mod = [int(v) for v in list(ctx.q[:s0]) + list(ctx.q[s1:L]) + list(ctx.p)]
            src = rand_residues(RNG, src_mod, N)
            out = ctx.modup(level, part, src).reshape(len(dst_mod), N)
            for c in range(N):
                x, Q = crt([src[i * N + c] for i in range(len(src_mod))], src_mod)
                xc = x - Q if x > Q // 2 else x  # centred lift
                for t, m in enumerate(dst_mod):
                    assert int(out[t, c]) == xc % m, (level, part, c, t)


def test_rescale_closed_form():
    ctx = small_ctx(N=64, nq=5, np_=1, dnum=1, qbits=45)
    N = ctx.N
    for level in (4, 2, 1):
        L = level + 1
        mod = [int(v) for v in ctx.q[:L]]
        polys = np.stack([rand_residues(RNG, mod, N) for _ in range(2)])
        out = ctx.rescale(level, polys)
        p = mod[level]
        for k in range(2):
            coeff = ctx.ntt_inv(polys[k], L).reshape(L, N)
            expect = np.zeros((level, N), np.uint64)
            for c in range(N):
                X, _ = crt(coeff[:, c], mod)
                r = (X + p // 2) // p
                for i in range(level):
                    expect[i, c] = r % mod[i]
            assert np.array_equal(out[k], ctx.ntt_fwd(expect.ravel(), level))


def test_automorphism_closed_form():
    N = 64
    ctx = small_ctx(N=N, nq=2, np_=1, dnum=1)
    q = int(ctx.q[0])
    for g in (O.galois_rotation(N, 1), O.galois_rotation(N, 7), O.galois_rotation(N, -3), 2 * N - 1):
        a = RNG.integers(0, q, N, dtype=np.uint64)
        ap = np.zeros(N, np.uint64)
        for j in range(N):
            k = j * g % (2 * N)
            if k < N:
                ap[k] = a[j]
            else:
                ap[k - N] = (q - int(a[j])) % q
        fa = ctx.ntt_fwd(np.concatenate([a, a]), 1)[:N]
        m = O.automorphism_map(N, g)
        assert np.array_equal(fa[m.astype(np.int64)], ctx.ntt_fwd(np.concatenate([ap, ap]), 1)[:N])


def test_tensor_is_negacyclic_product():
    ctx = small_ctx(N=32, nq=2, np_=1, dnum=1)
    N = ctx.N
    mod = [int(v) for v in ctx.q[:2]]
    b1, a1, b2, a2 = (rand_residues(RNG, mod, N) for _ in range(4))
    ob, oa, oc = ctx.tensor(1, b1, a1, b2, a2)
    for i, q in enumerate(mod):
        sl = slice(i * N, (i + 1) * N)
        f = lambda v: [int(x) for x in ctx.ntt_inv(np.concatenate([v] * 2), 1)[:N]] if i == 0 else None
        # closed form in the NTT domain: per-slot products
        assert np.array_equal(oc[sl], np.array([int(x) * int(y) % q for x, y in zip(a1[sl], a2[sl])], np.uint64))
        assert np.array_equal(ob[sl], np.array([int(x) * int(y) % q for x, y in zip(b1[sl], b2[sl])], np.uint64))
        assert np.array_equal(oa[sl], np.array([(int(x) * int(y) + int(u) * int(v)) % q
                                                for x, y, u, v in zip(a1[sl], b2[sl], b1[sl], a2[sl])], np.uint64))


# ----------------------------------------------- reference evaluator vectors
def _ev(tag):
    d = load_npz(tag + ".npz")
    outs = load_json(tag + "_outputs.json")
    N = len(d["m1"]) * 2
    L = len(d["q"])
    nparts = len(d["key_relin"]) // (2 * (L + len(d["p"])) * N)
    return d, outs, N, L, nparts


@pytest.mark.parametrize("tag", ["ev10", "ev10b"])
def test_evaluator_matches_reference(tag):
    d, outs, N, L, dnum = _ev(tag)
    ctx = O.Ctx(N, d["q"], d["p"], dnum)
    lv = L - 1
    ct1 = d["ct1"].reshape(2, L * N)
    ct2 = d["ct2"].reshape(2, L * N)
    b1, a1 = ct1
    b2, a2 = ct2
    got = {}
    mb, ma, mc = ctx.tensor(lv, b1, a1, b2, a2)
    got["mul"] = np.concatenate([mb, ma, mc])
    rb, ra = ctx.relin(lv, mb, ma, mc, d["key_relin"])
    got["mul_relin"] = np.concatenate([rb, ra])
    rs = ctx.rescale(lv, np.stack([rb, ra]))
    got["mul_relin_rescale"] = rs.ravel()
    cb, ca = ctx.rotate(lv - 1, rs[0], rs[1], 1, d["key_rot1"])
    got["chain_rot1"] = np.concatenate([cb, ca])
    sb, sa, sc = ctx.square(lv, b1, a1)
    got["square"] = np.concatenate([sb, sa, sc])
    got["rot1_out"] = np.concatenate(ctx.rotate(lv, b1, a1, 1, d["key_rot1"]))
    if "key_rot5" in d:
        got["rot5_out"] = np.concatenate(ctx.rotate(lv, b1, a1, 5, d["key_rot5"]))
    if "key_conj" in d:
        got["conj_out"] = np.concatenate(ctx.conjugate(lv, b1, a1, d["key_conj"]))
    r1 = ctx.rescale(lv, ct1)
    got["rescale_ct1"] = r1.ravel()
    got["rescale2_ct1"] = ctx.rescale(lv - 1, r1).ravel()
    got["add"] = ctx.add(lv, ct1, ct2).ravel()
    got["sub"] = ctx.sub(lv, ct1, ct2).ravel()
    got["mulint"] = ctx.mul_int(lv, ct1, -7).ravel()
    scale = float(d["ct1_meta"][2:3].view(np.float64)[0])
    res = np.array([int(round(3 * scale)) % int(q) for q in d["q"]], np.uint64)
    got["addconst"] = np.concatenate([ctx.add_scalar_residues(lv, b1, res), a1])
    got["mulplain"] = ctx.mul_plain(lv, ct1, d["pt_m2"]).ravel()
    l1 = ct1[:, : 2 * N]
    got["rot1_level1"] = np.concatenate(ctx.rotate(1, l1[0], l1[1], 1, d["key_rot1"]))
    for name, arr in got.items():
        assert digest(arr) == outs[name]["digest"], name


def test_keyswitch_decrypts_under_target_key():
    """Relinearisation invariant: b' + a' s == c2 s^2 (mod Q) up to small noise."""
    d, outs, N, L, dnum = _ev("ev10")
    ctx = O.Ctx(N, d["q"], d["p"], dnum)
    lv = L - 1
    sk = d["sk"].reshape(L, N)
    c2 = rand_residues(RNG, d["q"], N)
    zero = np.zeros(L * N, np.uint64)
    b, a = ctx.keyswitch(lv, c2, d["key_relin"], zero, zero, 1, 1)
    diff = np.zeros(L * N, np.uint64)
    for i, q in enumerate(int(v) for v in d["q"]):
        sl = slice(i * N, (i + 1) * N)
        s = sk[i].astype(object)
        lhs = (b[sl].astype(object) + a[sl].astype(object) * s) % q
        rhs = (c2[sl].astype(object) * s * s) % q
        diff[sl] = np.array((lhs - rhs) % q, np.uint64)
    coeff = ctx.ntt_inv(diff, L).reshape(L, N)
    for i, q in enumerate(int(v) for v in d["q"]):
        centred = np.where(coeff[i] > q // 2, coeff[i].astype(object) - q, coeff[i].astype(object))
        assert max(abs(int(v)) for v in centred) < 2 ** 30


# ------------------------------------- hoisting, QP switches, BSGS, bootstrap
def test_hoisting_and_qp_switches_match_reference():
    d = load_npz("lt10.npz")
    o = load_json("lt10_outputs.json")
    N = 1 << 10
    q, p = d["q"], d["p"]
    L = len(q)
    ctx = O.Ctx(N, q, p, 2)
    lv = L - 1
    b, a = d["ct"].reshape(2, L * N)
    digits = ctx.hoist(lv, a)
    assert digest(digits) == o["hoist_digits"]
    hb, ha = ctx.keyswitch_hoisted(lv, digits, d["key_rot1"], b, np.zeros_like(a), 0, 1,
                                   galois=O.galois_rotation(N, 1))
    assert digest(np.concatenate([hb, ha])) == o["hoisted_rot1"]
    qb, qa = ctx.keyswitch_hoisted_qp_rotated(lv, digits, d["key_rot2"], b, O.galois_rotation(N, 2))
    assert digest(np.concatenate([qb, qa])) == o["hoisted_qp_rot2"]
    sb, sa = ctx.keyswitch_qp_from_key_term(lv, a, d["key_rot4"])
    assert digest(np.concatenate([sb, sa])) == o["qp_from_key_term_rot4"]
    qp = np.concatenate([q, p])
    diags = lt_diags(N, qp, 6)
    keys = {1: d["key_rot1"], 2: d["key_rot2"], 4: d["key_rot4"]}
    ao, bo = ctx.lt_stage_qp(lv, a, b, o["diag_indices"], diags, o["baby_steps"], keys)
    assert digest(np.concatenate([bo, ao])) == o["lt"]
    r = ctx.rescale(lv, np.stack([bo, ao]))
    assert digest(r.ravel()) == o["lt_rescaled"]


def test_q0p0_and_modraise_match_reference():
    d = load_npz("bt10.npz")
    o = load_json("bt10_outputs.json")
    N = 1 << 10
    q, p = d["q"], d["p"]
    ctx = O.Ctx(N, q, p, 2)
    b, a = d["ct0"].reshape(2, N)
    sb, sa = ctx.q0p0_keyswitch(b, a, d["key_d2s"])
    assert digest(np.concatenate([sb, sa])) == o["q0p0"]
    lv = o["modraise_level"]
    rb, ra = ctx.modraise(lv, sb), ctx.modraise(lv, sa)
    assert digest(np.concatenate([rb, ra])) == o["modraise"]


def test_modraise_closed_form():
    """ModRaise = centred lift of the q0 coefficients, re-reduced mod each q_l."""
    ctx = small_ctx(N=64, nq=4, np_=1, dnum=1)
    N = ctx.N
    q0 = int(ctx.q[0])
    x = RNG.integers(0, q0, N, dtype=np.uint64)
    out = ctx.modraise(3, x)
    coeff = ctx.ntt_inv(x, 1)
    lifted = [int(v) - q0 if int(v) > q0 // 2 else int(v) for v in coeff]
    expect = np.concatenate([np.array([v % int(ql) for v in lifted], np.uint64) for ql in ctx.q[:4]])
    assert np.array_equal(out, ctx.ntt_fwd(expect, 4))


def test_lt_stage_zero_diagonals():
    """All-zero diagonals give an exactly zero stage output."""
    d = load_npz("lt10.npz")
    N = 1 << 10
    q, p = d["q"], d["p"]
    L = len(q)
    ctx = O.Ctx(N, q, p, 2)
    lv = L - 1
    b, a = d["ct"].reshape(2, L * N)
    qp = np.concatenate([q, p])
    zeros = np.zeros(6 * (L + len(p)) * N, np.uint64)
    keys = {1: d["key_rot1"], 2: d["key_rot2"], 4: d["key_rot4"]}
    ao, bo = ctx.lt_stage_qp(lv, a, b, list(range(6)), zeros, 2, keys)
    assert not ao.any() and not bo.any()


def test_walsh_product_sum_composition_special_cases():
    """The oracle's Walsh product sum (composed from the pinned scaled-add,
    constant, plaintext and tensor primitives) reduces to a single tensor for
    a unit coefficient table and is linear in the real coefficients."""
    N = 1 << 10
    p, q = O.find_primes([45], [40, 40, 40], N)
    ctx = O.Ctx(N, q, p, 1)
    rng = np.random.default_rng(17)
    lv = 2
    rnd = lambda: np.stack([rand_residues(rng, q, N), rand_residues(rng, q, N)])
    low = [rnd() for _ in range(15)]
    high = [rnd() for _ in range(15)]
    pt = rand_residues(rng, q, N)
    zero = np.zeros((16, 16), dtype=np.int64)
    consts0 = [(0, 0)] * 16
    cr = zero.copy()
    cr[5][9] = 1
    acc, raw = ctx.walsh_product_sum(lv, low, high, cr, zero, consts0, pt)
    tb, ta, tc = ctx.tensor(lv, low[8][0], low[8][1], high[4][0], high[4][1])
    assert not acc.any()
    assert np.array_equal(raw, np.stack([tb, ta, tc]))
    c1 = rng.integers(-64, 65, (16, 16))
    c2 = rng.integers(-64, 65, (16, 16))
    a1, r1 = ctx.walsh_product_sum(lv, low, high, c1, zero, consts0, pt)
    a2, r2 = ctx.walsh_product_sum(lv, low, high, c2, zero, consts0, pt)
    a12, r12 = ctx.walsh_product_sum(lv, low, high, c1 + c2, zero, consts0, pt)
    assert np.array_equal(a12, ctx.add(lv, a1, a2))
    assert np.array_equal(r12, ctx.add(lv, r1, r2))
