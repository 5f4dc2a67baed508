"""GPU tests of the C++ host API (CKKSParams / CKKSContext / CKKSCiphertext and
the evaluator free functions), which runs every ring operation through the
HIP C-ABI.

* With the reference's captured keys and ciphertexts installed, the API's
  operations reproduce the reference's outputs bit for bit (ev10 fixtures).
* With the API's own seeded keys, the configuration-1 chain (encrypt, HMult,
  relinearise, rescale, rotate by 1, decrypt) matches the oracle limb for limb
  and decrypts to the expected slots.
* Generated switching keys decrypt to P * s_src on their digit limbs.
"""
import numpy as np
import pytest

import oracle as O
from helpers import digest, host, load_json, load_npz, lt_diags

pytestmark = pytest.mark.gpu


def _ev_config(H, tag):
    c = H.CKKSParamsConfig()
    c.N = 1 << 10
    c.mult_depth = 3 if tag == "ev10" else 4
    c.q0_bit = 50
    c.q_bit = 40
    c.p_bits = [60, 60] if tag == "ev10" else [61, 61]
    c.dnum = 2 if tag == "ev10" else 3
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    c.rotate_values = [1, 5]
    return c


@pytest.mark.parametrize("tag", ["ev10", "ev10b"])
def test_host_api_reproduces_reference_evaluator(tag):
    H = host()
    d = load_npz(tag + ".npz")
    outs = load_json(tag + "_outputs.json")
    ctx = H.CKKSContext(H.CKKSParams(_ev_config(H, tag)))
    P = ctx.getParams()
    assert list(P.getModuli()) == list(d["q"]) and list(P.getPModuli()) == list(d["p"])
    ctx.setSecretKeyEval(d["sk"])
    ctx.setRelinKey(d["key_relin"])
    ctx.setRotationKey(1, d["key_rot1"])
    if "key_rot5" in d:
        ctx.setRotationKey(5, d["key_rot5"])
    if "key_conj" in d:
        ctx.setConjugationKey(d["key_conj"])
    N, L = 1 << 10, len(d["q"])
    lv, W = L - 1, L * N
    scale = float(d["ct1_meta"][2:3].view(np.float64)[0])

    def ct(arr):
        return H.CKKSCiphertext(ctx, [arr[:W], arr[W:2 * W]], scale, lv)

    def flat(x):
        return np.concatenate([x.getPoly(i) for i in range(x.getNumPolys())])

    def check(name, x):
        assert digest(flat(x)) == outs[name]["digest"], name
        assert x.getLevel() == outs[name]["level"], name
        assert x.getScale() == outs[name]["scale"], name

    ct1, ct2 = ct(d["ct1"]), ct(d["ct2"])
    enc = H.ckks_encode(ctx, d["m1"], lv, eval=True)
    assert np.array_equal(enc.getPlaintext(), d["encode_m1"])
    assert np.array_equal(H.ckks_decrypt(ctx, ct1), d["dec_ct1"])

    x = ct1.clone()
    H.ckks_mul_inplace(ctx, x, ct2)
    check("mul", x)
    H.ckks_relin_hybrid_inplace(ctx, x)
    check("mul_relin", x)
    H.ckks_rescale(ctx, x)
    check("mul_relin_rescale", x)
    H.ckks_rotate_inplace(ctx, x, 1)
    check("chain_rot1", x)
    x = ct1.clone()
    H.ckks_square_inplace(ctx, x)
    check("square", x)
    check("rot1_out", H.ckks_rotate(ctx, ct1, 1))
    if "key_rot5" in d:
        check("rot5_out", H.ckks_rotate(ctx, ct1, 5))
    if "key_conj" in d:
        check("conj_out", H.ckks_conjugate(ctx, ct1))
    x = ct1.clone()
    H.ckks_rescale(ctx, x)
    check("rescale_ct1", x)
    H.ckks_rescale(ctx, x)
    check("rescale2_ct1", x)
    x = ct1.clone()
    H.ckks_add_inplace(ctx, x, ct2)
    check("add", x)
    x = ct1.clone()
    H.ckks_sub_inplace(ctx, x, ct2)
    check("sub", x)
    x = ct1.clone()
    H.ckks_mul_int_inplace(ctx, x, -7)
    check("mulint", x)
    x = ct1.clone()
    H.ckks_add_integer_constant_inplace(ctx, x, 3)
    check("addconst", x)
    pt = H.ckks_encode(ctx, d["m2"], lv, eval=True)
    assert np.array_equal(pt.getPlaintext(), d["pt_m2"])
    x = ct1.clone()
    H.ckks_mul_plain_inplace(ctx, x, pt)
    check("mulplain", x)
    x = ct1.clone()
    H.ckks_drop_to_level_inplace(x, 1)
    H.ckks_rotate_inplace(ctx, x, 1)
    check("rot1_level1", x)


def _config1(H, seed=11):
    """Configuration 1: N=2^14, 8 limbs of 50-bit primes, P = 2 x 50 bits,
    dnum 4, Hamming weight 192, sigma 3.2, rotation key for 1."""
    c = H.CKKSParamsConfig()
    c.N = 1 << 14
    c.mult_depth = 7
    c.q0_bit = 50
    c.q_bit = 50
    c.p_bits = [50, 50]
    c.dnum = 4
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    c.rotate_values = [1]
    c.seed = seed
    return c


def test_host_config1_chain_matches_oracle_and_decrypts():
    H = host()
    ctx = H.CKKSContext(H.CKKSParams(_config1(H)))
    P = ctx.getParams()
    q, p = P.getModuli(), P.getPModuli()
    N, lv = 1 << 14, 7
    octx = O.Ctx(N, q, p, 4)
    rng = np.random.default_rng(2024)
    m1 = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    m2 = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    ct1 = H.ckks_encrypt(ctx, m1, lv)
    ct2 = H.ckks_encrypt(ctx, m2, lv)
    assert np.max(np.abs(H.ckks_decrypt(ctx, ct1) - m1)) < 1e-9
    b1, a1, b2, a2 = ct1.getPoly(0), ct1.getPoly(1), ct2.getPoly(0), ct2.getPoly(1)
    relin, rot1 = ctx.getRelinKey(), ctx.getRotationKey(1)

    H.ckks_mul_inplace(ctx, ct1, ct2)
    ob, oa, oc2 = octx.tensor(lv, b1, a1, b2, a2)
    assert np.array_equal(ct1.getPoly(0), ob) and np.array_equal(ct1.getPoly(1), oa)
    assert np.array_equal(ct1.getPoly(2), oc2)
    H.ckks_relin_hybrid_inplace(ctx, ct1)
    ob, oa = octx.relin(lv, ob, oa, oc2, relin)
    assert np.array_equal(ct1.getPoly(0), ob) and np.array_equal(ct1.getPoly(1), oa)
    H.ckks_rescale(ctx, ct1)
    r = octx.rescale(lv, np.stack([ob, oa]))
    assert np.array_equal(ct1.getPoly(0), r[0]) and np.array_equal(ct1.getPoly(1), r[1])
    H.ckks_rotate_inplace(ctx, ct1, 1)
    ob, oa = octx.rotate(lv - 1, r[0], r[1], 1, rot1)
    assert np.array_equal(ct1.getPoly(0), ob) and np.array_equal(ct1.getPoly(1), oa)
    assert ct1.getLevel() == lv - 1
    assert ct1.getScale() == 2.0 ** 100 / float(q[lv])
    dec = H.ckks_decrypt(ctx, ct1)
    want = np.roll(m1 * m2, -1)
    assert np.max(np.abs(dec - want)) < 1e-6


def _unfold(key_limb, m, K_mods, l, nq, fold, mont):
    """Undo the folding and Montgomery scalings of one key limb (python ints)."""
    f = 1
    if fold:
        if l < nq:
            Pm = 1
            for pv in K_mods:
                Pm = Pm * pv % m
            f = pow(Pm, -1, m)
        else:
            ph = 1
            for j, pv in enumerate(K_mods):
                if j != l - nq:
                    ph = ph * pv % m
            f = pow(ph, -1, m)
    if mont:
        f = f * (1 << 64) % m
    inv = pow(f, -1, m)
    return [int(v) * inv % m for v in key_limb]


def test_host_switching_keys_decrypt_to_their_messages():
    """b + a * s_tgt - P * s_src (on the digit's limbs) is a small error for the
    relinearisation (s^2 -> s), rotation (s -> pi^-1(s)) and conjugation keys."""
    H = host()
    c = _config1(H, seed=5)
    c.N = 1 << 11
    c.rotate_values = [3]
    ctx = H.CKKSContext(H.CKKSParams(c))
    P = ctx.getParams()
    q, p = [int(v) for v in P.getModuli()], [int(v) for v in P.getPModuli()]
    N, nq, K = 1 << 11, len(q), len(p)
    basis = q + p
    octx = O.Ctx(N, np.array(q, np.uint64), np.array(p, np.uint64), 4)
    s = [int(v) for v in ctx.getSecretKeyCoeffs()]
    assert sum(1 for v in s if v) == 192 and set(s) <= {-1, 0, 1}

    def coeff_auto(sv, g):
        out = [0] * N
        for i, v in enumerate(sv):
            j = i * g % (2 * N)
            if j >= N:
                out[j - N] = -v
            else:
                out[j] = v
        return out

    def ntt_qp(sv):
        limbs = np.concatenate([np.array([v % m for v in sv], np.uint64) for m in basis])
        bidx = np.arange(nq + K, dtype=np.int32)
        return octx.ntt_fwd_basis(limbs, bidx)

    s_eval = ntt_qp(s)
    s2_eval = np.concatenate([np.array([int(x) * int(x) % m for x in s_eval[l * N:(l + 1) * N]], np.uint64)
                              for l, m in enumerate(basis)])
    g3 = pow(5, 3, 2 * N)
    g3inv = pow(5, N // 2 - 3, 2 * N)
    cases = [("relin", ctx.getRelinKey(), s2_eval, s_eval),
             ("rot3", ctx.getRotationKey(3), s_eval, ntt_qp(coeff_auto(s, g3inv))),
             ("conj", ctx.getConjugationKey(), ntt_qp(coeff_auto(s, 2 * N - 1)), s_eval)]
    alpha = (nq + 3) // 4
    KL = nq + K
    Pq = 1
    for pv in p:
        Pq *= pv
    for name, key, src, tgt in cases:
        digits = len(key) // (2 * KL * N)
        assert digits == ctx.keyDigitCount()
        for d in range(digits):
            kb = key[(2 * d) * KL * N:(2 * d + 1) * KL * N]
            ka = key[(2 * d + 1) * KL * N:(2 * d + 2) * KL * N]
            err = np.zeros(KL * N, np.uint64)
            for l, m in enumerate(basis):
                sl = slice(l * N, (l + 1) * N)
                bl = _unfold(kb[sl], m, p, l, nq, True, True)
                al = _unfold(ka[sl], m, p, l, nq, True, True)
                in_digit = d * alpha <= l < min(d * alpha + alpha, nq)
                pm = Pq % m if in_digit else 0
                err[sl] = np.array([(bv + av * int(t) - pm * int(sv)) % m
                                    for bv, av, t, sv in zip(bl, al, tgt[sl], src[sl])], np.uint64)
            e = octx.ntt_inv_basis(err, np.arange(KL, dtype=np.int32))
            for l, m in enumerate(basis):
                c_ = [int(v) - m if int(v) > m // 2 else int(v) for v in e[l * N:(l + 1) * N]]
                assert max(abs(v) for v in c_) < 64, (name, d, l)
                if l:
                    assert c_ == prev, (name, d, l)  # one integer error reduced per limb
                prev = c_


def test_host_api_errors_and_memory_accounting():
    H = host()
    c = _config1(H, seed=3)
    c.N = 1 << 12
    ctx = H.CKKSContext(H.CKKSParams(c))
    live0 = ctx.memoryStats()[0]
    m = np.zeros(1 << 11, complex)
    ct = H.ckks_encrypt(ctx, m, 7)
    assert ctx.memoryStats()[0] > live0
    with pytest.raises(ValueError, match="ckks_rotate: rotation key is unavailable"):
        H.ckks_rotate_inplace(ctx, ct, 2)
    x = ct.clone()
    H.ckks_square_inplace(ctx, x)
    with pytest.raises(ValueError, match="expected 2-component"):
        H.ckks_square_inplace(ctx, x)
    with pytest.raises(ValueError, match="ckks_relin: expected 3-polynomial"):
        H.ckks_relin_hybrid_inplace(ctx, ct)
    y = ct.clone()
    H.ckks_drop_to_level_inplace(y, 0)
    with pytest.raises(ValueError, match="ckks_rescale: invalid level"):
        H.ckks_rescale(ctx, y)
    with pytest.raises(ValueError, match="ciphertext size mismatch"):
        H.ckks_add_inplace(ctx, y, ct)
    del x, y, ct
    import gc
    gc.collect()
    assert ctx.memoryStats()[0] == live0


def test_host_evalchebyshev_matches_reference():
    """Paterson-Stockmeyer Chebyshev evaluation of six coefficient sets
    (generic, even, odd, degree 63, 2, 1) reproduces the reference's output
    ciphertexts bit for bit, with the reference's key and input installed."""
    H = host()
    d = load_npz("ep10.npz")
    outs = load_json("ep10_outputs.json")
    c = H.CKKSParamsConfig()
    c.N = 1 << 10
    c.mult_depth = 8
    c.q0_bit = 50
    c.q_bit = 40
    c.p_bits = [60, 60, 60]
    c.dnum = 3
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    ctx = H.CKKSContext(H.CKKSParams(c))
    P = ctx.getParams()
    assert list(P.getModuli()) == list(d["q"]) and list(P.getPModuli()) == list(d["p"])
    ctx.setSecretKeyEval(d["sk"])
    ctx.setRelinKey(d["key_relin"])
    N, L = 1 << 10, len(d["q"])
    W = L * N
    x = H.CKKSCiphertext(ctx, [d["ct"][:W], d["ct"][W:]], outs["ct"]["scale"], L - 1)
    m = d["m"].real
    for k in range(6):
        co = d["coeffs%d" % k]
        out = H.ckks_evalchebyshev_preconditioned(ctx, x, list(co))
        flat = np.concatenate([out.getPoly(i) for i in range(out.getNumPolys())])
        ref = outs["out%d" % k]
        assert digest(flat) == ref["digest"], k
        assert out.getLevel() == ref["level"] and out.getScale() == ref["scale"], k
        # value check for the even and odd polynomials: the decrypted slots
        # follow sum c_i T_i(m).  (For polynomials with both parities the
        # reference's leaf evaluation adds each leaf's T_0 term only when the
        # polynomial is even, ckks_evalpoly.cpp:1011-1017, so those results are
        # not sum c_i T_i; the digests above pin that behaviour.)
        if not any(co[1::2]) or not any(co[0::2]):
            dec = H.ckks_decrypt(ctx, out).real
            want = np.polynomial.chebyshev.chebval(m, co)
            assert np.max(np.abs(dec - want)) < 1e-5, k


def _bt10_context(H):
    """The reference harness's bootstrapping-pieces context: N=2^10, chain
    [40, 25, 25, 25, 35, 25, 25, 25] (S2C groups, one mult limb, C2S groups),
    P = 2 x 60 bits, dnum 2, scale 2^35."""
    d = load_npz("bt10.npz")
    c = H.CKKSParamsConfig()
    c.N = 1 << 10
    c.mult_depth = 1
    c.q0_bit = 40
    c.q_bit = 35
    c.p_bits = [60, 60]
    c.dnum = 2
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    c.bts_keys = True
    bp = H.CKKSBootstrapParams()
    bp.sparse_secret_hamming_weight = 16
    bp.linear_transform_log_plaintext_scales = [[[25], [25], [25]], [[25], [25], [25]]]
    bp.linear_transform_depths = [3, 3]
    bp.linear_transform_use_qp = [True, True]
    c.bootstrap_params = bp
    ctx = H.CKKSContext(H.CKKSParams(c))
    P = ctx.getParams()
    assert list(P.getModuli()) == list(d["q"]) and list(P.getPModuli()) == list(d["p"])
    return ctx, d


def _s2c_literal(H, q):
    lit = H.CKKSBootstrapDftMatrixLiteral()
    lit.type = H.CKKSBootstrapDftType.SlotsToCoeffs
    lit.level = 3
    lit.stage_count = 3
    lit.log_plaintext_scale_groups = [[25], [25], [25]]
    lit.transform_scale = (float(q[0]) * 0.5) / 2.0 ** 35
    lit.plaintext_scale = float(q[3])
    return lit


def test_host_s2c_dft_plan_matches_reference_and_decodes():
    """The factorised S2C plan's QP diagonals equal the reference's bit for bit
    (bt10 digests); applying it maps slots v to coefficients 16 v
    (transform scale (q0/2)/2^35) and each stage matches the oracle's BSGS
    stage plus rescale."""
    H = host()
    ctx, d = _bt10_context(H)
    o = load_json("bt10_outputs.json")
    P = ctx.getParams()
    q, p = P.getModuli(), P.getPModuli()
    lit = _s2c_literal(H, q)
    plan = H.ckks_generate_bootstrap_dft_plan(ctx, lit)
    for pl in (plan, ctx.getEvalModS2CPlan()):
        assert pl.stageCount() == len(o["s2c_stages"])
        for i, ref in enumerate(o["s2c_stages"]):
            st = pl.stage(i)
            assert st.baby_step_count == ref["baby_steps"] and st.qp_limb_count == ref["qp_limbs"], i
            assert list(st.diagonal_indices) == ref["diag_indices"], i
            flat = np.concatenate([st.downloadDiagonalQp(ctx, t) for t in range(len(st.diagonal_indices))])
            assert digest(flat) == ref["digest"], i

    N = 1 << 10
    for i in range(plan.stageCount()):
        st = plan.stage(i)
        for dgi in st.diagonal_indices:
            for r in (dgi % st.baby_step_count, (dgi // st.baby_step_count) * st.baby_step_count):
                if r:
                    ctx.addRotationKey(int(r))
    rng = np.random.default_rng(8)
    v = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    ct = H.ckks_encrypt(ctx, v, 3)
    out = H.ckks_apply_linear_transform_plan(ctx, ct, plan)
    assert out.getLevel() == 0
    assert out.getMessageEncodingState() == H.CKKSMessageEncodingState.Coefficients
    dec = H.ckks_decrypt(ctx, out, H.CKKSDecryptDomain.Coefficients)
    assert np.max(np.abs(dec - lit.transform_scale * v)) < 1e-2

    # stage 0 against the oracle's BSGS stage and rescale
    octx = O.Ctx(N, q, p, 2)
    st = plan.stage(0)
    L = st.level + 1
    diags = np.concatenate([st.downloadDiagonalQp(ctx, t) for t in range(len(st.diagonal_indices))])
    keys = {}
    for dgi in st.diagonal_indices:
        for r in (dgi % st.baby_step_count, (dgi // st.baby_step_count) * st.baby_step_count):
            if r:
                keys[int(r)] = ctx.getRotationKey(int(r))
    ao, bo = octx.lt_stage_qp(st.level, ct.getPoly(1), ct.getPoly(0), list(st.diagonal_indices), diags,
                               st.baby_step_count, keys)
    r = octx.rescale(st.level, np.stack([bo, ao]))
    x = H.CKKSCiphertext(ctx, [ct.getPoly(0), ct.getPoly(1)], ct.getScale(), 3)
    y = H.ckks_apply_linear_transform_plan_stage(ctx, x, plan, 0)
    assert np.array_equal(y.getPoly(0), r[0]) and np.array_equal(y.getPoly(1), r[1])
    assert L == 4


def test_host_dense_sparse_switches_and_modraise_match_reference():
    """With the reference's dense-to-sparse key installed, the host API's
    switch and ModRaise reproduce the reference's outputs (bt10)."""
    H = host()
    ctx, d = _bt10_context(H)
    o = load_json("bt10_outputs.json")
    N = 1 << 10
    ctx.setDenseToSparseSwitchKey(d["key_d2s"])
    b, a = d["ct0"].reshape(2, N)
    ct = H.CKKSCiphertext(ctx, [b, a], 2.0 ** 35, 0)
    H.ckks_key_switch_dense_to_sparse_inplace(ctx, ct)
    assert ct.getSecretOwner() == H.CKKSSecretOwner.BootstrapSparse
    assert digest(np.concatenate([ct.getPoly(0), ct.getPoly(1)])) == o["q0p0"]
    H.ckks_modraise_inplace(ctx, ct, o["modraise_level"])
    assert ct.getLevel() == o["modraise_level"]
    assert digest(np.concatenate([ct.getPoly(0), ct.getPoly(1)])) == o["modraise"]


def test_host_linear_transform_bsgs_matches_reference_lq10():
    """ckks_apply_linear_transform_bsgs (Q evaluator, normal-form diagonals,
    rescale after) with the reference's keys reproduces its output limbs,
    level and scale; its missing-key error names the reference's message."""
    H = host()
    d = load_npz("lq10.npz")
    o = load_json("lq10_outputs.json")
    c = H.CKKSParamsConfig()
    c.N = 1 << 10
    c.mult_depth = 3
    c.q0_bit = 50
    c.q_bit = 40
    c.p_bits = [60, 60]
    c.dnum = 2
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    c.rotate_values = [1, 2, 3, 6]
    ctx = H.CKKSContext(H.CKKSParams(c))
    P = ctx.getParams()
    q = P.getModuli()
    assert list(q) == list(d["q"]) and list(P.getPModuli()) == list(d["p"])
    for r in (1, 2, 3, 6):
        ctx.setRotationKey(r, d[f"key_rot{r}"])
    N, L = 1 << 10, len(q)
    W = L * N
    scale = float(d["ct_meta"][2:3].view(np.float64)[0])
    ct = H.CKKSCiphertext(ctx, [d["ct"][:W], d["ct"][W:]], scale, L - 1)
    e = o["bsgs"]
    diags = lt_diags(N, q, 14, seed_base=e["diag_seed_base"]).reshape(14, W)
    out = H.ckks_apply_linear_transform_bsgs(ctx, ct, [diags[t] for t in e["terms"]], o["pt_scale"], False,
                                             e["diag_indices"], e["baby_steps"])
    assert digest(np.concatenate([out.getPoly(0), out.getPoly(1)])) == e["digest"]
    assert out.getLevel() == e["level"] and out.getScale() == e["scale"]
    with pytest.raises(ValueError, match="missing giant-step rotation key"):
        H.ckks_apply_linear_transform_bsgs(ctx, ct, [diags[0]], o["pt_scale"], False, [12], 3)


def test_host_bsgs_stage_error_vs_fp64_matvec_at_full_size():
    """The BSGS linear-transform workload at full size: N=2^16, 2^15 slots,
    level 24 of Q = 30 x 50-bit, P = 10 x 50-bit, dnum 3; 128 diagonals
    {16 g + b : g < 8, b < 16} with real and imaginary parts uniform in
    [-1, 1], encoded at scale 2^40.  Through both evaluators the decrypted
    result matches the FP64 matrix-vector product y_i = sum_d diag_d[i] x[i+d]
    within 1e-6 (the stated tolerance; the observed error is printed)."""
    H = host()
    c = H.CKKSParamsConfig()
    c.N = 1 << 16
    c.mult_depth = 29
    c.q0_bit = 50
    c.q_bit = 50
    c.p_bits = [50] * 10
    c.dnum = 3
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    c.rotate_values = list(range(1, 16)) + [16 * g for g in range(1, 8)]
    c.seed = 7
    ctx = H.CKKSContext(H.CKKSParams(c))
    n = 1 << 15
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    diags = {16 * g + b: rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n) for g in range(8) for b in range(16)}
    want = sum(v * np.roll(x, -d) for d, v in diags.items())
    ct = H.ckks_encrypt(ctx, x, 24)
    outs = []
    for qp in (True, False):
        plan = H.ckks_encode_linear_transform_plan(ctx, diags, 24, 16, 2.0 ** 40, qp)
        out = H.ckks_apply_linear_transform_plan(ctx, ct, plan)
        assert out.getLevel() == 23
        dec = H.ckks_decrypt(ctx, out)
        err = float(np.max(np.abs(dec - want)))
        print(f"BSGS stage N=2^16 level 24, 128 diagonals, {'QP' if qp else 'Q'} evaluator: "
              f"max |error| vs FP64 mat-vec = {err:.3e}")
        assert err < 1e-6, (qp, err)
        outs.append(dec)
    assert np.max(np.abs(outs[0] - outs[1])) < 1e-6


def _bits_bootstrap_context(H):
    """N=2^10 bootstrapping chain: q0 41 bits (one bit above the 2^40 scale, the
    relation the EvalMod C2S scaling assumes: bit b sits at b * q0 / 2 after
    S2C, and the C2S output keeps the 2^40 scale), S2C 3 x 30, one 40-bit mult limb,
    7 EvalMod limbs of 40 bits, C2S 3 x 30; P = 3 x 60 bits, dnum 4 (digits of
    at most 160 bits); sparse secret of Hamming weight 16 (EvalMod K = 9)."""
    c = H.CKKSParamsConfig()
    c.N = 1 << 10
    c.mult_depth = 1
    c.q0_bit = 41
    c.q_bit = 40
    c.p_bits = [60, 60, 60]
    c.dnum = 4
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    c.bts_keys = True
    c.seed = 99
    bp = H.CKKSBootstrapParams()
    bp.sparse_secret_hamming_weight = 16
    bp.linear_transform_log_plaintext_scales = [[[30], [30], [30]], [[30], [30], [30]]]
    bp.linear_transform_depths = [3, 3]
    bp.evalmod_depth = 7
    c.bootstrap_params = bp
    return H.CKKSContext(H.CKKSParams(c))


def test_host_bootstrap_refreshes_encrypted_bits():
    """End to end: two slot vectors of bits at level 3 go through S2C, the
    sparse-secret ModRaise, C2S and EvalMod and come back as fresh encryptions
    of the same bits.  The EvalMod polynomial is the test's own Chebyshev
    interpolant of 0.5 - 0.5 cos(2 pi K u) on [-1, 1] (the reference derives its
    coefficients with MPFR; those doubles are not reproduced here)."""
    H = host()
    ctx = _bits_bootstrap_context(H)
    P = ctx.getParams()
    K = P.getBootstrapEvalModK()
    assert K == 9
    deg = 100
    coeffs = np.polynomial.chebyshev.chebinterpolate(lambda u: 0.5 - 0.5 * np.cos(2 * np.pi * K * u), deg)
    coeffs[1::2] = 0.0
    ctx.setEvalModPolynomial(list(coeffs))
    N = 1 << 10
    rng = np.random.default_rng(17)
    b0 = rng.integers(0, 2, N // 2).astype(float)
    b1 = rng.integers(0, 2, N // 2).astype(float)
    s2c = P.getBootstrapS2CDepth()
    ct0 = H.ckks_encrypt(ctx, b0.astype(complex), s2c)
    ct1 = H.ckks_encrypt(ctx, b1.astype(complex), s2c)
    H.ckks_batch_bootstrap_evalmod_inplace(ctx, ct0, ct1)
    d0 = H.ckks_decrypt(ctx, ct0)
    d1 = H.ckks_decrypt(ctx, ct1)
    assert ct0.getLevel() > 0 and ct0.getScale() == P.getScale()
    assert np.max(np.abs(d0.real - b0)) < 0.05 and np.max(np.abs(d1.real - b1)) < 0.05


def test_host_serialization_round_trips_and_rejects_corruption():
    """Ciphertexts, plaintexts and evaluation keys survive a byte round trip
    bit for bit (keys loaded into a second context with another seed switch
    exactly as they do in the first); blobs from another chain, truncated or
    altered blobs are rejected."""
    H = host()
    c = _config1(H, seed=21)
    c.N = 1 << 12
    c.rotate_values = [1, 3]
    ctx = H.CKKSContext(H.CKKSParams(c))
    rng = np.random.default_rng(5)
    m = rng.uniform(-1, 1, 1 << 11) + 1j * rng.uniform(-1, 1, 1 << 11)
    ct = H.ckks_encrypt(ctx, m, 6)
    H.ckks_mul_inplace(ctx, ct, H.ckks_encrypt(ctx, m, 6))  # three polys
    blob = H.ckks_serialize_ciphertext(ct)
    back = H.ckks_deserialize_ciphertext(ctx, blob)
    assert back.getNumPolys() == 3 and back.getLevel() == 6 and back.getScale() == ct.getScale()
    for k in range(3):
        assert np.array_equal(back.getPoly(k), ct.getPoly(k))
    pt = H.ckks_encode(ctx, m, 4, eval=True)
    pback = H.ckks_deserialize_encoding(ctx, H.ckks_serialize_encoding(pt))
    assert np.array_equal(pback.getPlaintext(), pt.getPlaintext()) and pback.getLevel() == 4
    # keys: a second context (other seed, so other keys) given the first
    # context's keys switches exactly as the first one does
    keys = H.ckks_serialize_evaluation_keys(ctx)
    c2 = _config1(H, seed=22)
    c2.N = 1 << 12
    c2.rotate_values = [1, 3]
    ctx2 = H.CKKSContext(H.CKKSParams(c2))
    H.ckks_load_evaluation_keys(ctx2, keys)
    x1 = H.ckks_encrypt(ctx, m, 5)
    x2 = H.ckks_deserialize_ciphertext(ctx2, H.ckks_serialize_ciphertext(x1))
    H.ckks_rotate_inplace(ctx, x1, 3)
    H.ckks_rotate_inplace(ctx2, x2, 3)
    assert np.array_equal(x1.getPoly(0), x2.getPoly(0)) and np.array_equal(x1.getPoly(1), x2.getPoly(1))
    with pytest.raises(ValueError, match="checksum"):
        H.ckks_deserialize_ciphertext(ctx, blob[:100] + bytes([blob[100] ^ 1]) + blob[101:])
    with pytest.raises(ValueError, match="checksum|truncated"):
        H.ckks_deserialize_ciphertext(ctx, blob[: len(blob) // 2])
    other = _config1(H, seed=21)
    other.N = 1 << 12
    other.q_bit = 45
    ctx3 = H.CKKSContext(H.CKKSParams(other))
    with pytest.raises(ValueError, match="different ring or modulus chain"):
        H.ckks_deserialize_ciphertext(ctx3, blob)
    with pytest.raises(ValueError, match="wrong magic"):
        H.ckks_deserialize_encoding(ctx, blob)


@pytest.mark.parametrize("logn,mult_depth", [(10, 3), (13, 6), (16, 20)])
def test_host_device_encode_matches_host_restatement(logn, mult_depth):
    """ckks_encode runs the special FFT, rounding and residues on the device;
    its limbs equal the host restatement of encode_slots_custom_basis (which
    the reference's golden encodings pin) for small and large magnitudes and
    two scales."""
    H = host()
    c = H.CKKSParamsConfig()
    c.N = 1 << logn
    c.mult_depth = mult_depth
    c.q0_bit = 55
    c.q_bit = 45
    c.p_bits = [60] * (6 if mult_depth > 10 else 3)
    c.dnum = 3
    c.secure = False
    ctx = H.CKKSContext(H.CKKSParams(c))
    q = list(ctx.getParams().getModuli())
    rng = np.random.default_rng(logn)
    n = 1 << (logn - 1)
    for mag, scale, level in [(1.0, 2.0 ** 45, mult_depth), (1e3, 2.0 ** 30, 1), (1e-6, 2.0 ** 55, 0)]:
        v = mag * (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n))
        pt = H.ckks_encode(ctx, v, level, scale=scale, eval=False)
        want = H.encode_slots_custom_basis(ctx, v, scale, q[: level + 1])
        assert np.array_equal(pt.getPlaintext(), want), (mag, scale, level)
