"""The bootstrap driver bit for bit against the reference (bs10 vectors).

The reference harness (oracle/ref_harness.cpp, run_bootstrap_sequence) runs
the sequence of ckks_batch_bootstrap_evalmod_from_complex_packed
(ckks_bootstrap_evalmod_inplace.cpp:354-434) through the reference's own
public functions on an N = 2^10 bootstrapping context, with a fixed even
degree-6 Chebyshev polynomial in place of the MPFR fit: S2C, dense -> sparse
switch, ModRaise, sparse -> dense switch, C2S with the EvalMod-scaled plan,
the conjugate split into real and imaginary lanes, and EvalMod on each.
With the reference's keys installed, the host API reproduces every step's
limbs (SHA-256 digests of the reference outputs), the S2C plan's and the
EvalMod-scaled C2S plan's QP diagonals, and the driver's two outputs."""
import numpy as np
import pytest

from helpers import digest, host, load_json, load_npz

pytestmark = pytest.mark.gpu


def _bs10_context(H):
    """N = 2^10, chain [40 | 25 x 3 S2C | 35 x 3 mult | 25 x 3 C2S], P = 3 x
    60 bits, dnum 2, scale 2^35, sparse weight 16, QP evaluators, no EvalMod
    limbs (the mult limbs carry the Chebyshev evaluation)."""
    d = load_npz("bs10.npz")
    c = H.CKKSParamsConfig()
    c.N = 1 << 10
    c.mult_depth = 3
    c.q0_bit = 40
    c.q_bit = 35
    c.p_bits = [60, 60, 60]
    c.dnum = 2
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    c.bts_keys = True
    c.seed = 5
    bp = H.CKKSBootstrapParams()
    bp.sparse_secret_hamming_weight = 16
    bp.linear_transform_log_plaintext_scales = [[[25], [25], [25]], [[25], [25], [25]]]
    bp.linear_transform_depths = [3, 3]
    bp.evalmod_depth = 0
    bp.linear_transform_use_qp = [True, True]
    c.bootstrap_params = bp
    ctx = H.CKKSContext(H.CKKSParams(c))
    P = ctx.getParams()
    assert list(P.getModuli()) == list(d["q"]) and list(P.getPModuli()) == list(d["p"])
    ctx.setRelinKey(d["key_relin"])
    ctx.setConjugationKey(d["key_conj"])
    ctx.setDenseToSparseSwitchKey(d["key_d2s"])
    ctx.setSparseToDenseSwitchKey(d["key_s2d"])
    for r in d["rots"]:
        ctx.setRotationKey(int(r), d["key_rot%d" % int(r)])
    ctx.setEvalModPolynomial(list(d["coeffs"]))
    return ctx, d


def _check(ct, ref, name):
    polys = np.concatenate([ct.getPoly(k) for k in range(ct.getNumPolys())])
    assert ct.getNumPolys() == ref["npolys"] and ct.getLevel() == ref["level"], name
    assert ct.getScale() == ref["scale"], (name, ct.getScale(), ref["scale"])
    assert digest(polys) == ref["digest"], name


def _input(H, ctx, d, o):
    N = 1 << 10
    b, a = d["ct_in"].reshape(2, -1)
    return H.CKKSCiphertext(ctx, [b, a], o["in"]["scale"], o["in"]["level"])


def test_bootstrap_plans_match_reference():
    H = host()
    ctx, d = _bs10_context(H)
    o = load_json("bs10_outputs.json")
    assert ctx.getEvalModScaledC2STransformScale() == float(d["scales"][0])
    for plan, ref in ((ctx.getEvalModS2CPlan(), o["s2c_plan"]), (ctx.getEvalModScaledC2SPlan(), o["c2s_plan"])):
        assert plan.stageCount() == len(ref)
        for i, rs in enumerate(ref):
            st = plan.stage(i)
            assert st.baby_step_count == rs["baby_steps"] and st.qp_limb_count == rs["qp_limbs"], i
            assert list(st.diagonal_indices) == rs["diag_indices"], i
            flat = np.concatenate([st.downloadDiagonalQp(ctx, t) for t in range(len(st.diagonal_indices))])
            assert digest(flat) == rs["digest"], i


def test_bootstrap_steps_match_reference():
    H = host()
    ctx, d = _bs10_context(H)
    o = load_json("bs10_outputs.json")
    P = ctx.getParams()
    x = _input(H, ctx, d, o)
    s2c = H.ckks_s2c_generalized(ctx, x, P.getBootstrapS2CDepth())
    s2c.setScale(P.getScale())
    _check(s2c, o["s2c_out"], "s2c")
    H.ckks_key_switch_dense_to_sparse_inplace(ctx, s2c)
    _check(s2c, o["d2s_out"], "dense->sparse")
    H.ckks_modraise_inplace(ctx, s2c, P.getDepth())
    _check(s2c, o["modraise_out"], "modraise")
    H.ckks_key_switch_sparse_to_dense_inplace(ctx, s2c)
    _check(s2c, o["s2d_out"], "sparse->dense")
    c2s = H.ckks_c2s_generalized(ctx, s2c, ctx.getEvalModScaledC2STransformScale())
    _check(c2s, o["c2s_out"], "c2s")
    # the driver end to end
    out0, out1 = H.ckks_batch_bootstrap_evalmod_from_complex_packed(ctx, _input(H, ctx, d, o))
    _check(out0, o["out0"], "out0")
    _check(out1, o["out1"], "out1")


def test_batched_bootstrap_matches_reference_per_member():
    """Three copies of the reference input as one ciphertext batch through
    the driver: every member's two outputs equal the reference's."""
    H = host()
    ctx, d = _bs10_context(H)
    o = load_json("bs10_outputs.json")
    batch = H.ckks_stack(ctx, [_input(H, ctx, d, o) for _ in range(3)])
    out0, out1 = H.ckks_batch_bootstrap_evalmod_from_complex_packed(ctx, batch)
    for lane, ref in ((out0, o["out0"]), (out1, o["out1"])):
        for m, member in enumerate(H.ckks_unstack(ctx, lane)):
            _check(member, ref, ("member", m))
