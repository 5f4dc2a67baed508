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
    (1 << 11, 43, 3, 2, 32767, True, True),    # one-word sums at their bound: q < 2^43, extreme coefficients
    (1 << 12, 36, 4, 3, 32767, False, False),
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


def _walsh_context(H, seed=5):
    c = H.CKKSParamsConfig()
    c.N = 1 << 12
    c.mult_depth = 4
    c.q0_bit = 60
    c.q_bit = 40
    c.p_bits = [60, 60, 60]
    c.dnum = 2
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    c.rotate_values = [1]
    c.seed = seed
    return H.CKKSContext(H.CKKSParams(c))


def _signs(H, ctx, nibbles, level):
    """16 sign ciphertexts of one nibble per slot: mask 0 the constant 1
    (zero_like + 1, as unpack_mask_packed_signs builds it), mask m the
    encryption of (-1)^parity(m & nibble)."""
    out = []
    for m in range(16):
        vals = (1 - 2 * np.array([bin(m & int(v)).count("1") & 1 for v in nibbles])).astype(np.complex128)
        ct = H.ckks_encrypt(ctx, vals, level)
        if m == 0:
            H.ckks_mul_int_inplace(ctx, ct, 0)
            H.ckks_add_integer_constant_inplace(ctx, ct, 1)
        out.append(ct)
    return out


def test_walsh_sum_for_output_pair_matches_the_reference_sequence():
    """The fused output-pair sum equals, bit for bit and in scale, the
    reference's sequence of host operations: linear combinations by scaled
    additions and constants, the imaginary part times the encoded i, raw
    tensors summed, one relinearisation, the integer scale-up of the h = 0
    term and the final addition (ckks_walsh_sbox.cpp:288-328, 862-1043)."""
    from helpers import host
    H = host()
    ctx = _walsh_context(H)
    level = 3
    slots = ctx.getParams().getSlots()
    rng = np.random.default_rng(9)
    x = rng.integers(0, 256, slots)
    low, high = _signs(H, ctx, x & 15, level), _signs(H, ctx, x >> 4, level)
    enc_i = H.ckks_encode(ctx, np.full(slots, 1j), level, 1.0, True)
    table = np.array(H.ckks_aes_walsh_table(), dtype=np.int64)
    rb, ib = 2, 6
    fused = H.ckks_walsh_sum_for_output_pair(ctx, low, high, rb, ib, enc_i)

    S = low[0].getScale()

    def lc(coeffs):
        out = low[0].clone()
        H.ckks_mul_int_inplace(ctx, out, 0)
        if coeffs[0]:
            H.ckks_add_integer_constant_inplace(ctx, out, int(coeffs[0]))
        for m in range(1, 16):
            if coeffs[m]:
                term = low[m].clone()
                H.ckks_mul_int_inplace(ctx, term, int(coeffs[m]))
                H.ckks_add_inplace(ctx, out, term)
        return out

    def z(h):
        real, imag = lc(table[rb][:, h]), lc(table[ib][:, h])
        H.ckks_mul_plain_inplace(ctx, imag, enc_i)
        imag.setScale(S)
        H.ckks_add_inplace(ctx, real, imag)
        return real

    acc = z(0)
    raw = None
    for h in range(1, 16):
        zh = z(h)
        zh.setScale(S)
        H.ckks_mul_inplace(ctx, zh, high[h])
        if raw is None:
            raw = zh
        else:
            H.ckks_add_inplace(ctx, raw, zh)
    H.ckks_relin_hybrid_inplace(ctx, raw)
    factor = raw.getScale() / acc.getScale()
    assert factor == float(int(factor))
    H.ckks_mul_int_inplace(ctx, acc, int(factor))
    acc.setScale(raw.getScale())
    H.ckks_add_inplace(ctx, acc, raw)

    assert fused.getNumPolys() == 2 and fused.getLevel() == level
    assert fused.getScale() == acc.getScale()
    for k in range(2):
        assert np.array_equal(fused.getPoly(k), acc.getPoly(k)), k


def test_walsh_recombine_decrypts_to_sbox_bits():
    """Signs of random bytes through the four output pairs decrypt to the
    AES S-box output bits (real: bit p, imaginary: bit p + 4)."""
    from helpers import host
    H = host()
    ctx = _walsh_context(H, seed=6)
    level = 3
    slots = ctx.getParams().getSlots()
    rng = np.random.default_rng(10)
    x = rng.integers(0, 256, slots)
    low, high = _signs(H, ctx, x & 15, level), _signs(H, ctx, x >> 4, level)
    outs = H.ckks_walsh_recombine_signs_to_paired_outputs(ctx, low, high)
    assert len(outs) == 4
    table = np.array(H.ckks_aes_walsh_table(), dtype=np.int64)
    # S(x) from the table (pinned against FIPS-197 values on the CPU)
    par = np.array([[bin(m & int(v)).count("1") & 1 for v in range(256)] for m in range(256)])
    masks = np.arange(256)
    f = np.stack([4 * table[b][masks & 15, masks >> 4] for b in range(8)]) @ (1 - 2 * par)
    sbox = ((f < 0).astype(np.int64) * (1 << np.arange(8))[:, None]).sum(axis=0)
    sx = sbox[x]
    for p, ct in enumerate(outs):
        assert ct.getLevel() == level - 1
        dec = H.ckks_decrypt(ctx, ct)
        want = ((sx >> p) & 1) + 1j * ((sx >> (p + 4)) & 1)
        err = np.max(np.abs(dec - want))
        print(f"pair {p}: max error {err:.2e}", flush=True)
        assert err < 0.02, (p, err)


def _packing_context(H, bc, seed=7):
    c = H.CKKSParamsConfig()
    c.N = 1 << 12
    c.mult_depth = 2
    c.q0_bit = 60
    c.q_bit = 40
    c.p_bits = [60, 60]
    c.dnum = 2
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    c.rotate_values = [bc * s for s in range(1, 30)] + [-bc * s for s in range(1, 30)]
    c.seed = seed
    return H.CKKSContext(H.CKKSParams(c))


def test_walsh_parity_packing_matches_rotations_and_unpacks():
    """The mask-packed parity input (30 parity sums, 29 batched rotations,
    summed) equals the reference's sequence of clones, additions and
    one-at-a-time rotations bit for bit; unpacking a packed ciphertext (one
    hoist, 29 batched hoisted rotations) decrypts to each segment rotated
    back, with the constant-1 signs at mask 0."""
    from helpers import host
    H = host()
    bc = 64
    ctx = _packing_context(H, bc)
    level = 2
    slots = ctx.getParams().getSlots()
    rng = np.random.default_rng(12)
    vals = rng.integers(0, 2, (8, slots)).astype(np.float64)
    bits = [H.ckks_encrypt(ctx, v.astype(np.complex128), level) for v in vals]
    packed = H.ckks_walsh_build_mask_packed_parity_input(ctx, bits, bc)

    ref, want = None, np.zeros(slots, np.complex128)
    for mask in range(1, 16):
        for hi in (0, 1):
            sel = [4 * hi + k for k in range(4) if (mask >> k) & 1]
            t = bits[sel[0]].clone()
            for k in sel[1:]:
                H.ckks_add_inplace(ctx, t, bits[k])
            seg = 15 * hi + mask - 1
            if seg:
                H.ckks_rotate_inplace(ctx, t, seg * bc)
            want += np.roll(vals[sel].sum(axis=0), -seg * bc)
            if ref is None:
                ref = t
            else:
                H.ckks_add_inplace(ctx, ref, t)
    for k in range(2):
        assert np.array_equal(packed.getPoly(k), ref.getPoly(k)), k
    assert np.max(np.abs(H.ckks_decrypt(ctx, packed) - want)) < 1e-6

    v = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)
    src = H.ckks_encrypt(ctx, v, level)
    low, high = H.ckks_walsh_unpack_mask_packed_signs(ctx, src, bc)
    assert len(low) == 16 and len(high) == 16
    for signs in (low, high):
        assert np.max(np.abs(H.ckks_decrypt(ctx, signs[0]) - 1.0)) < 1e-9
    assert np.array_equal(low[1].getPoly(0), src.getPoly(0))
    for mask in range(1, 16):
        for hi, signs in ((0, low), (1, high)):
            seg = 15 * hi + mask - 1
            err = np.max(np.abs(H.ckks_decrypt(ctx, signs[mask]) - np.roll(v, seg * bc)))
            assert err < 1e-6, (hi, mask, err)


def test_aes_masking_matches_per_ciphertext_sequence():
    """The AES-CTR masking loops (one batched product and rescale over all
    ciphertexts) equal the reference's per-ciphertext sequence (encode the
    active-slot mask, multiply, rescale, restore the scale) bit for bit, in
    level and in scale (ckks_aes_ctr_walsh_sbox.cpp:217-329)."""
    from helpers import host
    H = host()
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
    c.bts_keys = True   # the S2C depth is read from the bootstrap parameters
    c.seed = 31
    bp = H.CKKSBootstrapParams()
    bp.sparse_secret_hamming_weight = 16
    bp.linear_transform_log_plaintext_scales = [[[30], [30], [30]], [[30], [30], [30]]]
    bp.linear_transform_depths = [3, 3]
    bp.evalmod_depth = 7
    c.bootstrap_params = bp
    ctx = H.CKKSContext(H.CKKSParams(c))
    P = ctx.getParams()
    q = list(P.getModuli())
    slots = P.getSlots()
    rng = np.random.default_rng(14)
    level = 6
    vals = [rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots) for _ in range(5)]
    state = [H.ckks_encrypt(ctx, v, level) for v in vals]
    ref = [s.clone() for s in state]
    target = state[0].getScale()

    def mask(count, lv, scale):
        m = np.zeros(slots, np.complex128)
        m[:count] = 1
        return H.ckks_encode(ctx, m, lv, scale, True)

    H.ckks_aes_mask_inactive_slots_inplace(ctx, state, 100)
    mk = mask(100, level, float(q[level]))
    for r in ref:
        H.ckks_mul_plain_inplace(ctx, r, mk)
        H.ckks_rescale(ctx, r)
        r.setScale(target)
    for s, r in zip(state, ref):
        assert s.getLevel() == r.getLevel() == level - 1 and s.getScale() == r.getScale() == target
        for k in range(2):
            assert np.array_equal(s.getPoly(k), r.getPoly(k))
    want = vals[0].copy()
    want[100:] = 0
    assert np.max(np.abs(H.ckks_decrypt(ctx, state[0]) - want)) < 1e-6

    # the final refresh, consuming the selection limb, then keeping the level
    for minimum, consume in ((2, True), (level - 1, False)):
        H.ckks_aes_mask_final_refresh_slots_inplace(ctx, state, 50, minimum)
        lv = ref[0].getLevel()
        mk = mask(50, lv, float(q[lv]) if consume else 1.0)
        for r in ref:
            H.ckks_mul_plain_inplace(ctx, r, mk)
            if consume:
                H.ckks_rescale(ctx, r)
                r.setScale(target)
        for s, r in zip(state, ref):
            assert s.getLevel() == r.getLevel() and s.getScale() == r.getScale()
            for k in range(2):
                assert np.array_equal(s.getPoly(k), r.getPoly(k)), (minimum, k)
    low = [H.ckks_encrypt(ctx, vals[0], 3)]
    with pytest.raises(ValueError, match="selection limb"):
        H.ckks_aes_mask_inactive_slots_inplace(ctx, low, 10)


def _sbox_context(H, bc):
    """The N = 2^10 bootstrapping chain of the host bootstrap test with the
    rotation keys of a bc-block mask packing, and the test's own Chebyshev
    interpolant as its EvalMod polynomial."""
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
    c.rotate_values = [bc * s for s in range(1, 30)] + [-bc * s for s in range(1, 30)]
    bp = H.CKKSBootstrapParams()
    bp.sparse_secret_hamming_weight = 16
    bp.linear_transform_log_plaintext_scales = [[[30], [30], [30]], [[30], [30], [30]]]
    bp.linear_transform_depths = [3, 3]
    bp.evalmod_depth = 7
    c.bootstrap_params = bp
    ctx = H.CKKSContext(H.CKKSParams(c))
    K = ctx.getParams().getBootstrapEvalModK()
    coeffs = np.polynomial.chebyshev.chebinterpolate(lambda u: 0.5 - 0.5 * np.cos(2 * np.pi * K * u), 100)
    coeffs[1::2] = 0.0
    ctx.setEvalModPolynomial(list(coeffs))
    return ctx


def test_walsh_sbox_end_to_end_on_a_bootstrapping_chain():
    """The whole mask-packed S-box (parity packing, complex-packed bootstrap
    to parity bits, signs, unpacking, Walsh output pairs of both lanes, split
    into 8 bits) on the N = 2^10 bootstrapping chain of the host bootstrap
    test: 16 blocks per lane, random bytes; the decrypted outputs round to
    the AES S-box bits of every block in both lanes.  The EvalMod polynomial
    is the test's own Chebyshev interpolant (the reference fits its
    coefficients with MPFR)."""
    from helpers import host
    H = host()
    bc = 16
    ctx = _sbox_context(H, bc)
    P = ctx.getParams()
    slots = P.getSlots()
    rng = np.random.default_rng(23)
    x0, x1 = rng.integers(0, 256, bc), rng.integers(0, 256, bc)
    level = P.getBootstrapS2CDepth()
    bits = []
    for k in range(8):
        v = np.zeros(slots, np.complex128)
        v[:bc] = ((x0 >> k) & 1) + 1j * ((x1 >> k) & 1)
        bits.append(H.ckks_encrypt(ctx, v, level))
    out = H.ckks_walsh_sbox_complex_blocks_mask_packed(ctx, bits, bc)
    assert len(out) == 8
    table = np.array(H.ckks_aes_walsh_table(), dtype=np.int64)
    par = np.array([[bin(m & int(v)).count("1") & 1 for v in range(256)] for m in range(256)])
    masks = np.arange(256)
    f = np.stack([4 * table[b][masks & 15, masks >> 4] for b in range(8)]) @ (1 - 2 * par)
    sbox = ((f < 0).astype(np.int64) * (1 << np.arange(8))[:, None]).sum(axis=0)
    worst = 0.0
    for k, ct in enumerate(out):
        dec = H.ckks_decrypt(ctx, ct)[:bc]
        want = ((sbox[x0] >> k) & 1) + 1j * ((sbox[x1] >> k) & 1)
        err = np.max(np.abs(dec - want))
        worst = max(worst, err)
        print(f"bit {k}: level {ct.getLevel()} max error {err:.3e}", flush=True)
    assert worst < 1e-4, worst


@pytest.mark.parametrize("streams", ["1", "0"])
def test_walsh_product_sum_batch_matches_per_item_calls(streams, monkeypatch):
    """mi_walsh_product_sum_batch over 3 members x 3 output pairs (9 items,
    each with its own coefficient tables and constants; the items of a member
    share its sign ciphertexts), on the library's side streams and on one
    stream: every output equal, bit for bit, to mi_walsh_product_sum of its
    item alone; overlapping outputs are refused."""
    import ctypes
    import torch
    monkeypatch.setenv("MICKKS_MULTI_STREAM", streams)
    M = product()
    N, level = 1 << 11, 2
    p, q = O.find_primes([51], [50] * 3, N)
    ctx = M.Context(N, q, p, 1)
    rng = np.random.default_rng(211)
    qs = q[: level + 1]
    W = (level + 1) * N
    rnd = lambda: [M.to_dev(rand_residues(rng, qs, N)), M.to_dev(rand_residues(rng, qs, N))]
    members, pairs = 3, 3
    signs = [([rnd() for _ in range(15)], [rnd() for _ in range(15)]) for _ in range(members)]
    pt = M.to_dev(rand_residues(rng, qs, N))
    items = []
    for i in range(members * pairs):
        cr = np.ascontiguousarray(rng.integers(-64, 65, (16, 16)).astype(np.int16))
        ci = np.ascontiguousarray(rng.integers(-64, 65, (16, 16)).astype(np.int16))
        consts = [(int(rng.integers(-2**62, 2**62)) * (2**20), -int(rng.integers(0, 2**50))) for _ in range(16)]
        absv = np.zeros(64, dtype=np.uint64)
        neg = np.zeros(32, dtype=np.int32)
        for h in range(16):
            for r in range(2):
                v, t = consts[h][r], 2 * h + r
                absv[2 * t], absv[2 * t + 1], neg[t] = abs(v) & (2**64 - 1), abs(v) >> 64, int(v < 0)
        items.append((i // pairs, cr, ci, consts, absv, neg))
    want = []
    for m, cr, ci, consts, _, _ in items:
        acc = [M.to_dev(np.zeros(W, np.uint64)) for _ in range(2)]
        raw = [M.to_dev(np.zeros(W, np.uint64)) for _ in range(3)]
        ctx.walsh_product_sum(level, signs[m][0], signs[m][1], cr, ci, consts, pt, acc, raw)
        want.append([M.from_dev(t) for t in acc + raw])
    K = len(items)
    outs = [[M.to_dev(np.zeros(W, np.uint64)) for _ in range(5)] for _ in range(K)]
    P = lambda xs: (ctypes.c_void_p * len(xs))(*xs)
    sig = lambda part, comp: P([signs[it[0]][part][h][comp].data_ptr() for it in items for h in range(15)])
    lib = M.lib()
    s = M._stream(torch.cuda.current_stream())

    def call(o):
        return lib.mi_walsh_product_sum_batch(
            ctx.h, level, K, sig(0, 0), sig(0, 1), sig(1, 0), sig(1, 1),
            P([it[1].ctypes.data for it in items]), P([it[2].ctypes.data for it in items]),
            P([it[4].ctypes.data for it in items]), P([it[5].ctypes.data for it in items]), pt.data_ptr(), 1,
            *[P([o[k][j].data_ptr() for k in range(K)]) for j in range(5)], s)

    print("walsh batch: 9 items, streams", streams, flush=True)
    M.check(call(outs))
    for k in range(K):
        for j in range(5):
            assert np.array_equal(M.from_dev(outs[k][j]), want[k][j]), (k, j)
    clash = [list(o) for o in outs]
    clash[1][0] = clash[0][0]
    assert call(clash) < 0
