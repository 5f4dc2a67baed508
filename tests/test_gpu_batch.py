"""Ciphertext batches: the same evaluation schedule over several ciphertexts
at once (the reference runs such schedules per ciphertext on OpenMP
threads: 16 S-boxes per AES round, ckks_aes_ctr_walsh_sbox.cpp:659-721).

* the list kernels a batch uses (constant add, zero fill, Montgomery
  plaintext add, listed tensor products and squares) against numpy with
  Python integers on the same residues;
* every evaluator function on a batch equals the same function on each
  member alone, limb for limb (tensor, square, relinearisation, rescale,
  rotation, conjugation, constants, plaintext products, Chebyshev
  evaluation, the bootstrap and the whole mask-packed Walsh S-box);
* the batched conjugation rejects a sparse-secret ciphertext as the
  one-at-a-time conjugation does."""
import ctypes

import numpy as np
import pytest

import oracle as O
from helpers import host, product, rand_residues

pytestmark = pytest.mark.gpu


def _arr(ptrs):
    return (ctypes.c_void_p * len(ptrs))(*ptrs)


@pytest.mark.parametrize("op", [5, 6, 7])
def test_new_list_ops_match_numpy(op):
    M = product()
    import torch
    N, level = 1 << 11, 2
    p, q = O.find_primes([60], [50, 45, 40], N)
    ctx = M.Context(N, q, p, 1)
    rng = np.random.default_rng(70 + op)
    L = level + 1
    count = 67  # more than one launch
    xs = [rand_residues(rng, q[:L], N) for _ in range(count)]
    ys = [rand_residues(rng, q[:L], N) for _ in range(count)]
    r = np.array([int(rng.integers(0, int(m))) for m in q[:L]], dtype=np.uint64)
    dx, dy = [M.to_dev(x) for x in xs], [M.to_dev(y) for y in ys]
    print(f"list op {op}: {count} polys", flush=True)
    M.check(M.lib().mi_poly_list_op(ctx.h, op, level, count, _arr([M.ptr(d) for d in dx]),
                                    _arr([M.ptr(d) for d in dy]) if op == 7 else None,
                                    r.ctypes.data_as(ctypes.c_void_p) if op == 5 else None,
                                    M._stream(torch.cuda.current_stream())))
    torch.cuda.synchronize()
    qv = np.repeat(np.array(q[:L], dtype=object), N)
    rv = np.repeat(r.astype(object), N)
    rinv = np.repeat(np.array([pow(1 << 64, -1, int(m)) for m in q[:L]], dtype=object), N)
    for x, y, d in zip(xs, ys, dx):
        xo, yo = x.astype(object), y.astype(object)
        if op == 5:
            want = (xo + rv) % qv
        elif op == 6:
            want = np.zeros(len(xo), dtype=object)
        else:
            want = (xo + yo * rinv) % qv
        assert np.array_equal(M.from_dev(d), np.array(want, dtype=np.uint64))


@pytest.mark.parametrize("square", [False, True])
def test_tensor_list_matches_numpy(square):
    M = product()
    import torch
    N, level = 1 << 10, 3
    p, q = O.find_primes([60], [60, 50, 45, 40], N)
    ctx = M.Context(N, q, p, 1)
    rng = np.random.default_rng(90 + square)
    L = level + 1
    count = 19  # more than one launch of 16
    polys = [[rand_residues(rng, q[:L], N) for _ in range(4)] for _ in range(count)]
    dev = [[M.to_dev(x) for x in item] for item in polys]
    c2 = [M.to_dev(np.zeros(L * N, np.uint64)) for _ in range(count)]
    b2 = None if square else _arr([M.ptr(d[2]) for d in dev])
    a2 = None if square else _arr([M.ptr(d[3]) for d in dev])
    M.check(M.lib().mi_tensor_list(ctx.h, level, count, _arr([M.ptr(d[0]) for d in dev]),
                                   _arr([M.ptr(d[1]) for d in dev]), b2, a2, _arr([M.ptr(d) for d in c2]),
                                   M._stream(torch.cuda.current_stream())))
    torch.cuda.synchronize()
    qv = np.repeat(np.array(q[:L], dtype=object), N)
    for item, d, dc in zip(polys, dev, c2):
        b1, a1, b2v, a2v = (x.astype(object) for x in item)
        if square:
            b2v, a2v = b1, a1
        want = [(b1 * b2v) % qv, (a1 * b2v + b1 * a2v) % qv, (a1 * a2v) % qv]
        got = [M.from_dev(d[0]), M.from_dev(d[1]), M.from_dev(dc)]
        for w, g in zip(want, got):
            assert np.array_equal(g, np.array(w, dtype=np.uint64))
    # overlapping outputs are refused
    one = _arr([M.ptr(dev[0][0])])
    assert M.lib().mi_tensor_list(ctx.h, level, 1, one, one, None, None, _arr([M.ptr(c2[0])]),
                                  M._stream(torch.cuda.current_stream())) < 0


@pytest.mark.parametrize("square", [False, True])
def test_tensor_list_into_matches_numpy(square):
    """mi_tensor_list_into: the inputs one level above the product's level
    (read at that level, their first limbs), the products in new polys, the
    inputs unchanged; an output on an input or on another output is refused."""
    M = product()
    import torch
    N, level = 1 << 10, 2
    p, q = O.find_primes([60], [60, 50, 45, 40], N)
    ctx = M.Context(N, q, p, 1)
    rng = np.random.default_rng(95 + square)
    L, Lin = level + 1, level + 2
    count = 19  # more than one launch of 16
    polys = [[rand_residues(rng, q[:Lin], N) for _ in range(4)] for _ in range(count)]
    dev = [[M.to_dev(x) for x in item] for item in polys]
    outs = [[M.to_dev(np.zeros(L * N, np.uint64)) for _ in range(3)] for _ in range(count)]
    b2 = None if square else _arr([M.ptr(d[2]) for d in dev])
    a2 = None if square else _arr([M.ptr(d[3]) for d in dev])
    s = M._stream(torch.cuda.current_stream())
    print("tensor_list_into: square", square, flush=True)
    M.check(M.lib().mi_tensor_list_into(ctx.h, level, count, _arr([M.ptr(d[0]) for d in dev]),
                                        _arr([M.ptr(d[1]) for d in dev]), b2, a2,
                                        *[_arr([M.ptr(o[j]) for o in outs]) for j in range(3)], s))
    torch.cuda.synchronize()
    qv = np.repeat(np.array(q[:L], dtype=object), N)
    for item, d, o in zip(polys, dev, outs):
        b1, a1, b2v, a2v = (x[: L * N].astype(object) for x in item)
        if square:
            b2v, a2v = b1, a1
        want = [(b1 * b2v) % qv, (a1 * b2v + b1 * a2v) % qv, (a1 * a2v) % qv]
        for w, g in zip(want, o):
            assert np.array_equal(M.from_dev(g), np.array(w, dtype=np.uint64))
        for x, dx in zip(item, d):
            assert np.array_equal(M.from_dev(dx), x)
    one = lambda t: _arr([M.ptr(t)])
    assert M.lib().mi_tensor_list_into(ctx.h, level, 1, one(dev[0][0]), one(dev[0][1]), None, None,
                                       one(dev[0][0]), one(outs[0][1]), one(outs[0][2]), s) < 0
    assert M.lib().mi_tensor_list_into(ctx.h, level, 1, one(dev[0][0]), one(dev[0][1]), None, None,
                                       one(outs[0][0]), one(outs[0][0]), one(outs[0][2]), s) < 0


def _chain_context(H, seed=21):
    c = H.CKKSParamsConfig()
    c.N = 1 << 12
    c.mult_depth = 4
    c.q0_bit = 50
    c.q_bit = 40
    c.p_bits = [60, 60]
    c.dnum = 3
    c.secure = False
    c.rotate_values = [1, 3, -7]
    c.seed = seed
    return H.CKKSContext(H.CKKSParams(c))


def _assert_members_equal(H, ctx, batch, singles):
    assert batch.getBatchSize() == len(singles)
    for m, s in enumerate(singles):
        assert batch.getLevel() == s.getLevel() and batch.getScale() == s.getScale()
        assert batch.getNumPolys() == s.getNumPolys()
        for k in range(s.getNumPolys()):
            assert np.array_equal(batch.getMemberPoly(k, m), s.getPoly(k)), (m, k)


def test_batched_evaluator_ops_equal_member_ops():
    H = host()
    ctx = _chain_context(H)
    slots = ctx.getParams().getSlots()
    rng = np.random.default_rng(5)
    B = 5
    xs = [rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots) for _ in range(B)]
    ys = [rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots) for _ in range(B)]
    cx = [H.ckks_encrypt(ctx, v, 4) for v in xs]
    cy = [H.ckks_encrypt(ctx, v, 4) for v in ys]
    bx, by = H.ckks_stack(ctx, cx), H.ckks_stack(ctx, cy)
    _assert_members_equal(H, ctx, bx, cx)
    pt = H.ckks_encode(ctx, rng.uniform(-1, 1, slots).astype(np.complex128), 4, 2.0 ** 20, True)

    def both(fn):
        fn(bx)
        for c in cx:
            fn(c)
        _assert_members_equal(H, ctx, bx, cx)

    print("tensor / relin / rescale", flush=True)
    both(lambda c: H.ckks_mul_inplace(ctx, c, by if c is bx else cy[cx.index(c)]))
    both(lambda c: H.ckks_relin_hybrid_inplace(ctx, c))
    both(lambda c: H.ckks_rescale(ctx, c))
    print("rotate / conjugate / constants", flush=True)
    both(lambda c: H.ckks_rotate_inplace(ctx, c, 3))
    both(lambda c: H.ckks_rotate_inplace(ctx, c, -7))
    both(lambda c: H.ckks_conjugate_inplace(ctx, c))
    both(lambda c: H.ckks_add_integer_constant_inplace(ctx, c, -3))
    both(lambda c: H.ckks_mul_int_inplace(ctx, c, 7))
    both(lambda c: H.ckks_mul_const_inplace(ctx, c, 0.5 - 0.25j, 2.0 ** 10))
    both(lambda c: H.ckks_rescale(ctx, c))
    print("square / plaintext", flush=True)
    both(lambda c: H.ckks_square_inplace(ctx, c))
    both(lambda c: H.ckks_relin_hybrid_inplace(ctx, c))
    both(lambda c: H.ckks_rescale(ctx, c))
    pt2 = H.ckks_encode(ctx, rng.uniform(-1, 1, slots).astype(np.complex128), bx.getLevel(), 2.0 ** 20, True)
    both(lambda c: H.ckks_add_plain_inplace(ctx, c, pt2))
    both(lambda c: H.ckks_mul_plain_inplace(ctx, c, pt2))
    del pt
    # the members come back out unchanged
    members = H.ckks_unstack(ctx, bx)
    for m, s in zip(members, cx):
        for k in range(2):
            assert np.array_equal(m.getPoly(k), s.getPoly(k))
    # members of differing metadata cannot be stacked
    with pytest.raises(ValueError, match="differ"):
        H.ckks_stack(ctx, [cy[0], H.ckks_encrypt(ctx, xs[0], 2)])


def test_batched_chebyshev_equals_member_evaluation():
    H = host()
    ctx = _chain_context(H, seed=23)
    slots = ctx.getParams().getSlots()
    rng = np.random.default_rng(8)
    B = 3
    cts = [H.ckks_encrypt(ctx, rng.uniform(-1, 1, slots).astype(np.complex128), 4) for _ in range(B)]
    coeffs = list(rng.uniform(-1, 1, 12))
    coeffs[1::2] = [0.0] * len(coeffs[1::2])
    batch = H.ckks_evalchebyshev_preconditioned(ctx, H.ckks_stack(ctx, cts), coeffs)
    singles = [H.ckks_evalchebyshev_preconditioned(ctx, c, coeffs) for c in cts]
    _assert_members_equal(H, ctx, batch, singles)


def test_conjugate_batch_of_stacked_ciphertexts_conjugates_every_member():
    """ckks_conjugate_batch over a list of stacked ciphertexts (two members
    each) conjugates all members of all elements, as ckks_conjugate does one
    ciphertext at a time."""
    H = host()
    ctx = _chain_context(H, seed=29)
    slots = ctx.getParams().getSlots()
    rng = np.random.default_rng(12)
    cts = [H.ckks_encrypt(ctx, rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots), 3) for _ in range(6)]
    stacked = [H.ckks_stack(ctx, cts[2 * i:2 * i + 2]) for i in range(3)]
    out = H.ckks_conjugate_batch(ctx, stacked)
    for i in range(3):
        _assert_members_equal(H, ctx, out[i], [H.ckks_conjugate(ctx, c) for c in cts[2 * i:2 * i + 2]])


def test_batched_conjugation_rejects_sparse_owner():
    """ckks_conjugate_batch checks each element as ckks_conjugate_inplace does."""
    H = host()
    from test_gpu_host import _bits_bootstrap_context
    ctx = _bits_bootstrap_context(H)
    P = ctx.getParams()
    v = np.zeros(P.getSlots(), np.complex128)
    ct = H.ckks_encrypt(ctx, v, 0)
    H.ckks_key_switch_dense_to_sparse_inplace(ctx, ct)
    dense = H.ckks_encrypt(ctx, v, 0)
    with pytest.raises(ValueError, match="sparse-secret"):
        H.ckks_conjugate_batch(ctx, [dense, ct])


def test_batched_walsh_sbox_equals_member_sboxes():
    """The whole mask-packed S-box on a batch of three inputs (bit k of every
    input stacked) equals three one-at-a-time S-boxes limb for limb, and each
    member decrypts to the AES S-box bits of its own input."""
    H = host()
    from test_gpu_walsh import _sbox_context
    bc = 16
    ctx = _sbox_context(H, bc)
    P = ctx.getParams()
    slots, level = P.getSlots(), P.getBootstrapS2CDepth()
    rng = np.random.default_rng(31)
    B = 3
    inputs = []
    for _ in range(B):
        x0, x1 = rng.integers(0, 256, bc), rng.integers(0, 256, bc)
        bits = []
        for k in range(8):
            v = np.zeros(slots, np.complex128)
            v[:bc] = ((x0 >> k) & 1) + 1j * ((x1 >> k) & 1)
            bits.append(H.ckks_encrypt(ctx, v, level))
        inputs.append((x0, x1, bits))
    stacked = [H.ckks_stack(ctx, [inp[2][k] for inp in inputs]) for k in range(8)]

    def same(batch, singles, what):
        assert batch.getLevel() == singles[0].getLevel() and batch.getScale() == singles[0].getScale(), what
        for m, s in enumerate(singles):
            for j in range(s.getNumPolys()):
                assert np.array_equal(batch.getMemberPoly(j, m), s.getPoly(j)), (what, m, j)

    # stage by stage: parity packing, bootstrap, unpacking, output pairs
    print("stages", flush=True)
    packed = H.ckks_walsh_build_mask_packed_parity_input(ctx, stacked, bc)
    packed1 = [H.ckks_walsh_build_mask_packed_parity_input(ctx, inp[2], bc) for inp in inputs]
    same(packed, packed1, "parity input")
    lanes = H.ckks_batch_bootstrap_evalmod_from_complex_packed(ctx, packed)
    lanes1 = [H.ckks_batch_bootstrap_evalmod_from_complex_packed(ctx, p) for p in packed1]
    for i in range(2):
        same(lanes[i], [l[i] for l in lanes1], ("bootstrap lane", i))
    low, high = H.ckks_walsh_unpack_mask_packed_signs(ctx, lanes[0], bc)
    for m in range(B):
        low1, high1 = H.ckks_walsh_unpack_mask_packed_signs(ctx, lanes1[m][0], bc)
        for t in range(16):
            for j in range(2):
                assert np.array_equal(low[t].getMemberPoly(j, m), low1[t].getPoly(j)), ("unpack low", t, m, j)
                assert np.array_equal(high[t].getMemberPoly(j, m), high1[t].getPoly(j)), ("unpack high", t, m, j)
    print("batched S-box", flush=True)
    out = H.ckks_walsh_sbox_complex_blocks_mask_packed(ctx, stacked, bc)
    sbox = np.array(H.aes_sbox_table(), dtype=np.int64)
    for m, (x0, x1, bits) in enumerate(inputs):
        print(f"member {m}", flush=True)
        single = H.ckks_walsh_sbox_complex_blocks_mask_packed(ctx, bits, bc)
        for k in range(8):
            assert out[k].getLevel() == single[k].getLevel() and out[k].getScale() == single[k].getScale()
            for j in range(2):
                assert np.array_equal(out[k].getMemberPoly(j, m), single[k].getPoly(j)), (m, k, j)
            dec = H.ckks_decrypt(ctx, single[k])[:bc]
            want = ((sbox[x0] >> k) & 1) + 1j * ((sbox[x1] >> k) & 1)
            assert np.max(np.abs(dec - want)) < 1e-3
