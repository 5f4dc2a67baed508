"""mi_poly_list_op (copy / add / sub of listed poly pairs in one launch)
against numpy on the same residues, including a list longer than one
launch (70 pairs), and its argument checks."""
import ctypes

import numpy as np
import pytest

import oracle as O
from helpers import product, rand_residues

pytestmark = pytest.mark.gpu


def _arr(ptrs):
    return (ctypes.c_void_p * len(ptrs))(*ptrs)


@pytest.mark.parametrize("op", [0, 1, 2])
def test_poly_list_op_matches_numpy(op):
    M = product()
    import torch
    N, level = 1 << 11, 3
    p, q = O.find_primes([60], [50, 45, 40, 55], N)
    ctx = M.Context(N, q, p, 1)
    rng = np.random.default_rng(40 + op)
    L = level + 1
    count = 70
    xs = [rand_residues(rng, q[:L], N) for _ in range(count)]
    ys = [rand_residues(rng, q[:L], N) for _ in range(count)]
    dx = [M.to_dev(x) for x in xs]
    dy = [M.to_dev(y) for y in ys]
    print(f"list op {op}: {count} pairs, N={N}, level {level}", flush=True)
    M.check(M.lib().mi_poly_list_op(ctx.h, op, level, count, _arr([M.ptr(d) for d in dx]),
                                    _arr([M.ptr(d) for d in dy]), None, M._stream(torch.cuda.current_stream())))
    torch.cuda.synchronize()
    qv = np.repeat(np.array(q[:L], dtype=object), N)
    for x, y, d in zip(xs, ys, dx):
        xo, yo = x.astype(object), y.astype(object)
        want = yo if op == 0 else (xo + yo) % qv if op == 1 else (xo - yo) % qv
        assert np.array_equal(M.from_dev(d), np.array(want, dtype=np.uint64))


def test_poly_list_op_rejects_bad_arguments():
    M = product()
    import torch
    N = 1 << 11
    p, q = O.find_primes([60], [50, 45], N)
    ctx = M.Context(N, q, p, 1)
    d = M.to_dev(np.zeros(2 * N, np.uint64))
    s = M._stream(torch.cuda.current_stream())
    one = _arr([M.ptr(d)])
    assert M.lib().mi_poly_list_op(ctx.h, 0, 1, 1, one, one, None, s) < 0       # copy onto itself
    assert M.lib().mi_poly_list_op(ctx.h, 8, 1, 1, one, one, None, s) < 0       # unknown operation
    assert M.lib().mi_poly_list_op(ctx.h, 3, 1, 1, one, one, None, s) < 0       # scaling without residues
    assert M.lib().mi_poly_list_op(ctx.h, 1, 2, 1, one, one, None, s) < 0       # level beyond the chain
    assert M.lib().mi_poly_list_op(ctx.h, 1, 1, 0, one, one, None, s) < 0       # empty list


@pytest.mark.parametrize("op", [3, 4])
def test_poly_list_scalar_ops_match_numpy(op):
    M = product()
    import torch
    N, level = 1 << 11, 2
    p, q = O.find_primes([60], [50, 45, 40], N)
    ctx = M.Context(N, q, p, 1)
    rng = np.random.default_rng(60 + op)
    L = level + 1
    xs = [rand_residues(rng, q[:L], N) for _ in range(3)]
    ys = [rand_residues(rng, q[:L], N) for _ in range(3)]
    r = np.array([int(rng.integers(0, int(m))) for m in q[:L]], dtype=np.uint64)
    dx, dy = [M.to_dev(x) for x in xs], [M.to_dev(y) for y in ys]
    M.check(M.lib().mi_poly_list_op(ctx.h, op, level, 3, _arr([M.ptr(d) for d in dx]), _arr([M.ptr(d) for d in dy]),
                                    r.ctypes.data_as(ctypes.c_void_p), M._stream(torch.cuda.current_stream())))
    torch.cuda.synchronize()
    qv = np.repeat(np.array(q[:L], dtype=object), N)
    rv = np.repeat(r.astype(object), N)
    for x, y, d in zip(xs, ys, dx):
        xo, yo = x.astype(object), y.astype(object)
        want = (xo * rv) % qv if op == 3 else (xo + yo * rv) % qv
        assert np.array_equal(M.from_dev(d), np.array(want, dtype=np.uint64))


def test_linear_combine_matches_numpy_and_checks_arguments():
    M = product()
    import torch
    N, level = 1 << 11, 2
    p, q = O.find_primes([60], [50, 45, 40], N)
    ctx = M.Context(N, q, p, 1)
    rng = np.random.default_rng(77)
    L = level + 1
    ins = [rand_residues(rng, q[:L], N) for _ in range(6)]
    din = [M.to_dev(x) for x in ins]
    nout, nterms = 14, 4     # two launches of up to 12 outputs
    dout = [M.to_dev(np.zeros(L * N, np.uint64)) for _ in range(nout)]
    idx = rng.integers(0, 6, (nout, nterms))
    coef = rng.integers(0, 9, (nout, nterms)).astype(np.uint32)
    src = _arr([M.ptr(din[i]) for i in idx.reshape(-1)])
    s = M._stream(torch.cuda.current_stream())
    print(f"linear combine: {nout} outputs x {nterms} terms", flush=True)
    M.check(M.lib().mi_linear_combine(ctx.h, level, nout, _arr([M.ptr(d) for d in dout]), nterms, src,
                                      coef.ctypes.data_as(ctypes.c_void_p), s))
    torch.cuda.synchronize()
    qv = np.repeat(np.array(q[:L], dtype=object), N)
    for k in range(nout):
        want = sum(int(coef[k, t]) * ins[idx[k, t]].astype(object) for t in range(nterms)) % qv
        assert np.array_equal(M.from_dev(dout[k]), np.array(want, dtype=np.uint64))
    one = _arr([M.ptr(din[0])])
    c1 = np.ones(1, np.uint32)
    assert M.lib().mi_linear_combine(ctx.h, level, 1, one, 1, one, c1.ctypes.data_as(ctypes.c_void_p), s) < 0  # aliasing
    c17 = np.ones(17, np.uint32)
    src17 = _arr([M.ptr(din[1])] * 17)
    assert M.lib().mi_linear_combine(ctx.h, level, 1, _arr([M.ptr(dout[0])]), 17, src17,
                                     c17.ctypes.data_as(ctypes.c_void_p), s) < 0                                # > 16 terms


@pytest.mark.parametrize("nlimbs,npolys,nterms", [(4, 40, 11), (40, 18, 16), (3, 70, 2), (5, 33, 30)])
def test_scaled_sum_matches_python_ints(nlimbs, npolys, nterms):
    """mi_scaled_sum (one Chebyshev leaf: constant + sum of per-limb scaled
    terms) against Python integers: terms in chunks (8 at 4 limbs, 5 at 40
    limbs) that accumulate, outputs in chunks of 16 to 64 polys by the
    chunk's term count, up to 30 terms, absent terms, a constant on the
    first polys, 61-bit moduli (128-bit sums)."""
    M = product()
    import torch
    N = 1 << 10
    level = nlimbs - 1
    p, q = O.find_primes([61], [61] * nlimbs, N)
    ctx = M.Context(N, q, p, 1)
    rng = np.random.default_rng(nlimbs * 100 + nterms)
    L = level + 1
    ins = [[rand_residues(rng, q[:L], N) if rng.random() > 0.15 else None for _ in range(npolys)]
           for _ in range(nterms)]
    din = [[M.to_dev(x) if x is not None else None for x in row] for row in ins]
    res = np.array([[int(rng.integers(0, int(q[l]))) for l in range(L)] for _ in range(nterms)], dtype=np.uint64)
    cst = np.array([int(rng.integers(0, int(q[l]))) for l in range(L)], dtype=np.uint64)
    nconst = npolys // 3
    dout = [M.to_dev(np.full(L * N, 7, np.uint64)) for _ in range(npolys)]
    src = _arr([M.ptr(d) if d is not None else None for row in din for d in row])
    s = M._stream(torch.cuda.current_stream())
    print(f"scaled sum: {npolys} polys x {nterms} terms x {L} limbs", flush=True)
    M.check(M.lib().mi_scaled_sum(ctx.h, level, npolys, _arr([M.ptr(d) for d in dout]), nterms, src,
                                  res.ctypes.data_as(ctypes.c_void_p), nconst, cst.ctypes.data_as(ctypes.c_void_p), s))
    torch.cuda.synchronize()
    qv = np.repeat(np.array(q[:L], dtype=object), N)
    for k in range(npolys):
        want = np.repeat(cst.astype(object), N) if k < nconst else np.zeros(L * N, dtype=object)
        for t in range(nterms):
            if ins[t][k] is not None:
                want = want + ins[t][k].astype(object) * np.repeat(res[t].astype(object), N)
        assert np.array_equal(M.from_dev(dout[k]), np.array(want % qv, dtype=np.uint64)), k
    one = _arr([M.ptr(din[0][0] if din[0][0] is not None else dout[0])])
    r1 = res[:1].copy()
    assert M.lib().mi_scaled_sum(ctx.h, level, 1, one, 1, one, r1.ctypes.data_as(ctypes.c_void_p), 0, None, s) < 0
    assert "overlaps" in M.lib().mi_last_error().decode()
    src65 = _arr([M.ptr(dout[1])] * 65)
    r65 = np.zeros((65, L), np.uint64)
    assert M.lib().mi_scaled_sum(ctx.h, level, 1, _arr([M.ptr(dout[0])]), 65, src65,
                                 r65.ctypes.data_as(ctypes.c_void_p), 0, None, s) < 0                # > 64 terms
    bad = res[:1].copy()
    bad[0, 0] = q[0]
    assert M.lib().mi_scaled_sum(ctx.h, level, 1, _arr([M.ptr(dout[0])]), 1, _arr([M.ptr(dout[1])]),
                                 bad.ctypes.data_as(ctypes.c_void_p), 0, None, s) < 0                # residue >= q


def test_conjugate_batch_equals_one_at_a_time():
    """ckks_conjugate_batch (mi_conjugate_batch) against ckks_conjugate of each
    ciphertext, limb for limb, on 5 ciphertexts."""
    from helpers import host
    H = host()
    c = H.CKKSParamsConfig()
    c.N = 1 << 12
    c.mult_depth = 3
    c.q0_bit = 50
    c.q_bit = 40
    c.p_bits = [60, 60]
    c.dnum = 2
    c.secure = False
    c.seed = 81
    ctx = H.CKKSContext(H.CKKSParams(c))
    rng = np.random.default_rng(82)
    slots = ctx.getParams().getSlots()
    cts = [H.ckks_encrypt(ctx, rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots), 2) for _ in range(5)]
    batch = H.ckks_conjugate_batch(ctx, cts)
    for ct, got in zip(cts, batch):
        want = H.ckks_conjugate(ctx, ct)
        for k in range(2):
            assert np.array_equal(np.asarray(got.getPoly(k)), np.asarray(want.getPoly(k)))
