"""HIP graph capture through the C-ABI (mi_graph_*): a recorded sequence of
evaluator calls replays bit-exactly against the oracle on every replay, with
eager calls (sharing the same workspaces) between the replays."""
import ctypes

import numpy as np
import pytest

import oracle as O
from helpers import product, rand_residues

pytestmark = pytest.mark.gpu


def _setup(N, nq, np_, dnum, seed):
    p, q = O.find_primes([55] * np_, [50] * nq, N)
    M = product()
    return M, M.Context(N, q, p, dnum), O.Ctx(N, q, p, dnum), q, p, np.random.default_rng(seed)


def _key(rng, q, p, N, parts):
    basis = list(q) + list(p)
    return np.concatenate([rand_residues(rng, basis, N) for _ in range(2 * parts)])


def test_graph_replays_match_oracle():
    """Mixed-level key-switch batch (independent level groups fork onto side
    streams and join back inside the graph), a rescale, a tensor, a zero fill
    and an addition, captured once and replayed three times."""
    import torch
    N = 1 << 13
    M, ctx, octx, q, p, rng = _setup(N, 10, 3, 3, seed=41)
    keys = [_key(rng, q, p, N, octx.parts(9)) for _ in range(3)]
    dkeys = [M.to_dev(k) for k in keys]
    levels = [9, 2, 5, 0, 7, 9]
    host = []
    for level in levels:
        host.append([rand_residues(rng, q[: level + 1], N) for _ in range(3)])  # kt, b, a
    # item 0's (b, a) live in one buffer so the rescale sees a 2-poly ciphertext
    W0 = 10 * N
    ct0 = M.to_dev(np.concatenate([host[0][1], host[0][2]]))
    dev = []
    for i, (kt, b, a) in enumerate(host):
        if i == 0:
            dev.append([M.to_dev(kt), ct0[:W0], ct0[W0:]])
        else:
            dev.append([M.to_dev(kt), M.to_dev(b), M.to_dev(a)])
    t2b, t2a = rand_residues(rng, q[:3], N), rand_residues(rng, q[:3], N)
    d_t2b, d_t2a = M.to_dev(t2b), M.to_dev(t2a)
    c2 = M.to_dev(np.zeros(3 * N, np.uint64))
    y = rand_residues(rng, q[:4], N)
    d_y, d_z = M.to_dev(y), M.to_dev(rand_residues(rng, q[:4], N))

    # expected, from the oracle
    exp = []
    for i, (kt, b, a) in enumerate(host):
        exp.append(list(octx.keyswitch(levels[i], kt, keys[i % 3], b, a, 0, 1)))
    r0 = octx.rescale(9, np.stack(exp[0]))
    tb, ta, tc = octx.tensor(2, exp[1][0], exp[1][1], t2b, t2a)

    originals = [[t.clone() for t in d] for d in dev]
    z_dirty = d_z.clone()

    st = torch.cuda.Stream()

    def sequence():
        items = [(levels[i], dev[i][0], dkeys[i % 3], dev[i][1], dev[i][2], 0, 1) for i in range(len(levels))]
        ctx.keyswitch_batch(items, stream=st)
        ctx.rescale(9, ct0, 2, stride=W0, stream=st)
        ctx.tensor(2, dev[1][1], dev[1][2], d_t2b, d_t2a, c2, stream=st)
        M.check(M.lib().mi_memset_zero(ctx.h, M.ptr(d_z), d_z.numel() * 8, M._stream(st)))
        ctx.add(3, d_z, d_y, 1, stream=st)

    def restore():
        with torch.cuda.stream(st):
            for d, o in zip(dev, originals):
                for t, s in zip(d, o):
                    t.copy_(s)
            d_z.copy_(z_dirty)
            c2.zero_()

    def check(tag):
        got0 = M.from_dev(ct0).reshape(2, W0)
        assert np.array_equal(got0[:, : 9 * N], r0), tag
        assert np.array_equal(M.from_dev(dev[1][1]), tb) and np.array_equal(M.from_dev(dev[1][2]), ta), tag
        assert np.array_equal(M.from_dev(c2), tc), tag
        for i in range(2, len(levels)):
            assert np.array_equal(M.from_dev(dev[i][1]), exp[i][0]), (tag, i)
            assert np.array_equal(M.from_dev(dev[i][2]), exp[i][1]), (tag, i)
        assert np.array_equal(M.from_dev(d_z), y), tag

    # eager once: sizes the workspaces of the stream and the side streams
    restore()
    sequence()
    st.synchronize()
    check("eager")

    restore()
    st.synchronize()
    ctx.graph_begin(st)
    sequence()
    graph = ctx.graph_end(st)
    try:
        # nothing ran during the capture
        assert np.array_equal(M.from_dev(dev[2][1]), host[2][1])
        other_kt, other_b, other_a = (rand_residues(rng, q[:8], N) for _ in range(3))
        for r in range(3):
            restore()
            graph.launch(st)
            st.synchronize()
            check(f"replay {r}")
            # eager work between replays, on the same stream and workspaces
            db, da = M.to_dev(other_b), M.to_dev(other_a)
            ctx.keyswitch(7, M.to_dev(other_kt), dkeys[r], db, da, 0, 1, stream=st)
            st.synchronize()
            ob, oa = octx.keyswitch(7, other_kt, keys[r], other_b, other_a, 0, 1)
            assert np.array_equal(M.from_dev(db), ob) and np.array_equal(M.from_dev(da), oa), r
    finally:
        graph.destroy()


def test_graph_capture_rules():
    """A call whose stream workspace would have to grow during a capture fails
    with a clear error; device blocks freed inside a capture are not handed
    out again while the graph is alive."""
    import torch
    N = 1 << 12
    M, ctx, octx, q, p, rng = _setup(N, 6, 2, 3, seed=43)
    key = M.to_dev(_key(rng, q, p, N, octx.parts(5)))
    kt, b, a = (M.to_dev(rand_residues(rng, q, N)) for _ in range(3))
    scratch = M.to_dev(np.zeros(N, np.uint64))
    st = torch.cuda.Stream()
    lib = M.lib()

    # a fresh stream has no workspace yet
    ctx.graph_begin(st)
    M.check(lib.mi_memset_zero(ctx.h, M.ptr(scratch), N * 8, M._stream(st)))
    with pytest.raises(M.MiError, match="outside the capture"):
        ctx.keyswitch(5, kt, key, b, a, 0, 1, stream=st)
    graph = ctx.graph_end(st)
    graph.destroy()

    # pool blocks
    size = 1 << 20
    blk = ctypes.c_void_p()
    M.check(lib.mi_alloc(ctx.h, size, ctypes.byref(blk)))
    M.check(lib.mi_free(ctx.h, blk))          # cached in the pool
    ctx.graph_begin(st)
    inner = ctypes.c_void_p()
    M.check(lib.mi_alloc(ctx.h, size, ctypes.byref(inner)))
    assert inner.value == blk.value           # reused outside any graph
    M.check(lib.mi_memset_zero(ctx.h, inner, size, M._stream(st)))
    M.check(lib.mi_free(ctx.h, inner))
    graph = ctx.graph_end(st)
    again = ctypes.c_void_p()
    M.check(lib.mi_alloc(ctx.h, size, ctypes.byref(again)))
    assert again.value != inner.value         # held while the graph lives
    graph.launch(st)
    st.synchronize()
    graph.destroy()
    M.check(lib.mi_free(ctx.h, again))
    back = ctypes.c_void_p()
    M.check(lib.mi_alloc(ctx.h, size, ctypes.byref(back)))
    assert back.value in (inner.value, again.value)  # both reusable again
    M.check(lib.mi_free(ctx.h, back))
