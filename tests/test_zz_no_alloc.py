"""mi_keyswitch_noalloc: the hybrid key switch with its scratch in the
library's fixed device array (no allocation) equals mi_keyswitch on the same
inputs bit for bit, on two calls in a row (the second at a lower level, so
smaller), and refuses a switch larger than the array with a clear error."""
import numpy as np
import pytest

import oracle as O
from helpers import product, rand_residues

pytestmark = pytest.mark.gpu


def _rand_key(rng, q, p, N, parts):
    basis = list(q) + list(p)
    return np.concatenate([rand_residues(rng, basis, N) for _ in range(2 * parts)])


def test_keyswitch_noalloc_matches_keyswitch():
    import torch
    M = product()
    lib = M.lib()
    # the walsh15 shape: N = 2^15, 13 Q limbs, 4 P limbs, dnum 3
    N = 1 << 15
    p, q = O.find_primes([61] * 4, [36, 61, 42, 42] + [38] * 7 + [50, 50], N)
    ctx = M.Context(N, q, p, 3)
    octx = O.Ctx(N, q, p, 3)
    rng = np.random.default_rng(404)
    key = _rand_key(rng, q, p, N, octx.parts(len(q) - 1))
    dkey = M.to_dev(key)
    s = M._stream(torch.cuda.current_stream())
    for level, bm, am in ((len(q) - 1, 0, 1), (5, 1, 0)):
        L = level + 1
        kt, b, a = (rand_residues(rng, q[:L], N) for _ in range(3))
        print(f"no-alloc key switch at level {level}", flush=True)
        ref_b, ref_a = M.to_dev(b), M.to_dev(a)
        ctx.keyswitch(level, M.to_dev(kt), dkey, ref_b, ref_a, bm, am)
        db, da, dkt = M.to_dev(b), M.to_dev(a), M.to_dev(kt)
        M.check(lib.mi_keyswitch_noalloc(ctx.h, level, M.ptr(dkt), M.ptr(dkey), M.ptr(db), M.ptr(da), bm, am, 1, s))
        torch.cuda.synchronize()
        assert torch.equal(db, ref_b) and torch.equal(da, ref_a), level
        assert np.array_equal(M.from_dev(dkt), kt)


def test_keyswitch_noalloc_refuses_beyond_its_scratch():
    import torch
    M = product()
    lib = M.lib()
    N = 1 << 16
    p, q = O.find_primes([50] * 10, [50] * 30, N)
    ctx = M.Context(N, q, p, 3)
    d = M.to_dev(np.zeros(30 * N, np.uint64))
    k = M.to_dev(np.zeros(2 * 3 * 40 * N, np.uint64))
    rc = lib.mi_keyswitch_noalloc(ctx.h, 29, M.ptr(d), M.ptr(k), M.ptr(d), M.ptr(d), 0, 1, 1,
                                  M._stream(torch.cuda.current_stream()))
    assert rc != 0
    assert "scratch" in lib.mi_last_error().decode()
