"""A batched walsh15-shaped S-box (parity packing, the complex-packed
bootstrap with EvalMod, unpacking, the Walsh output pairs, the bit split)
recorded once as a HIP graph (ckks_capture_graph) and replayed: three
replays on three different inputs, with eager work and synchronisation in
between, each equal limb for limb to the eager S-box of the same input."""
import numpy as np
import pytest

from helpers import host

pytestmark = pytest.mark.gpu


def _inputs(H, ctx, bc, B, seed):
    P = ctx.getParams()
    slots, level = P.getSlots(), P.getBootstrapS2CDepth()
    rng = np.random.default_rng(seed)
    members = [[] for _ in range(8)]
    for _ in range(B):
        x0, x1 = rng.integers(0, 256, bc), rng.integers(0, 256, bc)
        for k in range(8):
            v = np.zeros(slots, np.complex128)
            v[:bc] = ((x0 >> k) & 1) + 1j * ((x1 >> k) & 1)
            members[k].append(H.ckks_encrypt(ctx, v, level))
    return [H.ckks_stack(ctx, m) for m in members]


def test_sbox_graph_replays_equal_eager_evaluations():
    H = host()
    from test_gpu_walsh import _sbox_context
    bc, B = 16, 3
    ctx = _sbox_context(H, bc)
    inputs = _inputs(H, ctx, bc, B, 41)
    # eager once: workspaces, pool blocks and cached constants exist
    H.ckks_walsh_sbox_complex_blocks_mask_packed(ctx, [c.clone() for c in inputs], bc)
    outs = []

    def body():
        outs.extend(H.ckks_walsh_sbox_complex_blocks_mask_packed(ctx, [c.clone() for c in inputs], bc))

    graph = H.ckks_capture_graph(ctx, body)
    assert len(outs) == 8
    for trial in range(3):
        fresh = _inputs(H, ctx, bc, B, 100 + trial)
        for dst, src in zip(inputs, fresh):
            H.ckks_copy_into(ctx, dst, src)
        print(f"replay {trial}", flush=True)
        graph.launch()
        eager = H.ckks_walsh_sbox_complex_blocks_mask_packed(ctx, [c.clone() for c in fresh], bc)
        for k in range(8):
            assert outs[k].getLevel() == eager[k].getLevel()
            for m in range(B):
                for j in range(2):
                    assert np.array_equal(outs[k].getMemberPoly(j, m), eager[k].getMemberPoly(j, m)), (trial, k, m, j)
    graph.reset()
