"""The keyswitch-level C++ API (ckks_keyswitch.hpp:59-149) over the device
pipeline: ckks_key_switch_inplace followed by the rotation's NTT-domain map
(build_ntt_automorphism_map, rotation_key_generation.cpp:80-90, applied by
ckks_apply_ntt_automorphism_to_flat_rns_poly_inplace) is ckks_rotate_inplace
(ckks_rotation.cpp:41-96); ckks_hybrid_key_switch_inplace on host vectors
equals ckks_key_switch_inplace; the hoisted switch with the map equals the
rotation; the maps are ring automorphisms and bad maps are refused."""
import numpy as np
import pytest

from helpers import host

pytestmark = pytest.mark.gpu


def _ctx(H):
    c = H.CKKSParamsConfig()
    c.N = 1 << 12
    c.mult_depth = 3
    c.q0_bit = 50
    c.q_bit = 40
    c.p_bits = [60, 60]
    c.dnum = 2
    c.secure = False
    c.rotate_values = [1, 3]
    c.seed = 77
    return H.CKKSContext(H.CKKSParams(c))


def test_keyswitch_api_matches_rotation_and_itself():
    H = host()
    ctx = _ctx(H)
    P = ctx.getParams()
    N, slots = P.getN(), P.getSlots()
    rng = np.random.default_rng(3)
    ct = H.ckks_encrypt(ctx, rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots), 2)
    L = ct.getLevel() + 1
    mont = True
    rot = H.ckks_rotate(ctx, ct, 3)
    nmap = ctx.getRotationNttMap(3)
    # key switch + the map = the rotation
    ks = ct.clone()
    H.ckks_key_switch_inplace(ctx, ks, 3, mont)
    for k in range(2):
        moved = H.ckks_apply_ntt_automorphism_to_flat_rns_poly_inplace(ks.getPoly(k), N, L, nmap)
        assert np.array_equal(moved, rot.getPoly(k)), k
    # the hybrid switch on host vectors = ckks_key_switch_inplace
    b, a = ct.getPoly(0), ct.getPoly(1)
    b2, a2 = H.ckks_hybrid_key_switch_inplace(ctx, ct.getLevel(), a, 3, b, np.zeros_like(a), 0, 1, mont)
    assert np.array_equal(b2, ks.getPoly(0)) and np.array_equal(a2, ks.getPoly(1))
    # hoisted: the switch of the hoisted a with the map = the rotation
    h = H.ckks_hoist_ciphertext_a_qp_eval(ctx, ct)
    assert h.level == ct.getLevel() and h.q_limb_count == L and h.p_limb_count == 2
    b3, a3 = H.ckks_hybrid_key_switch_hoisted_inplace(ctx, h, 3, b, np.zeros_like(a), 0, 1, mont, nmap)
    assert np.array_equal(b3, rot.getPoly(0)) and np.array_equal(a3, rot.getPoly(1))
    # QP from a key term: b and a over Q and P limbs
    bqp, aqp = H.ckks_hybrid_key_switch_qp_from_key_term(ctx, ct.getLevel(), a, 3, mont)
    assert len(bqp) == (L + 2) * N and len(aqp) == (L + 2) * N
    # a map that is not an automorphism is refused
    bad = list(nmap)
    bad[1], bad[2] = bad[2], bad[1]
    with pytest.raises(ValueError, match="automorphism"):
        H.ckks_hybrid_key_switch_hoisted_inplace(ctx, h, 3, b, np.zeros_like(a), 0, 1, mont, bad)


def test_ntt_maps_compose_like_rotations():
    """The map of rotation 1 applied three times is the map of rotation 3,
    and the conjugation map is an involution."""
    H = host()
    ctx = _ctx(H)
    N = ctx.getParams().getN()
    m1, m3 = np.array(ctx.getRotationNttMap(1)), np.array(ctx.getRotationNttMap(3))
    x = np.arange(N, dtype=np.uint64)
    y = x[m1][m1][m1]
    assert np.array_equal(y, x[m3])
    mc = np.array(ctx.getConjugationNttMap())
    assert np.array_equal(x[mc][mc], x)
