"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The HIP product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u64 = ctypes.c_uint64
i64 = ctypes.c_int64
cint = ctypes.c_int
vp = ctypes.c_void_p


def build():
    """Compile liboracle.so (gcc) in place."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "oc_powmod": (u64, [u64, u64, u64]),
            "oc_mod_inverse": (u64, [u64, u64]),
            "oc_compute_shoup": (u64, [u64, u64]),
            "oc_barrett_mul": (u64, [u64, u64, u64]),
            "oc_barrett_reduce_u64": (u64, [u64, u64]),
            "oc_shoup_mul": (u64, [u64, u64, u64]),
            "oc_montgomery_neg_inverse": (u64, [u64]),
            "oc_montgomery_radix_mod_q": (u64, [u64]),
            "oc_montgomery_mul": (u64, [u64, u64, u64]),
            "oc_is_prime": (cint, [u64]),
            "oc_find_p_and_q_primes": (cint, [vp, cint, vp, cint, u64, vp, vp]),
            "oc_find_root_of_unity": (u64, [u64, u64]),
            "oc_ntt_tables": (cint, [u64, u64, vp, vp, vp, vp, vp]),
            "oc_ntt_forward": (None, [vp, u64, u64, vp, vp, cint]),
            "oc_ntt_inverse": (None, [vp, u64, u64, vp, vp, u64, cint]),
            "oc_ctx_create": (vp, [u64, vp, cint, vp, cint, cint]),
            "oc_ctx_destroy": (None, [vp]),
            "oc_ctx_alpha": (cint, [vp]),
            "oc_omp_threads": (cint, []),
            "oc_ntt_fwd_q": (None, [vp, vp, cint, cint]),
            "oc_ntt_inv_q": (None, [vp, vp, cint, cint]),
            "oc_ntt_fwd_basis": (None, [vp, vp, vp, cint, cint]),
            "oc_ntt_inv_basis": (None, [vp, vp, vp, cint, cint]),
            "oc_modup": (cint, [vp, cint, cint, vp, vp]),
            "oc_keyswitch": (cint, [vp, cint, vp, vp, vp, vp, cint, cint, cint]),
            "oc_rotation_galois_element": (u64, [u64, i64]),
            "oc_ntt_automorphism_map": (None, [u64, u64, vp]),
            "oc_apply_automorphism": (None, [vp, u64, cint, vp]),
            "oc_rotate": (cint, [vp, cint, vp, vp, i64, vp, cint]),
            "oc_conjugate": (cint, [vp, cint, vp, vp, vp, cint]),
            "oc_relin": (cint, [vp, cint, vp, vp, vp, vp, cint]),
            "oc_rescale": (cint, [vp, cint, vp, cint]),
            "oc_tensor": (cint, [vp, cint, vp, vp, vp, vp, vp]),
            "oc_square": (cint, [vp, cint, vp, vp, vp, cint]),
            "oc_mul_plain": (cint, [vp, cint, vp, cint, vp, cint]),
            "oc_mul_int": (cint, [vp, cint, vp, cint, i64, cint]),
            "oc_add": (cint, [vp, cint, vp, vp, cint]),
            "oc_sub": (cint, [vp, cint, vp, vp, cint]),
            "oc_add_plain": (cint, [vp, cint, vp, vp, cint]),
            "oc_add_scalar_residues": (cint, [vp, cint, vp, vp]),
            "oc_hoist": (cint, [vp, cint, vp, vp]),
            "oc_keyswitch_hoisted": (cint, [vp, cint, vp, vp, vp, vp, cint, cint, u64, cint]),
            "oc_keyswitch_hoisted_qp_rotated": (cint, [vp, cint, vp, vp, vp, vp, vp, u64, cint]),
            "oc_keyswitch_qp_from_key_term": (cint, [vp, cint, vp, vp, vp, vp, cint]),
            "oc_moddown_qp_to_q": (cint, [vp, cint, vp]),
            "oc_lt_stage_qp": (cint, [vp, cint, vp, vp, cint, vp, vp, cint, cint, cint, vp, vp, cint,
                                      cint, vp, vp]),
            "oc_modraise": (cint, [vp, cint, vp, vp]),
            "oc_q0p0_keyswitch": (cint, [vp, vp, vp, vp, cint]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data


def u64a(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


def find_primes(p_bits, q_bits, N):
    pb, qb = u64a(p_bits), u64a(q_bits)
    po = np.zeros(max(len(pb), 1), np.uint64)
    qo = np.zeros(max(len(qb), 1), np.uint64)
    rc = lib().oc_find_p_and_q_primes(_p(pb) if len(pb) else None, len(pb),
                                      _p(qb) if len(qb) else None, len(qb), N, _p(po), _p(qo))
    if rc:
        raise RuntimeError("prime search failed (%d)" % rc)
    return po[: len(pb)], qo[: len(qb)]


def ntt_tables(N, q):
    fwd = np.zeros(N, np.uint64); fs = np.zeros(N, np.uint64)
    inv = np.zeros(N, np.uint64); is_ = np.zeros(N + 1, np.uint64)
    ni = np.zeros(1, np.uint64)
    if lib().oc_ntt_tables(N, int(q), _p(fwd), _p(fs), _p(inv), _p(is_), _p(ni)):
        raise RuntimeError("no 2N-th root of unity")
    return fwd, fs, inv, is_, int(ni[0])


def ntt_forward(x, q, tables, lazy=False):
    a = u64a(x).copy()
    fwd, fs = tables[0], tables[1]
    lib().oc_ntt_forward(_p(a), len(a), int(q), _p(fwd), _p(fs), int(lazy))
    return a


def ntt_inverse(x, q, tables, lazy=False):
    a = u64a(x).copy()
    inv, is_, ni = tables[2], tables[3], tables[4]
    lib().oc_ntt_inverse(_p(a), len(a), int(q), _p(inv), _p(is_), ni, int(lazy))
    return a


class Ctx:
    """Oracle context: chain moduli q (limb 0..Lmax), special moduli p, dnum."""

    def __init__(self, N, q, p, dnum):
        self.N, self.q, self.p, self.dnum = int(N), u64a(q), u64a(p), int(dnum)
        self.h = lib().oc_ctx_create(self.N, _p(self.q), len(self.q), _p(self.p), len(self.p), self.dnum)
        if not self.h:
            raise RuntimeError("oc_ctx_create failed")
        self.alpha = lib().oc_ctx_alpha(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_ctx_destroy(self.h)
            self.h = None

    def parts(self, level):
        L = level + 1
        return min(self.dnum, (L + self.alpha - 1) // self.alpha)

    def ntt_fwd(self, data, limbs, lazy=False):
        a = u64a(data).copy(); lib().oc_ntt_fwd_q(self.h, _p(a), limbs, int(lazy)); return a

    def ntt_inv(self, data, limbs, lazy=False):
        a = u64a(data).copy(); lib().oc_ntt_inv_q(self.h, _p(a), limbs, int(lazy)); return a

    def ntt_fwd_basis(self, data, basis, lazy=False):
        """NTT over basis indices (idx < nq: q[idx], else p[idx - nq])."""
        a = u64a(data).copy()
        b = np.ascontiguousarray(np.asarray(basis, np.int32))
        lib().oc_ntt_fwd_basis(self.h, _p(a), b.ctypes.data, len(b), int(lazy))
        return a

    def ntt_inv_basis(self, data, basis, lazy=False):
        a = u64a(data).copy()
        b = np.ascontiguousarray(np.asarray(basis, np.int32))
        lib().oc_ntt_inv_basis(self.h, _p(a), b.ctypes.data, len(b), int(lazy))
        return a

    def modup(self, level, part, src):
        L = level + 1
        s0 = part * self.alpha
        s1 = min(s0 + self.alpha, L)
        nd = L - (s1 - s0) + len(self.p)
        out = np.zeros(nd * self.N, np.uint64)
        src = u64a(src)
        if lib().oc_modup(self.h, level, part, _p(src), _p(out)):
            raise RuntimeError("oc_modup")
        return out

    def keyswitch(self, level, key_term, keys, b, a, b_mode, a_mode, montgomery=True):
        kt, b, a, keys = u64a(key_term).copy(), u64a(b).copy(), u64a(a).copy(), u64a(keys)
        if lib().oc_keyswitch(self.h, level, _p(kt), _p(keys), _p(b), _p(a), b_mode, a_mode, int(montgomery)):
            raise RuntimeError("oc_keyswitch")
        return b, a

    def rotate(self, level, b, a, r, keys, montgomery=True):
        b, a, keys = u64a(b).copy(), u64a(a).copy(), u64a(keys)
        if lib().oc_rotate(self.h, level, _p(b), _p(a), int(r), _p(keys), int(montgomery)):
            raise RuntimeError("oc_rotate")
        return b, a

    def conjugate(self, level, b, a, keys, montgomery=True):
        b, a, keys = u64a(b).copy(), u64a(a).copy(), u64a(keys)
        if lib().oc_conjugate(self.h, level, _p(b), _p(a), _p(keys), int(montgomery)):
            raise RuntimeError("oc_conjugate")
        return b, a

    def relin(self, level, b, a, c2, keys, montgomery=True):
        b, a, c2, keys = u64a(b).copy(), u64a(a).copy(), u64a(c2).copy(), u64a(keys)
        if lib().oc_relin(self.h, level, _p(b), _p(a), _p(c2), _p(keys), int(montgomery)):
            raise RuntimeError("oc_relin")
        return b, a

    def rescale(self, level, polys):
        """polys: (npolys, (level+1)*N) -> (npolys, level*N)."""
        x = u64a(polys).copy()
        n = x.shape[0]
        if lib().oc_rescale(self.h, level, _p(x), n):
            raise RuntimeError("oc_rescale")
        return x[:, : level * self.N].copy()

    def tensor(self, level, b1, a1, b2, a2):
        b1, a1, b2, a2 = u64a(b1).copy(), u64a(a1).copy(), u64a(b2), u64a(a2)
        c2 = np.zeros_like(b1)
        lib().oc_tensor(self.h, level, _p(b1), _p(a1), _p(b2), _p(a2), _p(c2))
        return b1, a1, c2

    def square(self, level, b, a, montgomery=True):
        b, a = u64a(b).copy(), u64a(a).copy()
        c2 = np.zeros_like(b)
        lib().oc_square(self.h, level, _p(b), _p(a), _p(c2), int(montgomery))
        return b, a, c2

    def mul_plain(self, level, polys, pt, pt_mont=False):
        x, pt = u64a(polys).copy(), u64a(pt)
        lib().oc_mul_plain(self.h, level, _p(x), x.shape[0], _p(pt), int(pt_mont))
        return x

    def mul_int(self, level, polys, s, montgomery=True):
        x = u64a(polys).copy()
        lib().oc_mul_int(self.h, level, _p(x), x.shape[0], int(s), int(montgomery))
        return x

    def add(self, level, x, y):
        x, y = u64a(x).copy(), u64a(y)
        lib().oc_add(self.h, level, _p(x), _p(y), x.shape[0])
        return x

    def sub(self, level, x, y):
        x, y = u64a(x).copy(), u64a(y)
        lib().oc_sub(self.h, level, _p(x), _p(y), x.shape[0])
        return x

    def add_scalar_residues(self, level, b, res):
        b, res = u64a(b).copy(), u64a(res)
        lib().oc_add_scalar_residues(self.h, level, _p(b), _p(res))
        return b

    def walsh_product_sum(self, level, low, high, cr, ci, consts, pt, pt_mont=True, montgomery=True):
        """walsh_sum_for_output_pair (ckks_walsh_sbox.cpp:984-1043) up to its
        relinearisation, composed from this oracle's primitives in the
        reference's order: linear_combination_with_constant (:862-905; the
        constant by ckks_add_integer_constant_inplace's residues, each sign
        term by a scaled add, :827-860), combine_real_imag (:931-958: the
        imaginary sum times the encoded i, then added), mul_raw (:297-321: the
        tensor) and the raw-product accumulation (add_term_inplace, :288-295).
        low, high: 15 arrays (2, L*N) [b, a] for masks 1..15; cr, ci [16][16];
        consts: 16 pairs (V_real, V_imag) of integers.  Returns (acc (2, L*N),
        raw (3, L*N))."""
        L = level + 1
        W = L * self.N

        def lc(coeffs, v):
            out = np.zeros((2, W), dtype=np.uint64)
            if v:
                out[0] = self.add_scalar_residues(level, out[0], [v % int(self.q[l]) for l in range(L)])
            for m in range(1, 16):
                c = int(coeffs[m])
                if c:
                    out = self.add(level, out, self.mul_int(level, low[m - 1], c, montgomery))
            return out

        def z(h):
            real = lc(cr[h], consts[h][0])
            imag = self.mul_plain(level, lc(ci[h], consts[h][1]), pt, pt_mont)
            return self.add(level, real, imag)

        acc = z(0)
        raw = None
        for h in range(1, 16):
            zb, za = z(h)
            prod = np.stack(self.tensor(level, zb, za, high[h - 1][0], high[h - 1][1]))
            raw = prod if raw is None else self.add(level, raw, prod)
        return acc, raw


def _ctx_extra():
    """Methods for hoisting, QP key switching, BSGS, ModRaise and q0p0."""

    def qp_limbs(self, level):
        return level + 1 + len(self.p)

    def hoist(self, level, a_eval):
        a = u64a(a_eval)
        out = np.zeros(self.parts(level) * self.qp_limbs(level) * self.N, np.uint64)
        if lib().oc_hoist(self.h, level, _p(a), _p(out)):
            raise RuntimeError("oc_hoist")
        return out

    def keyswitch_hoisted(self, level, digits, keys, b, a, b_mode, a_mode, galois=0, montgomery=True):
        digits, keys, b, a = u64a(digits), u64a(keys), u64a(b).copy(), u64a(a).copy()
        if lib().oc_keyswitch_hoisted(self.h, level, _p(digits), _p(keys), _p(b), _p(a), b_mode, a_mode,
                                      int(galois), int(montgomery)):
            raise RuntimeError("oc_keyswitch_hoisted")
        return b, a

    def keyswitch_hoisted_qp_rotated(self, level, digits, keys, b_q, galois, montgomery=True):
        digits, keys, b_q = u64a(digits), u64a(keys), u64a(b_q)
        W = self.qp_limbs(level) * self.N
        bqp, aqp = np.zeros(W, np.uint64), np.zeros(W, np.uint64)
        if lib().oc_keyswitch_hoisted_qp_rotated(self.h, level, _p(digits), _p(keys), _p(b_q), _p(bqp),
                                                 _p(aqp), int(galois), int(montgomery)):
            raise RuntimeError("oc_keyswitch_hoisted_qp_rotated")
        return bqp, aqp

    def keyswitch_qp_from_key_term(self, level, key_term, keys, montgomery=True):
        kt, keys = u64a(key_term), u64a(keys)
        W = self.qp_limbs(level) * self.N
        bqp, aqp = np.zeros(W, np.uint64), np.zeros(W, np.uint64)
        if lib().oc_keyswitch_qp_from_key_term(self.h, level, _p(kt), _p(keys), _p(bqp), _p(aqp),
                                               int(montgomery)):
            raise RuntimeError("oc_keyswitch_qp_from_key_term")
        return bqp, aqp

    def moddown_qp_to_q(self, level, qp):
        x = u64a(qp).copy()
        if lib().oc_moddown_qp_to_q(self.h, level, _p(x)):
            raise RuntimeError("oc_moddown_qp_to_q")
        return x[: (level + 1) * self.N].copy()

    def lt_stage_qp(self, level, a, b, diag_idx, diags, baby_steps, rot_keys, pt_mont=True,
                    slots=None, montgomery=True):
        """rot_keys: {rotation amount: key array}.  Returns (a_out, b_out)."""
        a, b, diags = u64a(a), u64a(b), u64a(diags)
        idx = np.ascontiguousarray(np.asarray(diag_idx, np.int64))
        amounts = np.ascontiguousarray(np.asarray(sorted(rot_keys), np.int64))
        karr = [u64a(rot_keys[int(r)]) for r in amounts]
        kptr = (ctypes.c_void_p * max(len(karr), 1))(*[k.ctypes.data for k in karr])
        L = level + 1
        ao, bo = np.zeros(L * self.N, np.uint64), np.zeros(L * self.N, np.uint64)
        slots = self.N // 2 if slots is None else slots
        rc = lib().oc_lt_stage_qp(self.h, level, _p(a), _p(b), len(idx), idx.ctypes.data, _p(diags),
                                  int(pt_mont), int(baby_steps), int(slots), amounts.ctypes.data,
                                  ctypes.addressof(kptr), len(karr), int(montgomery), _p(ao), _p(bo))
        if rc:
            raise RuntimeError("oc_lt_stage_qp (%d)" % rc)
        return ao, bo

    def modraise(self, target_level, poly_q0):
        x = u64a(poly_q0)
        out = np.zeros((target_level + 1) * self.N, np.uint64)
        if lib().oc_modraise(self.h, target_level, _p(x), _p(out)):
            raise RuntimeError("oc_modraise")
        return out

    def q0p0_keyswitch(self, b, a, key, montgomery=True):
        b, a, key = u64a(b).copy(), u64a(a).copy(), u64a(key)
        if lib().oc_q0p0_keyswitch(self.h, _p(b), _p(a), _p(key), int(montgomery)):
            raise RuntimeError("oc_q0p0_keyswitch")
        return b, a

    for f in (qp_limbs, hoist, keyswitch_hoisted, keyswitch_hoisted_qp_rotated, keyswitch_qp_from_key_term,
              moddown_qp_to_q, lt_stage_qp, modraise, q0p0_keyswitch):
        setattr(Ctx, f.__name__, f)


_ctx_extra()


def galois_rotation(N, r):
    return lib().oc_rotation_galois_element(N, int(r))


def automorphism_map(N, g):
    m = np.zeros(N, np.uint64)
    lib().oc_ntt_automorphism_map(N, int(g), _p(m))
    return m


def apply_automorphism(data, N, limbs, galois):
    """NTT-domain automorphism of `limbs` limbs (returns a new array)."""
    out = u64a(data).copy()
    m = automorphism_map(N, galois)
    lib().oc_apply_automorphism(_p(out), N, limbs, _p(m))
    return out
