"""Python binding of libmickks.so (the HIP C-ABI in include/mickks.h).

Device memory and streams come from PyTorch (ROCm): buffers are torch int64
tensors on the GPU holding uint64 residues bit for bit.  There is no CPU
fallback: if the library is missing or no GPU is present, calls raise.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmickks.so")

_lib = None

vp = ctypes.c_void_p
cint = ctypes.c_int
u64 = ctypes.c_uint64
i64 = ctypes.c_int64

_SIGS = {
    "mi_last_error": (ctypes.c_char_p, []),
    "mi_context_create": (cint, [ctypes.POINTER(vp), u64, vp, cint, vp, cint, cint]),
    "mi_context_destroy": (cint, [vp]),
    "mi_context_info": (cint, [vp, vp]),
    "mi_context_device_bytes": (i64, [vp]),
    "mi_alloc": (cint, [vp, ctypes.c_size_t, vp]),
    "mi_free": (cint, [vp, vp]),
    "mi_memory_trim": (cint, [vp]),
    "mi_memory_stats": (cint, [vp, vp]),
    "mi_set_memory_limit": (cint, [vp, i64]),
    "mi_device_memory_info": (cint, [vp, vp, vp]),
    "mi_tune_report": (cint, [vp, vp, ctypes.c_size_t]),
    "mi_set_decoding_tables": (cint, [vp, vp, vp]),
    "mi_decode_batch": (cint, [vp, cint, vp, cint, vp, vp, cint, vp, vp]),
    "mi_tune_set": (cint, [vp, vp, cint]),
    "mi_probe_begin": (cint, [vp]),
    "mi_probe_end": (cint, [vp, vp]),
    "mi_memcpy": (cint, [vp, vp, vp, ctypes.c_size_t, cint, vp]),
    "mi_stream_synchronize": (cint, [vp, vp]),
    "mi_stream_create": (cint, [vp, ctypes.POINTER(vp)]),
    "mi_poly_list_op": (cint, [vp, cint, cint, cint, vp, vp, vp, vp]),
    "mi_linear_combine": (cint, [vp, cint, cint, vp, cint, vp, vp, vp]),
    "mi_scaled_sum": (cint, [vp, cint, cint, vp, cint, vp, vp, cint, vp, vp]),
    "mi_hoist_batch": (cint, [vp, cint, cint, vp, vp, ctypes.c_int64, vp]),
    "mi_modraise_batch": (cint, [vp, cint, cint, vp, vp, vp]),
    "mi_q0p0_keyswitch_batch": (cint, [vp, cint, vp, vp, vp, cint, vp]),
    "mi_rotate_hoisted_items": (cint, [vp, cint, cint, vp, ctypes.c_int64, cint, vp, vp, vp, vp, vp, cint, vp]),
    "mi_walsh_product_sum_batch": (cint, [vp, cint, cint] + [vp] * 9 + [cint] + [vp] * 5 + [vp]),
    "mi_conjugate_batch": (cint, [vp, cint, cint, vp, vp, vp, cint, vp]),
    "mi_stream_destroy": (cint, [vp, vp]),
    "mi_poly_op": (cint, [vp, cint, vp, vp, cint, cint, cint, i64, vp]),
    "mi_mul_scalar_add": (cint, [vp, cint, vp, vp, cint, i64, vp, vp]),
    "mi_memset_zero": (cint, [vp, vp, ctypes.c_size_t, vp]),
    "mi_poly_scalar": (cint, [vp, cint, vp, vp, cint, cint, cint, i64, vp]),
    "mi_ntt_forward": (cint, [vp, vp, cint, cint, i64, cint, vp]),
    "mi_ntt_inverse": (cint, [vp, vp, cint, cint, i64, cint, vp]),
    "mi_tensor": (cint, [vp, cint, vp, vp, vp, vp, vp, vp]),
    "mi_square": (cint, [vp, cint, vp, vp, vp, vp]),
    "mi_tensor_list": (cint, [vp, cint, cint, vp, vp, vp, vp, vp, vp]),
    "mi_tensor_list_into": (cint, [vp, cint, cint, vp, vp, vp, vp, vp, vp, vp, vp]),
    "mi_mul_plain": (cint, [vp, cint, vp, cint, i64, vp, cint, vp]),
    "mi_mul_scalar": (cint, [vp, cint, vp, cint, i64, vp, vp]),
    "mi_add": (cint, [vp, cint, vp, vp, cint, i64, vp]),
    "mi_sub": (cint, [vp, cint, vp, vp, cint, i64, vp]),
    "mi_add_plain": (cint, [vp, cint, vp, vp, cint, vp]),
    "mi_add_scalar": (cint, [vp, cint, vp, vp, vp]),
    "mi_automorphism": (cint, [vp, cint, u64, vp, vp, cint, i64, vp]),
    "mi_rescale": (cint, [vp, cint, vp, cint, i64, vp]),
    "mi_keyswitch": (cint, [vp, cint, vp, vp, vp, vp, cint, cint, cint, vp]),
    "mi_keyswitch_noalloc": (cint, [vp, cint, vp, vp, vp, vp, cint, cint, cint, vp]),
    "mi_keyswitch_batch": (cint, [vp, vp, cint, cint, vp]),
    "mi_relin": (cint, [vp, cint, vp, vp, vp, vp, cint, vp]),
    "mi_rotate": (cint, [vp, cint, vp, vp, u64, vp, cint, vp]),
    "mi_conjugate": (cint, [vp, cint, vp, vp, vp, cint, vp]),
    "mi_hoist": (cint, [vp, cint, vp, vp, vp]),
    "mi_keyswitch_hoisted": (cint, [vp, cint, vp, vp, vp, vp, cint, cint, u64, cint, vp]),
    "mi_keyswitch_hoisted_qp": (cint, [vp, cint, vp, vp, vp, vp, vp, u64, cint, vp]),
    "mi_rotate_hoisted_batch": (cint, [vp, cint, vp, vp, cint, vp, vp, vp, vp, cint, vp]),
    "mi_rotate_batch": (cint, [vp, cint, cint, vp, vp, vp, vp, cint, vp]),
    "mi_rescale_list": (cint, [vp, cint, cint, vp, vp]),
    "mi_mul_plain_rescale_list": (cint, [vp, cint, cint, vp, vp, cint, cint, vp]),
    "mi_keyswitch_qp": (cint, [vp, cint, vp, vp, vp, vp, cint, vp]),
    "mi_moddown_qp": (cint, [vp, cint, vp, vp]),
    "mi_linear_transform_qp": (cint, [vp, cint, vp, vp, cint, vp, vp, cint, cint, cint, cint, vp, vp,
                                      cint, vp, vp, vp]),
    "mi_linear_transform_qp_batch": (cint, [vp, cint, cint, vp, vp, cint, vp, vp, cint, cint, cint, cint,
                                            vp, vp, cint, vp, vp, vp]),
    "mi_linear_transform_q": (cint, [vp, cint, vp, vp, cint, vp, vp, cint, cint, cint, cint, vp, vp,
                                     cint, vp, vp, vp]),
    "mi_linear_transform_q_batch": (cint, [vp, cint, cint, vp, vp, cint, vp, vp, cint, cint, cint, cint,
                                           vp, vp, cint, vp, vp, vp]),
    "mi_set_encoding_tables": (cint, [vp, vp, vp]),
    "mi_encode_slots": (cint, [vp, vp, cint, ctypes.c_double, cint, cint, vp, i64, vp]),
    "mi_sample_uniform": (cint, [vp, vp, cint, cint, u64, u64, vp]),
    "mi_sample_gaussian": (cint, [vp, vp, cint, ctypes.c_double, u64, u64, vp]),
    "mi_modraise": (cint, [vp, cint, vp, vp, vp]),
    "mi_q0p0_keyswitch": (cint, [vp, vp, vp, vp, cint, vp]),
    "mi_walsh_product_sum": (cint, [vp, cint, vp, vp, vp, vp, vp, vp, vp, vp, vp, cint, vp, vp, vp, vp, vp, vp]),
    "mi_graph_begin": (cint, [vp, vp]),
    "mi_graph_end": (cint, [vp, vp, vp]),
    "mi_graph_launch": (cint, [vp, vp, vp]),
    "mi_graph_destroy": (cint, [vp, vp]),
}

EXPORTED = sorted(_SIGS)


class KeySwitchItem(ctypes.Structure):
    _fields_ = [("key_term", vp), ("key", vp), ("b_out", vp), ("a_out", vp),
                ("level", ctypes.c_int32), ("b_mode", ctypes.c_int32),
                ("a_mode", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "-j8"])


def lib():
    """Load libmickks.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        # Load PyTorch's HIP runtime first so that libmickks.so binds to the
        # same libamdhip64.so.7 (one runtime per process; a second copy would
        # hide the GPU from the first).
        import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmickks.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class MiError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise MiError(lib().mi_last_error().decode())


def _stream(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return vp(stream.cuda_stream)


def ptr(t):
    return vp(t.data_ptr())


class Context:
    """Device-resident CKKS tables for chain moduli q, special moduli p, dnum digits."""

    def __init__(self, N, q, p, dnum):
        import numpy as np
        self.N, self.dnum = int(N), int(dnum)
        self.q = np.ascontiguousarray(np.asarray(q, dtype=np.uint64))
        self.p = np.ascontiguousarray(np.asarray(p, dtype=np.uint64))
        h = vp()
        check(lib().mi_context_create(ctypes.byref(h), self.N, self.q.ctypes.data, len(self.q),
                                      self.p.ctypes.data, len(self.p), self.dnum))
        self.h = h
        info = (ctypes.c_int64 * 6)()
        check(lib().mi_context_info(self.h, ctypes.addressof(info)))
        self.alpha = int(info[4])
        self.nq, self.np = len(self.q), len(self.p)

    def close(self):
        if getattr(self, "h", None):
            lib().mi_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def parts(self, level):
        L = level + 1
        return min(self.dnum, (L + self.alpha - 1) // self.alpha)

    def device_bytes(self):
        return int(lib().mi_context_device_bytes(self.h))

    def memory_stats(self):
        """(live, peak, cached, limit) bytes of the context's device pool."""
        st = (ctypes.c_int64 * 4)()
        check(lib().mi_memory_stats(self.h, ctypes.addressof(st)))
        return tuple(int(v) for v in st)

    def poly_op(self, op, x, y, limbs, basis0=0, npolys=1, stride=None, stream=None):
        """op: 0 add, 1 sub, 2 mul, 3 Montgomery mul; limb l uses basis basis0 + l."""
        stride = limbs * self.N if stride is None else stride
        check(lib().mi_poly_op(self.h, op, ptr(x), ptr(y), limbs, basis0, npolys, stride, _stream(stream)))

    def poly_scalar(self, op, x, residues, limbs, basis0=0, npolys=1, stride=None, stream=None):
        import numpy as np
        r = np.ascontiguousarray(np.asarray(residues, dtype=np.uint64))
        stride = limbs * self.N if stride is None else stride
        check(lib().mi_poly_scalar(self.h, op, ptr(x), r.ctypes.data, limbs, basis0, npolys, stride,
                                   _stream(stream)))

    # ---- NTT
    def ntt_forward(self, data, npolys, limbs, stride=None, basis0=0, stream=None):
        stride = limbs * self.N if stride is None else stride
        check(lib().mi_ntt_forward(self.h, ptr(data), npolys, limbs, stride, basis0, _stream(stream)))

    def ntt_inverse(self, data, npolys, limbs, stride=None, basis0=0, stream=None):
        stride = limbs * self.N if stride is None else stride
        check(lib().mi_ntt_inverse(self.h, ptr(data), npolys, limbs, stride, basis0, _stream(stream)))

    # ---- element-wise
    def tensor(self, level, b1, a1, b2, a2, c2, stream=None):
        check(lib().mi_tensor(self.h, level, ptr(b1), ptr(a1), ptr(b2), ptr(a2), ptr(c2), _stream(stream)))

    def square(self, level, b, a, c2, stream=None):
        check(lib().mi_square(self.h, level, ptr(b), ptr(a), ptr(c2), _stream(stream)))

    def mul_plain(self, level, polys, npolys, pt, pt_montgomery=False, stride=None, stream=None):
        stride = (level + 1) * self.N if stride is None else stride
        check(lib().mi_mul_plain(self.h, level, ptr(polys), npolys, stride, ptr(pt), int(pt_montgomery),
                                 _stream(stream)))

    def mul_scalar(self, level, polys, npolys, residues, stride=None, stream=None):
        import numpy as np
        r = np.ascontiguousarray(np.asarray(residues, dtype=np.uint64))
        stride = (level + 1) * self.N if stride is None else stride
        check(lib().mi_mul_scalar(self.h, level, ptr(polys), npolys, stride, r.ctypes.data, _stream(stream)))

    def add(self, level, x, y, npolys, stride=None, stream=None):
        stride = (level + 1) * self.N if stride is None else stride
        check(lib().mi_add(self.h, level, ptr(x), ptr(y), npolys, stride, _stream(stream)))

    def sub(self, level, x, y, npolys, stride=None, stream=None):
        stride = (level + 1) * self.N if stride is None else stride
        check(lib().mi_sub(self.h, level, ptr(x), ptr(y), npolys, stride, _stream(stream)))

    def add_plain(self, level, b, pt, pt_montgomery=False, stream=None):
        check(lib().mi_add_plain(self.h, level, ptr(b), ptr(pt), int(pt_montgomery), _stream(stream)))

    def add_scalar(self, level, b, residues, stream=None):
        import numpy as np
        r = np.ascontiguousarray(np.asarray(residues, dtype=np.uint64))
        check(lib().mi_add_scalar(self.h, level, ptr(b), r.ctypes.data, _stream(stream)))

    def automorphism(self, limbs, galois, src, dst, npolys=1, stride=None, stream=None):
        stride = limbs * self.N if stride is None else stride
        check(lib().mi_automorphism(self.h, limbs, int(galois), ptr(src), ptr(dst), npolys, stride,
                                    _stream(stream)))

    def rescale(self, level, polys, npolys, stride=None, stream=None):
        stride = (level + 1) * self.N if stride is None else stride
        check(lib().mi_rescale(self.h, level, ptr(polys), npolys, stride, _stream(stream)))

    # ---- key switching
    def tune_report(self):
        import json
        buf = ctypes.create_string_buffer(4096)
        check(lib().mi_tune_report(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def tune_set(self, ntt_variant, keyswitch_group):
        arr = (ctypes.c_int * 6)(*ntt_variant)
        check(lib().mi_tune_set(self.h, arr, int(keyswitch_group)))

    def keyswitch(self, level, key_term, key, b_out, a_out, b_mode, a_mode, montgomery=True, stream=None):
        check(lib().mi_keyswitch(self.h, level, ptr(key_term), ptr(key), ptr(b_out), ptr(a_out), b_mode,
                                 a_mode, int(montgomery), _stream(stream)))

    def keyswitch_batch(self, items, montgomery=True, stream=None):
        arr = (KeySwitchItem * len(items))()
        for i, (lv, kt, key, b, a, bm, am) in enumerate(items):
            arr[i] = KeySwitchItem(kt.data_ptr(), key.data_ptr(), b.data_ptr(), a.data_ptr(), lv, bm, am, 0)
        check(lib().mi_keyswitch_batch(self.h, ctypes.addressof(arr), len(items), int(montgomery),
                                       _stream(stream)))

    def relin(self, level, b, a, c2, key, montgomery=True, stream=None):
        check(lib().mi_relin(self.h, level, ptr(b), ptr(a), ptr(c2), ptr(key), int(montgomery), _stream(stream)))

    def rotate(self, level, b, a, galois, key, montgomery=True, stream=None):
        check(lib().mi_rotate(self.h, level, ptr(b), ptr(a), int(galois), ptr(key), int(montgomery),
                              _stream(stream)))

    def conjugate(self, level, b, a, key, montgomery=True, stream=None):
        check(lib().mi_conjugate(self.h, level, ptr(b), ptr(a), ptr(key), int(montgomery), _stream(stream)))

    # ---- hoisting and QP-basis switching
    def qp_limbs(self, level):
        return level + 1 + self.np

    def hoist(self, level, a, digits, stream=None):
        """digits: parts(level) * (level+1+np) * N words."""
        check(lib().mi_hoist(self.h, level, ptr(a), ptr(digits), _stream(stream)))

    def keyswitch_hoisted(self, level, digits, key, b_out, a_out, b_mode, a_mode, galois=0, montgomery=True,
                          stream=None):
        check(lib().mi_keyswitch_hoisted(self.h, level, ptr(digits), ptr(key), ptr(b_out), ptr(a_out), b_mode,
                                         a_mode, int(galois), int(montgomery), _stream(stream)))

    def mul_plain_rescale_list(self, level, polys, pt, pt_montgomery=False, rescale=True, stream=None):
        """polys[k] *= pt, then (rescale) ckks_rescale of every listed poly."""
        P = vp * len(polys)
        arr = P(*[t.data_ptr() for t in polys])
        check(lib().mi_mul_plain_rescale_list(self.h, level, len(polys), ctypes.addressof(arr), ptr(pt),
                                              int(pt_montgomery), int(rescale), _stream(stream)))

    def rotate_batch(self, level, cts, galois, keys, montgomery=True, stream=None):
        """In-place rotations of independent ciphertexts: cts = [(b, a), ...]."""
        n = len(cts)
        P = vp * n
        b, a = P(*[c[0].data_ptr() for c in cts]), P(*[c[1].data_ptr() for c in cts])
        g = (u64 * n)(*[int(x) for x in galois])
        k = P(*[t.data_ptr() for t in keys])
        check(lib().mi_rotate_batch(self.h, level, n, ctypes.addressof(b), ctypes.addressof(a), ctypes.addressof(g),
                                    ctypes.addressof(k), int(montgomery), _stream(stream)))

    def rotate_hoisted_batch(self, level, digits, b, galois, keys, outs, montgomery=True, stream=None):
        """outs: list of (b_out, a_out) device tensors, one per Galois element."""
        n = len(galois)
        g = (u64 * max(n, 1))(*[int(x) for x in galois])
        kp = (vp * max(n, 1))(*[k.data_ptr() for k in keys])
        bo = (vp * max(n, 1))(*[o[0].data_ptr() for o in outs])
        ao = (vp * max(n, 1))(*[o[1].data_ptr() for o in outs])
        check(lib().mi_rotate_hoisted_batch(self.h, level, ptr(digits), ptr(b), n, ctypes.addressof(g),
                                            ctypes.addressof(kp), ctypes.addressof(bo), ctypes.addressof(ao),
                                            int(montgomery), _stream(stream)))

    def keyswitch_hoisted_qp(self, level, digits, key, b, b_qp, a_qp, galois, montgomery=True, stream=None):
        check(lib().mi_keyswitch_hoisted_qp(self.h, level, ptr(digits), ptr(key), ptr(b), ptr(b_qp), ptr(a_qp),
                                            int(galois), int(montgomery), _stream(stream)))

    def keyswitch_qp(self, level, key_term, key, b_qp, a_qp, montgomery=True, stream=None):
        check(lib().mi_keyswitch_qp(self.h, level, ptr(key_term), ptr(key), ptr(b_qp), ptr(a_qp),
                                    int(montgomery), _stream(stream)))

    def moddown_qp(self, level, qp, stream=None):
        check(lib().mi_moddown_qp(self.h, level, ptr(qp), _stream(stream)))

    def linear_transform_qp(self, level, a, b, diag_indices, diags, baby_steps, rot_keys, a_out, b_out,
                            pt_montgomery=True, slots=None, montgomery=True, stream=None):
        """diags: list of device tensors (one QP plaintext per term);
        rot_keys: {rotation amount: device key tensor}."""
        n = len(diag_indices)
        idx = (ctypes.c_int64 * max(n, 1))(*[int(d) for d in diag_indices])
        dp = (vp * max(n, 1))(*[d.data_ptr() for d in diags])
        rots = sorted(rot_keys)
        kr = (ctypes.c_int64 * max(len(rots), 1))(*[int(r) for r in rots])
        kp = (vp * max(len(rots), 1))(*[rot_keys[r].data_ptr() for r in rots])
        slots = self.N // 2 if slots is None else slots
        check(lib().mi_linear_transform_qp(self.h, level, ptr(a), ptr(b), n, ctypes.addressof(idx),
                                           ctypes.addressof(dp), int(pt_montgomery), int(baby_steps),
                                           int(slots), len(rots), ctypes.addressof(kr), ctypes.addressof(kp),
                                           int(montgomery), ptr(a_out), ptr(b_out), _stream(stream)))

    def linear_transform_q(self, level, a, b, diag_indices, diags, baby_steps, rot_keys, a_out, b_out,
                           pt_montgomery=True, slots=None, montgomery=True, stream=None):
        """Q-evaluator stage; diags: level+1 limb plaintexts."""
        self.linear_transform_qp_batch(level, [(a, b)], diag_indices, diags, baby_steps, rot_keys,
                                       [(a_out, b_out)], pt_montgomery, slots, montgomery, stream,
                                       q_evaluator=True)

    def linear_transform_qp_batch(self, level, cts, diag_indices, diags, baby_steps, rot_keys, outs,
                                  pt_montgomery=True, slots=None, montgomery=True, stream=None,
                                  q_evaluator=False):
        """cts / outs: lists of (a, b) device tensor pairs (level+1 limbs each);
        the stage is applied to every ciphertext of the batch (QP evaluator, or
        the Q evaluator with level+1 limb diagonals when q_evaluator)."""
        n = len(diag_indices)
        m = len(cts)
        if len(outs) != m:
            raise ValueError("one output pair per input ciphertext")
        idx = (ctypes.c_int64 * max(n, 1))(*[int(d) for d in diag_indices])
        dp = (vp * max(n, 1))(*[d.data_ptr() for d in diags])
        rots = sorted(rot_keys)
        kr = (ctypes.c_int64 * max(len(rots), 1))(*[int(r) for r in rots])
        kp = (vp * max(len(rots), 1))(*[rot_keys[r].data_ptr() for r in rots])
        pa = (vp * max(m, 1))(*[c[0].data_ptr() for c in cts])
        pb = (vp * max(m, 1))(*[c[1].data_ptr() for c in cts])
        oa = (vp * max(m, 1))(*[o[0].data_ptr() for o in outs])
        ob = (vp * max(m, 1))(*[o[1].data_ptr() for o in outs])
        slots = self.N // 2 if slots is None else slots
        fn = lib().mi_linear_transform_q_batch if q_evaluator else lib().mi_linear_transform_qp_batch
        check(fn(
            self.h, level, m, ctypes.addressof(pa), ctypes.addressof(pb), n, ctypes.addressof(idx),
            ctypes.addressof(dp), int(pt_montgomery), int(baby_steps), int(slots), len(rots),
            ctypes.addressof(kr), ctypes.addressof(kp), int(montgomery), ctypes.addressof(oa),
            ctypes.addressof(ob), _stream(stream)))

    # ---- sampling and encoding
    def sample_uniform(self, out, limbs, seed, counter=0, basis0=0, stream=None):
        check(lib().mi_sample_uniform(self.h, ptr(out), limbs, basis0, seed, counter, _stream(stream)))

    def sample_gaussian(self, out, limbs, sigma, seed, counter=0, stream=None):
        check(lib().mi_sample_gaussian(self.h, ptr(out), limbs, float(sigma), seed, counter, _stream(stream)))

    # ---- bootstrapping pieces
    def modraise(self, target_level, poly_q0, out, stream=None):
        check(lib().mi_modraise(self.h, target_level, ptr(poly_q0), ptr(out), _stream(stream)))

    def q0p0_keyswitch(self, b, a, key, montgomery=True, stream=None):
        check(lib().mi_q0p0_keyswitch(self.h, ptr(b), ptr(a), ptr(key), int(montgomery), _stream(stream)))

    # ---- Walsh S-box
    def walsh_product_sum(self, level, low, high, cr, ci, consts, pt, acc, raw, pt_montgomery=True, stream=None):
        """low, high: 15 (b, a) device tensors each (sign masks 1..15); cr, ci:
        int16 [16][16] ([high mask][low mask]); consts: 16 pairs (V_real, V_imag)
        of Python ints, the rounded scaled constants; pt: encoded i; acc: (b, a)
        outputs; raw: (b, a, c2) outputs."""
        import numpy as np
        if len(low) != 15 or len(high) != 15:
            raise ValueError("15 low and 15 high sign ciphertexts")
        P15 = vp * 15
        lb, la = P15(*[t[0].data_ptr() for t in low]), P15(*[t[1].data_ptr() for t in low])
        hb, ha = P15(*[t[0].data_ptr() for t in high]), P15(*[t[1].data_ptr() for t in high])
        cr = np.ascontiguousarray(np.asarray(cr, dtype=np.int16).reshape(16, 16))
        ci = np.ascontiguousarray(np.asarray(ci, dtype=np.int16).reshape(16, 16))
        absv = np.zeros(64, dtype=np.uint64)
        neg = np.zeros(32, dtype=np.int32)
        for h in range(16):
            for r in range(2):
                v = int(consts[h][r])
                t = 2 * h + r
                absv[2 * t] = abs(v) & (2**64 - 1)
                absv[2 * t + 1] = abs(v) >> 64
                neg[t] = 1 if v < 0 else 0
        check(lib().mi_walsh_product_sum(self.h, level, ctypes.addressof(lb), ctypes.addressof(la),
                                         ctypes.addressof(hb), ctypes.addressof(ha), cr.ctypes.data, ci.ctypes.data,
                                         absv.ctypes.data, neg.ctypes.data, ptr(pt), int(pt_montgomery),
                                         ptr(acc[0]), ptr(acc[1]), ptr(raw[0]), ptr(raw[1]), ptr(raw[2]),
                                         _stream(stream)))

    # ---- HIP graphs
    def graph_begin(self, stream):
        """Start recording the calls made on `stream` (a created torch stream)."""
        check(lib().mi_graph_begin(self.h, _stream(stream)))

    def graph_end(self, stream):
        """Stop recording on `stream`; returns a Graph to replay."""
        g = vp()
        check(lib().mi_graph_end(self.h, _stream(stream), ctypes.byref(g)))
        return Graph(self, g)


class Graph:
    """An instantiated capture of context calls (mi_graph_*)."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle

    def launch(self, stream):
        check(lib().mi_graph_launch(self.ctx.h, self.h, _stream(stream)))

    def destroy(self):
        if self.h:
            check(lib().mi_graph_destroy(self.ctx.h, self.h))
            self.h = None


def to_dev(arr):
    """numpy uint64 array -> torch int64 tensor on the GPU (same bits)."""
    import numpy as np
    import torch
    a = np.ascontiguousarray(np.asarray(arr, dtype=np.uint64))
    return torch.from_numpy(a.view(np.int64).copy()).cuda()


def from_dev(t):
    """torch int64 GPU tensor -> numpy uint64 array."""
    import numpy as np
    return t.detach().cpu().numpy().view(np.uint64).copy()
