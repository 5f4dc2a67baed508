"""Shared test helpers: deterministic inputs, digests, fixtures, big-int forms."""
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

M64 = (1 << 64) - 1


def splitmix64_stream(seed, n):
    """splitmix64 outputs for state `seed` (matches oracle/ref_harness.cpp)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def lt_diags(N, qp, nterms, seed_base=5000):
    """Diagonal QP plaintexts of the reference harness's BSGS stage
    (oracle/ref_harness.cpp): term t, limb l = splitmix64(seed_base+64t+l) mod m_l."""
    out = []
    for t in range(nterms):
        for l, m in enumerate(qp):
            out.append(splitmix64_stream(seed_base + 64 * t + l, N) % np.uint64(int(m)))
    return np.concatenate(out)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def load_npz(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def rand_residues(rng, moduli, N):
    """Uniform residues, limb-major, one limb per modulus."""
    return np.concatenate([rng.integers(0, int(q), N, dtype=np.uint64) for q in moduli])


def brev(x, bits):
    r = 0
    for i in range(bits):
        if x >> i & 1:
            r |= 1 << (bits - 1 - i)
    return r


def crt(residues, moduli):
    """Big-int CRT of one coefficient: value in [0, prod)."""
    Q = 1
    for m in moduli:
        Q *= int(m)
    x = 0
    for r, m in zip(residues, moduli):
        m = int(m)
        qh = Q // m
        x += int(r) * qh * pow(qh, -1, m)
    return x % Q, Q


ROOT = os.path.dirname(HERE)
PKG_DIR = os.path.join(ROOT, "aes-walsh-ckks_amd")
_product = None


def product():
    """The HIP product's Python binding (aes-walsh-ckks_amd/mickks.py)."""
    global _product
    if _product is None:
        import importlib.util
        spec = importlib.util.spec_from_file_location("mickks", os.path.join(PKG_DIR, "mickks.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _product = mod
    return _product


_host = None


def host():
    """The C++ host API binding (aes-walsh-ckks_amd/ckks_py*.so).  PyTorch is
    imported first so that the HIP runtime it carries is the one the library
    binds to (one runtime per process)."""
    global _host
    if _host is None:
        import importlib.util
        import sysconfig
        import torch  # noqa: F401
        path = os.path.join(PKG_DIR, "ckks_py" + sysconfig.get_config_var("EXT_SUFFIX"))
        spec = importlib.util.spec_from_file_location("ckks_py", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _host = mod
    return _host



def philox4x32_10(ctr, key):
    """Philox4x32-10 (Salmon et al., SC 2011) on Python ints: ctr = 4 words,
    key = 2 words; used to check the device samplers."""
    c0, c1, c2, c3 = ctr
    k0, k1 = key
    M = 0xFFFFFFFF
    for r in range(10):
        if r:
            k0 = (k0 + 0x9E3779B9) & M
            k1 = (k1 + 0xBB67AE85) & M
        p0 = 0xD2511F53 * c0
        p1 = 0xCD9E8D57 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & M, p1 & M, ((p0 >> 32) ^ c3 ^ k1) & M, p0 & M
    return c0, c1, c2, c3


def sampler_draw(seed, counter, limb, index):
    """The 128 random bits the device samplers use for (seed, counter, limb, index)."""
    v = philox4x32_10((index, limb, counter & 0xFFFFFFFF, counter >> 32), (seed & 0xFFFFFFFF, seed >> 32))
    lo = (v[1] << 32) | v[0]
    hi = (v[3] << 32) | v[2]
    return lo, hi
