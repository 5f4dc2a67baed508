"""The AES Walsh table of the host layer (ckks_aes_walsh_table, the
reference's build_walsh_table) read back through the inverse Walsh
transform: it reproduces the AES S-box at the FIPS-197 values and is a
permutation without fixed points."""
import numpy as np

from helpers import host


def test_aes_walsh_table_reconstructs_the_sbox():
    H = host()
    t = np.array(H.ckks_aes_walsh_table(), dtype=np.int64)  # [bit][low mask][high mask] = W / 4
    assert t.shape == (8, 16, 16)
    masks = np.arange(256)
    x = np.arange(256)
    par = np.array([[bin(m & v).count("1") & 1 for v in x] for m in masks])  # [mask][x]
    sign = 1 - 2 * par
    W = 4 * np.stack([t[b][masks & 15, masks >> 4] for b in range(8)])      # [bit][mask]
    # sum_mask W(mask) (-1)^(mask . x) = 256 (-1)^f(x)
    f = W @ sign
    assert np.all(np.abs(f) == 256)
    bits = (f < 0).astype(np.int64)
    sbox = (bits * (1 << np.arange(8))[:, None]).sum(axis=0)
    for v, s in {0x00: 0x63, 0x01: 0x7C, 0x10: 0xCA, 0x20: 0xB7, 0x53: 0xED, 0xFF: 0x16}.items():
        assert sbox[v] == s, hex(v)
    assert sorted(sbox) == list(range(256))
    assert not np.any(sbox == x) and not np.any(sbox == (x ^ 0xFF))
