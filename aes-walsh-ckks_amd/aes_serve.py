"""Served walsh15 AES-CTR transciphering with continuous batching: the
Python face of the native server (host/ckks_aes_serve.cpp, CKKSAesCtrServer).

A request is a run of AES-CTR blocks (public counter and ciphertext blocks)
to be decrypted homomorphically under the server's encrypted AES key; its
result is the plaintext bytes and the decrypted bit values.  The server keeps
the evaluation keys and round keys resident once, in a key-owning context,
and runs `lanes` evaluation lanes of it (own stream and working memory each)
on native threads; a free lane packs the waiting requests, first come first
served and whole, into the 1024 blocks of its next evaluation (ckks_aes_ctr_
walsh_sbox_complex_blocks, ckks_aes_ctr_walsh_sbox.cpp:1037-1051: logical
block j in complex slot j mod 512, real lane below 512, imaginary above), so
requests that arrive while evaluations run are admitted into the next free
lane and the lanes overlap on the device.  Admission is by device memory:
lanes open while their measured footprint fits the budget, and a request is
refused (alone) when it exceeds one evaluation or when the work queued and
in flight would exceed the open lanes' memory.
"""
import os
import sys

import numpy as np

BLOCKS_PER_EVALUATION = 1024


def _host():
    here = os.path.dirname(os.path.abspath(__file__))
    if here not in sys.path:
        sys.path.insert(0, here)
    import ckks_py
    return ckks_py


def check_request(counters, ciphertexts, max_blocks=BLOCKS_PER_EVALUATION):
    """uint8 [n, 16] counter and ciphertext blocks with 1 <= n <= max_blocks."""
    counters = np.ascontiguousarray(counters, dtype=np.uint8)
    ciphertexts = np.ascontiguousarray(ciphertexts, dtype=np.uint8)
    if counters.ndim != 2 or counters.shape[1] != 16 or counters.shape != ciphertexts.shape:
        raise ValueError("a request needs matching [n, 16] counter and ciphertext blocks")
    if not 1 <= counters.shape[0] <= max_blocks:
        raise ValueError(f"a request holds 1..{max_blocks} blocks")
    return counters, ciphertexts


def values_to_bytes(vals):
    """Bit values [n, 128] -> bytes [n, 16], bit b of byte k at value 8k + b
    (a bit is 1 where its value is >= 0.5, as aes_bench.cpp:178-183)."""
    bits = (vals >= 0.5).astype(np.uint8).reshape(vals.shape[0], 16, 8)
    return (bits << np.arange(8, dtype=np.uint8)).sum(axis=2).astype(np.uint8)


class AesCtrServer:
    """`lanes` lanes over one key-owning context of `params` (default
    AESWalsh15).  submit() returns a request id at once (shape errors raise);
    result(id) blocks until that request completed or was refused and
    returns its plaintexts, values, error and arrival / start / finish
    times (steady-clock seconds)."""

    def __init__(self, aes_key, lanes=2, memory_budget=None, params=None, max_queued_evaluations=4):
        self.H = _host()
        self.owner = self.H.CKKSContext(params if params is not None else self.H.CKKSParams.AESWalsh15())
        self.native = self.H.CKKSAesCtrServer(self.owner, list(aes_key), lanes, int(memory_budget or 0),
                                              BLOCKS_PER_EVALUATION, max_queued_evaluations)

    @property
    def lanes(self):
        return self.native.laneCount()

    @property
    def evaluation_bytes(self):
        return self.native.evaluationBytes()

    @property
    def evaluations(self):
        return self.native.evaluations()

    def submit(self, counters, ciphertexts):
        counters = np.ascontiguousarray(counters, dtype=np.uint8)
        ciphertexts = np.ascontiguousarray(ciphertexts, dtype=np.uint8)
        if counters.ndim != 2 or counters.shape[1] != 16 or counters.shape != ciphertexts.shape:
            raise ValueError("a request needs matching [n, 16] counter and ciphertext blocks")
        return self.native.submit(counters, ciphertexts)

    def result(self, request_id):
        return self.native.wait(request_id)

    def close(self):
        """Finishes the admitted requests and stops the lanes."""
        self.native = None
