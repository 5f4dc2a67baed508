"""Served walsh15 AES-CTR transciphering with continuous batching.

A request is a run of 1-32 AES-CTR blocks (public counter and ciphertext
blocks) to be decrypted homomorphically under the server's encrypted AES key;
its result is the plaintext bytes, decrypted from the output bit ciphertexts.
The evaluation is ckks_aes_ctr_walsh_sbox_complex_blocks
(ckks_aes_ctr_walsh_sbox.cpp:1037-1051), whose slot space holds 1024 blocks:
logical block j rides in complex slot j mod 512, real lane for j < 512 and
imaginary lane above (aes_bench.cpp:156-196 reads them the same way).

The server keeps `lanes` evaluation lanes resident on the device.  A lane is
a set of contexts built from the same parameters and seed (hence the same
keys; the S-boxes of a round are spread over them, one host thread and stream
each) with its own encrypted round keys.  A dispatcher thread packs the
waiting requests, first come first served and whole requests only, into the
1024 blocks of the next free lane's evaluation (unused blocks are zero), and
the lane evaluates, decrypts and completes them.  Requests that arrive while
evaluations run wait for the next free lane, so their work overlaps with the
running ones on the device.  Admission is by HBM: a lane is opened only if
its measured footprint (the peak bytes of its contexts' memory pools after a
warm-up evaluation) fits the memory budget beside the lanes already open.
"""
import os
import sys
import threading
import time
from collections import deque

import numpy as np

BLOCKS_PER_EVALUATION = 1024
MAX_REQUEST_BLOCKS = 32


def _host():
    here = os.path.dirname(os.path.abspath(__file__))
    if here not in sys.path:
        sys.path.insert(0, here)
    import ckks_py
    return ckks_py


class Request:
    """counters, ciphertexts: uint8 [n, 16] with 1 <= n <= 32."""

    def __init__(self, counters, ciphertexts, arrival=None):
        counters = np.ascontiguousarray(counters, dtype=np.uint8)
        ciphertexts = np.ascontiguousarray(ciphertexts, dtype=np.uint8)
        if counters.ndim != 2 or counters.shape[1] != 16 or counters.shape != ciphertexts.shape:
            raise ValueError("a request needs matching [n, 16] counter and ciphertext blocks")
        if not 1 <= counters.shape[0] <= MAX_REQUEST_BLOCKS:
            raise ValueError(f"a request holds 1..{MAX_REQUEST_BLOCKS} blocks")
        self.counters, self.ciphertexts = counters, ciphertexts
        self.arrival = time.perf_counter() if arrival is None else arrival
        self.done = threading.Event()
        self.plaintexts = None   # uint8 [n, 16]
        self.values = None       # decrypted bit values, float64 [n, 128]
        self.finish = None
        self.error = None

    @property
    def blocks(self):
        return self.counters.shape[0]

    @property
    def latency(self):
        return None if self.finish is None else self.finish - self.arrival


class Lane:
    """One evaluation lane: `contexts` contexts of the same keys and the
    round keys encrypted under them."""

    def __init__(self, H, params_fn, contexts, aes_key):
        self.H = H
        self.ctxs = [H.CKKSContext(params_fn()) for _ in range(max(1, contexts))]
        self.ctx = self.ctxs[0]
        P = self.ctx.getParams()
        self.complex_blocks = P.getAesWalshBlockCount()
        if 2 * self.complex_blocks != BLOCKS_PER_EVALUATION:
            raise ValueError("the serving layout expects 512 complex blocks per evaluation")
        self.keys = H.ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(
            self.ctx, H.aes128_expand_key(list(aes_key)), self.complex_blocks)

    def footprint(self):
        return sum(int(c.memoryStats()[1]) for c in self.ctxs)

    def evaluate(self, counters, ciphertexts):
        """AES-CTR of 1024 logical blocks; returns the decrypted bit values,
        float64 [1024, 128]."""
        out = self.H.ckks_aes_ctr_walsh_sbox_complex_blocks(self.ctx, self.keys, counters, ciphertexts,
                                                            sbox_contexts=self.ctxs[1:])
        cb = self.complex_blocks
        vals = np.empty((2 * cb, 128))
        for bit, v in enumerate(self.H.ckks_decrypt_batch(self.ctx, out)):
            v = np.asarray(v)[:cb]
            vals[:cb, bit] = v.real
            vals[cb:, bit] = v.imag
        return vals


def values_to_bytes(vals):
    """Bit values [n, 128] -> bytes [n, 16], bit b of byte k at value 8k + b
    (a bit is 1 where its value is >= 0.5, as aes_bench.cpp:178-183)."""
    bits = (vals >= 0.5).astype(np.uint8).reshape(vals.shape[0], 16, 8)
    return (bits << np.arange(8, dtype=np.uint8)).sum(axis=2).astype(np.uint8)


class AesCtrServer:
    """Continuous-batching server over `lanes` lanes of `contexts_per_lane`
    contexts each.  submit() queues a request; completed requests carry their
    plaintexts, decrypted values and finish time."""

    def __init__(self, aes_key, lanes=1, contexts_per_lane=16, memory_budget=None, params_fn=None):
        self.H = _host()
        params_fn = params_fn or self.H.CKKSParams.AESWalsh15
        self.lanes = []
        self.footprints = []
        budget = memory_budget
        zeros = np.zeros((BLOCKS_PER_EVALUATION, 16), np.uint8)
        for _ in range(max(1, lanes)):
            lane = Lane(self.H, params_fn, contexts_per_lane, aes_key)
            lane.evaluate(zeros, zeros)   # warm-up: sizes workspaces and pools
            need = lane.footprint()
            if budget is not None and self.lanes and sum(self.footprints) + need > budget:
                break
            self.lanes.append(lane)
            self.footprints.append(need)
        self.queue = deque()
        self.cv = threading.Condition()
        self.free = list(range(len(self.lanes)))
        self.stopping = False
        self.evaluations = 0
        self.dispatcher = threading.Thread(target=self._dispatch, daemon=True)
        self.dispatcher.start()

    def submit(self, request):
        with self.cv:
            if self.stopping:
                raise RuntimeError("server is stopping")
            self.queue.append(request)
            self.cv.notify_all()
        return request

    def close(self):
        with self.cv:
            self.stopping = True
            self.cv.notify_all()
        self.dispatcher.join()

    def _take_batch(self):
        batch, used = [], 0
        while self.queue and used + self.queue[0].blocks <= BLOCKS_PER_EVALUATION:
            r = self.queue.popleft()
            batch.append((r, used))
            used += r.blocks
        return batch

    def _dispatch(self):
        workers = []
        while True:
            with self.cv:
                while not ((self.queue and self.free) or (self.stopping and not self.queue)):
                    self.cv.wait()
                if self.stopping and not self.queue:
                    break
                lane = self.free.pop(0)
                batch = self._take_batch()
            t = threading.Thread(target=self._run, args=(lane, batch), daemon=True)
            t.start()
            workers.append(t)
        for t in workers:
            t.join()

    def _run(self, lane_idx, batch):
        lane = self.lanes[lane_idx]
        try:
            counters = np.zeros((BLOCKS_PER_EVALUATION, 16), np.uint8)
            cts = np.zeros((BLOCKS_PER_EVALUATION, 16), np.uint8)
            for r, off in batch:
                counters[off:off + r.blocks] = r.counters
                cts[off:off + r.blocks] = r.ciphertexts
            vals = lane.evaluate(counters, cts)
            now = time.perf_counter()
            for r, off in batch:
                r.values = vals[off:off + r.blocks].copy()
                r.plaintexts = values_to_bytes(r.values)
                r.finish = now
                r.done.set()
        except Exception as e:  # every request of the batch carries the error
            for r, _ in batch:
                r.error = e
                r.finish = time.perf_counter()
                r.done.set()
        finally:
            with self.cv:
                self.evaluations += 1
                self.free.append(lane_idx)
                self.cv.notify_all()
