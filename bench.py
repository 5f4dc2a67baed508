"""Benchmark of the CKKS evaluation hot path on one MI355X.

Default workload: walsh15 AES-CTR (the reference's AESWalsh15 parameters,
N = 2^15), 1024 blocks per evaluation (512 complex slots x 2 lanes) on the
reference bench's counters and plaintexts; one step = one homomorphic CTR
decryption of all 1024 blocks (10 AES rounds of 16 mask-packed Walsh
S-boxes, each with its bootstrap, then the final binary refresh), with the
S-boxes of a round spread over 16 contexts of the same keys (one host thread
and stream each).  The plain form (the S-boxes one after another on one
stream) is timed beside it and checked bit for bit; the outputs are
decrypted and compared with the plaintexts (0 mismatched blocks).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload aes|serve|ntt|keyswitch|lt|chain|hoisted|walsh|sbox]

Prints one JSON line.  Timing: W untimed steps, then K timed steps bracketed
by torch.cuda.synchronize(); per-launch durations come from HIP events on the
stream the library launches on.  The CPU baseline times the oracle (a C
restatement of the reference's arithmetic, OpenMP) on a bounded sample.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


def _host():
    sys.path.insert(0, os.path.join(ROOT, "aes-walsh-ckks_amd"))
    import ckks_py
    return ckks_py


def _primes(p_bits, q_bits, N):
    """The reference's prime search (find_p_and_q_primes, random.cpp:66-131),
    from the product's host layer: (P primes, Q primes) as uint64 arrays."""
    p, q = _host().find_p_and_q_primes(list(p_bits), list(q_bits), N)
    return np.array(p, dtype=np.uint64), np.array(q, dtype=np.uint64)


def _rand_limbs(torch, q, N, count_per_limb, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    L = len(q)
    out = torch.empty((count_per_limb, L, N), dtype=torch.int64, device="cuda")
    for i, qi in enumerate(q):
        out[:, i, :] = torch.randint(0, int(qi), (count_per_limb, N), generator=g, device="cuda",
                                     dtype=torch.int64)
    return out.reshape(-1)


def _profile_traffic(name):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get(name)
    return None


def bench_ntt(args, M, torch):
    N = 1 << 16
    L = 40
    ncts, npolys = 64, 2
    _, q = _primes([], [60] * L, N)
    p, _ = _primes([61], [], N)
    ctx = M.Context(N, q, p, 1)
    polys = ncts * npolys
    limbs = polys * L
    data = _rand_limbs(torch, q, N, polys, seed=2026)
    ref = data.clone()
    stream = torch.cuda.current_stream()

    def step():
        ctx.ntt_forward(data, polys, L)
        ctx.ntt_inverse(data, polys, L)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1, e2 in ev:
        e0.record(stream)
        ctx.ntt_forward(data, polys, L)
        e1.record(stream)
        ctx.ntt_inverse(data, polys, L)
        e2.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in ev]))
    inv_ms = float(np.mean([b.elapsed_time(c) for _, b, c in ev]))
    ok = bool(torch.equal(data, ref))
    ms_per_step = wall * 1e3 / args.steps
    value = 2 * limbs / (ms_per_step / 1e3)  # limb-NTTs per second (each direction counts)

    # plain form: one limb per launch pair, on a 2-ciphertext sample
    plain_rate = None
    if args.plain:
        plain_rate = _plain_ntt_rate(ctx, data, N, L, torch)


    algo_bytes = 16 * N * limbs  # read + write every word once per direction
    achieved = algo_bytes / (fwd_ms / 1e3)
    traffic = _profile_traffic("ntt16_forward")
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic,
                "kernel": "forward NTT (column + row pass) over 5120 limbs",
                "forward_ms": fwd_ms, "inverse_ms": inv_ms,
                "inverse_achieved_GBps": algo_bytes / (inv_ms / 1e3) / 1e9}

    cpu = cpu_baseline_ntt(N, q) if args.cpu_baseline else None
    return {
        "metric": "batched negacyclic NTT throughput (forward+inverse), N=2^16, 40 x 60-bit limbs",
        "value": value, "unit": "limb-NTT/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 (exact mod-q integer)", "data": "synthetic (uniform residues, seed 2026)",
        "config": {"workload": "batched NTT: N=2^16, 64 ct x 2 polys x 40 limbs (60-bit primes), "
                               "forward + inverse, one launch pair per direction",
                   "limbs": limbs, "round_trip_exact": ok,
                   "plain_one_limb_per_launch_limb_ntt_per_s": plain_rate,
                   "batched_over_plain": value / plain_rate if plain_rate else None,
                   "tuning": ctx.tune_report()},
        "roofline": roofline, "cpu_baseline": cpu,
    }


def _plain_ntt_rate(ctx, data, N, L, torch):
    """One limb per launch pair (the plain form), on a 2-ciphertext sample."""
    sample = 4 * L
    sub = data[: sample * N]
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        for k in range(sample):
            ctx.ntt_forward(sub[k * N:(k + 1) * N], 1, 1, basis0=k % L)
        for k in range(sample):
            ctx.ntt_inverse(sub[k * N:(k + 1) * N], 1, 1, basis0=k % L)
    torch.cuda.synchronize()
    return 2 * sample * reps / (time.perf_counter() - t0)


def cpu_baseline_ntt(N, q):
    O = _oracle()
    octx = O.Ctx(N, q[:40], [O.find_primes([61], [], N)[0][0]], 1)
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.integers(0, int(qi), N, dtype=np.uint64) for qi in q[:40]])
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 10.0:
        y = octx.ntt_fwd(x, 40)
        octx.ntt_inv(y, 40)
        done += 80
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "limb-NTT/s", "cores": int(O.lib().oc_omp_threads()),
            "kind": "port",
            "sample": "oracle forward+inverse NTT of one 40-limb N=2^16 poly, repeated for 10 s "
                      "(OpenMP over limbs)"}


def bench_keyswitch(args, M, torch):
    """Batched hybrid key switching at N=2^16: Q = 30 x 50-bit, P = 10 x 50-bit,
    dnum = 3 (alpha 10), 128 ciphertexts at levels drawn from 8..30 limbs,
    16 rotation keys (random residues; the arithmetic does not depend on them)."""
    N = 1 << 16
    p, q = _primes([50] * 10, [50] * 30, N)
    ctx = M.Context(N, q, p, 3)
    rng = np.random.default_rng(3)
    nkeys, nct = 16, 128
    KL = 40
    keys = [_rand_limbs(torch, list(q) + list(p), N, 2 * 3, seed=100 + k) for k in range(nkeys)]
    levels = rng.integers(7, 30, nct)
    levels[:3] = [29, 19, 7]
    items, inputs = [], []
    for i, lv in enumerate(levels):
        Lc = int(lv) + 1
        kt = _rand_limbs(torch, q[:Lc], N, 1, seed=1000 + i)
        b = _rand_limbs(torch, q[:Lc], N, 1, seed=2000 + i)
        a = _rand_limbs(torch, q[:Lc], N, 1, seed=3000 + i)
        items.append((int(lv), kt, keys[i % nkeys], b, a, 0, 1))
        inputs.append((b.clone(), a.clone()))
    # one switch of the batch checked against the oracle for items 0-2
    # (levels 29, 19, 7): the timed steps then reuse the same buffers
    ctx.keyswitch_batch(items)
    torch.cuda.synchronize()
    O = _oracle()
    octx = O.Ctx(N, q, p, 3)
    host = lambda t: t.detach().cpu().numpy().view(np.uint64)
    for i in range(3):
        lv, kt, key, b, a, _, _ = items[i]
        ob, oa = octx.keyswitch(lv, host(kt), host(key), host(inputs[i][0]), host(inputs[i][1]), 0, 1)
        if not (np.array_equal(host(b), ob) and np.array_equal(host(a), oa)):
            raise SystemExit(f"batched key switch item {i} (level {lv}) differs from the oracle")
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        ctx.keyswitch_batch(items)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        ctx.keyswitch_batch(items)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = wall * 1e3 / args.steps
    Ls = levels + 1
    algo = float(np.sum(8 * N * (4 * Ls))) + nkeys * 2 * 3 * KL * 8 * N
    cpu = cpu_baseline_keyswitch(N, q, p) if args.cpu_baseline else None
    return {
        "metric": "batched hybrid key switches/s, N=2^16, L in [8,30], K=10, dnum=3",
        "value": nct / (ms / 1e3), "unit": "keyswitch/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64 (exact mod-q integer)", "data": "synthetic (seeded)",
        "config": {"workload": "128 key switches, mixed levels, 16 keys", "algorithmic_GBps":
                   algo / (ms / 1e3) / 1e9, "oracle_checked": "items at levels 29, 19, 7 bit-exact"},
        "roofline": None, "cpu_baseline": cpu,
    }


def cpu_baseline_keyswitch(N, q, p):
    """The oracle's hybrid key switch at the batch's mean level (L = 19), repeated for ~10 s."""
    O = _oracle()
    octx = O.Ctx(N, q, p, 3)
    rng = np.random.default_rng(2)
    level = 18
    L = level + 1
    rnd = lambda mods: np.concatenate([rng.integers(0, int(m), N, dtype=np.uint64) for m in mods])
    key = np.concatenate([rnd(list(q) + list(p)) for _ in range(6)])
    kt, b, a = rnd(q[:L]), rnd(q[:L]), rnd(q[:L])
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 10.0:
        octx.keyswitch(level, kt, key, b, a, 0, 1)
        done += 1
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "keyswitch/s", "cores": int(O.lib().oc_omp_threads()),
            "kind": "port", "sample": "oracle key switch at L=19 (the batch's mean level), repeated for 10 s"}


def bench_lt(args, M, torch):
    """BSGS linear-transform stage with hoisted rotations at N=2^16, 2^15 slots,
    level 24 (25 Q limbs), P = 10 x 50-bit, dnum 3: 128 diagonals
    {16 g + b : g < 8, b < 16} (16 baby x 8 giant steps) over the QP basis, in
    Montgomery form, applied to 32 input ciphertexts.  Diagonals, keys and
    ciphertexts are seeded random residues (the integer work does not depend
    on their values)."""
    N = 1 << 16
    p, q = _primes([50] * 10, [50] * 30, N)
    ctx = M.Context(N, q, p, 3)
    level, bs, giants, nct = 24, 16, 8, 32
    L, K = level + 1, len(p)
    qp = list(q[:L]) + list(p)
    diag_idx = [bs * g + b for g in range(giants) for b in range(bs)]
    diags = [_rand_limbs(torch, qp, N, 1, seed=5000 + t) for t in range(len(diag_idx))]
    parts = ctx.parts(len(q) - 1)
    rots = list(range(1, bs)) + [bs * g for g in range(1, giants)]
    keys = {r: _rand_limbs(torch, list(q) + list(p), N, 2 * parts, seed=7000 + r) for r in rots}
    cts = [_rand_limbs(torch, q[:L], N, 2, seed=9000 + i) for i in range(nct)]
    outs = [torch.empty(2 * L * N, dtype=torch.int64, device="cuda") for _ in range(nct)]
    W = L * N
    pairs = [(c[W:], c[:W]) for c in cts]
    opairs = [(o[W:], o[:W]) for o in outs]
    stream = torch.cuda.Stream()

    def step():
        # all 32 ciphertexts through the stage in one call
        ctx.linear_transform_qp_batch(level, pairs, diag_idx, diags, bs, keys, opairs, stream=stream)

    def step_plain():
        for (a, b), (ao, bo) in zip(pairs, opairs):
            ctx.linear_transform_qp(level, a, b, diag_idx, diags, bs, keys, ao, bo, stream=stream)

    def timed(fn, n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / n

    with torch.cuda.stream(stream):
        for _ in range(max(1, args.warmup)):
            step()
    torch.cuda.synchronize()
    ref = torch.cat(outs).clone()
    eager_ms = timed(step, args.steps)
    plain_ms = timed(step_plain, 1) if args.plain else None
    if args.eager:
        return _lt_result(args, eager_ms, eager_ms, plain_ms, None, nct, N, q, p, L, K, diag_idx, rots, parts, level,
                          bs)
    # the batch captured once in a HIP graph and replayed
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        step()
    for o in outs:
        o.zero_()
    replay_ok = True
    for _ in range(2):
        graph.replay()
        torch.cuda.synchronize()
        replay_ok = replay_ok and bool(torch.equal(torch.cat(outs), ref))
        for o in outs:
            o.zero_()
    if not replay_ok:
        raise RuntimeError("graph replay differs from the eager batch")
    ms = timed(graph.replay, args.steps)
    torch.cuda.synchronize()
    if not torch.equal(torch.cat(outs), ref):
        raise RuntimeError("graph replay differs from the eager batch")
    return _lt_result(args, ms, eager_ms, plain_ms, replay_ok, nct, N, q, p, L, K, diag_idx, rots, parts, level, bs)


def _lt_result(args, ms, eager_ms, plain_ms, replay_ok, nct, N, q, p, L, K, diag_idx, rots, parts, level, bs):
    T = L + K
    # per ciphertext: ct in/out, diagonals and keys amortised over the batch
    algo = nct * (2 * L * 8 * N * 2) + len(diag_idx) * T * 8 * N + len(rots) * parts * 2 * (len(q) + K) * 8 * N
    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline_lt(N, q, p, level, diag_idx, bs, rots, parts)
    return {
        "metric": "BSGS linear-transform stage throughput, N=2^16, level 24, 128 diagonals",
        "value": nct / (ms / 1e3), "unit": "ciphertext/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64 (exact mod-q integer)", "data": "synthetic (seeded residues)",
        "config": {"workload": "32 ct x BSGS stage (16 baby x 8 giant), QP diagonals, 22 rotation keys, "
                               + ("one batched call (eager)" if replay_ok is None else
                                  "one batched call captured in a HIP graph"),
                   "algorithmic_GBps": algo / (ms / 1e3) / 1e9, "eager_batched_ms": eager_ms,
                   "plain_one_ct_per_call_ms": plain_ms,
                   "batched_over_plain": (plain_ms / ms) if plain_ms else None,
                   "graph_replays_match_eager": replay_ok},
        "roofline": None, "cpu_baseline": cpu,
    }


def cpu_baseline_lt(N, q, p, level, diag_idx, bs, rots, parts):
    """The oracle's BSGS stage on one ciphertext (bounded sample)."""
    O = _oracle()
    octx = O.Ctx(N, q, p, 3)
    rng = np.random.default_rng(1)
    L = level + 1
    qp = list(q[:L]) + list(p)
    rnd = lambda mods: np.concatenate([rng.integers(0, int(m), N, dtype=np.uint64) for m in mods])
    a, b = rnd(q[:L]), rnd(q[:L])
    diags = np.concatenate([rnd(qp) for _ in diag_idx])
    keys = {r: np.concatenate([rnd(list(q) + list(p)) for _ in range(2 * parts)]) for r in rots}
    t0 = time.perf_counter()
    octx.lt_stage_qp(level, a, b, diag_idx, diags, bs, keys)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "ciphertext/s", "cores": int(O.lib().oc_omp_threads()),
            "kind": "port", "sample": "oracle BSGS stage on one ciphertext (same shape)"}


def bench_hoisted(args, M, torch):
    """Hoisted rotations of one ciphertext at the walsh15 shape: N=2^15, the
    walsh15 chain (13 limbs: 36, 61, 42, 42, 38 x 7, 50, 50 bits) at its top
    level, P = 4 x 61 bits, dnum 3; 58 rotations (the Walsh S-box unpack
    count) by 512 s for s = 1..29 and their negatives, keys as seeded
    residues.  One step = hoist the a-part once, then all 58 rotations."""
    N = 1 << 15
    p, q = _primes([61] * 4, [36, 61, 42, 42] + [38] * 7 + [50, 50], N)
    ctx = M.Context(N, q, p, 3)
    L, K = len(q), len(p)
    level = L - 1
    parts = ctx.parts(level)
    slots = N // 2
    rots = [512 * s for s in range(1, 30)] + [slots - 512 * s for s in range(1, 30)]
    gal = [pow(5, r, 2 * N) for r in rots]
    keys = [_rand_limbs(torch, list(q) + list(p), N, 2 * parts, seed=300 + i) for i in range(len(rots))]
    ct = _rand_limbs(torch, q, N, 2, seed=77)
    W = L * N
    b, a = ct[:W], ct[W:]
    dig = torch.empty(parts * (L + K) * N, dtype=torch.int64, device="cuda")
    outs = [(torch.empty(W, dtype=torch.int64, device="cuda"), torch.empty(W, dtype=torch.int64, device="cuda"))
            for _ in rots]

    def step():
        ctx.hoist(level, a, dig)
        ctx.rotate_hoisted_batch(level, dig, b, gal, keys, outs)

    def step_plain():
        ctx.hoist(level, a, dig)
        for g, k, (bo, ao) in zip(gal, keys, outs):
            bo.copy_(b)
            ctx.keyswitch_hoisted(level, dig, k, bo, ao, 0, 1, galois=g)

    def timed(fn, n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / n

    for _ in range(args.warmup):
        step()
    ms = timed(step, args.steps)
    ref = [torch.cat(o).clone() for o in outs]
    plain_ms = None
    if args.plain:
        step_plain()
        plain_ms = timed(step_plain, max(1, args.steps // 4))
        if not all(torch.equal(torch.cat(o), r) for o, r in zip(outs, ref)):
            raise RuntimeError("batched and one-by-one hoisted rotations differ")
    cpu = cpu_baseline_hoisted(N, q, p, level, gal, parts) if args.cpu_baseline else None
    return {
        "metric": "hoisted rotations/s of one ciphertext, walsh15 shape (N=2^15, 13 + 4 limbs, dnum 3)",
        "value": len(rots) / (ms / 1e3), "unit": "rotation/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64 (exact mod-q integer)", "data": "synthetic (seeded residues)",
        "config": {"workload": "hoist once + 58 rotations (512 s, s = +-1..29) in one batched call",
                   "plain_one_rotation_per_call_ms": plain_ms,
                   "batched_over_plain": (plain_ms / ms) if plain_ms else None},
        "roofline": None, "cpu_baseline": cpu,
    }


def cpu_baseline_hoisted(N, q, p, level, gal, parts):
    """The oracle's hoist and hoisted key switches (a bounded sample of 8
    rotations, scaled to rotations/s)."""
    O = _oracle()
    octx = O.Ctx(N, q, p, 3)
    rng = np.random.default_rng(5)
    L = level + 1
    rnd = lambda mods: np.concatenate([rng.integers(0, int(m), N, dtype=np.uint64) for m in mods])
    a, b = rnd(q[:L]), rnd(q[:L])
    key = np.concatenate([rnd(list(q) + list(p)) for _ in range(2 * parts)])
    t0 = time.perf_counter()
    dig = octx.hoist(level, a)
    for g in gal[:8]:
        octx.keyswitch_hoisted(level, dig, key, b, np.zeros(L * N, np.uint64), 0, 1, galois=g)
    dt = time.perf_counter() - t0
    return {"value": 8 / dt, "unit": "rotation/s", "cores": int(O.lib().oc_omp_threads()), "kind": "port",
            "sample": "oracle hoist + 8 hoisted rotations at the same shape"}


def bench_walsh(args, M, torch):
    """The four output pairs of one Walsh S-box at the walsh15 shape: N=2^15,
    the walsh15 chain with the sign ciphertexts at level 2 (36-, 61- and
    42-bit limbs), P = 4 x 61 bits, dnum 3; 15 + 15 sign ciphertexts (seeded
    residues), the AES Walsh table, the encoded i in Montgomery form.  One
    step = for each pair, the fused product sum (linear combinations,
    i-combination, 15 raw 4+4 products), then the four relinearisations as
    one key-switch batch, captured in a HIP graph.  The comparison runs the reference's structure of the same
    pairs (ckks_walsh_sbox.cpp:862-1043: one scaled addition per sign term,
    a plaintext product per imaginary part, a tensor and an addition per
    product), also captured in a graph, and checks both agree bit for bit."""
    import ctypes
    H = _host()
    N = 1 << 15
    p, q = _primes([61] * 4, [36, 61, 42, 42] + [38] * 7 + [50, 50], N)
    ctx = M.Context(N, q, p, 3)
    level = 2
    L = level + 1
    W = L * N
    parts = ctx.parts(level)
    table = np.array(H.ckks_aes_walsh_table(), dtype=np.int64)
    signs = _rand_limbs(torch, q[:L], N, 60, seed=901).reshape(60, W)
    low = [(signs[2 * m], signs[2 * m + 1]) for m in range(15)]
    high = [(signs[30 + 2 * m], signs[31 + 2 * m]) for m in range(15)]
    pt = _rand_limbs(torch, q[:L], N, 1, seed=902)
    key = _rand_limbs(torch, list(q) + list(p), N, 2 * parts, seed=903)
    scale = 2.0 ** 38
    pairs = []
    for pr in range(4):
        cr = table[pr].T.copy()        # [h][m]
        ci = table[pr + 4].T.copy()
        consts = [(int(round(float(cr[h][0]) * scale)), int(round(float(ci[h][0]) * scale))) for h in range(16)]
        pairs.append((cr, ci, consts))
    new = lambda n: torch.empty(n * W, dtype=torch.int64, device="cuda")
    fused_out = [(new(2), new(3)) for _ in range(4)]
    ref_out = [(new(2), new(3)) for _ in range(4)]
    tmp_r, tmp_i, c2 = new(2), new(2), new(1)
    stream = torch.cuda.Stream()
    lib = M.lib()

    def res(v):
        return np.array([int(v) % int(qi) for qi in q[:L]], dtype=np.uint64)

    def fused():
        for (cr, ci, consts), (acc, raw) in zip(pairs, fused_out):
            ctx.walsh_product_sum(level, low, high, cr, ci, consts, pt, (acc[:W], acc[W:]),
                                  (raw[:W], raw[W:2 * W], raw[2 * W:]), pt_montgomery=True, stream=stream)
        # the four relinearisations as one key-switch batch: (b, a) += KS(c2)
        ctx.keyswitch_batch([(level, raw[2 * W:], key, raw[:W], raw[W:2 * W], 0, 0) for _, raw in fused_out],
                            stream=stream)

    def lc(out, coeffs, const):
        M.check(lib.mi_memset_zero(ctx.h, M.ptr(out), out.numel() * 8, M._stream(stream)))
        if const:
            ctx.add_scalar(level, out[:W], res(const), stream=stream)
        for m in range(1, 16):
            c = int(coeffs[m])
            if c:
                src = signs[2 * (m - 1)]  # (b, a) of mask m are consecutive rows
                r = res(c)
                M.check(lib.mi_mul_scalar_add(ctx.h, level, M.ptr(out), M.ptr(src), 2, W, r.ctypes.data,
                                              M._stream(stream)))

    def reference_structure():
        for (cr, ci, consts), (acc, raw) in zip(pairs, ref_out):
            for h in range(16):
                lc(tmp_r, cr[h], consts[h][0])
                lc(tmp_i, ci[h], consts[h][1])
                ctx.mul_plain(level, tmp_i, 2, pt, pt_montgomery=True, stride=W, stream=stream)
                ctx.add(level, tmp_r, tmp_i, 2, stride=W, stream=stream)
                if h == 0:
                    acc.copy_(tmp_r)
                    continue
                hb, ha = high[h - 1]
                ctx.tensor(level, tmp_r[:W], tmp_r[W:], hb, ha, c2, stream=stream)
                if h == 1:
                    raw[:2 * W].copy_(tmp_r)
                    raw[2 * W:].copy_(c2)
                else:
                    ctx.add(level, raw[:2 * W], tmp_r, 2, stride=W, stream=stream)
                    ctx.add(level, raw[2 * W:], c2, 1, stride=W, stream=stream)
            ctx.relin(level, raw[:W], raw[W:2 * W], raw[2 * W:], key, stream=stream)

    # rows 2m, 2m+1 of `signs` are low mask m+1 (b, a): lc reads them as one 2-poly block
    assert all(low[m][0].data_ptr() + W * 8 == low[m][1].data_ptr() for m in range(15))

    def capture(fn):
        with torch.cuda.stream(stream):
            fn()                      # eager once: sizes the workspaces
        stream.synchronize()
        ctx.graph_begin(stream)
        with torch.cuda.stream(stream):
            fn()
        return ctx.graph_end(stream)

    g_fused = capture(fused)
    if args.plain:
        g_ref = capture(reference_structure)
        stream.synchronize()
        for (a1, r1), (a2, r2) in zip(fused_out, ref_out):
            if not (torch.equal(a1, a2) and torch.equal(r1[:2 * W], r2[:2 * W])):
                raise RuntimeError("fused and reference-structured Walsh sums differ")

    def timed(g, n):
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            g.launch(stream)
        stream.synchronize()
        return (time.perf_counter() - t0) * 1e3 / n

    for _ in range(args.warmup):
        g_fused.launch(stream)
    ms = timed(g_fused, args.steps)
    ref_ms = timed(g_ref, max(1, args.steps // 4)) if args.plain else None
    g_fused.destroy()
    if args.plain:
        g_ref.destroy()
    cpu = cpu_baseline_walsh(N, q, p, level, pairs[0]) if args.cpu_baseline else None
    return {
        "metric": "Walsh S-box output-pair sums/s, walsh15 shape (N=2^15, signs at level 2)",
        "value": 4 / (ms / 1e3), "unit": "output-pair/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64 (exact mod-q integer)", "data": "synthetic (seeded residues)",
        "config": {"workload": "4 output pairs of one S-box: fused product sums + one batch of 4 "
                               "relinearisations, HIP graph",
                   "sbox_per_s": 1 / (ms / 1e3),
                   "reference_structure_ms": ref_ms,
                   "fused_over_reference_structure": (ref_ms / ms) if ref_ms else None},
        "roofline": None, "cpu_baseline": cpu,
    }


def cpu_baseline_walsh(N, q, p, level, pair):
    """The oracle's output-pair product sum (one pair, the same shape)."""
    O = _oracle()
    octx = O.Ctx(N, q, p, 3)
    rng = np.random.default_rng(6)
    L = level + 1
    rnd = lambda: np.stack([np.concatenate([rng.integers(0, int(m), N, dtype=np.uint64) for m in q[:L]])
                            for _ in range(2)])
    low = [rnd() for _ in range(15)]
    high = [rnd() for _ in range(15)]
    pt = rnd()[0]
    cr, ci, consts = pair
    t0 = time.perf_counter()
    octx.walsh_product_sum(level, low, high, cr, ci, consts, pt, True)
    dt = time.perf_counter() - t0
    return {"value": 1 / dt, "unit": "output-pair/s", "cores": int(O.lib().oc_omp_threads()), "kind": "port",
            "sample": "oracle product sum of one output pair (no relinearisation) at the same shape"}


def walsh15_context(H, seed=15):
    """The reference's AESWalsh15 parameters (ckks_aes_params.hpp:35-82) in
    the host API: N=2^15, q0 36 bits, mult limbs 42 x 2, scale 2^35, dnum 3,
    P = 4 x 61 bits, Hamming weight 256; bootstrap: sparse secret of weight
    32, S2C one stage of scale groups 18+20+23, C2S two stages of 25+25,
    EvalMod 7 x 38 bits with K from the BinBoot bound at 2^-16, degree 100,
    QP evaluators; the mask-packed S-box's 58 segment rotation keys
    (block_count 512).  The EvalMod coefficients are a Chebyshev interpolant
    of 0.5 - 0.5 cos(2 pi K u) on [-1, 1] (the reference fits its own with
    MPFR)."""
    c = H.CKKSParamsConfig()
    c.N = 1 << 15
    c.mult_depth = 2
    c.q0_bit = 36
    c.q_mult_bits = [42, 42]
    c.scale = 2.0 ** 35
    c.dnum = 3
    c.p_bits = [61, 61, 61, 61]
    c.secret_key_hamming_weight = 256
    c.secure = False
    c.bts_keys = True
    c.seed = seed
    c.rotate_values = [512 * s for s in range(1, 30)] + [-512 * s for s in range(1, 30)]
    bp = H.CKKSBootstrapParams()
    bp.sparse_secret_hamming_weight = 32
    bp.linear_transform_log_plaintext_scales = [[[18, 20, 23]], [[25, 25], [25, 25]]]
    bp.linear_transform_depths = [1, 2]
    bp.evalmod_depth = 7
    bp.evalmod_bit_size = 38
    bp.evalmod_k_failure_probability_log2 = -16.0
    bp.evalmod_polynomial_degree = 100
    bp.linear_transform_log_bsgs_ratios = [1, 1]
    bp.linear_transform_use_qp = [True, True]
    c.bootstrap_params = bp
    ctx = H.CKKSContext(H.CKKSParams(c))
    K = ctx.getParams().getBootstrapEvalModK()
    coeffs = np.polynomial.chebyshev.chebinterpolate(lambda u: 0.5 - 0.5 * np.cos(2 * np.pi * K * u), 100)
    coeffs[1::2] = 0.0
    ctx.setEvalModPolynomial(list(coeffs))
    return ctx


def bench_sbox(args, M, torch):
    """One round of AES S-boxes of the mask-packed Walsh circuit at the
    walsh15 parameters: 16 S-boxes as one ciphertext batch (bit k of 16
    bytes stacked), 8 bit ciphertexts at the S2C entry level, 512 blocks per
    lane, two lanes; parity packing, the complex-packed bootstrap, signs,
    unpacking, the Walsh output pairs and the split into 8 output bits,
    through the C++ host API.  Timed eagerly and as a HIP graph recorded
    once and replayed (ckks_capture_graph); the replayed outputs are
    compared limb for limb with the eager ones and decrypted against the
    AES S-box."""
    H = _host()
    t0 = time.perf_counter()
    ctx = walsh15_context(H)
    setup_s = time.perf_counter() - t0
    P = ctx.getParams()
    slots, bc, B = P.getSlots(), 512, 16
    rng = np.random.default_rng(29)
    level = P.getBootstrapS2CDepth()
    xs = [(rng.integers(0, 256, bc), rng.integers(0, 256, bc)) for _ in range(B)]
    members = [[] for _ in range(8)]
    for x0, x1 in xs:
        for k in range(8):
            v = np.zeros(slots, np.complex128)
            v[:bc] = ((x0 >> k) & 1) + 1j * ((x1 >> k) & 1)
            members[k].append(H.ckks_encrypt(ctx, v, level))
    bits = [H.ckks_stack(ctx, m) for m in members]

    def sbox():
        return H.ckks_walsh_sbox_complex_blocks_mask_packed(ctx, [c.clone() for c in bits], bc)

    for _ in range(max(1, args.warmup)):
        out = sbox()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = sbox()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    held = []
    graph = H.ckks_capture_graph(ctx, lambda: held.extend(sbox()))
    graph.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.launch()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    same = all(np.array_equal(held[k].getMemberPoly(j, m), out[k].getMemberPoly(j, m))
               for k in range(8) for m in range(B) for j in range(2))
    # the AES S-box from the Walsh table (its inverse transform)
    table = np.array(H.ckks_aes_walsh_table(), dtype=np.int64)
    par = np.array([[bin(m & int(v)).count("1") & 1 for v in range(256)] for m in range(256)])
    masks = np.arange(256)
    f = np.stack([4 * table[b][masks & 15, masks >> 4] for b in range(8)]) @ (1 - 2 * par)
    sbox_tab = ((f < 0).astype(np.int64) * (1 << np.arange(8))[:, None]).sum(axis=0)
    worst, wrong = 0.0, 0
    members_out = [H.ckks_unstack(ctx, ct) for ct in held]
    for k in range(8):
        dec = H.ckks_decrypt_batch(ctx, members_out[k])
        for m, (x0, x1) in enumerate(xs):
            d = np.asarray(dec[m])[:bc]
            want = ((sbox_tab[x0] >> k) & 1) + 1j * ((sbox_tab[x1] >> k) & 1)
            worst = max(worst, float(np.max(np.abs(d - want))))
            wrong += int(np.sum(np.round(d.real) != want.real) + np.sum(np.round(d.imag) != want.imag))
    graph.reset()
    return {
        "metric": "walsh15 AES S-box rounds/s (16 mask-packed Walsh S-boxes as one batch, 512 blocks x 2 lanes)",
        "value": 1e3 / graph_ms, "unit": "rounds/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": graph_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 (exact mod-q integer)", "data": "synthetic (random bytes, seeded keys)",
        "config": {"workload": "one AES round of S-boxes: 16 S-boxes as one batch, replayed as a HIP graph",
                   "eager_ms": ms, "graph_ms": graph_ms, "graph_equals_eager": same,
                   "byte_substitutions_per_s": B * 2 * bc * 1e3 / graph_ms, "output_level": out[0].getLevel(),
                   "max_abs_error": worst, "wrong_bits": wrong, "context_setup_s": setup_s,
                   "tuning": json.loads(ctx.tuneReport())},
        "roofline": None, "cpu_baseline": None,
    }


AES_KEY = bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c")


def aes_bench_blocks(H, count):
    """The blocks of the reference's AES bench (aes_bench.cpp:29-101): counter
    f0f1..ff + i (big-endian), plaintext 0 the SP 800-38A block and block i
    bytes (0x5a + 17 i + 29 b) & 0xff, ciphertext = AES-CTR under the test key."""
    base = int.from_bytes(bytes.fromhex("f0f1f2f3f4f5f6f7f8f9fafbfcfdfeff"), "big")
    counters = np.array([list((base + i).to_bytes(16, "big")) for i in range(count)], np.uint8)
    plains = np.array([list(bytes.fromhex("6bc1bee22e409f96e93d7e117393172a")) if i == 0 else
                       [(0x5A + 17 * i + 29 * b) & 0xFF for b in range(16)] for i in range(count)], np.uint8)
    cts = np.array([H.aes128_ctr_crypt_block(list(AES_KEY), list(c), list(p)) for c, p in zip(counters, plains)],
                   np.uint8)
    return counters, plains, cts


def aes_validate(H, ctx, bits, expected, cb):
    """validate_walsh_outputs (aes_bench.cpp:156-196): mismatched blocks, max
    and average error of the decrypted bits."""
    decoded = np.zeros((2 * cb, 16), np.uint8)
    errs = []
    for bit in range(128):
        v = np.asarray(H.ckks_decrypt(ctx, bits[bit]))[:cb]
        for lane, vals in enumerate((v.real, v.imag)):
            rows = np.arange(cb) + lane * cb
            want = (expected[rows, bit // 8] >> (bit % 8)) & 1
            decoded[rows, bit // 8] |= (vals >= 0.5).astype(np.uint8) << (bit % 8)
            errs.append(np.abs(vals - want))
    errs = np.concatenate(errs)
    return int(np.sum(np.any(decoded != expected, axis=1))), float(errs.max()), float(errs.mean())


def bench_aes(args, M, torch):
    """walsh15 AES-CTR (aes_bench.cpp:247-292): 1024 blocks (512 complex
    slots x 2 lanes) decrypted homomorphically per step through
    ckks_aes_ctr_walsh_sbox_complex_blocks; the context build and round-key
    encryption are outside the timed region, as in the reference."""
    H = _host()
    t0 = time.perf_counter()
    params = H.CKKSParams.AESWalsh15()  # one seed: every context holds the same keys
    ctxs = [H.CKKSContext(params) for _ in range(max(1, args.aes_contexts))]
    ctx, workers = ctxs[0], ctxs[1:]
    cb = ctx.getParams().getAesWalshBlockCount()
    counters, plains, cts = aes_bench_blocks(H, 2 * cb)
    keys = H.ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(ctx, H.aes128_expand_key(list(AES_KEY)), cb)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    print(f"# walsh15 context and round keys: {setup_s:.1f} s", file=sys.stderr, flush=True)
    batched = not workers  # one context: the 16 S-boxes of a round as one ciphertext batch

    def evaluate():
        return H.ckks_aes_ctr_walsh_sbox_complex_blocks(ctx, keys, counters, cts, sbox_contexts=workers,
                                                        batch_sboxes=batched)

    for _ in range(args.warmup):
        out = evaluate()
    torch.cuda.synchronize()
    lat = []
    for step in range(args.steps):
        t1 = time.perf_counter()
        out = evaluate()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
        print(f"# step {step}: {lat[-1]:.3f} s", file=sys.stderr, flush=True)
    ms = 1e3 * sum(lat) / len(lat)
    mism, max_err, avg_err = aes_validate(H, ctx, out, plains, cb)
    # the roofline kernel inside the evaluation itself: one more evaluation
    # with HIP events around every NTT launch pair on its stream (the probe
    # keeps the side-stream work on the caller's stream meanwhile, so each
    # pair's events time that pair alone, not pairs overlapping other work)
    ctx.nttProbeBegin()
    t1 = time.perf_counter()
    evaluate()
    torch.cuda.synchronize()
    probe_s = time.perf_counter() - t1
    ntt_ms, pairs, ntt_bytes, ntt_limbs = ctx.nttProbeEnd()
    roofline = aes_ntt_roofline(M, torch, ctx.getParams(), (ntt_ms, pairs, ntt_bytes, ntt_limbs, probe_s))
    plain_s, plain = None, None
    if args.plain:
        # the plain form: the same evaluation one ciphertext at a time (the
        # S-boxes one after another on one stream), limb for limb the same
        t1 = time.perf_counter()
        plain = H.ckks_aes_ctr_walsh_sbox_complex_blocks(ctx, keys, counters, cts, batch_sboxes=False)
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t1
        same = all(np.array_equal(np.asarray(plain[b].getPoly(k)), np.asarray(out[b].getPoly(k)))
                   for b in range(128) for k in range(2))
        if not same:
            raise SystemExit("batched and one-at-a-time S-box evaluations differ")
    res = {
        "metric": "walsh15 AES-CTR blocks/s (homomorphic CTR decryption, mask-packed Walsh S-box)",
        "value": 2 * cb * 1e3 / ms, "unit": "blocks/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 (exact mod-q integer)", "data": "synthetic (the reference bench's counters and plaintexts)",
        "config": {"workload": "walsh15 AES-CTR, 1024 blocks per evaluation", "blocks": 2 * cb,
                   "mismatches": mism, "max_error": max_err, "average_error": avg_err,
                   "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                   "form": "batched (16 S-boxes per launch sequence, one context)" if batched
                   else f"{len(ctxs)} S-box contexts on host threads",
                   "plain_blocks_per_s": None if plain_s is None else 2 * cb / plain_s,
                   "plain_bit_exact": None if plain_s is None else True,
                   "context_and_key_setup_s": setup_s,
                   "tuning": json.loads(ctx.tuneReport())},
        "roofline": roofline,
        "cpu_baseline": cpu_baseline_aes(H, ctx.getParams(), 2 * cb) if args.cpu_baseline else None,
    }
    if args.serve_requests > 0:
        del out, plain
        torch.cuda.synchronize()
        res["config"]["served"] = serve_leg(H, args.serve_requests, args.rate, args.lanes)
    return res


def serve_leg(H, requests, rate, lanes, seed=55, plain=False):
    """Served walsh15 AES-CTR through the native continuous-batching server
    (aes-walsh-ckks_amd/aes_serve.py over host/ckks_aes_serve.cpp): Poisson
    arrivals at `rate` requests/s from `seed`, each request 1-32 blocks of
    random plaintext under the test key, submitted at its arrival time;
    blocks/s over the run (first arrival to last result) and p50/p99 of
    arrival -> decrypted result.  `plain`: the first three requests again on
    a one-lane server, one request per evaluation, for comparison."""
    sys.path.insert(0, os.path.join(ROOT, "aes-walsh-ckks_amd"))
    import aes_serve
    rng = np.random.default_rng(seed)
    base = int.from_bytes(bytes.fromhex("f0f1f2f3f4f5f6f7f8f9fafbfcfdfeff"), "big")
    reqs, first = [], 0
    for _ in range(requests):
        n = int(rng.integers(1, 33))
        c = np.array([list((base + first + i).to_bytes(16, "big")) for i in range(n)], np.uint8)
        p = rng.integers(0, 256, (n, 16), dtype=np.uint8)
        x = np.array([H.aes128_ctr_crypt_block(list(AES_KEY), list(a), list(b)) for a, b in zip(c, p)], np.uint8)
        reqs.append((c, x, p))
        first += n
    gaps = rng.exponential(1.0 / rate, requests)
    t0 = time.perf_counter()
    srv = aes_serve.AesCtrServer(AES_KEY, lanes=lanes)
    setup_s = time.perf_counter() - t0
    print(f"# server: {srv.lanes} lanes, {srv.evaluation_bytes / 2**30:.2f} GiB per evaluation, "
          f"set-up {setup_s:.1f} s", file=sys.stderr, flush=True)
    ids = []
    start = time.perf_counter()
    due = start
    for (c, x, _), g in zip(reqs, gaps):
        due += g
        while time.perf_counter() < due:
            time.sleep(min(0.005, max(0.0, due - time.perf_counter())))
        ids.append(srv.submit(c, x))
    res = [srv.result(i) for i in ids]
    evaluations, lane_count = srv.evaluations, srv.lanes
    srv.close()
    bad = [r for r in res if r["refused"] or r["error"]]
    if bad:
        raise RuntimeError(f"served request {bad[0]['id']} failed: {bad[0]['error']}")
    mism = sum(int(np.sum(np.any(r["plaintexts"] != p, axis=1))) for r, (_, _, p) in zip(res, reqs))
    blocks = sum(len(c) for c, _, _ in reqs)
    t_first = min(r["arrival"] for r in res)
    t_last = max(r["finish"] for r in res)
    lat = np.array([r["finish"] - r["arrival"] for r in res])
    out = {"requests": requests, "rate_per_s": rate, "blocks": blocks, "mismatched_blocks": mism,
           "blocks_per_s": blocks / (t_last - t_first), "wall_s": t_last - t_first,
           "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
           "evaluations": evaluations, "lanes": lane_count, "server_setup_s": setup_s}
    if plain:
        one = aes_serve.AesCtrServer(AES_KEY, lanes=1)
        t1 = time.perf_counter()
        nb = 0
        for c, x, p in reqs[:3]:
            r = one.result(one.submit(c, x))
            nb += len(c)
            if r["refused"] or r["error"] or not np.array_equal(r["plaintexts"], p):
                raise RuntimeError("plain serving decrypted a wrong block")
        out["plain_one_request_per_evaluation_blocks_per_s"] = nb / (time.perf_counter() - t1)
        one.close()
    return out


def bench_serve(args, M, torch):
    """Served walsh15 AES-CTR (serve_leg) with --requests requests at --rate
    requests/s over --lanes lanes; value = served blocks/s."""
    H = _host()
    r = serve_leg(H, args.requests, args.rate, args.lanes, plain=args.plain)
    return {
        "metric": "served walsh15 AES-CTR blocks/s with p50/p99 latency (continuous batching)",
        "value": r["blocks_per_s"], "unit": "blocks/s", "n_gpus": 1, "steps": args.requests, "warmup": 0,
        "ms_per_step": 1e3 * r["wall_s"] / args.requests, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64 (exact mod-q integer)",
        "data": "synthetic (Poisson arrivals, random plaintexts, seed 55)",
        "config": dict({"workload": "served walsh15 AES-CTR, requests of 1-32 blocks"}, **r),
        "roofline": None, "cpu_baseline": None,
    }


def bench_ntt_qp(args, M, torch):
    """The AES workload's roofline kernel alone: the forward NTT of one QP
    ciphertext at walsh15 (2 polys x 17 limbs, N=2^15), for counter passes."""
    H = _host()
    r = aes_ntt_roofline(M, torch, H.CKKSParams.AESWalsh15(), (1.0, 1.0, 1.0, 1.0, 1.0))
    us = r["isolated_qp_pair_us"]
    N = 1 << 15
    r.update({"achieved": 16 * N * 34 / (us / 1e6) / 1e9, "frac": r["isolated_qp_pair_frac"],
              "traffic": _profile_traffic("ntt15_qp_forward"), "traffic_over_algorithmic": None, "valu_busy": None,
              "kernel": "forward NTT (column + row pass), N=2^15, 2 polys x 17 limbs (QP ciphertext)"})
    for k in ("launch_pairs", "limbs", "avg_pair_us", "avg_limbs_per_pair", "ntt_event_ms", "evaluation_ms"):
        r.pop(k)
    return {"metric": "forward NTT launch pair, walsh15 QP ciphertext (2 polys x 17 limbs, N=2^15)",
            "value": 1e6 / us, "unit": "launch-pairs/s", "n_gpus": 1, "steps": 50, "warmup": 3,
            "ms_per_step": us / 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64 (exact mod-q integer)", "data": "synthetic (seeded residues)",
            "config": {"workload": "forward NTT, N=2^15, 34 limbs"}, "roofline": r, "cpu_baseline": None}


def aes_ntt_roofline(M, torch, P, probe):
    """The NTT passes are the largest kernel family of the AES evaluation
    (about half of its kernel time; the rest is spread over the Walsh product
    sums, key-switch inner products, element-wise ops and copies).  `probe`
    is what the context's NTT probe measured over one whole evaluation, run
    with the side streams off so that no pair overlaps other work: HIP
    events on each launch pair's own stream around every forward and inverse
    NTT (column + row pass; standalone, inside key switches, rescales,
    ModRaise and the linear transforms), 16 N bytes per limb algorithmically
    (each limb read and written once).  achieved = algorithmic bytes of all
    pairs / their summed event time.  For comparison the forward NTT of one
    QP ciphertext at the top level alone (2 polys x 17 limbs, N=2^15, the
    key switches' ModUp / ModDown shape) is timed the same way."""
    ntt_ms, pairs, ntt_bytes, ntt_limbs, probe_s = probe
    N = P.getN()
    q, p = [int(x) for x in P.getModuli()], [int(x) for x in P.getPModuli()]
    ctx = M.Context(N, q, p, P.getDnum())
    limbs = len(q) + len(p)
    data = _rand_limbs(torch, q + p, N, 2, seed=77)
    stream = torch.cuda.current_stream()
    for _ in range(3):
        ctx.ntt_forward(data, 2, limbs)
    reps = 50
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        ctx.ntt_forward(data, 2, limbs)
    e1.record(stream)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    achieved = ntt_bytes / (ntt_ms / 1e3)
    # counters of the same launches (profiles/aes_walsh15_counters.md): HBM
    # bytes per pair as a ratio to the algorithmic bytes, VALU busy fraction
    ratio = _profile_traffic("ntt15_aes_inrun_ratio")
    avg_bytes = ntt_bytes / max(1.0, pairs)
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": None if ratio is None else ratio * avg_bytes,
            "traffic_over_algorithmic": ratio, "valu_busy": _profile_traffic("ntt15_aes_inrun_valu_busy"),
            "kernel": "NTT launch pairs (column + row pass) inside one walsh15 AES-CTR evaluation, N=2^15",
            "launch_pairs": int(pairs), "limbs": int(ntt_limbs),
            "avg_pair_us": 1e3 * ntt_ms / max(1.0, pairs), "avg_limbs_per_pair": ntt_limbs / max(1.0, pairs),
            "ntt_event_ms": ntt_ms, "evaluation_ms": 1e3 * probe_s,
            "isolated_qp_pair_us": us, "isolated_qp_pair_frac": 16 * N * 2 * limbs / (us / 1e6) / HBM_PEAK}


def cpu_baseline_aes(H, P, blocks):
    """The reference's own CPU path, built from its sources by oracle/Makefile
    (oracle/_ref/ref_harness, mode sbox15time): whole walsh15 S-boxes (parity
    packing, the complex-packed bootstrap with its degree-100 EvalMod, hoisted
    unpacking, Walsh product sums, real/imaginary split) through the
    reference's evaluator, `threads` of them at once, one per OpenMP thread,
    as its sub_bytes runs the S-boxes of a round (ckks_aes_ctr_walsh_sbox.cpp:
    659-721).  One 1024-block evaluation holds 160 S-boxes and 4 output
    refreshes (bootstraps, counted as S-boxes), so blocks/s = blocks x
    threads / (164 x wall).  The EvalMod coefficients come from this
    package's restated fit (the reference fits them with MPFR, absent here).
    Without the built harness: the oracle's Walsh product sum alone."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    threads = max(1, min(16, os.cpu_count() or 1))
    if os.path.exists(harness):
        import subprocess
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            coeffs = os.path.join(d, "evalmod_coeffs.bin")
            np.asarray(H.default_evalmod_polynomial(P), np.float64).tofile(coeffs)
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            r = subprocess.run([harness, "sbox15time", coeffs, str(threads)], capture_output=True, text=True,
                               timeout=600, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"reference S-box timing failed: {r.stderr[-2000:]}")
        t = json.loads(r.stdout.strip().splitlines()[-1])
        return {"value": blocks * t["concurrent_sboxes"] / (164 * t["wall_s"]), "unit": "blocks/s",
                "cores": t["omp_threads"], "kind": "reference",
                "sample": f"{t['concurrent_sboxes']} whole walsh15 S-boxes at once on {t['omp_threads']} OpenMP "
                          f"threads through the reference's evaluator: {t['wall_s']:.2f} s wall "
                          f"({t['sbox_mean_s']:.2f} s per S-box: bootstrap linear {t['bootstrap_linear_s']:.2f}, "
                          f"EvalMod {t['evalmod_s']:.2f}, Walsh products {t['walsh_products_s']:.2f}, unpack "
                          f"{t['unpack_s']:.2f}, packing {t['parity_pack_s']:.2f}, split {t['split_s']:.2f}); "
                          "scaled by the 164 S-box-sized steps of one 1024-block evaluation"}
    N = P.getN()
    q, p = [int(x) for x in P.getModuli()], [int(x) for x in P.getPModuli()]
    table = np.array(H.ckks_aes_walsh_table(), dtype=np.int64)
    cr, ci = table[0].T.copy(), table[4].T.copy()
    scale = 2.0 ** 38
    consts = [(int(round(float(cr[h][0]) * scale)), int(round(float(ci[h][0]) * scale))) for h in range(16)]
    pair = cpu_baseline_walsh(N, q, p, 2, (cr, ci, consts))
    return {"value": pair["value"] * blocks / 640, "unit": "blocks/s (upper bound)", "cores": pair["cores"],
            "kind": "port",
            "sample": f"oracle product sum of one Walsh output pair at the walsh15 shape ({1 / pair['value']:.2f} s); "
                      "scaled by the 640 pairs of one 1024-block evaluation, bootstraps left out"}


def bench_chain(args, M, torch):
    """The plain CKKS primitive chain through the C++ host API: N=2^14, 8 chain
    limbs of 50-bit primes, P = 2 x 50-bit, dnum 4, Hamming weight 192,
    sigma 3.2; 2^13 complex slots uniform in [-1, 1].  One step: encrypt two
    ciphertexts at level 7, HMult, relinearise, rescale, rotate by 1,
    decrypt (host encode/decode included)."""
    H = _host()
    c = H.CKKSParamsConfig()
    c.N = 1 << 14
    c.mult_depth = 7
    c.q0_bit = 50
    c.q_bit = 50
    c.p_bits = [50, 50]
    c.dnum = 4
    c.secret_key_hamming_weight = 192
    c.sigma = 3.2
    c.secure = False
    c.rotate_values = [1]
    c.seed = 11
    ctx = H.CKKSContext(H.CKKSParams(c))
    N, lv = 1 << 14, 7
    rng = np.random.default_rng(2024)
    m1 = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    m2 = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    want = np.roll(m1 * m2, -1)

    def evaluate(ct1, ct2):
        H.ckks_mul_inplace(ctx, ct1, ct2)
        H.ckks_relin_hybrid_inplace(ctx, ct1)
        H.ckks_rescale(ctx, ct1)
        H.ckks_rotate_inplace(ctx, ct1, 1)

    def step():
        ct1 = H.ckks_encrypt(ctx, m1, lv)
        ct2 = H.ckks_encrypt(ctx, m2, lv)
        evaluate(ct1, ct2)
        return H.ckks_decrypt(ctx, ct1)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dec = step()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    err = float(np.max(np.abs(dec - want)))
    # the evaluation part alone (device-resident inputs)
    cts = [(H.ckks_encrypt(ctx, m1, lv), H.ckks_encrypt(ctx, m2, lv)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in cts:
        evaluate(a, b)
    torch.cuda.synchronize()
    eval_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    cpu = cpu_baseline_chain(ctx, N) if args.cpu_baseline else None
    return {
        "metric": "CKKS primitive chain (encrypt x2, HMult, relinearise, rescale, rotate, decrypt), N=2^14",
        "value": 1e3 / ms, "unit": "chain/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 (exact mod-q integer); f64 encode/decode", "data": "synthetic (seeded slots)",
        "config": {"workload": "one chain per step, N=2^14, 8 + 2 limbs of 50 bits, dnum 4",
                   "evaluation_only_ms": eval_ms, "max_abs_error_vs_slot_product": err},
        "roofline": None, "cpu_baseline": cpu,
    }


def cpu_baseline_chain(ctx, N):
    """The oracle's tensor, relinearisation, rescale and rotation at the same
    shape with the context's keys, repeated for ~10 s (evaluation part)."""
    O = _oracle()
    P = ctx.getParams()
    q, p = np.array(P.getModuli(), np.uint64), np.array(P.getPModuli(), np.uint64)
    octx = O.Ctx(N, q, p, 4)
    rng = np.random.default_rng(3)
    lv = 7
    rnd = lambda: np.concatenate([rng.integers(0, int(m), N, dtype=np.uint64) for m in q[: lv + 1]])
    relin, rot1 = ctx.getRelinKey(), ctx.getRotationKey(1)
    b1, a1, b2, a2 = rnd(), rnd(), rnd(), rnd()
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 10.0:
        ob, oa, oc2 = octx.tensor(lv, b1, a1, b2, a2)
        ob, oa = octx.relin(lv, ob, oa, oc2, relin)
        r = octx.rescale(lv, np.stack([ob, oa]))
        octx.rotate(lv - 1, r[0], r[1], 1, rot1)
        done += 1
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "chain/s (evaluation part)", "cores": int(O.lib().oc_omp_threads()),
            "kind": "port", "sample": "oracle HMult + relinearise + rescale + rotate at N=2^14, repeated for 10 s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="aes", choices=["ntt", "keyswitch", "lt", "chain", "hoisted", "walsh", "sbox", "aes", "serve", "ntt_qp"])
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--requests", type=int, default=240, help="serve: number of requests")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="aes: requests of the short served leg after the timed region (0: none)")
    ap.add_argument("--lanes", type=int, default=2, help="serve: evaluation lanes (sharing one key set)")
    ap.add_argument("--rate", type=float, default=40.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--aes-contexts", type=int, default=1,
                    help="aes: 1 (default) evaluates the 16 S-boxes of a round as one ciphertext batch; more "
                         "spreads them over that many contexts (host threads and streams) one at a time")
    ap.add_argument("--no-plain", dest="plain", action="store_false",
                    help="skip the one-limb-per-launch comparison (profiling runs)")
    ap.add_argument("--eager", action="store_true",
                    help="lt: time the eager batched call, no HIP graph (counter collection does not "
                         "follow graph-launched kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              "(request sharding across GPUs is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    M = _load("mickks", os.path.join(ROOT, "aes-walsh-ckks_amd", "mickks.py"))
    M.lib()
    if not torch.cuda.is_available():
        raise SystemExit("bench.py: no GPU visible")
    out = {"ntt": bench_ntt, "keyswitch": bench_keyswitch, "lt": bench_lt, "chain": bench_chain,
           "hoisted": bench_hoisted, "walsh": bench_walsh, "sbox": bench_sbox, "aes": bench_aes, "serve": bench_serve, "ntt_qp": bench_ntt_qp}[args.workload](args, M, torch)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
