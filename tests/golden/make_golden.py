"""Pack the reference harness output into the committed golden fixtures.

Run in a container where /root/reference exists:
    make -C oracle ref && oracle/_ref/ref_harness oracle/_ref/raw [case ...]
    python tests/golden/make_golden.py [--raw DIR] [group ...]

Groups (all when none is named): primes ntt eval lt ltq boot evalpoly bootseq.  The
harness draws keys and encryption masks from the reference's own
unseedable randomness, so repack only the groups whose raw arrays were
regenerated together.

The harness (oracle/ref_harness.cpp) runs the reference's own C++ code; this
script only repacks its raw arrays.  Inputs are stored in full; large outputs
are stored as SHA-256 digests of their little-endian uint64 bytes (tests hash
their own results and compare), small ones in full.
"""
import argparse
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RAW = os.path.join(HERE, "..", "..", "oracle", "_ref", "raw")


def rd(name, dtype=np.uint64):
    return np.fromfile(os.path.join(RAW, name + ".bin"), dtype=dtype)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def pack_primes():
    primes = {}
    for tag in ["walsh15", "xboot13", "xboot14", "xboot15", "chain14", "ks16"]:
        primes[tag] = {"q": [int(v) for v in rd(tag + "_q")], "p": [int(v) for v in rd(tag + "_p")]}
    primes["ntt16"] = {"q": [int(v) for v in rd("ntt16_q")], "p": []}
    with open(os.path.join(HERE, "primes.json"), "w") as f:
        json.dump(primes, f, indent=1)


def pack_ntt():
    ntt = {}
    for tag, N in [("ntt10", 1 << 10), ("ntt12", 1 << 12), ("ntt16one", 1 << 16)]:
        q = rd(tag + "_q")
        ent = {"N": N, "q": [int(v) for v in q], "seed_base": 1000, "limbs": []}
        for i in range(len(q)):
            e = {
                "fwd_out": digest(rd(f"{tag}_fwdout{i}")),
                "fwd_lazy": digest(rd(f"{tag}_fwdlazy{i}")),
                "inv_out": digest(rd(f"{tag}_invout{i}")),
            }
            if tag == "ntt10":
                for t in ["fwd", "fwds", "inv", "invs"]:
                    e["tab_" + t] = digest(rd(f"{tag}_{t}{i}"))
            ent["limbs"].append(e)
        ntt[tag] = ent
    with open(os.path.join(HERE, "ntt.json"), "w") as f:
        json.dump(ntt, f, indent=1)
    # one full set of tables/outputs for diagnostics
    np.savez_compressed(
        os.path.join(HERE, "ntt10_limb0.npz"),
        fwd=rd("ntt10_fwd0"), fwds=rd("ntt10_fwds0"), inv=rd("ntt10_inv0"), invs=rd("ntt10_invs0"),
        fwd_out=rd("ntt10_fwdout0"), inv_out=rd("ntt10_invout0"))


def pack_eval():
    outs = ["mul", "mul_relin", "mul_relin_rescale", "chain_rot1", "square", "rot1_out", "rot5_out",
            "conj_out", "rescale_ct1", "rescale2_ct1", "add", "sub", "mulint", "addconst", "mulplain",
            "rot1_level1"]
    for tag, keys in [("ev10", ["relin", "rot1", "rot5", "conj"]), ("ev10b", ["relin", "rot1"])]:
        arr = {"q": rd(tag + "_q"), "p": rd(tag + "_p"), "sk": rd(tag + "_sk"),
               "ct1": rd(tag + "_ct1"), "ct2": rd(tag + "_ct2"),
               "ct1_meta": rd(tag + "_ct1_meta"), "m1": rd(tag + "_m1", np.complex128),
               "m2": rd(tag + "_m2", np.complex128), "encode_m1": rd(tag + "_encode_m1"), "pt_m2": rd(tag + "_pt_m2"),
               "dec_ct1": rd(tag + "_dec_ct1", np.complex128)}
        for k in keys:
            arr["key_" + k] = rd(f"{tag}_{k}")
        meta = {}
        for o in outs:
            m = rd(f"{tag}_{o}_meta")
            meta[o] = {"digest": digest(rd(f"{tag}_{o}")), "npolys": int(m[0]), "level": int(m[1]),
                       "scale": float(m[2:3].view(np.float64)[0])}
        np.savez_compressed(os.path.join(HERE, tag + ".npz"), **arr)
        with open(os.path.join(HERE, tag + "_outputs.json"), "w") as f:
            json.dump(meta, f, indent=1)


def pack_lt():
    # hoisted switches and a QP BSGS stage
    q, p = rd("lt10_q"), rd("lt10_p")
    N = 1 << 10
    L, K = len(q), len(p)
    qp = np.concatenate([q, p])
    digits = rd("lt10_hoist_digits_lazy").reshape(-1, L + K, N) % qp[None, :, None]
    lt_meta = {
        "hoist_digits": digest(digits),
        "hoisted_rot1": digest(rd("lt10_hoisted_rot1")),
        "hoisted_qp_rot2": digest(rd("lt10_hoisted_qp_rot2")),
        "qp_from_key_term_rot4": digest(rd("lt10_qp_from_key_term_rot4")),
        "lt": digest(rd("lt10_lt")),
        "lt_rescaled": digest(rd("lt10_lt_rescaled")),
        "lt_scale": float(rd("lt10_lt_meta")[2:3].view(np.float64)[0]),
        "diag_seed_base": 5000, "baby_steps": 2, "diag_indices": list(range(6)),
    }
    np.savez_compressed(os.path.join(HERE, "lt10.npz"), q=q, p=p, ct=rd("lt10_ct"),
                        key_rot1=rd("lt10_rot1"), key_rot2=rd("lt10_rot2"), key_rot4=rd("lt10_rot4"))
    with open(os.path.join(HERE, "lt10_outputs.json"), "w") as f:
        json.dump(lt_meta, f, indent=1)


def pack_ltq():
    # Q-evaluator BSGS stages (owned stage through a plan, and the stage-view API)
    meta = {}
    for o in ["ltq", "bsgs"]:
        m = rd(f"lq10_{o}_meta")
        meta[o] = {"digest": digest(rd(f"lq10_{o}")), "npolys": int(m[0]), "level": int(m[1]),
                   "scale": float(m[2:3].view(np.float64)[0])}
    meta["ltq"].update({"diag_seed_base": 7000, "terms": [0, 1, 2, 3, 4, 5], "diag_indices": [0, 1, 2, 4, 5, 7],
                        "baby_steps": 3, "montgomery": True, "rescale": False})
    meta["bsgs"].update({"diag_seed_base": 7000, "terms": [10, 11, 12, 13], "diag_indices": [1, 3, 5, 6],
                         "baby_steps": 3, "montgomery": False, "rescale": True})
    meta["pt_scale"] = 2.0 ** 40
    arr = {"q": rd("lq10_q"), "p": rd("lq10_p"), "ct": rd("lq10_ct"), "ct_meta": rd("lq10_ct_meta")}
    for r in (1, 2, 3, 6):
        arr[f"key_rot{r}"] = rd(f"lq10_rot{r}")
    np.savez_compressed(os.path.join(HERE, "lq10.npz"), **arr)
    with open(os.path.join(HERE, "lq10_outputs.json"), "w") as f:
        json.dump(meta, f, indent=1)


def pack_boot():
    # bootstrap pieces: dense->sparse q0p0 switch, ModRaise, S2C plan digests
    meta = rd("bt10_s2c_meta")
    stages, i = [], 0
    while i < len(meta):
        n, bs, qpl = int(meta[i]), int(meta[i + 1]), int(meta[i + 2])
        idx = [int(v) for v in meta[i + 3:i + 3 + n]]
        stages.append({"baby_steps": bs, "qp_limbs": qpl, "diag_indices": idx,
                       "digest": digest(rd("bt10_s2c_stage%d" % len(stages)))})
        i += 3 + n
    bt_meta = {"q0p0": digest(rd("bt10_q0p0")), "modraise": digest(rd("bt10_modraise")),
               "modraise_level": int(rd("bt10_modraise_meta")[1]), "s2c_stages": stages}
    np.savez_compressed(os.path.join(HERE, "bt10.npz"), q=rd("bt10_q"), p=rd("bt10_p"),
                        ct0=rd("bt10_ct0"), key_d2s=rd("bt10_d2s"))
    with open(os.path.join(HERE, "bt10_outputs.json"), "w") as f:
        json.dump(bt_meta, f, indent=1)


def ct_meta(name):
    m = rd(name + "_meta")
    return {"digest": digest(rd(name)), "npolys": int(m[0]), "level": int(m[1]),
            "scale": float(m[2:3].view(np.float64)[0])}


def pack_evalpoly():
    # Chebyshev evaluation of six coefficient sets on one encrypted input
    arr = {"q": rd("ep10_q"), "p": rd("ep10_p"), "sk": rd("ep10_sk"), "key_relin": rd("ep10_relin"),
           "ct": rd("ep10_ct"), "m": rd("ep10_m", np.complex128)}
    meta = {"ct": {"level": int(rd("ep10_ct_meta")[1]),
                   "scale": float(rd("ep10_ct_meta")[2:3].view(np.float64)[0])}}
    for k in range(6):
        arr["coeffs%d" % k] = rd("ep10_coeffs%d" % k, np.float64)
        meta["out%d" % k] = ct_meta("ep10_out%d" % k)
    np.savez_compressed(os.path.join(HERE, "ep10.npz"), **arr)
    with open(os.path.join(HERE, "ep10_outputs.json"), "w") as f:
        json.dump(meta, f, indent=1)


def plan_meta(tag):
    meta = rd(tag + "_meta")
    stages, i = [], 0
    while i < len(meta):
        n, bs, qpl = int(meta[i]), int(meta[i + 1]), int(meta[i + 2])
        idx = [int(v) for v in meta[i + 3:i + 3 + n]]
        stages.append({"baby_steps": bs, "qp_limbs": qpl, "diag_indices": idx,
                       "digest": digest(rd(tag + "_stage%d" % len(stages)))})
        i += 3 + n
    return stages


def pack_bootseq():
    # the bootstrap driver's sequence (S2C, dense->sparse, ModRaise,
    # sparse->dense, C2S, conjugate split, Chebyshev EvalMod) with its keys
    rots = rd("bs10_rots")
    arr = {"q": rd("bs10_q"), "p": rd("bs10_p"), "rots": rots, "key_relin": rd("bs10_relin"),
           "key_conj": rd("bs10_conj"), "key_d2s": rd("bs10_d2s"), "key_s2d": rd("bs10_s2d"),
           "ct_in": rd("bs10_in"), "coeffs": rd("bs10_coeffs", np.float64), "scales": rd("bs10_scales", np.float64)}
    for r in rots:
        arr["key_rot%d" % r] = rd("bs10_rot%d" % r)
    meta = {"in": ct_meta("bs10_in"), "s2c_plan": plan_meta("bs10_s2c"), "c2s_plan": plan_meta("bs10_c2s")}
    for step in ["s2c_out", "d2s_out", "modraise_out", "s2d_out", "c2s_out", "split0", "split1", "out0", "out1"]:
        meta[step] = ct_meta("bs10_" + step)
    np.savez_compressed(os.path.join(HERE, "bs10.npz"), **arr)
    with open(os.path.join(HERE, "bs10_outputs.json"), "w") as f:
        json.dump(meta, f, indent=1)


GROUPS = {"primes": pack_primes, "ntt": pack_ntt, "eval": pack_eval, "lt": pack_lt, "ltq": pack_ltq,
          "boot": pack_boot,
          "evalpoly": pack_evalpoly, "bootseq": pack_bootseq}


def main():
    global RAW
    ap = argparse.ArgumentParser()
    ap.add_argument("--raw", default=RAW)
    ap.add_argument("groups", nargs="*", choices=sorted(GROUPS) + [[]], default=[])
    args = ap.parse_args()
    RAW = args.raw
    for g in args.groups or sorted(GROUPS):
        GROUPS[g]()
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
