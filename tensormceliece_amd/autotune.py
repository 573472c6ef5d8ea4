"""Autotuner of the engine's launch shapes, per parameter set and batch size.

Knobs (tmce_tune_set): the MFMA tile shape of the syndrome GEMM and of the decapsulation front end,
whether the front end runs on MFMA or on the VALU, the sampler's form (one op per lane, or a lane
group per op with the SHAKE blocks pipelined) with its persistent grid, and the decoder's grid.  Each knob drives one kernel, so the tuner times that kernel alone (HIP events around its
launches, on device-resident inputs) for every candidate value and keeps the fastest; knobs are
independent of each other's kernels, so one pass over each suffices.  Results are cached in a JSON
file keyed by "<set>:<n>" and applied with apply().
"""
import json
import os

import numpy as np

import tensormceliece_amd as tm

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")

CANDIDATES = {
    "synd_mfma_cfg": list(range(8)),
    "front": [(0, 0)] + [(1, c) for c in range(8)],  # (front_mfma, front_cfg)
    # (gen_form, gen_blocks): the per-lane sampler keeps at most n/256 workgroups busy, the
    # pipelined one n * (lanes per op) / 256; larger grids are clamped to that
    "gen": [(0, 64), (0, 128), (0, 256), (0, 512), (1, 128), (1, 192), (1, 256), (1, 384), (1, 512), (1, 1024),
            (1, 2048)],
    "decode_blocks": [128, 256, 384, 512, 768, 1024],
}


def _kernel_ms(fn, names, reps):
    fn()
    tm.sync()
    tm.timing_enable(True)
    for _ in range(reps):
        fn()
    tm.sync()
    kt = tm.timing_read()
    tm.timing_enable(False)
    return sum(tot for k, (c, tot) in kt.items() if any(k.startswith(n) for n in names)) / reps


def tune(set_name, n=65536, reps=3, keypair=None, verbose=False):
    """Times every candidate of every knob for `set_name` at batch size `n`; returns the best values
    (also applied).  keypair: (pk, sk) bytes, or None for a seeded one."""
    p = tm.PARAMS[set_name]
    if keypair is None:
        keypair = tm.keypair_seeded(set_name, bytes(range(32)))
    pk, sk = keypair
    PK, SK = tm.PublicKey(set_name, pk), tm.SecretKey(set_name, sk)
    rng = np.random.default_rng(0)
    seeds = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    d_seeds, d_ct = tm.DeviceBuffer(n * 32), tm.DeviceBuffer(n * p["ct"])
    d_k, d_k2, d_e = tm.DeviceBuffer(n * 32), tm.DeviceBuffer(n * 32), tm.DeviceBuffer(n * p["n"] // 8)
    d_seeds.upload(seeds)
    e = np.zeros((n, p["n"] // 8), np.uint8)
    pos = rng.integers(0, p["n"], (n, p["t"]))  # weight <= t is enough for timing
    np.bitwise_or.at(e, (np.repeat(np.arange(n), p["t"]), (pos >> 3).ravel()), (1 << (pos & 7)).ravel().astype(np.uint8))
    d_e.upload(e)
    enc = lambda: tm.enc_device(PK, d_ct, d_k, None, d_seeds, n)
    dec = lambda: tm.dec_device(SK, d_k2, None, d_ct, n)
    enc()
    tm.sync()
    best, log = {}, {}

    def sweep(knob, values, setter, fn, names):
        times = {}
        for v in values:
            setter(v)
            times[v] = _kernel_ms(fn, names, reps)
        b = min(times, key=times.get)
        setter(b)
        log[knob] = {str(k): round(t, 4) for k, t in times.items()}
        if verbose:
            print(f"  {set_name} n={n} {knob}: best {b} ({times[b]:.3f} ms)", flush=True)
        return b

    best["synd_mfma_cfg"] = sweep("synd_mfma_cfg", CANDIDATES["synd_mfma_cfg"],
                                  lambda v: tm.tune_set(set_name, "synd_mfma_cfg", v),
                                  lambda: tm.syndrome_device(PK, d_ct, d_e, n, "mfma"), ["synd_mfma", "gf2_repack"])
    enc()  # restore valid ciphertexts after the syndrome runs overwrote d_ct
    tm.sync()

    def set_front(v):
        tm.tune_set(set_name, "front_mfma", v[0])
        tm.tune_set(set_name, "front_cfg", v[1])

    fb = sweep("front", CANDIDATES["front"], set_front, dec, ["dec_front", "gf2_repack"])
    best["front_mfma"], best["front_cfg"] = fb
    def set_gen(v):
        tm.tune_set(set_name, "gen_form", v[0])
        tm.tune_set(set_name, "gen_blocks", v[1])

    best["gen_form"], best["gen_blocks"] = sweep("gen", CANDIDATES["gen"], set_gen, enc, ["gen_e"])
    best["decode_blocks"] = sweep("decode_blocks", CANDIDATES["decode_blocks"],
                                  lambda v: tm.tune_set(set_name, "decode_blocks", v), dec, ["dec_decode"])
    for b in (d_seeds, d_ct, d_k, d_k2, d_e):
        b.free()
    return best, log


def apply(set_name, params):
    for k, v in params.items():
        tm.tune_set(set_name, k, int(v))


def load(path=CACHE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def apply_cached(set_name, n, path=CACHE):
    """Applies the cached knobs of the nearest tuned batch size for the set; returns them or None."""
    cache = load(path)
    cands = [(abs(np.log2(max(1, int(k.split(":")[1]))) - np.log2(max(1, n))), k) for k in cache
             if k.split(":")[0] == set_name]
    if not cands:
        return None
    params = cache[min(cands)[1]]["best"]
    apply(set_name, params)
    return params


def save(set_name, n, best, log, path=CACHE):
    cache = load(path)
    cache[f"{set_name}:{n}"] = {"best": {k: int(v) for k, v in best.items()}, "times_ms": log}
    with open(path, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="tune the engine's launch shapes and cache them")
    ap.add_argument("--sets", default="348864,460896,6688128,6960119,8192128")
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--out", default=CACHE)
    args = ap.parse_args()
    tm.init(0)
    for s in args.sets.split(","):
        best, log = tune(s, args.n, verbose=True)
        save(s, args.n, best, log, args.out)
        print(s, best, flush=True)
