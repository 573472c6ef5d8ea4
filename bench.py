#!/usr/bin/env python3
"""Benchmark of the batched Classic McEliece KEM engine on one MI355X.

Headline step = one batch of N seeded encapsulations (inputs resident in HBM) followed by one batch
of N decapsulations of the resulting ciphertexts, for one parameter set (default mceliece348864,
N = 65536, the reference's largest batch).  Steps run on two streams: the encapsulation of step i+1
overlaps the decapsulation of step i (the sampler's tail leaves most of the GPU to the decoder);
the same steps serialised on one stream are timed beside it.  Prints ONE JSON line: value = KEM
operations per second over the timed steps (N encaps + N decaps per step), with encaps/s and
decaps/s of each direction alone, per-kernel HIP-event
timings, the plain one-launch-per-stage GPU form on the same inputs, a roofline object for the
dominant kernel and an all-core CPU baseline (the C oracle, bounded sample).

Unless --quick, the line also carries the other workloads of the scope (key "suite"):
  syndrome  mceliece348864 syndrome alone (PKE encryption of given weight-64 errors), N = 8192 and
            65536: INT8-MFMA 0/1-byte GEMM vs bit-packed AND+popcount VALU GEMM vs sparse column
            gather, with the MFMA fraction of the INT8 peak;
  per_set   every parameter set at N = 65536: engine vs plain GPU form vs CPU oracle, enc and dec;
  c1        mceliece348864 KEM at N = 256 (latency of one batch);
  c3        mceliece460896 and mceliece6688128 decapsulation at N = 8192;
  c4        mceliece6960119 uneven sub-batches (sizes log-uniform in [1, 8192], total 65536) in one
            graph-captured multi-batch pipeline run per direction, autotuned launch shapes;
  c5        mceliece8192128 serving loop: 16 device-resident keypairs (GPU-batched generation),
            Poisson arrivals of requests of log-uniform size in [1, 4096], 50/50 enc/dec,
            HBM-footprint admission; throughput at saturation and p50/p99 latency at two loads;
  keygen    batched key generation (systematic form on the GPU) vs the host generator on all
            cores, mceliece348864 and mceliece8192128, byte-compared.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0  # MI355X dense INT8 MFMA
SETS = ["348864", "460896", "6688128", "6960119", "8192128"]


def xorshift_seeds(n, state=0x9E3779B97F4A7C15):
    """Seed bytes as the reference harness draws them (mceliece-gpu-params/main.cu:75-76,184-188)."""
    out = np.empty(n * 32, np.uint8)
    x = state
    M = (1 << 64) - 1
    for i in range(n * 32):
        x ^= (x << 13) & M
        x ^= x >> 7
        x ^= (x << 17) & M
        out[i] = x & 0xFF
    return out.reshape(n, 32), x


def algorithmic_bytes(kernel, p, n):
    """HBM bytes one launch must move at minimum (inputs once, outputs once), per kernel."""
    ct, e8 = p["ct"], p["n"] // 8
    msg = ((1 + e8 + ct) // 136 + 1) * 136
    splanes = 4 * p["m"] * ((2 * p["t"] + 31) // 32)
    per_op = {
        "gen_e": 32 + 2 * p["t"] + msg,           # seed in; support list + message row out
        "enc_finish": 2 * p["t"] + msg + ct + 32,  # support + message in; ct + key out
        "dec_front": ct + splanes,
        "dec_front_mfma": ct + splanes,
        "dec_decode": splanes + ct + msg,
        "dec_hash": msg + ct + 32 + 4,
    }[kernel]
    return per_op * n


def timed_kernels(tm, fn, reps):
    tm.sync()
    tm.timing_enable(True)
    for _ in range(reps):
        fn()
    tm.sync()
    kt = tm.timing_read()
    tm.timing_enable(False)
    return {k: {"launches": c, "avg_ms": tot / c} for k, (c, tot) in kt.items() if c}


def wall(tm, fn, reps):
    fn()
    tm.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    tm.sync()
    return (time.perf_counter() - t0) / reps


def section_keygen(tm, cores, log):
    """Batched keypairs (systematic form on the GPU, the rest on host threads) against the host
    generator on all cores, same seeds, byte-compared."""
    out = {}
    for s, n in (("348864", 128), ("8192128", 32)):
        seeds, _ = xorshift_seeds(n, 0x6B65)
        tm.keypair_batch(s, seeds[:2])  # warm-up: first launch, buffers
        t0 = time.perf_counter()
        pk, sk = tm.keypair_batch(s, seeds)
        tg = time.perf_counter() - t0
        nh = min(n, cores)
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=cores) as ex:
            host = list(ex.map(lambda i: tm.keypair_seeded(s, seeds[i].tobytes()), range(nh)))
        th = time.perf_counter() - t0
        exact = all(np.array_equal(pk[i], host[i][0]) and np.array_equal(sk[i], host[i][1]) for i in range(nh))
        out[s] = {"keys": n, "batch_keypairs_per_s": round(n / tg, 1), "host_keypairs_per_s": round(nh / th, 1),
                  "host_threads": cores, "host_sample_keys": nh, "speedup": round((n / tg) / (nh / th), 2),
                  "bit_exact_vs_host": exact}
        log(f"keygen {s}: {out[s]}")
    return out


def fixed_weight(p, n, rng):
    e = np.zeros((n, p["n"] // 8), np.uint8)
    for c0 in range(0, n, 4096):
        c1 = min(n, c0 + 4096)
        pos = np.argpartition(rng.random((c1 - c0, p["n"]), dtype=np.float32), p["t"], axis=1)[:, :p["t"]]
        rows = np.repeat(np.arange(c0, c1), p["t"])
        np.bitwise_or.at(e, (rows, pos.ravel() >> 3), (1 << (pos.ravel() & 7)).astype(np.uint8))
    return e


class DevKem:
    """Device-resident inputs/outputs of one set and batch size, engine and plain form timings."""

    def __init__(self, tm, s, n, keypair, seeds):
        self.tm, self.s, self.n = tm, s, n
        self.p = tm.PARAMS[s]
        self.pk, self.sk = keypair
        self.PK, self.SK = tm.PublicKey(s, self.pk), tm.SecretKey(s, self.sk)
        self.seeds = seeds
        self.d_seeds = tm.DeviceBuffer(n * 32)
        self.d_ct = tm.DeviceBuffer(n * self.p["ct"])
        self.d_k = tm.DeviceBuffer(n * 32)
        self.d_st = tm.DeviceBuffer(n * 4)
        self.d_k2 = tm.DeviceBuffer(n * 32)
        self.d_st2 = tm.DeviceBuffer(n * 4)
        self.d_seeds.upload(seeds)

    def enc(self):
        self.tm.enc_device(self.PK, self.d_ct, self.d_k, self.d_st, self.d_seeds, self.n)

    def dec(self):
        self.tm.dec_device(self.SK, self.d_k2, self.d_st2, self.d_ct, self.n)

    def rates(self, reps):
        te, td = wall(self.tm, self.enc, reps), wall(self.tm, self.dec, reps)
        return self.n / te, self.n / td

    def outputs(self):
        n, p = self.n, self.p
        return (self.d_ct.download((n, p["ct"])), self.d_k.download((n, 32)), self.d_k2.download((n, 32)),
                self.d_st.download(n, np.int32), self.d_st2.download(n, np.int32))

    def plain(self, reps):
        tm = self.tm
        tm.set_form("plain")
        try:
            e, d = self.rates(reps)
            kern = timed_kernels(tm, lambda: (self.enc(), self.dec()), 1)
            out = self.outputs()
        finally:
            tm.set_form("engine")
        return e, d, kern, out

    def free(self):
        for b in (self.d_seeds, self.d_ct, self.d_k, self.d_st, self.d_k2, self.d_st2):
            b.free()


class OverlapKem:
    """The headline step on two streams: the encapsulation of step i + 1 (its own stream) runs
    concurrently with the decapsulation of step i (another stream).  Ciphertext and key buffers
    are double-buffered; events order each decapsulation after its encapsulation and each
    encapsulation after the decapsulation that last read its ciphertext buffer.  The engine keeps
    separate scratch for the two directions, so the two streams share nothing else."""

    def __init__(self, dk):
        tm, n, ctb = dk.tm, dk.n, dk.p["ct"]
        self.dk, self.tm = dk, tm
        self.se, self.sd = tm.Stream(), tm.Stream()
        self.ct = [tm.DeviceBuffer(n * ctb) for _ in range(2)]
        self.k = [tm.DeviceBuffer(n * 32) for _ in range(2)]
        self.st = [tm.DeviceBuffer(n * 4) for _ in range(2)]
        self.k2 = [tm.DeviceBuffer(n * 32) for _ in range(2)]
        self.st2 = [tm.DeviceBuffer(n * 4) for _ in range(2)]
        self.ev_enc = [tm.Event() for _ in range(2)]
        self.ev_dec = [tm.Event() for _ in range(2)]
        self.i = 0

    def step(self):
        dk, tm, b = self.dk, self.tm, self.i & 1
        if self.i >= 2:
            self.se.wait(self.ev_dec[b])  # decapsulation i - 2 has read ct[b]
        tm.enc_device(dk.PK, self.ct[b], self.k[b], self.st[b], dk.d_seeds, dk.n, stream=self.se)
        self.ev_enc[b].record(self.se)
        self.sd.wait(self.ev_enc[b])
        tm.dec_device(dk.SK, self.k2[b], self.st2[b], self.ct[b], dk.n, stream=self.sd)
        self.ev_dec[b].record(self.sd)
        self.i += 1

    def sync(self):
        self.se.sync()
        self.sd.sync()

    def outputs(self):
        """(ct, k enc, k dec) of the last step"""
        n, p, b = self.dk.n, self.dk.p, (self.i - 1) & 1
        return (self.ct[b].download((n, p["ct"])), self.k[b].download((n, 32)), self.k2[b].download((n, 32)))

    def free(self):
        self.sync()
        for bufs in (self.ct, self.k, self.st, self.k2, self.st2):
            for x in bufs:
                x.free()
        for e in self.ev_enc + self.ev_dec:
            e.free()
        self.se.free()
        self.sd.free()


def cpu_rates(orc, s, pk, sk, seeds, cores, seconds):
    """Oracle enc and dec on all host cores over a bounded sample (about `seconds` in total)."""
    ne = max(cores, 32)
    t0 = time.perf_counter()
    ct_c, k_c, _ = orc.enc_batch(s, pk, seeds[:ne], cores)
    te = time.perf_counter() - t0
    ne2 = int(min(len(seeds), max(ne, ne * (seconds / 2) / max(te, 1e-3))))
    t0 = time.perf_counter()
    ct_c, k_c, _ = orc.enc_batch(s, pk, seeds[:ne2], cores)
    te = time.perf_counter() - t0
    nd = max(cores, 32)
    t0 = time.perf_counter()
    orc.dec_batch(s, sk, ct_c[:nd], cores)
    td = time.perf_counter() - t0
    nd2 = int(min(ne2, max(nd, nd * (seconds / 2) / max(td, 1e-3))))
    t0 = time.perf_counter()
    k_d, _ = orc.dec_batch(s, sk, ct_c[:nd2], cores)
    td = time.perf_counter() - t0
    assert np.array_equal(k_d, k_c[:nd2])
    return ne2 / te, nd2 / td, ne2, nd2, ct_c, k_c


# ------------------------------------------------------------------------------- suite sections
def section_syndrome(tm, orc, log):
    s = "348864"
    p = tm.PARAMS[s]
    pk, _ = tm.keypair_seeded(s, bytes(range(32)))
    K = tm.PublicKey(s, pk)
    out = {}
    rng = np.random.default_rng(4)
    for n in (8192, 65536):
        e = fixed_weight(p, n, rng)
        d_e, d_ct = tm.DeviceBuffer(e.nbytes), tm.DeviceBuffer(n * p["ct"])
        d_e.upload(e)
        ops = 2 * p["m"] * p["t"] * (p["n"] - p["m"] * p["t"]) * n  # int8 MAC-ops of the dense product
        res, outs = {}, {}
        for meth in ("sparse", "valu"):
            kt = timed_kernels(tm, lambda: tm.syndrome_device(K, d_ct, d_e, n, meth), 5)
            res[meth] = sum(v["avg_ms"] for v in kt.values())
            outs[meth] = d_ct.download((n, p["ct"]))
        best_cfg, best_ms = None, None
        for cfg in range(6):
            tm.tune_set(s, "synd_mfma_cfg", cfg)
            kt = timed_kernels(tm, lambda: tm.syndrome_device(K, d_ct, d_e, n, "mfma"), 5)
            ms = sum(v["avg_ms"] for v in kt.values())
            if best_ms is None or ms < best_ms:
                best_cfg, best_ms = cfg, ms
        tm.tune_set(s, "synd_mfma_cfg", best_cfg)
        tm.syndrome_device(K, d_ct, d_e, n, "mfma")
        tm.sync()
        res["mfma"] = best_ms
        outs["mfma"] = d_ct.download((n, p["ct"]))
        same = all(np.array_equal(outs["sparse"], outs[m]) for m in ("mfma", "valu"))
        idx = rng.choice(n, 32, replace=False)
        same = same and bool(np.array_equal(outs["mfma"][idx], orc.syndrome_batch(s, pk, e[idx])))
        tops = ops / (best_ms * 1e-3) / 1e12
        out[str(n)] = {
            "ms": {k: round(v, 4) for k, v in res.items()},
            "columns_per_s": {k: round(n / (v * 1e-3), 1) for k, v in res.items()},
            "mfma_cfg": best_cfg, "bit_exact": same,
            "mfma_roofline": {"bound": "mfma", "achieved": round(tops, 1), "peak": INT8_PEAK_TOPS, "unit": "TOPS",
                              "frac": round(tops / INT8_PEAK_TOPS, 4),
                              "work": "2*768*2720 int8 ops per column (0/1 bytes, int32 accumulate)"},
            "valu_lane_ops_per_column": 768 * 85 * 3 // 2,
        }
        log(f"syndrome n={n}: {out[str(n)]['ms']}")
        d_e.free()
        d_ct.free()
    return out


def section_variants(tm, orc, log, n=65536, reps=3):
    """The pc variants (plaintext confirmation: one more SHAKE256 over e in each direction, 32 more
    ciphertext bytes) against their base sets at N=65536, and batched key generation of an f variant
    (semi-systematic form) against its base set.  Outputs round-trip and a sample is byte-compared
    with the oracle."""
    out = {}
    for base in ("348864", "8192128"):
        kseed, _ = xorshift_seeds(1, 0x7661)
        pk, sk = tm.keypair_seeded(base, kseed[0].tobytes())  # the same key serves the pc variant
        seeds = np.random.default_rng(int(base)).integers(0, 256, (n, 32), dtype=np.uint8)
        row = {}
        for v in (base, base + "pc"):
            dk = DevKem(tm, v, n, (pk, sk), seeds)
            ee, ed = dk.rates(reps)
            ct, k1, k2, st1, st2 = dk.outputs()
            idx = np.random.default_rng(2).choice(n, 8, replace=False)
            ct_o, k_o, _ = orc.enc_batch(v, pk, seeds[idx])
            exact = bool(np.array_equal(k1, k2) and not st1.any() and not st2.any()
                         and np.array_equal(ct[idx], ct_o) and np.array_equal(k1[idx], k_o))
            dk.free()
            row[v] = {"encaps_per_s": round(ee, 1), "decaps_per_s": round(ed, 1), "bit_exact": exact}
        row["pc_cost"] = {"encaps": round(row[base]["encaps_per_s"] / row[base + "pc"]["encaps_per_s"], 3),
                          "decaps": round(row[base]["decaps_per_s"] / row[base + "pc"]["decaps_per_s"], 3)}
        out[base] = row
        log(f"variants {base}: {row}")
    kg = {}
    for v in ("348864", "348864f"):
        seeds, _ = xorshift_seeds(64, 0x6B66)
        tm.keypair_batch(v, seeds[:2])
        t0 = time.perf_counter()
        pkb, skb = tm.keypair_batch(v, seeds)
        kg[v] = {"keys": 64, "batch_keypairs_per_s": round(64 / (time.perf_counter() - t0), 1)}
    pk_h, sk_h = tm.keypair_seeded("348864f", seeds[0].tobytes())
    kg["348864f"]["bit_exact_vs_host_first_key"] = bool(np.array_equal(pkb[0], pk_h) and np.array_equal(skb[0], sk_h))
    out["keygen"] = kg
    log(f"variants keygen: {kg}")
    return out


def section_per_set(tm, orc, keys, cores, cpu_seconds, log, n=65536, reps=3):
    out = {}
    for s in SETS:
        pk, sk = keys[s]
        seeds = np.random.default_rng(int(s)).integers(0, 256, (n, 32), dtype=np.uint8)
        dk = DevKem(tm, s, n, (pk, sk), seeds)
        ee, ed = dk.rates(reps)
        ct, k1, k2, st1, st2 = dk.outputs()
        pe, pd, _, pout = dk.plain(max(1, reps - 1))
        exact = bool(np.array_equal(k1, k2) and not st1.any() and not st2.any()
                     and all(np.array_equal(a, b) for a, b in zip((ct, k1, k2), pout[:3])))
        idx = np.random.default_rng(1).choice(n, 16, replace=False)
        ct_o, k_o, _ = orc.enc_batch(s, pk, seeds[idx])
        exact = exact and bool(np.array_equal(ct[idx], ct_o) and np.array_equal(k1[idx], k_o))
        ce, cd, ne, nd, _, _ = cpu_rates(orc, s, pk, sk, seeds, cores, cpu_seconds)
        dk.free()
        out[s] = {"engine": {"encaps_per_s": round(ee, 1), "decaps_per_s": round(ed, 1)},
                  "plain_gpu_form": {"encaps_per_s": round(pe, 1), "decaps_per_s": round(pd, 1)},
                  "cpu_oracle": {"encaps_per_s": round(ce, 1), "decaps_per_s": round(cd, 1), "cores": cores,
                                 "sample": f"{ne} enc + {nd} dec"},
                  "speedup_vs_plain": {"encaps": round(ee / pe, 2), "decaps": round(ed / pd, 2)},
                  "speedup_vs_cpu": {"encaps": round(ee / ce, 1), "decaps": round(ed / cd, 1)},
                  "bit_exact": exact}
        log(f"per_set {s}: {out[s]['engine']} plain {out[s]['plain_gpu_form']} cpu {out[s]['cpu_oracle']}")
    return out


def section_c1(tm, orc, keys, cores, log):
    s, n = "348864", 256
    pk, sk = keys[s]
    seeds = xorshift_seeds(n, 0xC1)[0]
    dk = DevKem(tm, s, n, (pk, sk), seeds)
    te = wall(tm, dk.enc, 20)
    td = wall(tm, dk.dec, 20)
    ct, k1, k2, st1, st2 = dk.outputs()
    tm.set_form("plain")
    try:
        pte, ptd = wall(tm, dk.enc, 10), wall(tm, dk.dec, 10)
    finally:
        tm.set_form("engine")
    ct_o, k_o, _ = orc.enc_batch(s, pk, seeds, cores)
    t0 = time.perf_counter()
    orc.enc_batch(s, pk, seeds, cores)
    cte = time.perf_counter() - t0
    t0 = time.perf_counter()
    k_d, _ = orc.dec_batch(s, sk, ct_o, cores)
    ctd = time.perf_counter() - t0
    exact = bool(np.array_equal(ct, ct_o) and np.array_equal(k1, k_o) and np.array_equal(k2, k_d)
                 and not st1.any() and not st2.any())
    dk.free()
    out = {"batch_ms": {"engine_enc": round(te * 1e3, 4), "engine_dec": round(td * 1e3, 4),
                        "plain_enc": round(pte * 1e3, 4), "plain_dec": round(ptd * 1e3, 4),
                        "cpu_enc": round(cte * 1e3, 3), "cpu_dec": round(ctd * 1e3, 3)},
           "cpu_cores": cores, "bit_exact_vs_oracle": exact}
    log(f"c1: {out}")
    return out


def section_c3(tm, orc, keys, log, n=8192):
    out = {}
    for s in ("460896", "6688128"):
        pk, sk = keys[s]
        seeds = np.random.default_rng(33).integers(0, 256, (n, 32), dtype=np.uint8)
        dk = DevKem(tm, s, n, (pk, sk), seeds)
        dk.enc()
        td = wall(tm, dk.dec, 5)
        kern = timed_kernels(tm, dk.dec, 3)
        ct, k1, k2, st1, st2 = dk.outputs()
        tm.set_form("plain")
        try:
            ptd = wall(tm, dk.dec, 3)
        finally:
            tm.set_form("engine")
        idx = np.random.default_rng(2).choice(n, 16, replace=False)
        k_o, _ = orc.dec_batch(s, sk, ct[idx])
        exact = bool(np.array_equal(k1, k2) and not st2.any() and np.array_equal(k2[idx], k_o))
        dk.free()
        out[s] = {"decaps_per_s": round(n / td, 1), "plain_decaps_per_s": round(n / ptd, 1),
                  "kernels_ms": {k: round(v["avg_ms"], 4) for k, v in kern.items()}, "bit_exact": exact}
        log(f"c3 {s}: {out[s]}")
    return out


def section_c4(tm, keys, log, total=65536):
    from tensormceliece_amd import autotune
    s = "6960119"
    p = tm.PARAMS[s]
    pk, sk = keys[s]
    best, _ = autotune.tune(s, total, reps=2, keypair=(pk, sk))
    PK, SK = tm.PublicKey(s, pk), tm.SecretKey(s, sk)
    rng = np.random.default_rng(44)
    sizes = []
    while sum(sizes) < total:
        v = int(np.exp(rng.uniform(0, np.log(8193))))
        sizes.append(max(1, min(8192, v, total - sum(sizes))))
    seeds = rng.integers(0, 256, (total, 32), dtype=np.uint8)
    d_seeds, d_ct = tm.DeviceBuffer(total * 32), tm.DeviceBuffer(total * p["ct"])
    d_k, d_st = tm.DeviceBuffer(total * 32), tm.DeviceBuffer(total * 4)
    d_k2, d_st2 = tm.DeviceBuffer(total * 32), tm.DeviceBuffer(total * 4)
    d_seeds.upload(seeds)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    enc_subs = [(PK, d_seeds.ptr + 32 * o, d_ct.ptr + p["ct"] * o, d_k.ptr + 32 * o, d_st.ptr + 4 * o, m)
                for o, m in zip(offs, sizes)]
    dec_subs = [(SK, d_ct.ptr + p["ct"] * o, None, d_k2.ptr + 32 * o, d_st2.ptr + 4 * o, m)
                for o, m in zip(offs, sizes)]
    ep = tm.Pipeline(s, "enc", total, max_sub=len(sizes), use_graph=True)
    dp = tm.Pipeline(s, "dec", total, max_sub=len(sizes), use_graph=True)

    def run(pipe, subs):
        pipe.run(subs)
        pipe.sync()

    for _ in range(2):  # eager run, then capture
        run(ep, enc_subs)
        run(dp, dec_subs)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        run(ep, enc_subs)
    te = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        run(dp, dec_subs)
    td = (time.perf_counter() - t0) / reps
    ct = d_ct.download((total, p["ct"]))
    k1, k2 = d_k.download((total, 32)), d_k2.download((total, 32))
    st1, st2 = d_st.download(total, np.int32), d_st2.download(total, np.int32)
    # the same ops as one flat engine batch, and as the plain per-stage form
    dk = DevKem(tm, s, total, (pk, sk), seeds)
    fe, fd = dk.rates(3)
    fct, fk1, fk2, _, _ = dk.outputs()
    pe, pd, _, _ = dk.plain(2)
    dk.free()
    exact = bool(np.array_equal(ct, fct) and np.array_equal(k1, fk1) and np.array_equal(k2, fk2)
                 and np.array_equal(k1, k2) and not st1.any() and not st2.any())
    for b in (d_seeds, d_ct, d_k, d_st, d_k2, d_st2):
        b.free()
    out = {"sub_batches": len(sizes), "total_ops": total, "tuned": best,
           "pipeline_graph": {"encaps_per_s": round(total / te, 1), "decaps_per_s": round(total / td, 1)},
           "flat_engine_batch": {"encaps_per_s": round(fe, 1), "decaps_per_s": round(fd, 1)},
           "plain_gpu_form": {"encaps_per_s": round(pe, 1), "decaps_per_s": round(pd, 1)},
           "speedup_vs_plain": {"encaps": round(total / te / pe, 2), "decaps": round(total / td / pd, 2)},
           "bit_exact_vs_flat_engine": exact}
    log(f"c4: {out}")
    return out


def verify_serving(wl, reqs, outputs, sample=32):
    """Outputs of a serving run: every decapsulation returns the session keys of the pool
    encapsulations its ciphertexts came from, with status 0; a sample of requests of each direction
    equals the same request run alone through the plain batched calls."""
    dec_ok = all(not st.any() and np.array_equal(k, wl.pool_k[key][off:off + inp.shape[0]])
                 for (d, key, inp, off), (_, k, st) in zip(reqs, outputs) if d == "dec")
    rng = np.random.default_rng(9)
    alone_ok, checked = True, 0
    for i in rng.choice(len(reqs), min(sample, len(reqs)), replace=False):
        d, key, inp, _ = reqs[i]
        ct, k, st = outputs[i]
        if d == "enc":
            ct_a, k_a, st_a = wl.pk[key].enc(inp)
            alone_ok &= bool(np.array_equal(ct, ct_a) and np.array_equal(k, k_a) and np.array_equal(st, st_a))
        else:
            k_a, st_a = wl.sk[key].dec(inp)
            alone_ok &= bool(np.array_equal(k, k_a) and np.array_equal(st, st_a))
        checked += 1
    n_dec = sum(r[0] == "dec" for r in reqs)
    return {"bit_exact": bool(dec_ok and alone_ok), "dec_requests_checked_vs_pool": n_dec,
            "requests_checked_vs_alone": checked}


def section_c5(tm, log, nkeys=16, n_requests=6000):
    from tensormceliece_amd import serving
    s = "8192128"
    seeds, _ = xorshift_seeds(nkeys, 0x51ED)
    pks, sks = tm.keypair_batch(s, seeds)  # GPU-batched key generation
    kps = list(zip(pks, sks))
    wl = serving.Workload(s, kps, pool_per_key=4096)
    server = tm.Server(s, slot_capacity=32768, slots_per_direction=2, hbm_budget=8 << 30)
    try:
        reqs = serving.make_requests(wl, n_requests, seed=5)
        serving.run_poisson(server, wl, reqs[:200], 0.0)  # warm-up (first runs, graph capture)
        sat = serving.run_poisson(server, wl, reqs, 0.0, keep_outputs=True)
        verified = verify_serving(wl, reqs, sat["outputs"])
        sat["outputs"] = None
        mean_ops = float(np.mean([r[2].shape[0] for r in reqs]))
        req_rate_sat = sat["ops_per_s"] / mean_ops
        loads = {}
        for frac in (0.5, 0.8):
            r = serving.run_poisson(server, wl, reqs, frac * req_rate_sat, keep_outputs=True)
            v = verify_serving(wl, reqs, r["outputs"], sample=8)
            loads[f"{int(frac * 100)}%"] = {"offered_requests_per_s": round(frac * req_rate_sat, 1),
                                            "ops_per_s": round(r["ops_per_s"], 1), "p50_ms": round(r["p50_ms"], 3),
                                            "p99_ms": round(r["p99_ms"], 3), "bit_exact": v["bit_exact"]}
        # the same requests with inputs and outputs in pinned slabs: served in place (zero-copy)
        zreqs, inb, outb = serving.to_slab(reqs, tm.PARAMS[s])
        zsat = serving.run_poisson(server, wl, zreqs, 0.0, slab=outb)
        z80 = serving.run_poisson(server, wl, zreqs, 0.8 * req_rate_sat, slab=outb)
        zero_copy = {"saturation": {"ops_per_s": round(zsat["ops_per_s"], 1), "p50_ms": round(zsat["p50_ms"], 3),
                                    "p99_ms": round(zsat["p99_ms"], 3)},
                     "poisson_80%": {"ops_per_s": round(z80["ops_per_s"], 1), "p50_ms": round(z80["p50_ms"], 3),
                                     "p99_ms": round(z80["p99_ms"], 3)}}
        st = server.stats()
        zero_copy["zero_copy_requests"] = st["zero_copy_requests"]
        inb.free()
        outb.free()
    finally:
        server.close()
    out = {"keys": nkeys, "requests_per_run": n_requests, "mean_request_ops": round(mean_ops, 1),
           "key_device_bytes": wl.device_bytes(),
           "saturation": {"ops_per_s": round(sat["ops_per_s"], 1), "p50_ms": round(sat["p50_ms"], 3),
                          "p99_ms": round(sat["p99_ms"], 3)},
           "verified": verified,
           "poisson": loads, "zero_copy_slabs": zero_copy, "batches": st["batches"], "requests": st["requests"],
           "peak_in_flight_bytes": st["peak_in_flight_bytes"], "hbm_budget_bytes": 8 << 30}
    log(f"c5: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--set", default="348864")
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bounded CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-stage GPU form")
    ap.add_argument("--quick", action="store_true", help="headline only (no suite)")
    ap.add_argument("--sections", default="syndrome,per_set,c1,c3,c4,c5,keygen,variants")
    ap.add_argument("--one-stream", action="store_true",
                    help="headline steps serialised on one stream only (profiling: no concurrent kernels)")
    ap.add_argument("--step-gen-blocks", type=int, default=64,
                    help="sampler grid (workgroups) during the two-stream headline step: a small grid keeps "
                         "its lane groups busy from the work queue and leaves most SIMDs to the decoder")
    ap.add_argument("--tune", action="append", default=[], metavar="KNOB=VALUE",
                    help="set an engine launch knob for the headline set (tmce_tune_set), e.g. gen_blocks=128")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (calibrates profiler byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(every op is independent: G GPUs would each run this step on N/G ops)", file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    import tensormceliece_amd as tm
    from oracle import oracle as orc
    tm.init(0)
    s, n = args.set, args.n
    p = tm.PARAMS[s]
    cores = len(os.sched_getaffinity(0))

    for kv in args.tune:
        knob, val = kv.split("=")
        tm.tune_set(s, knob, int(val))
    kseed, st = xorshift_seeds(1)
    pk, sk = tm.keypair_seeded(s, kseed[0].tobytes())
    seeds, _ = xorshift_seeds(n, st)
    dk = DevKem(tm, s, n, (pk, sk), seeds)

    # ---- timed region: K steps of (N encaps + N decaps), device-resident inputs, with the
    #      encapsulation of step i + 1 concurrent with the decapsulation of step i (two streams)
    t_step, ov_out = None, None
    if not args.one_stream:
        ov = OverlapKem(dk)
        gen_blocks_default = tm.tune_get(s, "gen_blocks")
        if args.step_gen_blocks:
            tm.tune_set(s, "gen_blocks", args.step_gen_blocks)
        for _ in range(args.warmup):
            ov.step()
        ov.sync()
        tm.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ov.step()
        ov.sync()
        t_step = (time.perf_counter() - t0) / args.steps
        log(f"headline {s} n={n}: {t_step * 1e3:.3f} ms per step (two streams, sampler grid {args.step_gen_blocks})")
        ov_out = ov.outputs()
        ov.free()
        tm.tune_set(s, "gen_blocks", gen_blocks_default)

    # ---- the same steps serialised on one stream (encapsulation, then decapsulation)
    for _ in range(args.warmup):
        dk.enc()
        dk.dec()
    tm.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dk.enc()
        dk.dec()
    tm.sync()
    t_serial = (time.perf_counter() - t0) / args.steps
    log(f"headline {s} n={n}: {t_serial * 1e3:.3f} ms per step (one stream)")
    if t_step is None:
        t_step = t_serial

    # ---- per-phase and per-kernel timing (HIP events on the engine stream)
    reps = max(3, args.steps // 2)
    t_enc = wall(tm, dk.enc, reps)
    t_dec = wall(tm, dk.dec, reps)
    kernels = timed_kernels(tm, lambda: (dk.enc(), dk.dec()), reps)

    if args.calib_copy:
        nb = args.calib_copy << 20
        src, dst = tm.DeviceBuffer(nb), tm.DeviceBuffer(nb)
        dst.copy_from(src, nb)
        tm.sync()
        src.free()
        dst.free()

    ct, k1, k2, st1, st2 = dk.outputs()
    overlap_same = None if ov_out is None else all(np.array_equal(a, b) for a, b in zip((ct, k1, k2), ov_out))
    plain = None
    if not args.no_plain:
        pe, pd, pkern, pout = dk.plain(max(2, reps // 2))
        same = all(np.array_equal(a, b) for a, b in zip((ct, k1, k2), pout[:3]))
        plain = {"encaps_per_s": round(pe, 1), "decaps_per_s": round(pd, 1),
                 "kem_ops_per_s": round(2 / (1 / pe + 1 / pd), 1), "identical_to_engine": same,
                 "kernels": pkern}

    verified = None
    if not args.no_verify:
        ok = bool(np.array_equal(k1, k2) and not st1.any() and not st2.any())
        idx = np.random.default_rng(0).choice(n, min(n, 256), replace=False)
        ct_ref, k_ref, _ = orc.enc_batch(s, pk, seeds[idx])
        verified = ok and overlap_same is not False and bool(np.array_equal(ct[idx], ct_ref) and np.array_equal(k1[idx], k_ref))

    # ---- roofline of the dominant kernel (HBM; algorithmic bytes per launch / launch time)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    abytes = algorithmic_bytes(dom, p, n)
    achieved = abytes / (kernels[dom]["avg_ms"] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(tpath):  # measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (see that file)
        with open(tpath) as f:
            traffic = json.load(f).get(f"{s}:{n}", {}).get(dom)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": dom,
                "algorithmic_bytes_per_launch": abytes,
                "note": "the decoder is bound by its LDS table gathers (GF(2^m) log/exp lookups, bank "
                        "conflicts) and VALU issue, not by HBM; rocprofv3 SQ counters in the committed profile"}

    cpu = None
    if not args.no_cpu:
        ce, cd, ne, nd, _, _ = cpu_rates(orc, s, pk, sk, seeds, cores, args.cpu_seconds)
        cpu = {"value": round(2.0 / (1.0 / ce + 1.0 / cd), 1), "unit": "KEM ops/s", "cores": cores,
               "kind": "port", "encaps_per_s": round(ce, 1), "decaps_per_s": round(cd, 1),
               "sample": f"oracle C restatement, {ne} encaps + {nd} decaps of the same seeded {s} inputs, "
                         f"{cores} threads"}
    dk.free()

    suite = None
    if not args.quick:
        sections = args.sections.split(",")
        suite = {}
        keys = {}
        if {"per_set", "c1", "c3", "c4"} & set(sections):
            with ThreadPoolExecutor(max_workers=5) as ex:
                kp = list(ex.map(lambda x: tm.keypair_seeded(x, xorshift_seeds(1)[0][0].tobytes()), SETS))
            keys = dict(zip(SETS, kp))
        if "syndrome" in sections:
            suite["syndrome_348864"] = section_syndrome(tm, orc, log)
        if "per_set" in sections:
            suite["per_set_n65536"] = section_per_set(tm, orc, keys, cores, min(4.0, args.cpu_seconds), log)
        if "c1" in sections:
            suite["c1_348864_n256"] = section_c1(tm, orc, keys, cores, log)
        if "c3" in sections:
            suite["c3_decap_n8192"] = section_c3(tm, orc, keys, log)
        if "c4" in sections:
            suite["c4_6960119_uneven"] = section_c4(tm, keys, log)
        if "c5" in sections:
            suite["c5_8192128_serving"] = section_c5(tm, log)
        if "keygen" in sections:
            suite["keygen_batch"] = section_keygen(tm, cores, log)
        if "variants" in sections:
            suite["variants_n65536"] = section_variants(tm, orc, log)

    kem_ops_per_s = 2 * n / t_step
    line = {
        "metric": "encaps/s & decaps/s at N=65536 vs plain GPU form (bit-exact); serving p50/p99",
        "value": round(kem_ops_per_s, 1),
        "unit": "KEM ops/s (encaps+decaps)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "step_streams": "one stream" if args.one_stream else
                        "encapsulation of step i+1 concurrent with decapsulation of step i (two streams)",
        "step_sampler_grid": None if args.one_stream else args.step_gen_blocks,
        "ms_per_step_one_stream": round(t_serial * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8/u16 GF(2^m)", "data": "synthetic (seeded xorshift64 keypair and encapsulation seeds)",
        "config": {"workload": f"mceliece{s} KEM: {n} seeded encapsulations + {n} decapsulations per step",
                   "set": s, "batch": n},
        "encaps_per_s": round(n / t_enc, 1), "decaps_per_s": round(n / t_dec, 1),
        "bit_exact": verified,
        "two_stream_outputs_identical": overlap_same,
        "plain_gpu_form": plain,
        "speedup_vs_plain": None if plain is None else {
            "encaps": round((n / t_enc) / plain["encaps_per_s"], 2),
            "decaps": round((n / t_dec) / plain["decaps_per_s"], 2)},
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "suite": suite,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
