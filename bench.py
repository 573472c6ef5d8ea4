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
CUS, CLOCK_GHZ = 256, 2.4
# VALU: one wave64 instruction per cycle per CU (4 SIMDs x 16 lanes, 4 cycles per wave64 instruction)
VALU_PEAK_G = CUS * CLOCK_GHZ  # G wave-instructions/s
# LDS: a conflict-free wave64 access of <= 32 bits per lane takes 2 cycles per CU (32 banks x 4 bytes)
LDS_PEAK_G = CUS * CLOCK_GHZ / 2  # G wave-instructions/s
SETS = ["348864", "460896", "6688128", "6960119", "8192128"]
COUNTERS_JSON = os.path.join(HERE, "profiles", "counters.json")


def load_counters(key):
    """Per-kernel rocprofv3 counters of one workload ("<set>:<n>"), from the committed profile
    summary (profiles/counters.json): instructions, LDS cycles and HBM bytes per launch."""
    if not os.path.exists(COUNTERS_JSON):
        return {}
    with open(COUNTERS_JSON) as f:
        return json.load(f).get(key, {})


def hbm_bytes_per_op(kernel, p):
    """Algorithmic HBM bytes per op: what a launch must read and write at minimum.  The decoder
    counts the decapsulation's own inputs and outputs, SYND_BYTES + 32 per op; the intermediate
    rows the engine passes between its launches are not algorithmic."""
    ct = p["ct"]
    return {"gen_e": 32, "enc_finish": ct + 32, "dec_front": ct, "dec_front_mfma": ct,
            "dec_decode": ct + 32, "dec_hash": 32, "enc_total": 32 + ct + 32, "dec_total": ct + 32}.get(kernel)


def kernel_roofline(kernel, avg_ms, n, p, ctr):
    """Roofline of one kernel launch: the live average duration (HIP events) against the work one
    launch does -- VALU and LDS wave-instructions from the committed counters (rocprofv3 SQ_INSTS_*,
    same workload), algorithmic HBM bytes -- and the chip's peaks.  `bound` is the resource with
    the highest fraction."""
    sec = avg_ms * 1e-3
    out = {"kernel": kernel, "avg_ms": round(avg_ms, 4)}
    hb = hbm_bytes_per_op(kernel, p)
    if hb:
        gbs = hb * n / sec / 1e9
        out["hbm"] = {"achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(gbs / HBM_PEAK_GBS, 5), "algorithmic_bytes_per_op": hb}
    c = ctr.get(kernel)
    if c:
        v = c["valu_insts"] / sec / 1e9
        out["valu"] = {"achieved": round(v, 1), "peak": VALU_PEAK_G, "unit": "G VALU wave-instr/s",
                       "frac": round(v / VALU_PEAK_G, 4), "insts_per_launch": c["valu_insts"]}
        l = c["lds_insts"] / sec / 1e9
        out["lds"] = {"achieved": round(l, 1), "peak": LDS_PEAK_G, "unit": "G LDS wave-instr/s",
                      "frac": round(l / LDS_PEAK_G, 4), "insts_per_launch": c["lds_insts"]}
        if c.get("lds_idx_active"):
            cyc = sec * CLOCK_GHZ * 1e9 * CUS
            out["lds"]["busy_frac"] = round(c["lds_idx_active"] / cyc, 4)
            out["lds"]["bank_conflict_frac_of_busy"] = round(c.get("lds_bank_conflict", 0) / c["lds_idx_active"], 4)
        if c.get("traffic_bytes"):
            out["traffic"] = c["traffic_bytes"]
        out["counters_workload"] = c.get("workload")
    fr = {k: out[k]["frac"] for k in ("valu", "lds", "hbm") if k in out}
    if fr:
        out["bound"] = max(fr, key=fr.get)
    return out


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


def cpu_leg(orc, s, pk, sk, seeds, cts, cores, seconds):
    """The CPU baseline: the word-level port (oracle FastKey) on all `cores` host threads, over a
    bounded sample of the same inputs (about `seconds` of work in total, split between the
    directions).  seeds -> encapsulations; cts -> decapsulations (None: decapsulate the sample's
    own ciphertexts).  Returns rates and the sample's outputs, which the caller compares with the
    GPU's for the same inputs."""
    fk = orc.FastKey(s, pk=pk, sk=sk)
    out = {}
    if seeds is not None:
        ne = min(len(seeds), max(cores * 8, 64))
        t0 = time.perf_counter()
        fk.enc(seeds[:ne], cores)
        te = time.perf_counter() - t0
        ne = int(min(len(seeds), max(ne, ne * (seconds / 2) / max(te, 1e-4))))
        t0 = time.perf_counter()
        ct_c, k_c, st_c = fk.enc(seeds[:ne], cores)
        te = time.perf_counter() - t0
        out.update(enc_per_s=ne / te, n_enc=ne, ct=ct_c, k_enc=k_c, st_enc=st_c)
    if sk is not None:
        src = cts if cts is not None else out["ct"]
        nd = min(len(src), max(cores * 8, 64))
        t0 = time.perf_counter()
        fk.dec(src[:nd], cores)
        td = time.perf_counter() - t0
        nd = int(min(len(src), max(nd, nd * (seconds / 2) / max(td, 1e-4))))
        t0 = time.perf_counter()
        k_d, st_d = fk.dec(src[:nd], cores)
        td = time.perf_counter() - t0
        out.update(dec_per_s=nd / td, n_dec=nd, k_dec=k_d, st_dec=st_d)
    fk.close()
    return out


ENC_KERNELS = ("gen_e", "enc_finish")
DEC_KERNELS = ("dec_front_mfma", "dec_front", "dec_decode", "dec_hash")


def direction_roofline(s, n, rates, concurrent=True, sampler_key="gen_e"):
    """Whole-direction VALU and LDS issue fractions at the given op rates (ops/s per direction):
    the VALU / LDS wave-instructions one op costs in each direction (sum over the direction's
    kernels, from the committed counters of the same set at n ops per launch) times the rate,
    against the chip's issue peaks.  concurrent: the directions ran at these rates at the same time
    (their issue rates add); otherwise they were timed one after the other, and the combined figure
    is all instructions over the summed time, with each direction's own fraction beside it.
    sampler_key: the counters entry of the sampler launch shape that ran ("gen_e_step": the
    two-stream step's 64-workgroup grid, whose instruction count differs from the full grid's)."""
    ctr = load_counters(f"{s}:{n}")
    if not ctr:
        return None
    if sampler_key != "gen_e":
        if sampler_key not in ctr:
            return None
        ctr = dict(ctr)
        ctr["gen_e"] = ctr[sampler_key]

    def frac(v, l):
        return ({"achieved": round(v / 1e9, 1), "peak": VALU_PEAK_G, "unit": "G VALU wave-instr/s",
                 "frac": round(v / 1e9 / VALU_PEAK_G, 4)},
                {"achieved": round(l / 1e9, 1), "peak": LDS_PEAK_G, "unit": "G LDS wave-instr/s",
                 "frac": round(l / 1e9 / LDS_PEAK_G, 4)})

    out, v_tot, l_tot, v_ops, l_ops, t_ops = {}, 0.0, 0.0, 0.0, 0.0, 0.0
    for d, names in (("enc", ENC_KERNELS), ("dec", DEC_KERNELS)):
        if d not in rates:
            continue
        v = sum(ctr[k]["valu_insts"] for k in names if k in ctr) / n
        l = sum(ctr[k]["lds_insts"] for k in names if k in ctr) / n
        v_tot += v * rates[d]
        l_tot += l * rates[d]
        v_ops, l_ops, t_ops = v_ops + v, l_ops + l, t_ops + 1.0 / rates[d]
        out[f"{d}_valu_insts_per_op"] = round(v, 1)
        out[f"{d}_lds_insts_per_op"] = round(l, 1)
        if not concurrent:
            dv, dl = frac(v * rates[d], l * rates[d])
            out[f"{d}_valu_frac"], out[f"{d}_lds_frac"] = dv["frac"], dl["frac"]
    if concurrent:
        out["valu"], out["lds"] = frac(v_tot, l_tot)
    else:
        out["valu"], out["lds"] = frac(v_ops / t_ops, l_ops / t_ops)
    out["directions"] = "concurrent" if concurrent else "timed one after the other"
    out["counters_workload"] = f"{s}:{n}"
    return out


def cpu_summary(c, cores, what):
    r = {"cores": cores, "kind": "port",
         "sample": f"word-level C port (oracle FastKey: column-gather syndrome, log-domain BM, additive-FFT "
                   f"roots), {what}, {cores} threads"}
    if "enc_per_s" in c:
        r["encaps_per_s"] = round(c["enc_per_s"], 1)
    if "dec_per_s" in c:
        r["decaps_per_s"] = round(c["dec_per_s"], 1)
    return r


# ------------------------------------------------------------------------------- suite sections
def section_syndrome(tm, orc, cores, log):
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
        for cfg in range(8):
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
        # linearity H(e1 + e2) = H(e1) + H(e2) on a sample, through the MFMA form
        h = min(n, 2048)
        e2 = e[:h] ^ e[h:2 * h] if n >= 2 * h else e[:h] ^ e[::-1][:h]
        other = outs["mfma"][h:2 * h] if n >= 2 * h else outs["mfma"][::-1][:h]
        d_e2, d_c2 = tm.DeviceBuffer(e2.nbytes), tm.DeviceBuffer(h * p["ct"])
        d_e2.upload(np.ascontiguousarray(e2))
        tm.syndrome_device(K, d_c2, d_e2, h, "mfma")
        tm.sync()
        same = same and bool(np.array_equal(d_c2.download((h, p["ct"])), outs["mfma"][:h] ^ other))
        d_e2.free()
        d_c2.free()
        tops = ops / (best_ms * 1e-3) / 1e12
        out[str(n)] = {
            "ms": {k: round(v, 4) for k, v in res.items()},
            "columns_per_s": {k: round(n / (v * 1e-3), 1) for k, v in res.items()},
            "mfma_cfg": best_cfg, "bit_exact": same,
            "mfma_roofline": {"bound": "mfma", "achieved": round(tops, 1), "peak": INT8_PEAK_TOPS, "unit": "TOPS",
                              "frac": round(tops / INT8_PEAK_TOPS, 4),
                              "work": "2*768*2720 int8 ops per column (0/1 bytes, int32 accumulate)"},
            "valu_lane_ops_per_column": 768 * 85 * 3 // 2,
            "check": "three GPU forms identical; linearity H(e1+e2) = H(e1)+H(e2) on 2048 columns",
        }
        log(f"syndrome n={n}: {out[str(n)]['ms']}")
        d_e.free()
        d_ct.free()
    # CPU context (the CPU path does not expose the syndrome alone): the port's encapsulation,
    # dominated by the syndrome and SHAKE256, on all cores
    c = cpu_leg(orc, s, pk, None, np.random.default_rng(5).integers(0, 256, (65536, 32), dtype=np.uint8), None,
                cores, 4.0)
    out["cpu_context"] = cpu_summary(c, cores, f"{c['n_enc']} encapsulations (syndrome + SHAKE256)")
    return out


def section_variants(tm, orc, cores, log, n=65536, reps=3):
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
            c = cpu_leg(orc, v, pk, sk, seeds, None, cores, 2.0)
            ne, nd = c["n_enc"], c["n_dec"]
            exact = bool(np.array_equal(k1, k2) and not st1.any() and not st2.any()
                         and np.array_equal(ct[:ne], c["ct"]) and np.array_equal(k1[:ne], c["k_enc"])
                         and np.array_equal(k2[:nd], c["k_dec"]))
            dk.free()
            row[v] = {"encaps_per_s": round(ee, 1), "decaps_per_s": round(ed, 1), "bit_exact": exact,
                      "cpu_baseline": cpu_summary(c, cores, f"{ne} encaps + {nd} decaps of the same inputs")}
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
        c = cpu_leg(orc, s, pk, sk, seeds, ct, cores, cpu_seconds)
        ne, nd = c["n_enc"], c["n_dec"]
        exact = exact and bool(np.array_equal(ct[:ne], c["ct"]) and np.array_equal(k1[:ne], c["k_enc"])
                               and np.array_equal(k2[:nd], c["k_dec"]) and not c["st_dec"].any())
        ce, cd = c["enc_per_s"], c["dec_per_s"]
        dk.free()
        out[s] = {"engine": {"encaps_per_s": round(ee, 1), "decaps_per_s": round(ed, 1)},
                  "plain_gpu_form": {"encaps_per_s": round(pe, 1), "decaps_per_s": round(pd, 1)},
                  "cpu_baseline": cpu_summary(c, cores, f"{ne} encaps + {nd} decaps of the same inputs"),
                  "speedup_vs_plain": {"encaps": round(ee / pe, 2), "decaps": round(ed / pd, 2)},
                  "speedup_vs_cpu": {"encaps": round(ee / ce, 1), "decaps": round(ed / cd, 1)},
                  "bit_exact": exact}
        log(f"per_set {s}: {out[s]['engine']} plain {out[s]['plain_gpu_form']} cpu {out[s]['cpu_baseline']}")
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
    # the CPU baseline on the whole batch: all cores, one batch of 256 each way (best of 3)
    fk = orc.FastKey(s, pk=pk, sk=sk)
    cte, ctd = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        ct_o, k_o, _ = fk.enc(seeds, cores)
        cte.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        k_d, _ = fk.dec(ct_o, cores)
        ctd.append(time.perf_counter() - t0)
    fk.close()
    cte, ctd = min(cte), min(ctd)
    exact = bool(np.array_equal(ct, ct_o) and np.array_equal(k1, k_o) and np.array_equal(k2, k_d)
                 and not st1.any() and not st2.any())
    dk.free()
    out = {"batch_ms": {"engine_enc": round(te * 1e3, 4), "engine_dec": round(td * 1e3, 4),
                        "plain_enc": round(pte * 1e3, 4), "plain_dec": round(ptd * 1e3, 4),
                        "cpu_enc": round(cte * 1e3, 3), "cpu_dec": round(ctd * 1e3, 3)},
           "cpu_baseline": {"cores": cores, "kind": "port", "encaps_per_s": round(n / cte, 1),
                            "decaps_per_s": round(n / ctd, 1), "sample": "the whole batch (256 each way)"},
           "bit_exact_vs_cpu": exact}
    log(f"c1: {out}")
    return out


def section_c3(tm, orc, keys, cores, log, n=8192):
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
        c = cpu_leg(orc, s, None, sk, None, ct, cores, 3.0)
        nd = c["n_dec"]
        exact = bool(np.array_equal(k1, k2) and not st2.any() and np.array_equal(k2[:nd], c["k_dec"]))
        dk.free()
        ctr = load_counters(f"{s}:{n}")
        dom = max(kern, key=lambda k: kern[k]["avg_ms"])
        out[s] = {"decaps_per_s": round(n / td, 1), "plain_decaps_per_s": round(n / ptd, 1),
                  "kernels_ms": {k: round(v["avg_ms"], 4) for k, v in kern.items()},
                  "roofline": kernel_roofline(dom, kern[dom]["avg_ms"], n, tm.PARAMS[s], ctr),
                  "hbm_decap_total": {"bytes_per_op": hbm_bytes_per_op("dec_total", tm.PARAMS[s]),
                                      "achieved_gbs": round(hbm_bytes_per_op("dec_total", tm.PARAMS[s]) * n / td / 1e9, 2)},
                  "cpu_baseline": cpu_summary(c, cores, f"{nd} decaps of the same ciphertexts"),
                  "bit_exact": exact}
        log(f"c3 {s}: {out[s]}")
    return out


def section_c4(tm, orc, keys, cores, log, total=65536):
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
    c = cpu_leg(orc, s, pk, sk, seeds, ct, cores, 4.0)
    ne, nd = c["n_enc"], c["n_dec"]
    exact_cpu = bool(np.array_equal(ct[:ne], c["ct"]) and np.array_equal(k1[:ne], c["k_enc"])
                     and np.array_equal(k2[:nd], c["k_dec"]))
    out = {"sub_batches": len(sizes), "total_ops": total, "tuned": best,
           "pipeline_graph": {"encaps_per_s": round(total / te, 1), "decaps_per_s": round(total / td, 1)},
           "flat_engine_batch": {"encaps_per_s": round(fe, 1), "decaps_per_s": round(fd, 1)},
           "plain_gpu_form": {"encaps_per_s": round(pe, 1), "decaps_per_s": round(pd, 1)},
           "speedup_vs_plain": {"encaps": round(total / te / pe, 2), "decaps": round(total / td / pd, 2)},
           "roofline": direction_roofline(s, total, {"enc": total / te, "dec": total / td}, concurrent=False),
           "cpu_baseline": cpu_summary(c, cores, f"{ne} encaps + {nd} decaps of the same inputs"),
           "speedup_vs_cpu": {"encaps": round(total / te / c["enc_per_s"], 1),
                              "decaps": round(total / td / c["dec_per_s"], 1)},
           "bit_exact_vs_flat_engine": exact, "bit_exact_vs_cpu": exact_cpu}
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


def cpu_request_mix(orc, s, kps, reqs, outputs, cores, seconds):
    """The port on the serving workload's requests (in order, until about `seconds`), whole
    requests on `cores` threads; returns (summary, outputs identical to the served ones)."""
    keys = [orc.FastKey(s, pk=pk, sk=sk) for pk, sk in kps]

    def one(i):
        d, k, inp, _ = reqs[i]
        return keys[k].enc(inp, 1) if d == "enc" else keys[k].dec(inp, 1)

    done, ops, t0, same = 0, 0, time.perf_counter(), True
    with ThreadPoolExecutor(max_workers=cores) as ex:
        while done < len(reqs) and time.perf_counter() - t0 < seconds:
            idx = list(range(done, min(len(reqs), done + 4 * cores)))
            for i, r in zip(idx, ex.map(one, idx)):
                ct, k, st = outputs[i]
                if reqs[i][0] == "enc":
                    same &= bool(np.array_equal(r[0], ct) and np.array_equal(r[1], k) and np.array_equal(r[2], st))
                else:
                    same &= bool(np.array_equal(r[0], k) and np.array_equal(r[1], st))
                ops += reqs[i][2].shape[0]
            done = idx[-1] + 1
    el = time.perf_counter() - t0
    for fk in keys:
        fk.close()
    return {"ops_per_s": round(ops / el, 1), "cores": cores, "kind": "port",
            "sample": f"word-level C port, the first {done} requests of the mix ({ops} ops), whole requests "
                      f"on {cores} threads"}, same


def section_c5(tm, orc, cores, log, nkeys=16, n_requests=6000):
    from tensormceliece_amd import serving
    s = "8192128"
    seeds, _ = xorshift_seeds(nkeys, 0x51ED)
    pks, sks = tm.keypair_batch(s, seeds)  # GPU-batched key generation
    kps = list(zip(pks, sks))
    wl = serving.Workload(s, kps, pool_per_key=4096)
    server = tm.Server(s, slot_capacity=65536, slots_per_direction=3, hbm_budget=8 << 30)
    try:
        reqs = serving.make_requests(wl, n_requests, seed=5)
        # warm-up: one untimed pass of the whole mix, so that every slot has run and captured its
        # pipeline graph (200 requests had left some of the six slots uncaptured)
        serving.run_poisson(server, wl, reqs, 0.0)
        sat = serving.run_poisson(server, wl, reqs, 0.0, keep_outputs=True)
        verified = verify_serving(wl, reqs, sat["outputs"])
        sat_outputs = sat["outputs"]
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
    # the CPU baseline on the same request mix: requests spread over all cores (one request per
    # thread at a time), each against its own key's tables; outputs compared with the served ones
    cpu, cpu_exact = cpu_request_mix(orc, s, kps, reqs, sat_outputs, cores, 4.0)
    enc_ops = sum(r[2].shape[0] for r in reqs if r[0] == "enc")
    frac_enc = enc_ops / sum(r[2].shape[0] for r in reqs)
    rates = {"enc": sat["ops_per_s"] * frac_enc, "dec": sat["ops_per_s"] * (1 - frac_enc)}
    out = {"keys": nkeys, "requests_per_run": n_requests, "mean_request_ops": round(mean_ops, 1),
           "key_device_bytes": wl.device_bytes(),
           "saturation": {"ops_per_s": round(sat["ops_per_s"], 1), "p50_ms": round(sat["p50_ms"], 3),
                          "p99_ms": round(sat["p99_ms"], 3)},
           "verified": verified,
           "poisson": loads, "zero_copy_slabs": zero_copy, "batches": st["batches"], "requests": st["requests"],
           "peak_in_flight_bytes": st["peak_in_flight_bytes"], "hbm_budget_bytes": 8 << 30,
           "max_batch_keys": st["max_batch_keys"],
           "roofline_at_saturation": direction_roofline(s, 65536, rates),
           "cpu_baseline": cpu, "cpu_outputs_identical": cpu_exact,
           "saturation_vs_cpu": round(sat["ops_per_s"] / cpu["ops_per_s"], 1)}
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
    ap.add_argument("--step-gen-blocks", type=int, default=96,
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

    # ---- checks without the CPU: every decapsulation returns its encapsulation's key, the plain
    #      form and the two-stream step give the same bytes
    verified = None
    if not args.no_verify:
        verified = bool(np.array_equal(k1, k2) and not st1.any() and not st2.any() and overlap_same is not False
                        and (plain is None or plain["identical_to_engine"]))

    # ---- roofline of the dominant kernel: live HIP-event duration against the work one launch
    #      does (VALU / LDS wave-instructions from the committed counters of this workload,
    #      algorithmic HBM bytes); the binding resource is the one with the highest fraction
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    roofline = kernel_roofline(dom, kernels[dom]["avg_ms"], n, p, load_counters(f"{s}:{n}"))
    rb = roofline.get(roofline.get("bound", "hbm"), {})
    roofline.update({"achieved": rb.get("achieved"), "peak": rb.get("peak"), "unit": rb.get("unit"),
                     "frac": rb.get("frac")})
    roofline.setdefault("traffic", None)
    # the two-stream step's sampler runs on args.step_gen_blocks workgroups; the committed counters
    # hold that launch shape ("gen_e_step", with its grid) for the default grid
    if args.one_stream:
        step_roofline = direction_roofline(s, n, {"enc": n / t_step, "dec": n / t_step})
    elif load_counters(f"{s}:{n}").get("gen_e_step", {}).get("grid_workgroups") == args.step_gen_blocks:
        step_roofline = direction_roofline(s, n, {"enc": n / t_step, "dec": n / t_step}, sampler_key="gen_e_step")
    else:
        step_roofline = None

    # ---- CPU baseline: the word-level port on every host core, a bounded sample of the same
    #      inputs; its outputs are the byte-for-byte check of the GPU's
    cpu = None
    if not args.no_cpu:
        c = cpu_leg(orc, s, pk, sk, seeds, ct, cores, args.cpu_seconds)
        ne, nd = c["n_enc"], c["n_dec"]
        cpu = {"value": round(2.0 / (1.0 / c["enc_per_s"] + 1.0 / c["dec_per_s"]), 1), "unit": "KEM ops/s",
               "cores": cores, "kind": "port", "encaps_per_s": round(c["enc_per_s"], 1),
               "decaps_per_s": round(c["dec_per_s"], 1),
               "sample": f"word-level C port (oracle FastKey), {ne} encaps + {nd} decaps of the same seeded {s} "
                         f"inputs, {cores} threads",
               "outputs_identical_to_gpu": bool(np.array_equal(ct[:ne], c["ct"]) and np.array_equal(k1[:ne], c["k_enc"])
                                                and np.array_equal(k2[:nd], c["k_dec"]))}
        if verified is not None:
            verified = verified and cpu["outputs_identical_to_gpu"]
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
            suite["syndrome_348864"] = section_syndrome(tm, orc, cores, log)
        if "per_set" in sections:
            suite["per_set_n65536"] = section_per_set(tm, orc, keys, cores, min(4.0, args.cpu_seconds), log)
        if "c1" in sections:
            suite["c1_348864_n256"] = section_c1(tm, orc, keys, cores, log)
        if "c3" in sections:
            suite["c3_decap_n8192"] = section_c3(tm, orc, keys, cores, log)
        if "c4" in sections:
            suite["c4_6960119_uneven"] = section_c4(tm, orc, keys, cores, log)
        if "c5" in sections:
            suite["c5_8192128_serving"] = section_c5(tm, orc, cores, log)
        if "keygen" in sections:
            suite["keygen_batch"] = section_keygen(tm, cores, log)
        if "variants" in sections:
            suite["variants_n65536"] = section_variants(tm, orc, cores, log)

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
        "step_issue": step_roofline,
        "cpu_baseline": cpu,
        "suite": suite,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
