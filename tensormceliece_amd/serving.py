"""Serving workload driver: Poisson arrivals of encapsulation / decapsulation requests against
device-resident keys, through the library's serving loop (tmce_server_*), with latency percentiles.

The workload follows the serving configuration of the reference's scope: request sizes log-uniform in
[1, 4096] ops, 50/50 enc/dec, key chosen uniformly among the resident keypairs, exponential
inter-arrival times at a given request rate.  Decapsulation requests take ciphertexts from a pool
made by the engine beforehand (valid ciphertexts of the chosen key).
"""
import numpy as np

import tensormceliece_amd as tm


def log_uniform_sizes(rng, n, lo=1, hi=4096):
    return np.minimum(hi, np.maximum(lo, np.exp(rng.uniform(np.log(lo), np.log(hi + 1), n)).astype(np.int64)))


class Workload:
    """Resident keys (public and secret handles) and per-key ciphertext pools for decapsulation."""

    def __init__(self, set_name, keypairs, pool_per_key=4096, seed=1):
        self.set_name = set_name
        self.pk = [tm.PublicKey(set_name, pk) for pk, _ in keypairs]
        self.sk = [tm.SecretKey(set_name, sk) for _, sk in keypairs]
        rng = np.random.default_rng(seed)
        self.pool_seeds, self.pool_ct, self.pool_k = [], [], []
        for K in self.pk:
            sd = rng.integers(0, 256, (pool_per_key, 32), dtype=np.uint8)
            ct, k, st = K.enc(sd)
            assert not st.any()
            self.pool_seeds.append(sd)
            self.pool_ct.append(ct)
            self.pool_k.append(k)

    def device_bytes(self):
        return sum(k.device_bytes() for k in self.pk + self.sk)


def make_requests(wl, n_requests, seed=2, lo=1, hi=4096):
    """(direction, key index, inputs, pool offset) per request; enc seeds are fresh random bytes."""
    rng = np.random.default_rng(seed)
    sizes = log_uniform_sizes(rng, n_requests, lo, hi)
    dirs = rng.integers(0, 2, n_requests)
    keys = rng.integers(0, len(wl.pk), n_requests)
    reqs = []
    pool = wl.pool_ct[0].shape[0]
    for n, d, k in zip(sizes, dirs, keys):
        n = int(n)
        if d == 0:
            reqs.append(("enc", int(k), rng.integers(0, 256, (n, 32), dtype=np.uint8), None))
        else:
            off = int(rng.integers(0, pool - n + 1))
            reqs.append(("dec", int(k), wl.pool_ct[k][off:off + n], off))
    return reqs


def to_slab(reqs, params):
    """The requests' inputs copied into a pinned HostBuffer, and a second one sized for their
    outputs, so that the server serves them in place (zero-copy).  Returns (requests with slab
    inputs, input slab, output slab); pass the output slab to run_poisson."""
    inb = tm.HostBuffer(sum(r[2].nbytes + 16 for r in reqs) + 16)
    outb = tm.HostBuffer(sum(r[2].shape[0] * (32 + 4 + params["ct"]) + 48 for r in reqs) + 16)
    out = []
    for d, k, inp, off in reqs:
        a = inb.array(inp.shape, np.uint8)
        a[...] = inp
        out.append((d, k, a, off))
    return out, inb, outb


def run_poisson(server, wl, reqs, rate, keep_outputs=False, slab=None):
    """Submits `reqs` at Poisson arrival times (`rate` requests/s; rate <= 0: all at once) and
    collects them, through the library's native load driver (tmce_server_run_load: a native thread
    submits at the arrival times and completions are collected as they happen, no Python thread in
    the loop).  Returns dict with latencies (s), throughput (ops/s) and per-request outputs
    (ct, k, status).  With `slab` (the output HostBuffer of to_slab) the outputs are allocated in it:
    with slab inputs too, the requests are zero-copy."""
    rng = np.random.default_rng(3)
    if slab is not None:
        slab.reset()
    gaps = rng.exponential(1.0 / rate, len(reqs)) if rate > 0 else np.zeros(len(reqs))
    arrivals = np.cumsum(gaps)
    ctb = server.params["ct"]
    # outputs written before the run (np.full touches every page; np.zeros maps pages lazily and the
    # server's copier threads would take the first-touch page faults inside the timed run: ~0.5 GB
    # of fresh output arrays for the serving configuration's 6000 requests)
    new = (lambda shape, dt: slab.array(shape, dt)) if slab is not None else (lambda shape, dt: np.full(shape, 0, dt))
    items, outputs = [], []
    for d, k, inp, _ in reqs:
        n = inp.shape[0]
        ct = new((n, ctb), np.uint8) if d == "enc" else None
        kk = new((n, 32), np.uint8)
        st = new(n, np.int32)
        key = wl.pk[k] if d == "enc" else wl.sk[k]
        items.append((d, key, inp, ct, kk, st))
        outputs.append((ct, kk, st))
    lat_s, elapsed = server.run_load(items, arrivals)
    ops = int(sum(r[2].shape[0] for r in reqs))
    return {"latency_s": lat_s, "elapsed_s": elapsed, "ops": ops, "ops_per_s": ops / elapsed,
            "p50_ms": float(np.percentile(lat_s, 50) * 1e3), "p99_ms": float(np.percentile(lat_s, 99) * 1e3),
            "outputs": outputs if keep_outputs else [None] * len(reqs)}
