"""Multi-batch pipelines (tmce_pipeline_*): uneven sub-batches, several keys per run, graph replay.

Every result of a run must be byte-identical to what the same op gives in a plain batched call
(tmce_enc_batch / tmce_dec_batch, themselves checked against the oracle) -- "the same alone and in
the mix" -- and a sample is checked against the oracle directly.  Graph-captured pipelines are
replayed several times with different sub-batch layouts, each replay checked.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_KEYS = {}


@pytest.fixture(scope="module")
def tm():
    import tensormceliece_amd as tm
    tm.init(0)
    return tm


def keys(orc, s, idx=0):
    if (s, idx) not in _KEYS:
        kseed, st = orc.xorshift_seeds(1 + idx)
        pk, sk = orc.keypair_seeded(s, kseed[idx].tobytes())
        _KEYS[(s, idx)] = (pk, sk)
    return _KEYS[(s, idx)]


def log_uniform_sizes(rng, lo, hi, total):
    """Sizes drawn log-uniform in [lo, hi] until they sum to `total` (the last one truncated)."""
    out = []
    while sum(out) < total:
        v = int(np.exp(rng.uniform(np.log(lo), np.log(hi + 1))))
        out.append(max(lo, min(hi, v, total - sum(out))))
    return out


def run_pipeline(tm, pipe, plan, in_rows, row_bytes, ct_bytes):
    """plan: list of (key object, [op indices]); in_rows: [n][row_bytes] inputs.  Lays the sub-batches
    out in separate device regions, runs, returns (ct, k, status) per op index."""
    n = len(in_rows)
    d_in = tm.DeviceBuffer(max(1, n * row_bytes))
    d_ct = tm.DeviceBuffer(max(1, n * max(ct_bytes, 2)))
    d_k = tm.DeviceBuffer(n * 32)
    d_st = tm.DeviceBuffer(n * 4)
    order = [i for _, idx in plan for i in idx]
    d_in.upload(np.ascontiguousarray(in_rows[order]))
    subs, pos = [], 0
    for key, idx in plan:
        m = len(idx)
        subs.append((key, d_in.ptr + pos * row_bytes, d_ct.ptr + pos * ct_bytes if ct_bytes else None,
                     d_k.ptr + pos * 32, d_st.ptr + pos * 4, m))
        pos += m
    pipe.run(subs)
    pipe.sync()
    ct = np.zeros((n, max(ct_bytes, 1)), np.uint8)
    k = np.zeros((n, 32), np.uint8)
    st = np.zeros(n, np.int32)
    if ct_bytes:
        ct[order] = d_ct.download((n, ct_bytes))
    k[order] = d_k.download((n, 32))
    st[order] = d_st.download(n, np.int32)
    return ct, k, st


@pytest.mark.parametrize("s,total", [("6960119", 6000), ("348864", 5000)])
def test_uneven_single_key_graph_replays(tm, orc, s, total):
    """Uneven sub-batches (log-uniform sizes), one key: enc then dec through graph-captured pipelines,
    three runs with different layouts; every op equals the plain batched call and the oracle sample."""
    pk, sk = keys(orc, s)
    p = orc.params(s)
    PK, SK = tm.PublicKey(s, pk), tm.SecretKey(s, sk)
    rng = np.random.default_rng(21)
    seeds = rng.integers(0, 256, (total, 32), dtype=np.uint8)
    ct_ref, k_ref, st_ref = PK.enc(seeds)
    assert not st_ref.any()
    sample = rng.choice(total, 32, replace=False)
    o_ct, o_k, _ = orc.enc_batch(s, pk, seeds[sample])
    assert np.array_equal(ct_ref[sample], o_ct) and np.array_equal(k_ref[sample], o_k)
    enc_pipe = tm.Pipeline(s, "enc", total + 4096, use_graph=True)
    dec_pipe = tm.Pipeline(s, "dec", total + 4096, use_graph=True)
    for run in range(3):
        perm = rng.permutation(total)
        sizes = log_uniform_sizes(rng, 1, 2048, total)
        plan, pos = [], 0
        for m in sizes:
            plan.append((PK, perm[pos:pos + m]))
            pos += m
        print(f"pipeline {s} run {run}: {len(sizes)} sub-batches", flush=True)
        ct, k, st = run_pipeline(tm, enc_pipe, plan, seeds, 32, p.synd_bytes)
        assert not st.any()
        assert np.array_equal(ct, ct_ref) and np.array_equal(k, k_ref), f"enc run {run}"
        dplan = [(SK, idx) for _, idx in plan[::-1]]
        _, k2, st2 = run_pipeline(tm, dec_pipe, dplan, ct_ref, p.synd_bytes, 0)
        assert not st2.any()
        assert np.array_equal(k2, k_ref), f"dec run {run}"


def test_c4_full_size_6960119(tm, orc):
    """The uneven-batch configuration at full size: mceliece6960119, sub-batches log-uniform in
    [1, 8192] totalling 65536 ops, through graph-captured pipelines (the second and third runs replay
    the graph with new layouts).  Every op: the pipeline's ct and k equal one flat batched call's,
    decapsulation returns the encapsulated key with status 0, and a sample equals the oracle."""
    s, total = "6960119", 65536
    pk, sk = keys(orc, s)
    p = orc.params(s)
    PK, SK = tm.PublicKey(s, pk), tm.SecretKey(s, sk)
    rng = np.random.default_rng(41)
    seeds = rng.integers(0, 256, (total, 32), dtype=np.uint8)
    ct_ref, k_ref, st_ref = PK.enc(seeds)
    assert not st_ref.any()
    sample = rng.choice(total, 12, replace=False)
    o_ct, o_k, _ = orc.enc_batch(s, pk, seeds[sample])
    assert np.array_equal(ct_ref[sample], o_ct) and np.array_equal(k_ref[sample], o_k)
    o_k2, o_st2 = orc.dec_batch(s, sk, ct_ref[sample[:6]])
    assert np.array_equal(o_k2, k_ref[sample[:6]]) and not o_st2.any()
    enc_pipe = tm.Pipeline(s, "enc", total, use_graph=True)
    dec_pipe = tm.Pipeline(s, "dec", total, use_graph=True)
    for run in range(3):
        perm = rng.permutation(total)
        sizes = log_uniform_sizes(rng, 1, 8192, total)
        assert sum(sizes) == total and max(sizes) <= 8192
        plan, pos = [], 0
        for m in sizes:
            plan.append((PK, perm[pos:pos + m]))
            pos += m
        print(f"pipeline {s} full size run {run}: {len(sizes)} sub-batches", flush=True)
        ct, k, st = run_pipeline(tm, enc_pipe, plan, seeds, 32, p.synd_bytes)
        assert not st.any()
        assert np.array_equal(ct, ct_ref) and np.array_equal(k, k_ref), f"enc run {run}"
        dplan = [(SK, idx) for _, idx in plan[::-1]]
        _, k2, st2 = run_pipeline(tm, dec_pipe, dplan, ct, p.synd_bytes, 0)
        assert not st2.any()
        assert np.array_equal(k2, k), f"dec run {run}"


def test_mixed_keys_enc_dec(tm, orc):
    """Three keys interleaved in one run (each key's group padded separately): results equal the
    per-key plain calls.  Decapsulation adds tampered and random ciphertexts (implicit rejection)."""
    s = "348864"
    p = orc.params(s)
    K = [keys(orc, s, i) for i in range(3)]
    PKs = [tm.PublicKey(s, pk) for pk, _ in K]
    SKs = [tm.SecretKey(s, sk) for _, sk in K]
    rng = np.random.default_rng(5)
    n = 3000
    owner = rng.integers(0, 3, n)
    seeds = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    ct_ref = np.zeros((n, p.synd_bytes), np.uint8)
    k_ref = np.zeros((n, 32), np.uint8)
    for i in range(3):
        idx = np.nonzero(owner == i)[0]
        ct_ref[idx], k_ref[idx], _ = PKs[i].enc(seeds[idx])
    # sub-batches: runs of consecutive ops of one owner, in a shuffled order
    runs, start = [], 0
    for i in range(1, n + 1):
        if i == n or owner[i] != owner[start] or i - start >= 300:
            runs.append((owner[start], np.arange(start, i)))
            start = i
    rng.shuffle(runs)
    enc_pipe = tm.Pipeline(s, "enc", n + 3 * 256, use_graph=True)
    for rep in range(2):
        ct, k, st = run_pipeline(tm, enc_pipe, [(PKs[o], idx) for o, idx in runs], seeds, 32, p.synd_bytes)
        assert not st.any() and np.array_equal(ct, ct_ref) and np.array_equal(k, k_ref), rep
    # decapsulation: valid, one-byte-tampered and random ciphertexts
    cts = ct_ref.copy()
    tam = rng.choice(n, 200, replace=False)
    cts[tam, rng.integers(0, p.synd_bytes, 200)] ^= rng.integers(1, 256, 200).astype(np.uint8)
    rnd = rng.choice(n, 100, replace=False)
    cts[rnd] = rng.integers(0, 256, (100, p.synd_bytes), dtype=np.uint8)
    k_dref = np.zeros((n, 32), np.uint8)
    st_dref = np.zeros(n, np.int32)
    for i in range(3):
        idx = np.nonzero(owner == i)[0]
        k_dref[idx], st_dref[idx] = SKs[i].dec(cts[idx])
    dec_pipe = tm.Pipeline(s, "dec", n + 3 * 256, use_graph=True)
    for rep in range(2):
        _, k2, st2 = run_pipeline(tm, dec_pipe, [(SKs[o], idx) for o, idx in runs], cts, p.synd_bytes, 0)
        assert np.array_equal(k2, k_dref) and np.array_equal(st2, st_dref), rep
    idx = np.concatenate([tam[:16], rnd[:16]])
    for i in range(3):
        sel = idx[owner[idx] == i]
        if len(sel):
            k_o, r_o = orc.dec_batch(s, K[i][1], cts[sel])
            assert np.array_equal(k_dref[sel], k_o) and np.array_equal(st_dref[sel], r_o)


def test_6960119_bad_public_key_in_mix(tm, orc):
    """A 6960119 public key with nonzero padding bits inside a mixed run: its ops return -1 with ct
    and k cleared; the other key's ops are unaffected."""
    s = "6960119"
    p = orc.params(s)
    pk, _ = keys(orc, s)
    bad = pk.copy()
    bad[p.pk_row_bytes - 1] |= 0x80
    good_K, bad_K = tm.PublicKey(s, pk), tm.PublicKey(s, bad)
    rng = np.random.default_rng(6)
    seeds = rng.integers(0, 256, (700, 32), dtype=np.uint8)
    ct_ref, k_ref, _ = good_K.enc(seeds[:400])
    pipe = tm.Pipeline(s, "enc", 2048, use_graph=False)
    ct, k, st = run_pipeline(tm, pipe, [(good_K, np.arange(0, 200)), (bad_K, np.arange(400, 700)),
                                        (good_K, np.arange(200, 400))], seeds, 32, p.synd_bytes)
    assert not st[:400].any() and np.array_equal(ct[:400], ct_ref) and np.array_equal(k[:400], k_ref)
    assert (st[400:] == -1).all() and not ct[400:].any() and not k[400:].any()


def test_pipeline_argument_errors(tm, orc):
    s = "348864"
    pk, sk = keys(orc, s)
    pipe = tm.Pipeline(s, "enc", 256, max_sub=4, use_graph=False)
    d = tm.DeviceBuffer(4096)
    with pytest.raises(tm.TmceError):  # capacity exceeded (two keys -> two padded groups)
        pipe.run([(tm.PublicKey(s, pk), d.ptr, d.ptr, d.ptr, None, 200),
                  (tm.PublicKey(s, keys(orc, s, 1)[0]), d.ptr, d.ptr, d.ptr, None, 10)])
    with pytest.raises(tm.TmceError):  # secret key in an encapsulation pipeline
        pipe.run([(tm.SecretKey(s, sk), d.ptr, d.ptr, d.ptr, None, 1)])
    with pytest.raises(tm.TmceError):  # more sub-batches than the pipeline holds
        pipe.run([(tm.PublicKey(s, pk), d.ptr, d.ptr, d.ptr, None, 1)] * 5)
