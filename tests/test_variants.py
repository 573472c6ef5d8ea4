"""GPU parity of the f and pc variants (LM:crypto_kem/<set>{f,pc,pcf}) against the CPU oracle.

The oracle's variants are pinned by libmceliece's own KAT checksums (tests/test_oracle.py).  Here:
- pc: ciphertext = syndrome || SHAKE256(2 || e, 32) and the longer ciphertext in the session-key
  hash (LM:crypto_kem/<set>pc/vec/kem_enc.c:32-37); decapsulation rejects implicitly when the
  confirmation differs (kem_dec.c:43-59) -- bit-exact ciphertexts, keys and statuses, including
  random and tampered ciphertexts (syndrome part and confirmation part);
- f: the semi-systematic key generation (LM:crypto_kem/<set>f/vec/pk_gen.c mov_columns) of the
  batched GPU generator, byte-identical to the host generator and the oracle;
- the variants through the libmceliece single-operation entry points and the multi-batch pipeline.
"""
import numpy as np
import pytest

from conftest import SETS

pytestmark = pytest.mark.gpu

_KEYS = {}


@pytest.fixture(scope="module")
def tm():
    import tensormceliece_amd as tm
    tm.init(0)
    return tm


def keys(orc, v):
    if v not in _KEYS:
        kseed, st = orc.xorshift_seeds(1, 0xC0FFEE + len(v))
        pk, sk = orc.keypair_seeded(v, kseed[0].tobytes())
        _KEYS[v] = (pk, sk, st)
    return _KEYS[v]


def tampered(rng, ct, lo, hi):
    """copies of ct with one byte in [lo, hi) changed"""
    out = ct.copy()
    idx = rng.integers(lo, hi, size=len(ct))
    out[np.arange(len(ct)), idx] += rng.integers(1, 256, size=len(ct)).astype(np.uint8)
    return out


@pytest.mark.parametrize("s", SETS)
def test_pc_enc_dec_parity(tm, orc, s):
    v = s + "pc"
    pk, sk, st = keys(orc, v)
    synd, ctb = orc.params(v).synd_bytes, orc.ct_bytes(v)
    assert ctb == synd + 32 == tm.PARAMS[v]["ct"]
    seeds = orc.xorshift_seeds(257, st)[0]  # ragged: 257 ops
    print(f"pc enc {v} n=257", flush=True)
    ct, k, st_ = tm.PublicKey(v, pk).enc(seeds)
    ct_ref, k_ref, ret = orc.enc_batch(v, pk, seeds)
    assert not st_.any() and not ret.any()
    assert np.array_equal(ct, ct_ref) and np.array_equal(k, k_ref)
    rng = np.random.default_rng(99)
    cts = np.concatenate([ct_ref,                                          # valid
                          rng.integers(0, 256, (64, ctb), dtype=np.uint8),  # random
                          tampered(rng, ct_ref[:96], 0, synd),              # syndrome part changed
                          tampered(rng, ct_ref[96:192], synd, ctb)])        # confirmation changed
    print(f"pc dec {v} n={len(cts)}", flush=True)
    k2, st2 = tm.SecretKey(v, sk).dec(cts)
    k2_ref, ret2 = orc.dec_batch(v, sk, cts)
    assert np.array_equal(st2, ret2)
    assert np.array_equal(k2, k2_ref)
    assert np.array_equal(k2[:len(seeds)], k_ref)
    # a changed confirmation is rejected: the key differs from the valid one
    assert not (k2[len(seeds) + 64 + 96:] == k_ref[96:192]).all(axis=1).any()


def test_6960119pc_padding(tm, orc):
    v = "6960119pc"
    pk, sk, st = keys(orc, v)
    p = orc.params(v)
    bad = pk.copy()
    bad[p.pk_row_bytes - 1] |= 0x80
    seeds = orc.xorshift_seeds(8, st)[0]
    ct, k, st_ = tm.PublicKey(v, bad).enc(seeds)
    ct_ref, k_ref, ret = orc.enc_batch(v, bad, seeds)
    assert (st_ == -1).all() and (ret == -1).all()
    assert np.array_equal(ct, ct_ref) and np.array_equal(k, k_ref) and not ct.any() and not k.any()
    ct_ok, _, _ = orc.enc_batch(v, pk, seeds)
    ct_bad = ct_ok.copy()
    ct_bad[:, p.synd_bytes - 1] |= 0x10
    k, st_ = tm.SecretKey(v, sk).dec(ct_bad)
    k_ref, ret = orc.dec_batch(v, sk, ct_bad)
    assert (st_ == -1).all() and (ret == -1).all()
    assert np.array_equal(k, k_ref) and (k == 0xFF).all()


@pytest.mark.parametrize("v,n", [("348864pc", 65536), ("8192128pc", 16384)])
def test_pc_roundtrip_large(tm, orc, v, n):
    pk, sk, _ = keys(orc, v)
    rng = np.random.default_rng(8)
    seeds = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
    print(f"pc roundtrip {v} n={n}", flush=True)
    ct, k, st = tm.PublicKey(v, pk).enc(seeds)
    k2, st2 = tm.SecretKey(v, sk).dec(ct)
    assert not st.any() and not st2.any()
    assert np.array_equal(k, k2)
    sample = rng.choice(n, 48, replace=False)
    ct_ref, k_ref, _ = orc.enc_batch(v, pk, seeds[sample])
    assert np.array_equal(ct[sample], ct_ref) and np.array_equal(k[sample], k_ref)


@pytest.mark.parametrize("s", SETS)
def test_f_keypair_batch(tm, orc, s):
    """Batched GPU key generation of the f variant: byte-identical to the host generator; one key
    checked against the oracle; the key works for encapsulation and decapsulation."""
    v = s + "f"
    n = 6 if s == "348864" else 2
    seeds = orc.xorshift_seeds(n, 0xF1F1 + len(s))[0]
    print(f"f keypair batch {v} n={n}", flush=True)
    pk, sk = tm.keypair_batch(v, seeds)
    for i in range(n):
        pk_h, sk_h = tm.keypair_seeded(v, seeds[i].tobytes())
        assert np.array_equal(pk[i], pk_h) and np.array_equal(sk[i], sk_h), i
    pk_o, sk_o = orc.keypair_seeded(v, seeds[0].tobytes())
    assert np.array_equal(pk[0], pk_o) and np.array_equal(sk[0], sk_o)
    piv = [int.from_bytes(sk[i, 32:40].tobytes(), "little") for i in range(n)]
    if s == "348864":
        assert any(x != 0xFFFFFFFF for x in piv)  # some key needed column swaps
    enc_seeds = orc.xorshift_seeds(64, 0x5EED)[0]
    ct, k, _ = tm.PublicKey(v, pk[0]).enc(enc_seeds)
    ct_ref, k_ref, _ = orc.enc_batch(v, pk_o, enc_seeds)
    assert np.array_equal(ct, ct_ref) and np.array_equal(k, k_ref)
    k2, st2 = tm.SecretKey(v, sk[0]).dec(ct)
    assert not st2.any() and np.array_equal(k2, k_ref)


@pytest.mark.parametrize("v", ["348864pcf", "6960119pcf"])
def test_pcf_keypair_batch(tm, orc, v):
    seeds = orc.xorshift_seeds(3, 0xABC)[0]
    print(f"pcf keypair batch {v}", flush=True)
    pk, sk = tm.keypair_batch(v, seeds)
    for i in range(3):
        pk_o, sk_o = orc.keypair_seeded(v, seeds[i].tobytes())
        assert np.array_equal(pk[i], pk_o) and np.array_equal(sk[i], sk_o), i
    ct, k, _ = tm.PublicKey(v, pk[1]).enc(seeds)
    k2, _ = tm.SecretKey(v, sk[1]).dec(ct)
    assert np.array_equal(k, k2)


@pytest.mark.parametrize("v", ["348864f", "460896pc", "6688128pcf", "8192128pc"])
def test_variant_single_op_api(tm, orc, v):
    """mceliece_kem_<variant>_enc/_dec (libmceliece interface) on one op."""
    pk, sk, _ = keys(orc, v)
    r, ct, k = tm.kem_enc(v, pk)
    assert r == 0 and ct.size == orc.ct_bytes(v)
    r2, k2 = tm.kem_dec(v, ct, sk)
    assert r2 == 0 and np.array_equal(k, k2)
    k_ref, _ = orc.dec_batch(v, sk, ct[None, :])
    assert np.array_equal(k2, k_ref[0])


def test_pc_pipeline(tm, orc):
    """pc sub-batches of two keys through a graph-captured pipeline, both directions: the same
    bytes as plain batched calls."""
    v = "6960119pc"
    ctb = orc.ct_bytes(v)
    pk0, sk0, _ = keys(orc, v)
    kseed, _ = orc.xorshift_seeds(2, 0xBEEF)
    pk1, sk1 = orc.keypair_seeded(v, kseed[1].tobytes())
    PK = [tm.PublicKey(v, pk0), tm.PublicKey(v, pk1)]
    SK = [tm.SecretKey(v, sk0), tm.SecretKey(v, sk1)]
    rng = np.random.default_rng(3)
    sizes = [37, 300, 1, 555, 90]
    total = sum(sizes)
    seeds = rng.integers(0, 256, (total, 32), dtype=np.uint8)
    owner = np.repeat(np.arange(len(sizes)) % 2, sizes)
    ref_ct = np.zeros((total, ctb), np.uint8)
    ref_k = np.zeros((total, 32), np.uint8)
    for j in (0, 1):
        idx = np.flatnonzero(owner == j)
        ref_ct[idx], ref_k[idx], _ = PK[j].enc(seeds[idx])
    enc_pipe = tm.Pipeline(v, "enc", 2048, use_graph=True)
    dec_pipe = tm.Pipeline(v, "dec", 2048, use_graph=True)
    d_seeds, d_ct = tm.DeviceBuffer(total * 32), tm.DeviceBuffer(total * ctb)
    d_k, d_k2 = tm.DeviceBuffer(total * 32), tm.DeviceBuffer(total * 32)
    d_st, d_st2 = tm.DeviceBuffer(total * 4), tm.DeviceBuffer(total * 4)
    d_seeds.upload(seeds)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    for rep in range(3):  # eager run, capture, replay
        print(f"pc pipeline run {rep}", flush=True)
        enc_pipe.run([(PK[i % 2], d_seeds.ptr + 32 * o, d_ct.ptr + ctb * o, d_k.ptr + 32 * o, d_st.ptr + 4 * o, m)
                      for i, (o, m) in enumerate(zip(starts, sizes))])
        enc_pipe.sync()
        dec_pipe.run([(SK[i % 2], d_ct.ptr + ctb * o, None, d_k2.ptr + 32 * o, d_st2.ptr + 4 * o, m)
                      for i, (o, m) in enumerate(zip(starts, sizes))])
        dec_pipe.sync()
        assert np.array_equal(d_ct.download((total, ctb)), ref_ct)
        assert np.array_equal(d_k.download((total, 32)), ref_k)
        assert np.array_equal(d_k2.download((total, 32)), ref_k)
        assert not d_st2.download(total, np.int32).any()
    enc_pipe.close()
    dec_pipe.close()


def test_pc_plain_form_rejected(tm, orc):
    """The plain per-stage form has no pc path: a pc call in that form is an error, not a silent
    non-pc result."""
    v = "348864pc"
    pk, _, _ = keys(orc, v)
    K = tm.PublicKey(v, pk)
    tm.set_form("plain")
    try:
        with pytest.raises(tm.TmceError):
            K.enc(np.zeros((4, 32), np.uint8))
    finally:
        tm.set_form("engine")
