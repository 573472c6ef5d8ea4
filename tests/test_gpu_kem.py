"""GPU parity of the batched KEM engine (libtmce.so) against the CPU oracle, all five sets.

Bar: bit-exact ciphertexts and session keys.  Keys and seeds follow the reference harness's
xorshift64 stream (mceliece-gpu-params/main.cu:75-76,184-188).
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


def keys(orc, s):
    if s not in _KEYS:
        kseed, st = orc.xorshift_seeds(1)
        pk, sk = orc.keypair_seeded(s, kseed[0].tobytes())
        _KEYS[s] = (pk, sk, st)
    return _KEYS[s]


def seeds_for(orc, s, n):
    _, _, st = keys(orc, s)
    return orc.xorshift_seeds(n, st)[0]


@pytest.mark.parametrize("s", SETS)
def test_enc_parity(tm, orc, s):
    pk, sk, _ = keys(orc, s)
    seeds = seeds_for(orc, s, 256)
    print(f"enc parity {s} n=256", flush=True)
    ct, k, st = tm.PublicKey(s, pk).enc(seeds)
    ct_ref, k_ref, ret = orc.enc_batch(s, pk, seeds)
    assert not st.any() and not ret.any()
    assert np.array_equal(ct, ct_ref)
    assert np.array_equal(k, k_ref)


@pytest.mark.parametrize("s", SETS)
def test_dec_parity_valid(tm, orc, s):
    pk, sk, _ = keys(orc, s)
    seeds = seeds_for(orc, s, 256)
    ct_ref, k_ref, _ = orc.enc_batch(s, pk, seeds)
    print(f"dec parity {s} n=256", flush=True)
    k, st = tm.SecretKey(s, sk).dec(ct_ref)
    assert not st.any()
    assert np.array_equal(k, k_ref)


@pytest.mark.parametrize("s", SETS)
def test_dec_parity_invalid(tm, orc, s):
    """Implicit rejection: random ciphertexts and valid ones with one byte changed."""
    pk, sk, _ = keys(orc, s)
    rng = np.random.default_rng(1234)
    nbytes = orc.params(s).synd_bytes
    rnd = rng.integers(0, 256, size=(128, nbytes), dtype=np.uint8)
    ct_ref, _, _ = orc.enc_batch(s, pk, seeds_for(orc, s, 128))
    tam = ct_ref.copy()
    idx = rng.integers(0, nbytes, size=128)
    tam[np.arange(128), idx] += rng.integers(1, 256, size=128).astype(np.uint8)
    cts = np.concatenate([rnd, tam])
    print(f"dec invalid {s} n={len(cts)}", flush=True)
    k, st = tm.SecretKey(s, sk).dec(cts)
    k_ref, ret = orc.dec_batch(s, sk, cts)
    assert np.array_equal(st, ret)
    assert np.array_equal(k, k_ref)


def test_6960119_padding(tm, orc):
    s = "6960119"
    pk, sk, _ = keys(orc, s)
    p = orc.params(s)
    # public key with a nonzero padding bit: enc returns -1 and clears ct and k
    bad = pk.copy()
    bad[p.pk_row_bytes - 1] |= 0x80
    seeds = seeds_for(orc, s, 8)
    ct, k, st = tm.PublicKey(s, bad).enc(seeds)
    ct_ref, k_ref, ret = orc.enc_batch(s, bad, seeds)
    assert (st == -1).all() and (ret == -1).all()
    assert np.array_equal(ct, ct_ref) and np.array_equal(k, k_ref) and not ct.any() and not k.any()
    # ciphertext with a nonzero padding bit: dec returns -1 and k = 0xFF..FF
    ct_ok, _, _ = orc.enc_batch(s, pk, seeds)
    ct_bad = ct_ok.copy()
    ct_bad[:, -1] |= 0x10
    k, st = tm.SecretKey(s, sk).dec(ct_bad)
    k_ref, ret = orc.dec_batch(s, sk, ct_bad)
    assert (st == -1).all() and (ret == -1).all()
    assert np.array_equal(k, k_ref) and (k == 0xFF).all()


@pytest.mark.parametrize("s,n", [("348864", 65536), ("460896", 16384), ("6688128", 16384), ("6960119", 16384),
                                 ("8192128", 65536)])
def test_roundtrip_large(tm, orc, s, n):
    """At full batch sizes: dec(enc(seed)) recovers the session key for every op, every e has
    weight t (checked through a sample against the oracle), and results do not depend on N."""
    pk, sk, _ = keys(orc, s)
    rng = np.random.default_rng(7)
    seeds = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
    print(f"roundtrip {s} n={n}", flush=True)
    ct, k, st = tm.PublicKey(s, pk).enc(seeds)
    k2, st2 = tm.SecretKey(s, sk).dec(ct)
    assert not st.any() and not st2.any()
    assert np.array_equal(k, k2)
    sample = rng.choice(n, 64, replace=False)
    ct_ref, k_ref, _ = orc.enc_batch(s, pk, seeds[sample])
    assert np.array_equal(ct[sample], ct_ref) and np.array_equal(k[sample], k_ref)
    # batch-size independence: the sample alone gives the same bytes
    ct_s, k_s, _ = tm.PublicKey(s, pk).enc(seeds[sample])
    assert np.array_equal(ct_s, ct_ref) and np.array_equal(k_s, k_ref)


@pytest.mark.parametrize("s", ["348864", "460896", "6960119"])
def test_decoder_batch_shapes_agree(tm, orc, s):
    """Decapsulation at N = 65536 runs eight-op batches (Berlekamp-Massey on interleaved 8-lane
    rows); in pieces of 16384 it runs four-op batches (16-lane rows).  Both give the same keys and
    statuses on valid, tampered and random ciphertexts, and equal the oracle on a sample of each
    kind.  460896 keeps its top-down FFT's second buffer outside the syndrome rows (2t < 2D),
    6960119 (odd t) inside them."""
    n = 65536
    pk, sk, _ = keys(orc, s)
    p = orc.params(s)
    rng = np.random.default_rng(67)
    seeds = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
    ct, _, _ = tm.PublicKey(s, pk).enc(seeds)
    tam = rng.choice(n, 4000, replace=False)
    ct[tam, rng.integers(0, p.synd_bytes, 4000)] ^= rng.integers(1, 256, 4000).astype(np.uint8)
    rnd = rng.choice(np.setdiff1d(np.arange(n), tam), 2000, replace=False)
    ct[rnd] = rng.integers(0, 256, (2000, p.synd_bytes), dtype=np.uint8)
    SK = tm.SecretKey(s, sk)
    print(f"batch shapes {s} n={n}", flush=True)
    k_full, st_full = SK.dec(ct)
    parts = [SK.dec(ct[i:i + 16384]) for i in range(0, n, 16384)]
    assert np.array_equal(k_full, np.concatenate([x[0] for x in parts]))
    assert np.array_equal(st_full, np.concatenate([x[1] for x in parts]))
    sample = np.concatenate([tam[:24], rnd[:24], rng.choice(n, 16, replace=False)])
    k_ref, _ = orc.dec_batch(s, sk, ct[sample])
    assert np.array_equal(k_full[sample], k_ref)


@pytest.mark.parametrize("s", SETS)
def test_single_op_api(tm, orc, s):
    """mceliece_kem_<set>_enc/_dec entry points (libmceliece interface) on one op."""
    pk, sk, _ = keys(orc, s)
    r, ct, k = tm.kem_enc(s, pk)
    assert r == 0
    r2, k2 = tm.kem_dec(s, ct, sk)
    assert r2 == 0 and np.array_equal(k, k2)
    k_ref, _ = orc.dec_batch(s, sk, ct[None, :])
    assert np.array_equal(k2, k_ref[0])


@pytest.mark.parametrize("s", SETS)
def test_plain_form_parity(tm, orc, s):
    """The plain one-launch-per-stage form gives the same bytes as the oracle (and the engine)."""
    pk, sk, _ = keys(orc, s)
    seeds = seeds_for(orc, s, 192)
    ct_ref, k_ref, _ = orc.enc_batch(s, pk, seeds)
    rng = np.random.default_rng(5)
    rnd = rng.integers(0, 256, size=(64, orc.params(s).synd_bytes), dtype=np.uint8)
    print(f"plain form {s}", flush=True)
    tm.set_form("plain")
    try:
        ct, k, st = tm.PublicKey(s, pk).enc(seeds)
        k2, st2 = tm.SecretKey(s, sk).dec(np.concatenate([ct_ref, rnd]))
    finally:
        tm.set_form("engine")
    assert not st.any()
    assert np.array_equal(ct, ct_ref) and np.array_equal(k, k_ref)
    k2_ref, ret = orc.dec_batch(s, sk, np.concatenate([ct_ref, rnd]))
    assert np.array_equal(k2, k2_ref) and np.array_equal(st2, ret)


@pytest.mark.parametrize("s", SETS)
def test_sampler_forms(tm, orc, s):
    """Both samplers (one op per lane; pipelined lane groups) give the oracle's bytes, on a ragged
    batch, and the same bytes as each other on a larger one."""
    pk, sk, _ = keys(orc, s)
    rng = np.random.default_rng(17)
    seeds = rng.integers(0, 256, (3001, 32), dtype=np.uint8)
    K = tm.PublicKey(s, pk)
    old = tm.tune_get(s, "gen_form")
    out = {}
    try:
        for form in (0, 1):
            tm.tune_set(s, "gen_form", form)
            print(f"sampler form {form} {s}", flush=True)
            out[form] = K.enc(seeds)
    finally:
        tm.tune_set(s, "gen_form", old)
    for form in (0, 1):
        assert not out[form][2].any()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    idx = rng.choice(len(seeds), 48, replace=False)
    ct_ref, k_ref, _ = orc.enc_batch(s, pk, seeds[idx])
    assert np.array_equal(out[1][0][idx], ct_ref) and np.array_equal(out[1][1][idx], k_ref)


@pytest.mark.parametrize("s", SETS)
def test_decoder_fft_forms_agree(tm, orc, s):
    """The decoder's two bottom-up FFT forms -- bitsliced GF(2^m) arithmetic on the VALU and
    log/exp table lookups in LDS -- give the same keys and statuses on valid, tampered and random
    ciphertexts (ragged batch), and equal the oracle on a sample."""
    pk, sk, _ = keys(orc, s)
    p = orc.params(s)
    seeds = seeds_for(orc, s, 1500)
    ct, k_enc, _ = tm.PublicKey(s, pk).enc(seeds)
    rng = np.random.default_rng(61)
    cts = ct.copy()
    tam = rng.choice(len(cts), 200, replace=False)
    cts[tam, rng.integers(0, p.synd_bytes, 200)] ^= rng.integers(1, 256, 200).astype(np.uint8)
    rnd = rng.choice(len(cts), 100, replace=False)
    cts[rnd] = rng.integers(0, 256, (100, p.synd_bytes), dtype=np.uint8)
    cts = cts[:1337]
    SK = tm.SecretKey(s, sk)
    form0 = tm.tune_get(s, "fft_form")
    out = {}
    try:
        for form in (0, 1):
            tm.tune_set(s, "fft_form", form)
            print(f"decoder fft_form={form} {s}", flush=True)
            out[form] = SK.dec(cts)
    finally:
        tm.tune_set(s, "fft_form", form0)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    valid = np.setdiff1d(np.arange(1337), np.concatenate([tam, rnd]))
    assert np.array_equal(out[1][0][valid], k_enc[valid])
    idx = np.concatenate([tam[tam < 1337][:6], rnd[rnd < 1337][:6], valid[:4]])
    k_o, st_o = orc.dec_batch(s, sk, cts[idx])
    assert np.array_equal(out[1][0][idx], k_o) and np.array_equal(out[1][1][idx], st_o)
