"""Edge cases of the batched and single-op entry points on the GPU, against the CPU oracle.

* empty batches (n = 0) through the host and device entry points: return 0, write nothing;
* ragged batch sizes that leave a partial last batch of the decoder's four- and eight-op batches,
  a partial sampler lane group and a partial hash workgroup;
* full-size ragged batches that select the eight-op decoder batches (348864 from 65529 ops, the
  m = 13 sets from 49145) with a partial last batch: round trip for every op, the last ops and a
  sample against the oracle;
* libmceliece's aliasing rules for dec (k may overlap ct, and k may overlap sk:
  LM:command/mceliece-test.c checks both for every implementation).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_KEYS = {}


@pytest.fixture(scope="module")
def tm():
    import tensormceliece_amd as tm
    tm.init(0)
    return tm


def keys(orc, s):
    if s not in _KEYS:
        kseed, st = orc.xorshift_seeds(1, 0xED6E)
        _KEYS[s] = orc.keypair_seeded(s, kseed[0].tobytes())
    return _KEYS[s]


@pytest.mark.parametrize("s", ["348864", "6960119"])
def test_empty_batches(tm, orc, s):
    pk, sk = keys(orc, s)
    ctb = tm.PARAMS[s]["ct"]
    PK, SK = tm.PublicKey(s, pk), tm.SecretKey(s, sk)
    ct, k, st = PK.enc(np.zeros((0, 32), np.uint8))
    assert ct.shape == (0, ctb) and k.shape == (0, 32) and st.shape == (0,)
    k2, st2 = SK.dec(np.zeros((0, ctb), np.uint8))
    assert k2.shape == (0, 32) and st2.shape == (0,)
    # device entry points: n = 0 leaves the output buffers untouched
    d_ct, d_k, d_st, d_seeds = (tm.DeviceBuffer(64 * ctb), tm.DeviceBuffer(64 * 32), tm.DeviceBuffer(64 * 4),
                                tm.DeviceBuffer(64 * 32))
    marker = np.full(64 * 32, 0xA5, np.uint8)
    d_k.upload(marker)
    tm.enc_device(PK, d_ct, d_k, d_st, d_seeds, 0)
    tm.dec_device(SK, d_k, d_st, d_ct, 0)
    tm.sync()
    assert np.array_equal(d_k.download(64 * 32), marker)


@pytest.mark.parametrize("s,n", [("348864", 1), ("348864", 67), ("460896", 5), ("6688128", 3),
                                 ("6960119", 67), ("8192128", 9)])
def test_ragged_small(tm, orc, s, n):
    pk, sk = keys(orc, s)
    rng = np.random.default_rng(n)
    seeds = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    print(f"ragged {s} n={n}", flush=True)
    ct, k, st = tm.PublicKey(s, pk).enc(seeds)
    ct_ref, k_ref, _ = orc.enc_batch(s, pk, seeds)
    assert not st.any()
    assert np.array_equal(ct, ct_ref) and np.array_equal(k, k_ref)
    # every other ciphertext replaced by random bytes (implicit rejection)
    ct2 = ct.copy()
    ct2[1::2] = rng.integers(0, 256, ct2[1::2].shape, dtype=np.uint8)
    if s.startswith("6960119"):
        ct2[1::2, -1] &= 0x07  # keep the padding bits zero: a decoding failure, not a padding error
    k2, st2 = tm.SecretKey(s, sk).dec(ct2)
    k2_ref, ret = orc.dec_batch(s, sk, ct2)
    assert np.array_equal(k2, k2_ref) and not st2.any() and not ret.any()
    assert np.array_equal(k2[0::2], k[0::2])


@pytest.mark.parametrize("s,n", [("348864", 65531), ("6960119", 49150)])
def test_ragged_eight_op_batches(tm, orc, s, n):
    """Batch sizes just past the eight-op threshold, with a partial last batch (3 and 6 ops)."""
    pk, sk = keys(orc, s)
    rng = np.random.default_rng(n)
    seeds = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    print(f"ragged eight-op {s} n={n}", flush=True)
    ct, k, st = tm.PublicKey(s, pk).enc(seeds)
    assert not st.any()
    ct2 = ct.copy()
    bad = rng.choice(n, 512, replace=False)
    ct2[bad, 0] ^= 0x10  # tampered ciphertexts (implicit rejection)
    k2, st2 = tm.SecretKey(s, sk).dec(ct2)
    assert not st2.any()
    good = np.ones(n, bool)
    good[bad] = False
    assert np.array_equal(k2[good], k[good])
    sample = np.unique(np.concatenate([np.arange(n - 8, n), bad[:8], rng.choice(n, 16, replace=False)]))
    ct_ref, k_ref, _ = orc.enc_batch(s, pk, seeds[sample])
    assert np.array_equal(ct[sample], ct_ref) and np.array_equal(k[sample], k_ref)
    k2_ref, _ = orc.dec_batch(s, sk, ct2[sample])
    assert np.array_equal(k2[sample], k2_ref)


@pytest.mark.parametrize("s", ["348864", "8192128"])
def test_single_dec_aliasing(tm, orc, s):
    """mceliece_kem_<set>_dec with k overlapping ct, and with k overlapping sk."""
    pk, sk = keys(orc, s)
    r, ct, k = tm.kem_enc(s, pk)
    assert r == 0
    f = getattr(tm.load(), f"mceliece_kem_{s}_dec")
    f.argtypes = [ctypes.c_void_p] * 3
    f.restype = ctypes.c_int
    for tamper in (False, True):
        c = ct.copy()
        if tamper:
            c[3] ^= 1
        k_ref, _ = orc.dec_batch(s, sk, c[None, :])
        buf = np.zeros(max(len(c), 32), np.uint8)  # k written over the ciphertext's first bytes
        buf[: len(c)] = c
        assert f(buf.ctypes.data, buf.ctypes.data, sk.ctypes.data) == 0
        assert np.array_equal(buf[:32], k_ref[0])
        skb = sk.copy()  # k written over the secret key's first bytes
        assert f(skb.ctypes.data, c.ctypes.data, skb.ctypes.data) == 0
        assert np.array_equal(skb[:32], k_ref[0])
        if not tamper:
            assert np.array_equal(k_ref[0], k)
