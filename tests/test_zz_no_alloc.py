"""tmce_dec_batch_device_noalloc: the decapsulation entry point that allocates nothing (its scratch
is a static device array) gives exactly what tmce_dec_batch_device gives, called through the C ABI,
on two calls in a row (the second smaller, so the first call's scratch contents are still there),
for valid, random and (6960119) badly padded ciphertexts, and rejects batches beyond its limit.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_OPS = 16384  # TMCE_NOALLOC_MAX_OPS (include/tmce.h)


@pytest.fixture(scope="module")
def tm():
    import tensormceliece_amd as tm
    tm.init(0)
    return tm


def _fn(tm):
    f = tm.load().tmce_dec_batch_device_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_size_t, ctypes.c_void_p]
    return f


def _ciphertexts(tm, s, PK, n, rng):
    """Valid ciphertexts, with about 10% random ones (implicit rejection) and, for 6960119, some
    with nonzero padding bits (status -1)."""
    ctb = tm.PARAMS[s]["ct"]
    ct, _, st = PK.enc(rng.integers(0, 256, (n, 32), dtype=np.uint8))
    assert not st.any()
    bad = rng.choice(n, max(1, n // 10), replace=False)
    ct[bad] = rng.integers(0, 256, (len(bad), ctb), dtype=np.uint8)
    if s.startswith("6960119"):
        ct[bad[: len(bad) // 2], 193] |= 0x80  # padding bits of the last syndrome byte
    return ct


@pytest.mark.parametrize("s,n1,n2", [("348864", MAX_OPS, 1000), ("6960119", 5000, 77), ("8192128pc", MAX_OPS, 300)])
def test_noalloc_equals_dec_batch_device(tm, orc, s, n1, n2):
    seeds, _ = orc.xorshift_seeds(1, 0x4E41)
    pk, sk = tm.keypair_batch(s, seeds)
    PK, SK = tm.PublicKey(s, pk[0]), tm.SecretKey(s, sk[0])
    f = _fn(tm)
    ctb = tm.PARAMS[s]["ct"]
    rng = np.random.default_rng(n1 + n2)
    for n in (n1, n2):
        print(f"noalloc {s} n={n}", flush=True)
        ct = _ciphertexts(tm, s, PK, n, rng)
        d_ct = tm.DeviceBuffer(n * ctb)
        d_ct.upload(ct)
        d_k, d_st = tm.DeviceBuffer(n * 32), tm.DeviceBuffer(n * 4)
        d_k2, d_st2 = tm.DeviceBuffer(n * 32), tm.DeviceBuffer(n * 4)
        tm.dec_device(SK, d_k, d_st, d_ct, n)
        tm.sync()
        rc = f(SK.handle, d_k2.ptr, d_st2.ptr, d_ct.ptr, n, None)
        assert rc == 0, tm.load().tmce_last_error()
        tm.sync()
        k_ref, st_ref = d_k.download((n, 32)), d_st.download(n, np.int32)
        k, st = d_k2.download((n, 32)), d_st2.download(n, np.int32)
        assert np.array_equal(k, k_ref) and np.array_equal(st, st_ref)
        if s.startswith("6960119"):
            assert (st_ref == -1).any()
        for b in (d_ct, d_k, d_st, d_k2, d_st2):
            b.free()


def test_noalloc_rejects_oversized_batch(tm, orc):
    s = "348864"
    seeds, _ = orc.xorshift_seeds(1, 0x4E42)
    _, sk = tm.keypair_batch(s, seeds)
    SK = tm.SecretKey(s, sk[0])
    d = tm.DeviceBuffer(64)
    rc = _fn(tm)(SK.handle, d.ptr, d.ptr, d.ptr, MAX_OPS + 1, None)
    assert rc == -22  # TMCE_EINVAL, before any device work
    assert b"TMCE_NOALLOC_MAX_OPS" in tm.load().tmce_last_error()
    d.free()
