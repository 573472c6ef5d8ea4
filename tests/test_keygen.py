"""GPU: batched key generation (tmce_keypair_batch) against the host generator and the oracle.

Bar: byte-identical public and secret keys.  Key i of a batch must equal the host keypair on the
same seed (the seeded chain of tmce_keypair_seeded), including keys whose first candidates are
singular (the chain continues into the next GPU round), and the oracle's keypair, which
libmceliece's KAT checksums pin.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import SETS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tm():
    import tensormceliece_amd as tm
    tm.init(0)
    return tm


@pytest.mark.parametrize("s", SETS)
def test_keypair_batch_matches_host_and_oracle(tm, orc, s):
    rng = np.random.default_rng(int(s))
    seeds = rng.integers(0, 256, (3, 32), dtype=np.uint8)
    print(f"keypair batch {s} n=3", flush=True)
    pk, sk = tm.keypair_batch(s, seeds)
    for i in range(3):
        hpk, hsk = tm.keypair_seeded(s, seeds[i].tobytes())
        assert np.array_equal(pk[i], hpk), f"pk {i}"
        assert np.array_equal(sk[i], hsk), f"sk {i}"
    opk, osk = orc.keypair_seeded(s, seeds[0].tobytes())
    assert np.array_equal(pk[0], opk) and np.array_equal(sk[0], osk)


def test_keypair_batch_over_several_launches_and_usable(tm, orc):
    s = "348864"
    n = 150  # more candidates than one launch takes
    rng = np.random.default_rng(3)
    seeds = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    print(f"keypair batch {s} n={n}", flush=True)
    pk, sk = tm.keypair_batch(s, seeds)
    # every key against the host generator (threads: the host call releases the GIL)
    with ThreadPoolExecutor(max_workers=16) as ex:
        host = list(ex.map(lambda i: tm.keypair_seeded(s, seeds[i].tobytes()), range(n)))
    for i in range(n):
        assert np.array_equal(pk[i], host[i][0]) and np.array_equal(sk[i], host[i][1]), f"key {i}"
    # every key works: a seeded encapsulation decapsulates to the same session key
    es = rng.integers(0, 256, (8, 32), dtype=np.uint8)
    for i in (1, 100):
        ct, k, st = tm.PublicKey(s, pk[i]).enc(es)
        k2, st2 = tm.SecretKey(s, sk[i]).dec(ct)
        assert not st.any() and not st2.any() and np.array_equal(k, k2)
        ct_o, k_o, _ = orc.enc_batch(s, pk[i], es)
        assert np.array_equal(ct, ct_o) and np.array_equal(k, k_o)


def test_keypair_batch_staging_chunks(tm):
    """8192128: more winners in one launch (about 60 of 64 keys) than the pinned staging buffer
    holds public keys (49), so their downloads go in two chunks; keys spread over both chunks
    against the host generator."""
    s, n = "8192128", 64
    rng = np.random.default_rng(64)
    seeds = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    print(f"keypair batch {s} n={n}", flush=True)
    pk, sk = tm.keypair_batch(s, seeds)
    idx = list(range(0, n, 4)) + [n - 1]
    with ThreadPoolExecutor(max_workers=16) as ex:
        host = list(ex.map(lambda i: tm.keypair_seeded(s, seeds[i].tobytes()), idx))
    for i, (hpk, hsk) in zip(idx, host):
        assert np.array_equal(pk[i], hpk) and np.array_equal(sk[i], hsk), f"key {i}"


def test_keypair_batch_empty(tm):
    pk, sk = tm.keypair_batch("8192128", np.zeros((0, 32), np.uint8))
    assert pk.shape == (0, tm.PARAMS["8192128"]["pk"]) and sk.shape[0] == 0
