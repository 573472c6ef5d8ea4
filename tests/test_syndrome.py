"""Syndrome alone (PKE encryption c = H e over GF(2)): the oracle is pinned on the CPU, the three GPU
forms (sparse column gather, INT8 MFMA GEMM, bit-packed VALU GEMM) are checked
bit-exact against it and against each other.

Pins of oracle_syndrome_batch (LM:crypto_kem/<set>/vec/encrypt.c syndrome, :94-132):
- an independent formulation: numpy unpacks pk rows and e to bits and takes the integer matrix
  product mod 2 (c = e[:mt] + T e[mt:], little-endian bit order as libmceliece's bytes);
- the KAT-pinned encapsulation: for a seeded op, the ciphertext of oracle enc equals the syndrome of
  the error vector that the same seeded randombytes chain samples (gen_e).
"""
import numpy as np
import pytest

from conftest import SETS

_KEYS = {}


def keys(orc, s):
    if s not in _KEYS:
        kseed, st = orc.xorshift_seeds(1)
        pk, sk = orc.keypair_seeded(s, kseed[0].tobytes())
        _KEYS[s] = (pk, sk, st)
    return _KEYS[s]


def fixed_weight_errors(p, n, rng):
    """n error vectors of weight t with uniform supports."""
    e = np.zeros((n, p.n // 8), np.uint8)
    for i in range(n):
        pos = rng.choice(p.n, p.t, replace=False)
        np.bitwise_or.at(e[i], pos >> 3, (1 << (pos & 7)).astype(np.uint8))
    return e


def numpy_syndrome(p, pk, e):
    T = np.unpackbits(pk.reshape(p.mt, p.pk_row_bytes), axis=1, bitorder="little")[:, :p.k].astype(np.int64)
    E = np.unpackbits(e, axis=1, bitorder="little")[:, :p.n].astype(np.int64)
    s = (E[:, p.mt:] @ T.T) & 1
    c = s ^ E[:, :p.mt]
    pad = np.zeros((e.shape[0], 8 * p.synd_bytes - p.mt), np.int64)
    return np.packbits(np.concatenate([c, pad], axis=1).astype(np.uint8), axis=1, bitorder="little")


@pytest.mark.parametrize("s", SETS)
def test_oracle_syndrome_vs_numpy(orc, s):
    p = orc.params(s)
    pk, _, _ = keys(orc, s)
    rng = np.random.default_rng(11)
    e = np.concatenate([fixed_weight_errors(p, 6, rng), rng.integers(0, 256, (3, p.n // 8), dtype=np.uint8),
                        np.zeros((1, p.n // 8), np.uint8)])
    assert np.array_equal(orc.syndrome_batch(s, pk, e), numpy_syndrome(p, pk, e))


@pytest.mark.parametrize("s", SETS)
def test_oracle_syndrome_is_enc_ciphertext(orc, s):
    pk, _, st = keys(orc, s)
    seeds, _ = orc.xorshift_seeds(3, st)
    ct, _, _ = orc.enc_batch(s, pk, seeds)
    for i in range(3):
        e, _ = orc.gen_e_seeded(s, seeds[i].tobytes())
        assert np.array_equal(orc.syndrome_batch(s, pk, e[None, :])[0], ct[i])


# ------------------------------------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def tm():
    import tensormceliece_amd as tm
    tm.init(0)
    return tm


METHODS = ["sparse", "mfma", "valu"]


@pytest.mark.gpu
@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("s", SETS)
def test_syndrome_parity(tm, orc, s, method):
    """Ragged batch (not a multiple of any tile) of weight-t, dense random, zero and all-ones errors."""
    p = orc.params(s)
    pk, _, _ = keys(orc, s)
    rng = np.random.default_rng(3)
    e = np.concatenate([fixed_weight_errors(p, 300, rng), rng.integers(0, 256, (37, p.n // 8), dtype=np.uint8),
                        np.zeros((1, p.n // 8), np.uint8), np.full((1, p.n // 8), 0xFF, np.uint8)])
    print(f"syndrome {s} {method} n={len(e)}", flush=True)
    ct = tm.PublicKey(s, pk).syndrome(e, method)
    assert np.array_equal(ct, orc.syndrome_batch(s, pk, e))


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", range(8))
@pytest.mark.parametrize("s", ["348864", "6960119"])
def test_syndrome_mfma_tile_shapes(tm, orc, s, cfg):
    """Every MFMA tile shape offered to the autotuner gives the oracle's bytes."""
    p = orc.params(s)
    pk, _, _ = keys(orc, s)
    rng = np.random.default_rng(4)
    e = np.concatenate([fixed_weight_errors(p, 200, rng), rng.integers(0, 256, (57, p.n // 8), dtype=np.uint8)])
    old = tm.tune_get(s, "synd_mfma_cfg")
    tm.tune_set(s, "synd_mfma_cfg", cfg)
    try:
        print(f"syndrome mfma {s} cfg={cfg}", flush=True)
        ct = tm.PublicKey(s, pk).syndrome(e, "mfma")
    finally:
        tm.tune_set(s, "synd_mfma_cfg", old)
    assert np.array_equal(ct, orc.syndrome_batch(s, pk, e))


@pytest.mark.gpu
@pytest.mark.parametrize("s,n", [("348864", 65536), ("460896", 8192), ("6688128", 8192), ("6960119", 8192),
                                 ("8192128", 8192)])
def test_syndrome_large(tm, orc, s, n):
    """Full batch sizes: the three forms agree, H is linear (H(e1+e2) = He1 + He2), and a sample
    matches the oracle."""
    p = orc.params(s)
    pk, _, _ = keys(orc, s)
    rng = np.random.default_rng(8)
    e = rng.integers(0, 256, (n, p.n // 8), dtype=np.uint8)
    e[: n // 2] = fixed_weight_errors(p, n // 2, rng) if n <= 8192 else e[: n // 2] & rng.integers(
        0, 256, (n // 2, p.n // 8), dtype=np.uint8) & rng.integers(0, 256, (n // 2, p.n // 8), dtype=np.uint8)
    K = tm.PublicKey(s, pk)
    print(f"syndrome large {s} n={n}", flush=True)
    out = {m: K.syndrome(e, m) for m in METHODS}
    assert np.array_equal(out["sparse"], out["mfma"]) and np.array_equal(out["sparse"], out["valu"])
    half = n // 2
    lin = K.syndrome(e[:half] ^ e[half:2 * half], "mfma")
    assert np.array_equal(lin, out["sparse"][:half] ^ out["sparse"][half:2 * half])
    idx = rng.choice(n, 48, replace=False)
    assert np.array_equal(out["mfma"][idx], orc.syndrome_batch(s, pk, e[idx]))


@pytest.mark.gpu
@pytest.mark.parametrize("s", SETS)
def test_front_end_forms(tm, orc, s):
    """Decapsulation with the front end S = F c on the VALU and on INT8 MFMA (every tile shape):
    session keys and status identical to the oracle for valid, tampered and random ciphertexts."""
    pk, sk, st = keys(orc, s)
    p = orc.params(s)
    seeds, _ = orc.xorshift_seeds(160, st)
    ct_ok, _, _ = orc.enc_batch(s, pk, seeds)
    rng = np.random.default_rng(9)
    tam = ct_ok[:60].copy()
    tam[np.arange(60), rng.integers(0, p.synd_bytes, 60)] ^= rng.integers(1, 256, 60).astype(np.uint8)
    cts = np.concatenate([ct_ok, tam, rng.integers(0, 256, (41, p.synd_bytes), dtype=np.uint8)])
    k_ref, ret = orc.dec_batch(s, sk, cts)
    SK = tm.SecretKey(s, sk)
    old = (tm.tune_get(s, "front_mfma"), tm.tune_get(s, "front_cfg"))
    try:
        for mf, cfg in [(0, 0)] + [(1, c) for c in range(8)]:
            tm.tune_set(s, "front_mfma", mf)
            tm.tune_set(s, "front_cfg", cfg)
            print(f"front {s} mfma={mf} cfg={cfg}", flush=True)
            k, stt = SK.dec(cts)
            assert np.array_equal(k, k_ref) and np.array_equal(stt, ret), (mf, cfg)
    finally:
        tm.tune_set(s, "front_mfma", old[0])
        tm.tune_set(s, "front_cfg", old[1])
