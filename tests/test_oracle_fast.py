"""The word-level CPU port (oracle_fast_*, the CPU baseline timed beside the GPU) gives the same
bytes as the oracle, whose own pins are in test_oracle.py and test_seeded_chain.py.

- its additive FFT equals Horner evaluation (numpy, log/exp tables from the oracle's pinned field
  product) at every field element, for random polynomials of every degree the decoder uses;
- seeded encapsulation: ct, K and status equal oracle enc for every set and a pc variant, and the
  6960119 rule for a public key with nonzero padding bits (ct and K zero, -1);
- decapsulation: K and status equal oracle dec for valid, one-byte-tampered and random
  ciphertexts, 6960119 ciphertexts with padding bits set, and pc ciphertexts whose confirmation is
  tampered.
"""
import numpy as np
import pytest

from conftest import SETS

_KP = {}


def keypair(orc, s):
    base = s.replace("pc", "")
    if base not in _KP:
        seeds, _ = orc.xorshift_seeds(1, 0xFA57)
        _KP[base] = orc.keypair_seeded(base, seeds[0].tobytes())
    return _KP[base]


def field_tables(orc, m):
    q = (1 << m) - 1
    ex = np.zeros(2 * q, np.int64)
    x = 1
    g = next(c for c in range(2, 1 << m) if _order(orc, m, c) == q)
    for i in range(2 * q):
        ex[i] = x
        x = orc.gf_mul(m, x, g)
    lg = np.zeros(1 << m, np.int64)
    lg[ex[:q]] = np.arange(q)
    return lg, ex


def _order(orc, m, g):
    x, k = g, 1
    while x != 1 and k <= (1 << m):
        x = orc.gf_mul(m, x, g)
        k += 1
    return k


@pytest.mark.parametrize("m,deg", [(12, 64), (12, 10), (13, 96), (13, 119), (13, 127), (12, 0)])
def test_fft_equals_horner(orc, m, deg):
    lg, ex = field_tables(orc, m)
    xs = np.arange(1 << m)

    def mul(a, b):
        r = ex[(lg[a] + lg[b])]
        return np.where((a == 0) | (b == 0), 0, r)

    rng = np.random.default_rng(m * 1000 + deg)
    for _ in range(2):
        f = rng.integers(0, 1 << m, deg + 1)
        want = np.zeros(1 << m, np.int64)
        for c in f[::-1]:
            want = mul(want, xs) ^ c
        assert np.array_equal(orc.fast_eval_all(m, f).astype(np.int64), want)


@pytest.mark.parametrize("s", SETS + ["348864pc", "6960119pc"])
def test_fast_enc_equals_oracle(orc, s):
    pk, _ = keypair(orc, s)
    seeds, _ = orc.xorshift_seeds(48, 0x5EED)
    fk = orc.FastKey(s, pk=pk)
    got = fk.enc(seeds, nthreads=4)
    want = orc.enc_batch(s, pk, seeds)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    fk.close()


def test_fast_enc_bad_pk_padding(orc):
    s = "6960119"
    pk, _ = keypair(orc, s)
    p = orc.params(s)
    bad = pk.copy()
    bad[p.pk_row_bytes - 1] |= 0x80
    seeds, _ = orc.xorshift_seeds(4, 0xBAD)
    ct, k, st = orc.FastKey(s, pk=bad).enc(seeds)
    ct_o, k_o, st_o = orc.enc_batch(s, bad, seeds)
    assert (st == -1).all() and not ct.any() and not k.any()
    assert np.array_equal(ct, ct_o) and np.array_equal(k, k_o) and np.array_equal(st, st_o)


@pytest.mark.parametrize("s", SETS + ["348864pc", "8192128pc"])
def test_fast_dec_equals_oracle(orc, s):
    pk, sk = keypair(orc, s)
    p = orc.params(s)
    ctb = orc.ct_bytes(s)
    seeds, _ = orc.xorshift_seeds(24, 0xDEC)
    ct, _, _ = orc.enc_batch(s, pk, seeds)
    rng = np.random.default_rng(7)
    cts = np.concatenate([ct, ct[:8].copy(), rng.integers(0, 256, (6, ctb), dtype=np.uint8)])
    cts[24:32, rng.integers(0, p.synd_bytes, 8)] ^= rng.integers(1, 256, 8).astype(np.uint8)
    if s.endswith("pc"):  # tampered confirmations of otherwise valid ciphertexts
        extra = ct[:4].copy()
        extra[:, p.synd_bytes + 3] ^= 0x10
        cts = np.concatenate([cts, extra])
    if s == "6960119":  # nonzero padding bits of the last syndrome byte
        extra = ct[:4].copy()
        extra[:, p.synd_bytes - 1] |= 0x80
        cts = np.concatenate([cts, extra])
    fk = orc.FastKey(s, sk=sk)
    k, st = fk.dec(cts, nthreads=4)
    k_o, st_o = orc.dec_batch(s, sk, cts)
    assert np.array_equal(k, k_o) and np.array_equal(st, st_o)
    if s == "6960119":
        assert (st[-4:] == -1).all()
    fk.close()
