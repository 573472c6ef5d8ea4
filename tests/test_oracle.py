"""Pins the CPU oracle (oracle/tmce_oracle.c) before it is trusted as the GPU checker.

- whole KEM: libmceliece-20250507's own KAT checksums (LM:command/mceliece-test.c, e.g.
  kem_348864 :4567-4570), replayed through the oracle with the test driver's knownrandombytes;
  they cover keypair, enc, dec and three tampered ciphertexts per loop;
- SHAKE256: the test driver's xof_shake256 checksums (:487-490) and hashlib (FIPS 202);
- GF(2^m): brute-force polynomial products modulo the field polynomial;
- control bits / support: Benes layers applied to the identity reproduce the permutation;
- decoding: Berlekamp-Massey + root search recovers the seeded error vector exactly.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import SETS

KAT8 = {  # LM:command/mceliece-test.c kem_<set>_checksums[0] (8 loops)
    "348864": "2f36e1825115c5accc9a6ea609dec708f83be85e7fe448637bd502de50a1dd91",
    "460896": "006c110df7fd3ff1852c9096a6514855ef3e715469ecc21d5b6ba14d31057417",
    "6688128": "a3446acecf9aff0e757dfafc5402e365eb3ba485dd1f7b052f3ca6eabfdcb6c4",
    "6960119": "008c0538526a1c037e64f0543628b312a71e879c1f882f65493fd0d3a225d5d8",
    "8192128": "debad73db5027e8239b7e97d04c5a98b7521caba8ea4e58581c90efb50dd4b9e",
}
KAT64 = {  # kem_<set>_checksums[1] (64 loops)
    "348864": "df7965299cab8d4669eb29ebe87a2f61a539aa016a8dcda268a8cd595bf6d8da",
    "460896": "604d5d0be1ec220a2a914735fa5469622cd1f66737cf7e0709b7eaaa1e382f66",
    "6688128": "b2ee78aaa23c0612155239b1d92e9e1af589cb810bdb9d8c44eb85dbfd2e46d4",
    "6960119": "15c6727d4f4867487a54bbb208c06c1142a56a7937d7b598c05f63801cbdd277",
    "8192128": "5e64e0aeb5ab566a1efbb2b1127dcbf909ee1b395945f066cd5f097a85f2e135",
}


# the f (semi-systematic key generation) and pc (plaintext confirmation) variants:
# kem_<variant>_checksums[0], e.g. kem_348864f :4794-4796, kem_348864pc :5021-5023,
# kem_348864pcf :5248-5250, kem_6960119pcf :1616-1618
KAT8_VARIANTS = {
    "348864f": "5ad76eff6b5ccaedc1615e9795880cfd79b44c4247db4b8a13684c507a5db437",
    "348864pc": "ddfa65df30bee9c97e153089d24cce158a7979aa01ebdbb7bf374da13bd81feb",
    "348864pcf": "4dcfc56d5e162873f227befd1e7c51f2b317f9ccf811e931e86e2fcdeea6c10e",
    "460896f": "1418b90ac6aa9d36bd3c42c7e8c4ba299eafa45e93a903a383041629fb47f043",
    "460896pc": "e27bb3f53c37eaaef460488d21750a80baf3874b0b26b23e510325e45c31c560",
    "460896pcf": "1980e3f38d8258eca9d8d4300ac8dd6a6cde2f2c10c5e7b589971972531c8505",
    "6688128f": "daec95f7a2db91986d2d1ca7312d66314149b066db0f467935ecf027aa1f4403",
    "6688128pc": "f6d44759e9e52330923b5dc2957cd080c10b31990a96e0540c52ba42017fa849",
    "6688128pcf": "80e0655938457c4dcd6aa6323d60f2495efda8a2d4a13b1958e8ca8fd2e848b6",
    "6960119f": "8d0f85c0ffa074aff7cf1c5d2f8a048174de4ee62a74556e04e8da326e8b640d",
    "6960119pc": "f164d7a16ea6175870f34056b211d4d45a2bc95085759e2f688721f3fbf326c8",
    "6960119pcf": "873616887e3e83679d5a708b3fbd8eb2301d18ea21f0f7c238db87178596267a",
    "8192128f": "2d50dd3ea54c764b5f79505d181c08cefc2276a28c791681cce3a45888c687cc",
    "8192128pc": "a9d56a17b9b870e832cddde5c5787b55ada2189992d03eb75e23e710bcb98ccd",
    "8192128pcf": "fdf47ca26ac3e1a3ae8c2859619cce789420a4af8533ded9e1104d232bd62f5e",
}
KAT64_VARIANTS = {  # kem_<variant>_checksums[1]
    "348864pcf": "48fffebf6cd92a688b4724d88cf989ab199066040c7c62202dab6f30ef188d68",
    "6960119pcf": "d709f666d84fe4450ec59759ef20982df8d614afe788db2b7d0f91b0c0e02254",
}


@pytest.mark.parametrize("s", SETS)
def test_kat_checksum_8_loops(orc, s):
    assert orc.kat_checksum(s, 8) == KAT8[s]


@pytest.fixture(scope="module")
def variant_kats(orc):
    """The 15 variant checksums, computed on threads (the oracle's randomness state is per thread)."""
    from concurrent.futures import ThreadPoolExecutor
    names = list(KAT8_VARIANTS)
    with ThreadPoolExecutor(max_workers=max(1, min(8, os.cpu_count() or 1))) as ex:
        return dict(zip(names, ex.map(lambda v: orc.kat_checksum(v, 8), names)))


@pytest.mark.parametrize("v", list(KAT8_VARIANTS))
def test_kat_checksum_8_loops_variant(variant_kats, v):
    assert variant_kats[v] == KAT8_VARIANTS[v]


def test_kat_checksum_64_loops_348864pcf(orc):
    assert orc.kat_checksum("348864pcf", 64) == KAT64_VARIANTS["348864pcf"]


@pytest.mark.skipif(not os.environ.get("TMCE_SLOW_TESTS"), reason="about 25 s; set TMCE_SLOW_TESTS=1")
def test_kat_checksum_64_loops_6960119pcf(orc):
    assert orc.kat_checksum("6960119pcf", 64) == KAT64_VARIANTS["6960119pcf"]


def test_kat_checksum_64_loops_348864(orc):
    assert orc.kat_checksum("348864", 64) == KAT64["348864"]


@pytest.mark.skipif(not os.environ.get("TMCE_SLOW_TESTS"), reason="about 4 minutes; set TMCE_SLOW_TESTS=1")
@pytest.mark.parametrize("s", SETS[1:])
def test_kat_checksum_64_loops(orc, s):
    assert orc.kat_checksum(s, 64) == KAT64[s]


def test_shake_checksums(orc):
    assert orc.shake_checksum(64, 128) == "a3c4908bf4412d650a1f0a69587778b72aeea66db5a9991a4f0cf2d782ca65b9"
    assert orc.shake_checksum(512, 16384) == "cfec0f87445678928e802011140a1df31a1692ec574364da56ddaaf3d4a7d69e"


@pytest.mark.parametrize("n,outlen", [(0, 32), (33, 256), (135, 64), (136, 136), (137, 300), (1233, 32)])
def test_shake256_vs_hashlib(orc, n, outlen):
    data = bytes((i * 37 + 11) & 0xFF for i in range(n))
    assert orc.shake256(data, outlen) == hashlib.shake_256(data).digest(outlen)


def _polymul_mod(m, a, b):
    poly = 0x1009 if m == 12 else 0x201B
    r = 0
    for i in range(m):
        if (b >> i) & 1:
            r ^= a << i
    for i in range(2 * m - 2, m - 1, -1):
        if (r >> i) & 1:
            r ^= poly << (i - m)
    return r


@pytest.mark.parametrize("m", [12, 13])
def test_gf_mul_bruteforce(orc, m):
    rng = np.random.default_rng(m)
    for a, b in rng.integers(0, 1 << m, size=(2000, 2)):
        assert orc.gf_mul(m, int(a), int(b)) == _polymul_mod(m, int(a), int(b))
    for a in (1, 2, 3, (1 << m) - 1):
        assert orc.gf_mul(m, orc.lib().oracle_gf_inv(m, a), a) == 1


@pytest.mark.parametrize("m", [4, 7, 12, 13])
def test_controlbits_roundtrip(orc, m):
    rng = np.random.default_rng(100 + m)
    pi = rng.permutation(1 << m).astype(np.int16)
    cb = orc.controlbits(m, pi)
    assert np.array_equal(orc.perm_from_controlbits(m, cb), pi)


@pytest.mark.parametrize("s", SETS)
def test_support_is_injective_and_pk_consistent(orc, s):
    seed, _ = orc.xorshift_seeds(1, 99)
    pk, sk = orc.keypair_seeded(s, seed[0].tobytes())
    alpha = orc.support(s, sk)
    assert len(set(alpha.tolist())) == orc.params(s).n
    pk2 = np.zeros_like(pk)
    assert orc.lib().oracle_pk_from_sk(orc.SET_IDS[s], orc._p(sk), orc._p(pk2)) == 0
    assert np.array_equal(pk, pk2)


@pytest.mark.parametrize("s", SETS)
def test_decoder_recovers_error(orc, s):
    """Seeded e -> c = H e; BM + root search recovers e bit-for-bit (unique decoding)."""
    seed, st = orc.xorshift_seeds(1, 7)
    pk, sk = orc.keypair_seeded(s, seed[0].tobytes())
    p = orc.params(s)
    for i in range(3):
        es, st = orc.xorshift_seeds(1, st)
        e, _ = orc.gen_e_seeded(s, es[0].tobytes())
        assert int(np.unpackbits(e).sum()) == p.t
        ct, _, _ = orc.enc_batch(s, pk, es)
        r, e2 = orc.decrypt(s, sk, ct[0])
        assert r == 0 and np.array_equal(e, e2)
        # locator roots are exactly the support positions of e
        loc = orc.bm(s, orc.synd_ct(s, sk, ct[0]))
        assert loc[p.t] == 1


def _golden():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "kem_vectors.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("s", SETS)
def test_oracle_matches_golden(orc, s):
    g = _golden()[s]
    pk, sk = orc.keypair_seeded(s, bytes.fromhex(g["keypair_seed"]))
    assert hashlib.sha256(pk.tobytes()).hexdigest() == g["pk_sha256"]
    assert hashlib.sha256(sk.tobytes()).hexdigest() == g["sk_sha256"]
    seeds = np.array([np.frombuffer(bytes.fromhex(v["seed"]), np.uint8) for v in g["enc"]])
    ct, k, _ = orc.enc_batch(s, pk, seeds)
    for i, v in enumerate(g["enc"]):
        assert ct[i].tobytes().hex() == v["ct"] and k[i].tobytes().hex() == v["k"]
    rct = np.array([np.frombuffer(bytes.fromhex(v["ct"]), np.uint8) for v in g["dec_random"]])
    rk, ret = orc.dec_batch(s, sk, rct)
    for i, v in enumerate(g["dec_random"]):
        assert rk[i].tobytes().hex() == v["k"] and int(ret[i]) == v["ret"]
