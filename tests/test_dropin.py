"""libtmce.so as a drop-in for libmceliece: its mceliece_kem_<set>_{keypair,enc,dec} entry points,
driven by libmceliece's own KAT loop (LM:command/mceliece-test.c test_kem_<set>_impl, restated in
the oracle as oracle_kat_checksum_fns) with the driver's knownrandombytes installed as the
library's randombytes(), reproduce libmceliece's published checksums (kem_<set>_checksums[0], 8
loops: keypair, enc, dec and three tampered ciphertexts per loop).  This pins the library's
randomness consumption too: keypair draws 32 bytes, enc draws 4t bytes per gen_e attempt (2t for
8192128), exactly as libmceliece.
"""
import pytest

from test_oracle import KAT8, KAT8_VARIANTS

pytestmark = pytest.mark.gpu

KATS = dict(KAT8, **KAT8_VARIANTS)


@pytest.fixture(scope="module")
def tm():
    import tensormceliece_amd as tm
    tm.init(0)
    yield tm
    tm.set_randombytes(None)


@pytest.mark.parametrize("v", ["348864", "348864pcf", "460896f", "6960119pc", "8192128"])
def test_library_reproduces_libmceliece_kat(tm, orc, v):
    print(f"drop-in KAT {v}", flush=True)
    tm.set_randombytes(orc.randombytes_address())
    try:
        got = orc.kat_checksum_fns(v, 8, tm.entry_address(v, "keypair"), tm.entry_address(v, "enc"),
                                   tm.entry_address(v, "dec"))
    finally:
        tm.set_randombytes(None)
    assert got == KATS[v]
