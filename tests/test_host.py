"""CPU tests of libtmce.so: exported ABI and the host-side key generator (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, SETS


def header_functions():
    txt = open(os.path.join(ROOT, "include", "tmce.h")).read()
    names = set(re.findall(r"\b(tmce_[a-z0-9_]+)\s*\(", txt))
    for s in SETS:
        for f in ("keypair", "enc", "dec", "implementation", "compiler"):
            names.add(f"mceliece_kem_{s}_{f}")
    return sorted(names)


def test_library_exports_header_symbols():
    import tensormceliece_amd as tm
    lib = tm.load()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_header_sizes_match_params(orc):
    txt = open(os.path.join(ROOT, "include", "tmce.h")).read()
    for s in SETS:
        p = orc.params(s)
        for macro, val in (("PUBLICKEYBYTES", p.pk_bytes), ("SECRETKEYBYTES", p.sk_bytes),
                           ("CIPHERTEXTBYTES", p.synd_bytes), ("BYTES", 32)):
            m = re.search(rf"#define mceliece_kem_{s}_{macro} (\d+)", txt)
            assert m and int(m.group(1)) == val, (s, macro)


@pytest.mark.parametrize("s", SETS)
def test_host_keypair_matches_oracle(orc, s):
    """The library's host keypair (randombytes = seeded chain) is byte-identical to the oracle's."""
    import tensormceliece_amd as tm
    seed, _ = orc.xorshift_seeds(1, 0x1234567 + len(s))
    pk, sk = tm.keypair_seeded(s, seed[0].tobytes())
    pk_ref, sk_ref = orc.keypair_seeded(s, seed[0].tobytes())
    assert np.array_equal(sk, sk_ref)
    assert np.array_equal(pk, pk_ref)
