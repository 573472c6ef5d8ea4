"""CPU tests of libtmce.so: exported ABI and the host-side key generator (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, SETS


VARIANTS = [s + v for s in SETS for v in ("", "f", "pc", "pcf")]


def header_functions():
    txt = open(os.path.join(ROOT, "include", "tmce.h")).read()
    names = set(re.findall(r"\b(tmce_[a-z0-9_]+)\s*\(", txt))
    declared = set(re.findall(r"TMCE_DECLARE_SET\(([0-9a-z]+)\)", txt))
    assert declared == set(VARIANTS), sorted(set(VARIANTS) ^ declared)
    for s in declared:
        for f in ("keypair", "enc", "dec", "implementation", "compiler"):
            names.add(f"mceliece_kem_{s}_{f}")
            names.add(f"mceliece_dispatch_kem_{s}_{f}")
        names.add(f"mceliece_numimpl_kem_{s}")
    return sorted(names)


def test_dispatch_entry_points():
    """libmceliece's implementation dispatch: one implementation, whose function pointers are the
    entry points themselves (no GPU call)."""
    import tensormceliece_amd as tm
    lib = tm.load()
    for s in VARIANTS:
        numimpl = getattr(lib, f"mceliece_numimpl_kem_{s}")
        numimpl.restype = ctypes.c_longlong
        assert numimpl() == 1
        for f in ("keypair", "enc", "dec"):
            disp = getattr(lib, f"mceliece_dispatch_kem_{s}_{f}")
            disp.restype, disp.argtypes = ctypes.c_void_p, [ctypes.c_longlong]
            direct = ctypes.cast(getattr(lib, f"mceliece_kem_{s}_{f}"), ctypes.c_void_p).value
            assert disp(-1) == direct and disp(0) == direct
        impl = getattr(lib, f"mceliece_dispatch_kem_{s}_implementation")
        impl.restype, impl.argtypes = ctypes.c_char_p, [ctypes.c_longlong]
        assert impl(0) == b"tmce-gfx950"


def test_library_exports_header_symbols():
    import tensormceliece_amd as tm
    lib = tm.load()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_header_sizes_match_params(orc):
    """Size macros of all 20 variants against the oracle (whose sizes the KAT checksums pin), and
    the short aliases of LM:hdef."""
    txt = open(os.path.join(ROOT, "include", "tmce.h")).read()
    for s in VARIANTS:
        p = orc.params(s)
        for macro, val in (("PUBLICKEYBYTES", p.pk_bytes), ("SECRETKEYBYTES", p.sk_bytes),
                           ("CIPHERTEXTBYTES", orc.ct_bytes(s)), ("BYTES", 32)):
            m = re.search(rf"#define mceliece_kem_{s}_{macro} (\d+)", txt)
            assert m and int(m.group(1)) == val, (s, macro)
            assert f"#define mceliece{s}_{macro} mceliece_kem_{s}_{macro}" in txt, (s, macro)
        for f in ("keypair", "enc", "dec"):
            assert f"#define mceliece{s}_{f} mceliece_kem_{s}_{f}" in txt, (s, f)


def test_python_params_match_oracle(orc):
    import tensormceliece_amd as tm
    assert sorted(tm.PARAMS) == sorted(VARIANTS)
    for s in VARIANTS:
        p = orc.params(s)
        assert (tm.PARAMS[s]["pk"], tm.PARAMS[s]["sk"], tm.PARAMS[s]["ct"]) == (p.pk_bytes, p.sk_bytes, orc.ct_bytes(s))
        assert tm.SETS[s] == orc.SET_IDS[s]


@pytest.mark.parametrize("s", SETS + [s + "f" for s in SETS] + ["348864pcf"])
def test_host_keypair_matches_oracle(orc, s):
    """The library's host keypair (randombytes = seeded chain) is byte-identical to the oracle's."""
    import tensormceliece_amd as tm
    seed, _ = orc.xorshift_seeds(1, 0x1234567 + len(s))
    pk, sk = tm.keypair_seeded(s, seed[0].tobytes())
    pk_ref, sk_ref = orc.keypair_seeded(s, seed[0].tobytes())
    assert np.array_equal(sk, sk_ref)
    assert np.array_equal(pk, pk_ref)


def test_host_f_keypairs_with_column_swaps(orc):
    """f variants: keys whose column step swaps columns (pivot mask other than the identity
    0xFFFFFFFF) match the oracle byte for byte."""
    import tensormceliece_amd as tm
    seeds, _ = orc.xorshift_seeds(6, 0xF00D)
    swapped = 0
    for sd in seeds:
        pk, sk = tm.keypair_seeded("348864f", sd.tobytes())
        pk_ref, sk_ref = orc.keypair_seeded("348864f", sd.tobytes())
        assert np.array_equal(sk, sk_ref) and np.array_equal(pk, pk_ref)
        swapped += int.from_bytes(sk[32:40].tobytes(), "little") != 0xFFFFFFFF
    assert swapped >= 1
