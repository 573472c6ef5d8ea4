"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline, as the checker; the shipped library never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
BASE_IDS = {"348864": 0, "460896": 1, "6688128": 2, "6960119": 3, "8192128": 4}
SETS = list(BASE_IDS)
# variant ids: base | 8 ("f": semi-systematic key generation) | 16 ("pc": plaintext confirmation)
VARIANT_FLAGS = {"": 0, "f": 8, "pc": 16, "pcf": 24}
SET_IDS = {b + v: i | f for b, i in BASE_IDS.items() for v, f in VARIANT_FLAGS.items()}
VARIANTS = list(SET_IDS)


class OrcParams(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("m", ctypes.c_int), ("n", ctypes.c_int), ("t", ctypes.c_int),
                ("mt", ctypes.c_int), ("k", ctypes.c_int), ("pk_row_bytes", ctypes.c_int),
                ("synd_bytes", ctypes.c_int), ("irr_bytes", ctypes.c_int), ("cond_bytes", ctypes.c_int),
                ("pk_bytes", ctypes.c_long), ("sk_bytes", ctypes.c_long)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, I, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_long
        L.oracle_params.restype = ctypes.POINTER(OrcParams)
        L.oracle_params.argtypes = [I]
        L.oracle_ct_bytes.restype = I
        L.oracle_ct_bytes.argtypes = [I]
        for name, args, res in [
            ("oracle_keypair_seeded", [I, P, P, P], None),
            ("oracle_enc_seeded", [I, P, P, P, P], I),
            ("oracle_dec", [I, P, P, P], I),
            ("oracle_enc_batch", [I, P, P, P, P, P, LL, I], I),
            ("oracle_dec_batch", [I, P, P, P, P, LL, I], I),
            ("oracle_syndrome_batch", [I, P, P, P, LL, I], I),
            ("oracle_shake256", [P, ctypes.c_longlong, P, ctypes.c_longlong], None),
            ("oracle_gf_mul", [I, ctypes.c_uint16, ctypes.c_uint16], ctypes.c_uint16),
            ("oracle_gf_inv", [I, ctypes.c_uint16], ctypes.c_uint16),
            ("oracle_support", [I, P, P], None),
            ("oracle_perm_from_controlbits", [I, P, P], None),
            ("oracle_controlbits", [I, P, P], None),
            ("oracle_synd_ct", [I, P, P, P], None),
            ("oracle_bm", [I, P, P], None),
            ("oracle_decrypt", [I, P, P, P], I),
            ("oracle_pk_from_sk", [I, P, P], I),
            ("oracle_kat_checksum", [I, I, ctypes.c_char_p], None),
            ("oracle_kat_checksum_fns", [I, I, ctypes.c_char_p, P, P, P], None),
            ("oracle_kat_shake_checksum", [I, I, ctypes.c_char_p], None),
            ("oracle_rng_seeded", [P], None),
            ("oracle_gen_e", [I, P, P], None),
            ("oracle_fast_key", [I, P, P], P),
            ("oracle_fast_key_free", [P], None),
            ("oracle_fast_enc_batch", [P, P, P, P, P, LL, I], I),
            ("oracle_fast_dec_batch", [P, P, P, P, LL, I], I),
            ("oracle_fast_eval_all", [I, P, I, P], None),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def params(set_name):
    return lib().oracle_params(SET_IDS[set_name]).contents


def ct_bytes(set_name):
    return lib().oracle_ct_bytes(SET_IDS[set_name])


def keypair_seeded(set_name, seed):
    p = params(set_name)
    pk = np.zeros(p.pk_bytes, np.uint8)
    sk = np.zeros(p.sk_bytes, np.uint8)
    sd = np.frombuffer(bytes(seed), np.uint8).copy()
    lib().oracle_keypair_seeded(SET_IDS[set_name], _p(pk), _p(sk), _p(sd))
    return pk, sk


def enc_batch(set_name, pk, seeds, nthreads=0):
    p = params(set_name)
    seeds = np.ascontiguousarray(seeds, np.uint8).reshape(-1, 32)
    n = seeds.shape[0]
    ct = np.zeros((n, ct_bytes(set_name)), np.uint8)
    k = np.zeros((n, 32), np.uint8)
    ret = np.zeros(n, np.int32)
    pk = np.ascontiguousarray(pk, np.uint8)
    lib().oracle_enc_batch(SET_IDS[set_name], _p(ct), _p(k), _p(ret), _p(pk), _p(seeds), n, nthreads)
    return ct, k, ret


def dec_batch(set_name, sk, ct, nthreads=0):
    p = params(set_name)
    ct = np.ascontiguousarray(ct, np.uint8).reshape(-1, ct_bytes(set_name))
    n = ct.shape[0]
    k = np.zeros((n, 32), np.uint8)
    ret = np.zeros(n, np.int32)
    sk = np.ascontiguousarray(sk, np.uint8)
    lib().oracle_dec_batch(SET_IDS[set_name], _p(k), _p(ret), _p(ct), _p(sk), n, nthreads)
    return k, ret


class FastKey:
    """The word-level CPU port (oracle_fast_*): per-key tables built once, then batches of seeded
    encapsulations / decapsulations on `nthreads` threads (0 = every core).  Same results as
    enc_batch / dec_batch."""

    def __init__(self, set_name, pk=None, sk=None):
        self.set_name = set_name
        self._pk = None if pk is None else np.ascontiguousarray(pk, np.uint8)
        self._sk = None if sk is None else np.ascontiguousarray(sk, np.uint8)
        self._h = lib().oracle_fast_key(SET_IDS[set_name], _p(self._pk) if self._pk is not None else None,
                                        _p(self._sk) if self._sk is not None else None)

    def enc(self, seeds, nthreads=0):
        seeds = np.ascontiguousarray(seeds, np.uint8).reshape(-1, 32)
        n = seeds.shape[0]
        ct = np.zeros((n, ct_bytes(self.set_name)), np.uint8)
        k = np.zeros((n, 32), np.uint8)
        ret = np.zeros(n, np.int32)
        lib().oracle_fast_enc_batch(self._h, _p(ct), _p(k), _p(ret), _p(seeds), n, nthreads)
        return ct, k, ret

    def dec(self, ct, nthreads=0):
        ct = np.ascontiguousarray(ct, np.uint8).reshape(-1, ct_bytes(self.set_name))
        n = ct.shape[0]
        k = np.zeros((n, 32), np.uint8)
        ret = np.zeros(n, np.int32)
        lib().oracle_fast_dec_batch(self._h, _p(k), _p(ret), _p(ct), n, nthreads)
        return k, ret

    def close(self):
        if self._h:
            lib().oracle_fast_key_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fast_eval_all(m, f):
    """f (coefficients, degree < 128) evaluated at every element of GF(2^m) by the port's FFT."""
    f = np.ascontiguousarray(f, np.uint16)
    vals = np.zeros(1 << m, np.uint16)
    lib().oracle_fast_eval_all(m, _p(f), len(f) - 1, _p(vals))
    return vals


def syndrome_batch(set_name, pk, e, nthreads=0):
    """c_i = e_i[0:mt) + T e_i[mt:n) for rows of e (n/8 bytes each)."""
    p = params(set_name)
    e = np.ascontiguousarray(e, np.uint8).reshape(-1, p.n // 8)
    n = e.shape[0]
    ct = np.zeros((n, p.synd_bytes), np.uint8)
    pk = np.ascontiguousarray(pk, np.uint8)
    lib().oracle_syndrome_batch(SET_IDS[set_name], _p(ct), _p(pk), _p(e), n, nthreads)
    return ct


def kat_checksum(set_name, loops):
    buf = ctypes.create_string_buffer(65)
    lib().oracle_kat_checksum(SET_IDS[set_name], loops, buf)
    return buf.value.decode()


def kat_checksum_fns(set_name, loops, keypair_fn, enc_fn, dec_fn):
    """The driver's KAT loop over another implementation's entry points (addresses); that
    implementation must draw its randomness through randombytes_address()."""
    buf = ctypes.create_string_buffer(65)
    lib().oracle_kat_checksum_fns(SET_IDS[set_name], loops, buf, keypair_fn, enc_fn, dec_fn)
    return buf.value.decode()


def randombytes_address():
    """Address of oracle_randombytes(out, len): the randomness the KAT loop's driver provides."""
    return ctypes.cast(lib().oracle_randombytes, ctypes.c_void_p).value


def shake_checksum(loops, maxtest):
    buf = ctypes.create_string_buffer(65)
    lib().oracle_kat_shake_checksum(loops, maxtest, buf)
    return buf.value.decode()


def shake256(data, outlen):
    data = np.frombuffer(bytes(data), np.uint8).copy()
    out = np.zeros(outlen, np.uint8)
    lib().oracle_shake256(_p(out), outlen, _p(data) if data.size else None, data.size)
    return out.tobytes()


def gf_mul(m, a, b):
    return lib().oracle_gf_mul(m, a, b)


def support(set_name, sk):
    p = params(set_name)
    out = np.zeros(p.n, np.uint16)
    lib().oracle_support(SET_IDS[set_name], _p(np.ascontiguousarray(sk, np.uint8)), _p(out))
    return out


def synd_ct(set_name, sk, ct):
    p = params(set_name)
    out = np.zeros(2 * p.t, np.uint16)
    lib().oracle_synd_ct(SET_IDS[set_name], _p(np.ascontiguousarray(sk, np.uint8)),
                         _p(np.ascontiguousarray(ct, np.uint8)), _p(out))
    return out


def bm(set_name, s):
    p = params(set_name)
    out = np.zeros(p.t + 1, np.uint16)
    lib().oracle_bm(SET_IDS[set_name], _p(np.ascontiguousarray(s, np.uint16)), _p(out))
    return out


def decrypt(set_name, sk, ct):
    p = params(set_name)
    e = np.zeros(p.n // 8, np.uint8)
    r = lib().oracle_decrypt(SET_IDS[set_name], _p(np.ascontiguousarray(sk, np.uint8)),
                             _p(np.ascontiguousarray(ct, np.uint8)), _p(e))
    return r, e


def perm_from_controlbits(m, cb):
    pi = np.zeros(1 << m, np.int16)
    lib().oracle_perm_from_controlbits(m, _p(np.ascontiguousarray(cb, np.uint8)), _p(pi))
    return pi


def controlbits(m, pi):
    n = 1 << m
    cb = np.zeros(((2 * m - 1) * n // 2 + 7) // 8, np.uint8)
    lib().oracle_controlbits(m, _p(np.ascontiguousarray(pi, np.int16)), _p(cb))
    return cb


def gen_e_seeded(set_name, seed):
    p = params(set_name)
    e = np.zeros(p.n // 8, np.uint8)
    att = ctypes.c_int(0)
    lib().oracle_rng_seeded(_p(np.frombuffer(bytes(seed), np.uint8).copy()))
    lib().oracle_gen_e(SET_IDS[set_name], _p(e), ctypes.cast(ctypes.pointer(att), ctypes.c_void_p))
    return e, att.value


def xorshift_seeds(n, state=0x9E3779B97F4A7C15):
    """Seeds as the reference harness draws them (mceliece-gpu-params/main.cu:75-76,184-188):
    each seed byte is the low byte of one xorshift64 (13, 7, 17) step.  Returns (seeds, state)."""
    out = np.zeros(n * 32, np.uint8)
    x = state
    M = (1 << 64) - 1
    for i in range(n * 32):
        x ^= (x << 13) & M
        x ^= x >> 7
        x ^= (x << 17) & M
        out[i] = x & 0xFF
    return out.reshape(n, 32), x
