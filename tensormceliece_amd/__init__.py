"""tensormceliece_amd -- Python binding of libtmce.so, the gfx950 batched Classic McEliece KEM.

The shared library is the product; this module is a thin ctypes layer over its C ABI
(include/tmce.h) mirroring libmceliece's crypto_kem interface (keypair / enc / dec per parameter
set, LM:doc/api.md) plus the batched, device-resident entry points.  There is no CPU fallback:
importing works anywhere, but every KEM call needs the built library and a gfx950 GPU and raises
otherwise.
"""
import ctypes
import threading
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TMCE_LIBRARY selects another build of the library (for example one compiled with different
# launch-shape macros) in place of the in-tree libtmce.so
LIB_PATH = os.environ.get("TMCE_LIBRARY") or os.path.join(_HERE, "libtmce.so")

BASE_SETS = {"348864": 0, "460896": 1, "6688128": 2, "6960119": 3, "8192128": 4}
# variant flags (TMCE_VARIANT_F / TMCE_VARIANT_PC): f = semi-systematic key generation,
# pc = plaintext confirmation (ciphertext + 32 bytes); LM:crypto_kem/<set>{,f,pc,pcf}
VARIANT_F, VARIANT_PC = 8, 16
_VARIANT_FLAGS = {"": 0, "f": VARIANT_F, "pc": VARIANT_PC, "pcf": VARIANT_PC | VARIANT_F}
# (m, n, t) and byte sizes (synd = syndrome bytes, ct = ciphertext bytes); sizes as LM:api
_BASE_PARAMS = {
    "348864": dict(m=12, n=3488, t=64, pk=261120, sk=6492, synd=96),
    "460896": dict(m=13, n=4608, t=96, pk=524160, sk=13608, synd=156),
    "6688128": dict(m=13, n=6688, t=128, pk=1044992, sk=13932, synd=208),
    "6960119": dict(m=13, n=6960, t=119, pk=1047319, sk=13948, synd=194),
    "8192128": dict(m=13, n=8192, t=128, pk=1357824, sk=14120, synd=208),
}
# every parameter set and variant by libmceliece's name ("348864", "348864f", "348864pc", ...)
SETS = {b + v: i | fl for b, i in BASE_SETS.items() for v, fl in _VARIANT_FLAGS.items()}
PARAMS = {b + v: dict(_BASE_PARAMS[b], ct=_BASE_PARAMS[b]["synd"] + (32 if "pc" in v else 0))
          for b in BASE_SETS for v in _VARIANT_FLAGS}
KEY_BYTES = 32
# syndrome methods (tmce_syndrome_batch)
SYND_METHODS = {"sparse": 0, "mfma": 1, "valu": 2}
# tuning parameters (tmce_tune_set / tmce_tune_get)
TUNE_PARAMS = {"synd_mfma_cfg": 0, "gen_blocks": 1, "decode_blocks": 2, "front_mfma": 3, "front_cfg": 4, "gen_form": 5,
               "fft_form": 6}


class TmceError(RuntimeError):
    pass


_lib = None


def load():
    """Loads libtmce.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TmceError(f"{LIB_PATH} is missing: run tensormceliece_amd/build.py (or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P, S, I, V = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, None
    sig = {
        "tmce_init": (I, [I]),
        "tmce_last_error": (ctypes.c_char_p, []),
        "tmce_key_from_pk": (I, [I, P, ctypes.POINTER(P)]),
        "tmce_key_from_sk": (I, [I, P, ctypes.POINTER(P)]),
        "tmce_key_free": (V, [P]),
        "tmce_key_device_bytes": (S, [P]),
        "tmce_keypair_seeded": (I, [I, P, P, P]),
        "tmce_keypair_batch": (I, [I, P, P, P, S]),
        "tmce_enc_batch": (I, [P, P, P, P, P, S]),
        "tmce_dec_batch": (I, [P, P, P, P, S]),
        "tmce_enc_batch_device": (I, [P, P, P, P, P, S, P]),
        "tmce_dec_batch_device": (I, [P, P, P, P, S, P]),
        "tmce_dev_alloc": (P, [S]),
        "tmce_dev_free": (V, [P]),
        "tmce_host_alloc": (P, [S]),
        "tmce_set_randombytes": (V, [P]),
        "tmce_host_free": (I, [P]),
        "tmce_memcpy_h2d": (I, [P, P, S]),
        "tmce_memcpy_d2h": (I, [P, P, S]),
        "tmce_memcpy_d2d": (I, [P, P, S]),
        "tmce_sync": (I, []),
        "tmce_timing_enable": (I, [I]),
        "tmce_set_form": (I, [I]),
        "tmce_syndrome_batch": (I, [P, P, P, S, I]),
        "tmce_syndrome_batch_device": (I, [P, P, P, S, I, P]),
        "tmce_tune_set": (I, [I, I, I]),
        "tmce_tune_get": (I, [I, I]),
        "tmce_pipeline_create": (I, [I, I, S, S, I, P, ctypes.POINTER(P)]),
        "tmce_pipeline_run": (I, [P, P, S]),
        "tmce_pipeline_sync": (I, [P]),
        "tmce_pipeline_free": (V, [P]),
        "tmce_pipeline_device_bytes": (S, [P]),
        "tmce_pipeline_capacity": (S, [P]),
        "tmce_pipeline_stream": (P, [P]),
        "tmce_server_create": (I, [I, S, I, S, ctypes.POINTER(P)]),
        "tmce_server_submit": (I, [P, P, ctypes.POINTER(ctypes.c_ulonglong)]),
        "tmce_server_wait": (I, [P, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_double)]),
        "tmce_server_poll": (I, [P, P, P, I]),
        "tmce_server_get_stats": (I, [P, P]),
        "tmce_server_free": (V, [P]),
        "tmce_server_run_load": (I, [P, P, P, S, P, P]),
        "tmce_timing_read": (I, [I, P, P, P]),
        "tmce_stream_create": (I, [ctypes.POINTER(P)]),
        "tmce_stream_destroy": (I, [P]),
        "tmce_stream_sync": (I, [P]),
        "tmce_event_create": (I, [ctypes.POINTER(P)]),
        "tmce_event_destroy": (I, [P]),
        "tmce_event_record": (I, [P, P]),
        "tmce_stream_wait_event": (I, [P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc, what):
    if rc != 0:
        msg = load().tmce_last_error()
        raise TmceError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def init(device=0):
    _check(load().tmce_init(device), "tmce_init")


def set_randombytes(fn_address=None):
    """randombytes() of the libmceliece entry points (kem_keypair / kem_enc): the address of a C
    function void(void *out, long long len), or None for getrandom()."""
    load().tmce_set_randombytes(fn_address)


def entry_address(set_name, fn):
    """Address of mceliece_kem_<set>_<fn> (keypair, enc, dec) in libtmce.so."""
    return ctypes.cast(getattr(load(), f"mceliece_kem_{set_name}_{fn}"), ctypes.c_void_p).value


def keypair_seeded(set_name, seed):
    """libmceliece keypair() with randombytes = the seeded chain on `seed` (32 bytes)."""
    p = PARAMS[set_name]
    pk = np.zeros(p["pk"], np.uint8)
    sk = np.zeros(p["sk"], np.uint8)
    sd = np.frombuffer(bytes(seed), np.uint8).copy()
    assert sd.size == 32
    _check(load().tmce_keypair_seeded(SETS[set_name], _ptr(pk), _ptr(sk), _ptr(sd)), "keypair")
    return pk, sk


def keypair_batch(set_name, seeds):
    """Batched keypairs on the GPU: row i equals keypair_seeded(set_name, seeds[i]) byte for byte.
    seeds: (n, 32) uint8.  Returns (pk (n, PUBLICKEYBYTES), sk (n, SECRETKEYBYTES))."""
    p = PARAMS[set_name]
    sd = np.ascontiguousarray(seeds, dtype=np.uint8).reshape(-1, 32)
    n = sd.shape[0]
    pk = np.zeros((n, p["pk"]), np.uint8)
    sk = np.zeros((n, p["sk"]), np.uint8)
    _check(load().tmce_keypair_batch(SETS[set_name], _ptr(pk), _ptr(sk), _ptr(sd), n), "tmce_keypair_batch")
    return pk, sk


class _Key:
    def __init__(self, set_name, raw, kind):
        self.set_name = set_name
        self.params = PARAMS[set_name]
        self._h = ctypes.c_void_p()
        arr = np.ascontiguousarray(np.frombuffer(bytes(raw), np.uint8) if not isinstance(raw, np.ndarray) else raw,
                                   dtype=np.uint8)
        fn = load().tmce_key_from_pk if kind == "pk" else load().tmce_key_from_sk
        _check(fn(SETS[set_name], _ptr(arr), ctypes.byref(self._h)), f"tmce_key_from_{kind}")

    @property
    def handle(self):
        return self._h

    def device_bytes(self):
        return load().tmce_key_device_bytes(self._h)

    def close(self):
        if self._h:
            load().tmce_key_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PublicKey(_Key):
    """Device-resident public key; enc(seeds) runs batched encapsulation."""

    def __init__(self, set_name, pk):
        super().__init__(set_name, pk, "pk")

    def enc(self, seeds):
        seeds = np.ascontiguousarray(seeds, dtype=np.uint8).reshape(-1, 32)
        n = seeds.shape[0]
        ct = np.zeros((n, self.params["ct"]), np.uint8)
        k = np.zeros((n, KEY_BYTES), np.uint8)
        st = np.zeros(n, np.int32)
        _check(load().tmce_enc_batch(self._h, _ptr(ct), _ptr(k), _ptr(st), _ptr(seeds), n), "tmce_enc_batch")
        return ct, k, st


    def syndrome(self, e, method="sparse"):
        """ct_i = H e_i for rows of e (n/8 bytes each, any weight) -- the PKE encryption."""
        e = np.ascontiguousarray(e, dtype=np.uint8).reshape(-1, self.params["n"] // 8)
        n = e.shape[0]
        ct = np.zeros((n, self.params["synd"]), np.uint8)
        _check(load().tmce_syndrome_batch(self._h, _ptr(ct), _ptr(e), n, SYND_METHODS.get(method, method)),
               "tmce_syndrome_batch")
        return ct


class SecretKey(_Key):
    """Device-resident secret key; dec(ct) runs batched decapsulation."""

    def __init__(self, set_name, sk):
        super().__init__(set_name, sk, "sk")

    def dec(self, ct):
        ct = np.ascontiguousarray(ct, dtype=np.uint8).reshape(-1, self.params["ct"])
        n = ct.shape[0]
        k = np.zeros((n, KEY_BYTES), np.uint8)
        st = np.zeros(n, np.int32)
        _check(load().tmce_dec_batch(self._h, _ptr(k), _ptr(st), _ptr(ct), n), "tmce_dec_batch")
        return k, st


class HostBuffer:
    """Pinned host memory mapped into the GPU's address space (tmce_host_alloc).  Serving requests
    whose arrays are views of a HostBuffer (array()) are served in place, without staging copies.
    Keep it alive, and unfreed, while requests using it are in flight."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = load().tmce_host_alloc(self.nbytes)
        if not self.ptr:
            raise TmceError(f"tmce_host_alloc failed: {load().tmce_last_error().decode()}")
        self._off = 0

    def array(self, shape, dtype=np.uint8, align=16):
        """The next unused bytes of the buffer as a numpy array (aligned to `align` bytes)."""
        dt = np.dtype(dtype)
        nb = int(np.prod(shape)) * dt.itemsize
        off = (self._off + align - 1) // align * align
        if off + nb > self.nbytes:
            raise TmceError("HostBuffer exhausted")
        self._off = off + nb
        raw = (ctypes.c_uint8 * nb).from_address(self.ptr + off)
        return np.frombuffer(raw, dtype=dt).reshape(shape)

    def reset(self):
        self._off = 0

    def free(self):
        if self.ptr:
            load().tmce_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory owned through the library (for benchmarks with inputs resident in HBM)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = load().tmce_dev_alloc(self.nbytes)
        if not self.ptr:
            raise TmceError("tmce_dev_alloc failed")

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(load().tmce_memcpy_h2d(self.ptr, _ptr(arr), arr.nbytes), "h2d")

    def download(self, shape, dtype=np.uint8):
        out = np.zeros(shape, dtype)
        _check(load().tmce_memcpy_d2h(_ptr(out), self.ptr, out.nbytes), "d2h")
        return out

    def copy_from(self, other, nbytes):
        _check(load().tmce_memcpy_d2d(self.ptr, other.ptr, nbytes), "d2d")

    def free(self):
        if self.ptr:
            load().tmce_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def enc_device(pk, d_ct, d_k, d_status, d_seeds, n, stream=None):
    _check(load().tmce_enc_batch_device(pk.handle, d_ct.ptr, d_k.ptr, d_status.ptr if d_status else None,
                                        d_seeds.ptr, n, stream), "tmce_enc_batch_device")


def dec_device(sk, d_k, d_status, d_ct, n, stream=None):
    _check(load().tmce_dec_batch_device(sk.handle, d_k.ptr, d_status.ptr if d_status else None, d_ct.ptr, n,
                                        stream), "tmce_dec_batch_device")


def syndrome_device(pk, d_ct, d_e, n, method="sparse", stream=None):
    _check(load().tmce_syndrome_batch_device(pk.handle, d_ct.ptr, d_e.ptr, n, SYND_METHODS.get(method, method),
                                             stream), "tmce_syndrome_batch_device")


def tune_set(set_name, param, value):
    _check(load().tmce_tune_set(SETS[set_name], TUNE_PARAMS[param], int(value)), f"tmce_tune_set({param})")


def tune_get(set_name, param):
    v = load().tmce_tune_get(SETS[set_name], TUNE_PARAMS[param])
    if v < 0:
        _check(v, f"tmce_tune_get({param})")
    return v


class SubBatchC(ctypes.Structure):
    """tmce_sub_batch (include/tmce.h)."""
    _fields_ = [("key", ctypes.c_void_p), ("inp", ctypes.c_void_p), ("ct", ctypes.c_void_p),
                ("k", ctypes.c_void_p), ("status", ctypes.c_void_p), ("n", ctypes.c_size_t)]


class Pipeline:
    """Multi-batch pipeline (tmce_pipeline_*): sub-batches of mixed sizes under up to 16 keys of one
    parameter set per run, one set of launches, optionally replayed from a captured HIP graph.
    direction: "enc" or "dec".  Runs are asynchronous on the pipeline's stream."""

    def __init__(self, set_name, direction, capacity, max_sub=4096, use_graph=True, stream=None):
        self.set_name = set_name
        self.direction = {"enc": 0, "dec": 1}[direction]
        self._h = ctypes.c_void_p()
        _check(load().tmce_pipeline_create(SETS[set_name], self.direction, int(capacity), int(max_sub),
                                           1 if use_graph else 0, stream, ctypes.byref(self._h)),
               "tmce_pipeline_create")

    @property
    def capacity(self):
        return load().tmce_pipeline_capacity(self._h)

    def device_bytes(self):
        return load().tmce_pipeline_device_bytes(self._h)

    @property
    def stream(self):
        return load().tmce_pipeline_stream(self._h)

    def run(self, subs):
        """subs: iterable of (key handle object, in_ptr, ct_ptr or None, k_ptr, status_ptr or None, n)
        with device addresses as ints."""
        subs = list(subs)
        arr = (SubBatchC * max(1, len(subs)))()
        addr = lambda x: None if x is None else int(x)
        for i, (key, inp, ct, k, st, n) in enumerate(subs):
            arr[i] = SubBatchC(key.handle, addr(inp), addr(ct), addr(k), addr(st), int(n))
        _check(load().tmce_pipeline_run(self._h, ctypes.cast(arr, ctypes.c_void_p), len(subs)), "tmce_pipeline_run")

    def sync(self):
        _check(load().tmce_pipeline_sync(self._h), "tmce_pipeline_sync")

    def close(self):
        if self._h:
            load().tmce_pipeline_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RequestC(ctypes.Structure):
    """tmce_request (include/tmce.h)."""
    _fields_ = [("direction", ctypes.c_int), ("key", ctypes.c_void_p), ("n", ctypes.c_size_t),
                ("inp", ctypes.c_void_p), ("ct", ctypes.c_void_p), ("k", ctypes.c_void_p),
                ("status", ctypes.c_void_p)]


class ServerStatsC(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_ulonglong), ("requests", ctypes.c_ulonglong), ("ops", ctypes.c_ulonglong),
                ("pending", ctypes.c_ulonglong), ("in_flight_bytes", ctypes.c_size_t),
                ("peak_in_flight_bytes", ctypes.c_size_t), ("enc_bytes_per_op", ctypes.c_size_t),
                ("dec_bytes_per_op", ctypes.c_size_t), ("zero_copy_requests", ctypes.c_ulonglong),
                ("max_batch_keys", ctypes.c_ulonglong)]


class Server:
    """Serving loop (tmce_server_*): continuous batching of enc/dec requests of one parameter set.
    submit() takes host numpy arrays; they are kept alive here until the request is collected."""

    def __init__(self, set_name, slot_capacity=16384, slots_per_direction=2, hbm_budget=0):
        self.set_name = set_name
        self.params = PARAMS[set_name]
        self._h = ctypes.c_void_p()
        self._live = {}
        # submit() registers a request's arrays under this lock in the same critical section as the
        # C submit, so a poll() on another thread cannot collect the id before it is registered
        self._mu = threading.Lock()
        _check(load().tmce_server_create(SETS[set_name], int(slot_capacity), int(slots_per_direction),
                                         int(hbm_budget), ctypes.byref(self._h)), "tmce_server_create")

    def submit(self, direction, key, inp, ct=None, k=None, status=None, slab=None):
        """direction "enc" (inp = seeds [n][32]) or "dec" (inp = ciphertexts); returns (id, outputs).
        With `slab` (a HostBuffer) the outputs not given are allocated in it; when inp and every
        output lie in slabs the request is served in place (zero-copy)."""
        d = {"enc": 0, "dec": 1}[direction]
        inp = np.ascontiguousarray(inp, np.uint8)
        n = inp.shape[0]
        new = (lambda shape, dt: slab.array(shape, dt)) if slab is not None else (lambda shape, dt: np.zeros(shape, dt))
        if k is None:
            k = new((n, KEY_BYTES), np.uint8)
        if status is None:
            status = new(n, np.int32)
        if d == 0 and ct is None:
            ct = new((n, self.params["ct"]), np.uint8)
        rq = RequestC(d, key.handle.value, n, inp.ctypes.data, ct.ctypes.data if ct is not None else None,
                      k.ctypes.data, status.ctypes.data)
        rid = ctypes.c_ulonglong()
        with self._mu:
            _check(load().tmce_server_submit(self._h, ctypes.byref(rq), ctypes.byref(rid)), "tmce_server_submit")
            self._live[rid.value] = (key, inp, ct, k, status)  # the key handle must outlive the request
        return rid.value, (ct, k, status)

    def wait(self, rid):
        lat = ctypes.c_double()
        _check(load().tmce_server_wait(self._h, rid, ctypes.byref(lat)), "tmce_server_wait")
        with self._mu:
            self._live.pop(rid, None)
        return lat.value

    def poll(self, max_items=4096):
        ids = np.zeros(max_items, np.uint64)
        lats = np.zeros(max_items, np.float64)
        with self._mu:
            n = load().tmce_server_poll(self._h, _ptr(ids), _ptr(lats), max_items)
            if n < 0:
                _check(n, "tmce_server_poll")
            for i in ids[:n]:
                self._live.pop(int(i), None)
        return ids[:n].tolist(), lats[:n].tolist()

    def run_load(self, requests, arrival_s):
        """Open-loop load through the native driver (tmce_server_run_load): requests[i] =
        (direction, key, inputs, ct, k, status) with output arrays given, submitted at arrival_s[i]
        seconds after the start (sorted).  Blocks until all complete; returns (latency_s array,
        elapsed_s)."""
        n = len(requests)
        arr = (RequestC * max(n, 1))()
        keep = []
        for i, (direction, key, inp, ct, k, status) in enumerate(requests):
            inp = np.ascontiguousarray(inp, np.uint8)
            keep.append((key, inp, ct, k, status))
            arr[i] = RequestC({"enc": 0, "dec": 1}[direction], key.handle.value, inp.shape[0], inp.ctypes.data,
                              ct.ctypes.data if ct is not None else None, k.ctypes.data,
                              status.ctypes.data if status is not None else None)
        arrivals = np.ascontiguousarray(arrival_s, np.float64)
        lat = np.zeros(n, np.float64)
        el = ctypes.c_double()
        _check(load().tmce_server_run_load(self._h, ctypes.cast(arr, ctypes.c_void_p), _ptr(arrivals), n, _ptr(lat),
                                           ctypes.byref(el)), "tmce_server_run_load")
        del keep
        return lat, el.value

    def stats(self):
        st = ServerStatsC()
        _check(load().tmce_server_get_stats(self._h, ctypes.byref(st)), "tmce_server_get_stats")
        return {f: getattr(st, f) for f, _ in ServerStatsC._fields_}

    def close(self):
        if self._h:
            load().tmce_server_free(self._h)
            self._h = ctypes.c_void_p()
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_form(form):
    """0 = engine form (default), 1 = plain one-launch-per-stage form (the baseline)."""
    _check(load().tmce_set_form({"engine": 0, "plain": 1}.get(form, form)), "tmce_set_form")


def sync():
    _check(load().tmce_sync(), "tmce_sync")


class Stream:
    """A non-blocking HIP stream (default priority) for the asynchronous device entry points; pass
    it as `stream=` (its handle converts implicitly)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(load().tmce_stream_create(ctypes.byref(h)), "tmce_stream_create")
        self._as_parameter_ = h

    def sync(self):
        _check(load().tmce_stream_sync(self._as_parameter_), "tmce_stream_sync")

    def wait(self, event):
        """Later work on this stream waits for the work recorded in `event`."""
        _check(load().tmce_stream_wait_event(self._as_parameter_, event._as_parameter_), "tmce_stream_wait_event")

    def free(self):
        if self._as_parameter_:
            load().tmce_stream_destroy(self._as_parameter_)
            self._as_parameter_ = None


class Event:
    """A HIP event (no timing) recorded on a stream."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(load().tmce_event_create(ctypes.byref(h)), "tmce_event_create")
        self._as_parameter_ = h

    def record(self, stream=None):
        _check(load().tmce_event_record(self._as_parameter_, stream), "tmce_event_record")

    def free(self):
        if self._as_parameter_:
            load().tmce_event_destroy(self._as_parameter_)
            self._as_parameter_ = None


def timing_enable(on=True):
    _check(load().tmce_timing_enable(1 if on else 0), "timing_enable")


def timing_read():
    """{kernel name: (launch count, total ms)} of the engine's kernels since timing_enable(True)."""
    mx = 64
    names = ctypes.create_string_buffer(48 * mx)
    counts = np.zeros(mx, np.int64)
    tot = np.zeros(mx, np.float64)
    n = load().tmce_timing_read(mx, names, _ptr(counts), _ptr(tot))
    if n < 0:
        raise TmceError("timing_read failed")
    raw = names.raw
    out = {}
    for i in range(n):
        nm = raw[48 * i:48 * (i + 1)].split(b"\0", 1)[0].decode()
        out[nm] = (int(counts[i]), float(tot[i]))
    return out


# ---- libmceliece-style single-operation entry points (mceliece_kem_<set>_{keypair,enc,dec})
def kem_keypair(set_name):
    p = PARAMS[set_name]
    pk = np.zeros(p["pk"], np.uint8)
    sk = np.zeros(p["sk"], np.uint8)
    f = getattr(load(), f"mceliece_kem_{set_name}_keypair")
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    f.restype = None
    f(_ptr(pk), _ptr(sk))
    return pk, sk


def kem_enc(set_name, pk):
    p = PARAMS[set_name]
    ct = np.zeros(p["ct"], np.uint8)
    k = np.zeros(KEY_BYTES, np.uint8)
    pk = np.ascontiguousarray(pk, np.uint8)
    f = getattr(load(), f"mceliece_kem_{set_name}_enc")
    f.argtypes = [ctypes.c_void_p] * 3
    f.restype = ctypes.c_int
    r = f(_ptr(ct), _ptr(k), _ptr(pk))
    return r, ct, k


def kem_dec(set_name, ct, sk):
    k = np.zeros(KEY_BYTES, np.uint8)
    ct = np.ascontiguousarray(ct, np.uint8)
    sk = np.ascontiguousarray(sk, np.uint8)
    f = getattr(load(), f"mceliece_kem_{set_name}_dec")
    f.argtypes = [ctypes.c_void_p] * 3
    f.restype = ctypes.c_int
    r = f(_ptr(k), _ptr(ct), _ptr(sk))
    return r, k
