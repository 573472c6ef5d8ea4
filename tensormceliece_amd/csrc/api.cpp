// api.cpp -- the C ABI of libtmce.so (declared in include/tmce.h): engine context, device-resident
// key handles, batched encapsulation/decapsulation and the libmceliece-compatible entry points.
#include <hip/hip_runtime.h>
#include <sys/random.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"

namespace tmce {
std::atomic<KernelTimer *> g_ktimer{nullptr};
thread_local bool g_ktimer_off = false;
}

using namespace tmce;

namespace tmce {
thread_local std::string g_err;
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
Engine *g_engine = nullptr;
static std::mutex g_engine_mu;

int engine_init(int device) {
  std::lock_guard<std::mutex> lk(g_engine_mu);
  if (g_engine) return g_engine->device == device ? 0 : fail(TMCE_EINVAL, "engine already bound to another device");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return fail(TMCE_ENODEV, "no HIP device");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(TMCE_ENODEV, std::string("libtmce is built for gfx950, device is ") + prop.gcnArchName);
  HIPCHK(hipSetDevice(device));
  auto e = std::make_unique<Engine>();
  e->device = device;
  e->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&e->use_enc.done, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&e->use_dec.done, hipEventDisableTiming));
  for (int w = 0; w < 2; ++w) {
    const HostField &F = host_field(12 + w);
    std::vector<uint16_t> ex(F.ex.begin(), F.ex.begin() + F.q);
    if (int r = upload(&e->lg[w], F.lg)) return r;
    if (int r = upload(&e->ex[w], ex)) return r;
  }
  for (int s = 0; s < NUM_SETS; ++s) {
    const SetInfo &si = set_info(s);
    const int D = si.t == 64 ? 64 : 128;
    FftTables t = build_fft_tables(si.m, D);
    if (int r = upload(&e->st[s].twl, t.twl)) return r;
    if (int r = upload(&e->st[s].plo, t.plo)) return r;
    if (int r = upload(&e->st[s].phi, t.phi)) return r;
    if (si.t == D) {  // a -> a^t is additive for t = 2^k: a^t = (lane)^t + (64 r)^t
      const HostField &F = host_field(si.m);
      std::vector<uint16_t> lo(64), hi(size_t(1) << (si.m - 6));
      for (size_t l = 0; l < 64; ++l) lo[l] = F.pw((uint16_t)l, (uint32_t)si.t);
      for (size_t r = 0; r < hi.size(); ++r) hi[r] = F.pw((uint16_t)(64 * r), (uint32_t)si.t);
      if (int r = upload(&e->st[s].ptlo, lo)) return r;
      if (int r = upload(&e->st[s].pthi, hi)) return r;
    }
  }
  if (int r = noalloc_init(device)) return r;
  g_engine = e.release();
  return 0;
}

Engine *engine() {
  if (!g_engine && engine_init(0) != 0) return nullptr;
  return g_engine;
}

// pinned host slabs: host range -> device address of its start
struct Slab {
  uintptr_t hi;
  uint8_t *dev;
};
static std::mutex g_slab_mu;
static std::map<uintptr_t, Slab> g_slabs;  // keyed by the slab's first host address

uint8_t *slab_device_ptr(const void *p, size_t len) {
  if (!p) return nullptr;
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  std::lock_guard<std::mutex> lk(g_slab_mu);
  auto it = g_slabs.upper_bound(a);
  if (it == g_slabs.begin()) return nullptr;
  --it;
  if (a < it->first || a + len > it->second.hi || a + len < a) return nullptr;
  return it->second.dev + (a - it->first);
}

}  // namespace tmce

namespace {


// ---------------------------------------------------------------- randomness
tmce_randombytes_fn g_randombytes = nullptr;

void sys_randombytes(void *out, long long len) {
  uint8_t *p = static_cast<uint8_t *>(out);
  while (len > 0) {
    ssize_t r = getrandom(p, (size_t)(len > 256 ? 256 : len), 0);
    if (r < 0) {
      if (errno == EINTR) continue;
      std::abort();
    }
    p += r;
    len -= r;
  }
}

void randombytes(void *out, long long len) {
  if (g_randombytes) g_randombytes(out, len);
  else sys_randombytes(out, len);
}

void rnd_global(uint8_t *out, long long len, void *) { randombytes(out, len); }

// the seeded chain: state = 0x40||seed; out = SHAKE256(state,len); state[1..32] = out[0..31]
struct SeedChain { uint8_t state[33]; };
void rnd_chain(uint8_t *out, long long len, void *ctx) {
  SeedChain *c = static_cast<SeedChain *>(ctx);
  shake256(out, len, c->state, 33);
  std::memcpy(c->state + 1, out, (size_t)(len < 32 ? len : 32));
}

// ---------------------------------------------------------------- batched operations
// Orders a call's use of a scratch set on stream s after the set's previous use on another stream;
// scratch_done records the use.
int scratch_wait(ScratchUse &u, hipStream_t s) {
  if (u.last && u.last != s) HIPCHK(hipStreamWaitEvent(s, u.done, 0));
  return 0;
}
int scratch_done(ScratchUse &u, hipStream_t s) {
  HIPCHK(hipEventRecord(u.done, s));
  u.last = s;
  return 0;
}

int enc_device(tmce_key *k, uint8_t *d_ct, uint8_t *d_k, int *d_status, const uint8_t *d_seeds, size_t n,
               hipStream_t s) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  if (!k || k->kind != 1) return fail(TMCE_EINVAL, "encapsulation needs a public-key handle");
  if (n == 0) return 0;
  if (n > (size_t)INT32_MAX / 2) return fail(TMCE_EINVAL, "batch too large");
  const SetInfo &si = set_info(k->set);
  if (!s) s = e->stream;
  if (k->pc && e->form == 1) return fail(TMCE_EINVAL, "the plain form covers the variants without plaintext confirmation");
  if (k->pad_bad) {  // LM:crypto_kem/6960119/vec/kem_enc.c: outputs cleared, return -1
    HIPCHK(launch_fill_u8(d_ct, 0, n * key_ct_bytes(k), s));
    HIPCHK(launch_fill_u8(d_k, 0, n * 32, s));
    if (d_status) HIPCHK(launch_fill_i32(d_status, -1, n, s));
    return 0;
  }
  const int MSGW = si.msg_words;
  if (int r = scratch_wait(e->use_enc, s)) return r;
  if (e->form == 1) {
    const int NEW = (si.n + 31) / 32, KWE = (si.k + 31) / 32, CW = (si.mt + 31) / 32;
    if (int r = e->p_erow_enc.ensure(n * NEW * 4)) return r;
    if (int r = e->p_ens.ensure(n * KWE * 4)) return r;
    if (int r = e->p_s.ensure(n * CW * 4)) return r;
    if (int r = e->msg_enc.ensure(n * MSGW * sizeof(uint32_t))) return r;
    PlainEncArgs pa{d_seeds, k->Trows, d_ct, d_k, d_status, e->p_erow_enc.as<uint32_t>(), e->p_ens.as<uint32_t>(),
                    e->p_s.as<uint32_t>(), e->msg_enc.as<uint32_t>(), (int)n};
    HIPCHK(launch_plain_enc_set(k->set, pa, s));
    return scratch_done(e->use_enc, s);
  }
  if (int r = e->ind.ensure(n * ((si.t + 1) & ~1) * sizeof(uint16_t))) return r;
  if (int r = e->msg_enc.ensure(n * MSGW * sizeof(uint32_t))) return r;
  if (int r = e->queue.ensure(64)) return r;
  EncArgs a{};
  a.seeds = d_seeds;
  a.cols = k->cols;
  a.ct = d_ct;
  a.key = d_k;
  a.status = d_status;
  a.ind = e->ind.as<uint16_t>();
  a.msg = e->msg_enc.as<uint32_t>();
  a.attempts = nullptr;
  a.queue = e->queue.as<unsigned>();
  a.nops = (int)n;
  a.gen_blocks = e->gen_blocks[k->set];
  a.gen_form = e->gen_form[k->set];
  a.pc = k->pc ? 1 : 0;
  HIPCHK(launch_enc_set(k->set, a, s));
  return scratch_done(e->use_enc, s);
}

}  // namespace

namespace tmce {

// Device scratch of one decapsulation batch of n ops of `set` (pc: the variants' syndrome copy),
// laid out back to back at 256-byte boundaries from a base address: the decoder's work counters,
// the syndrome planes, the message rows, the pc syndrome copy and the MFMA front end's repacked
// ciphertext words.  Returns the total bytes; with `base` non-null fills `x`.
size_t dec_scratch_layout(int set, bool pc, bool front_mfma, size_t n, uint8_t *base, DecScratch *x) {
  const SetInfo &si = set_info(set);
  const size_t TT = ((2 * (size_t)si.t + 31) / 32) * 32;
  const size_t RW = (size_t)si.m * TT / 32;
  auto up = [](size_t a) { return (a + 255) & ~(size_t)255; };
  size_t off = 0;
  const size_t q = off;
  off = up(off + 1024);
  const size_t sS = off;
  off = up(off + n * RW * sizeof(uint32_t));
  const size_t sm = off;
  off = up(off + n * (size_t)si.msg_words * sizeof(uint32_t));
  const size_t sc = off;
  if (pc) off = up(off + n * (size_t)si.synd_bytes);
  const size_t xb = front_mfma ? gf2_scratch_bytes(set, GF2_FRONT, n) : 0;
  const size_t sx = off;
  off = up(off + xb);
  if (base && x) {
    x->queue = reinterpret_cast<unsigned *>(base + q);
    x->S = reinterpret_cast<uint32_t *>(base + sS);
    x->msg = reinterpret_cast<uint32_t *>(base + sm);
    x->ct_synd = pc ? base + sc : nullptr;
    x->xs = xb ? reinterpret_cast<uint32_t *>(base + sx) : nullptr;
  }
  return off;
}

// Launches one decapsulation batch on stream s with the given scratch (engine form when form is
// 0, the plain per-stage form when 1: its extra scratch in `px`).  No allocation.
int dec_launch(Engine *e, tmce_key *k, uint8_t *d_k, int *d_status, const uint8_t *d_ct, size_t n, hipStream_t s,
               const DecScratch &x, int form, const PlainDecScratch *px) {
  const SetInfo &si = set_info(k->set);
  DecArgs a{};
  a.ct = d_ct;
  a.pc = k->pc ? 1 : 0;
  a.ct_synd = x.ct_synd;
  a.F = k->F;
  a.Fsw = k->Fsw;
  a.front_mfma = e->front_mfma[k->set];
  a.front_cfg = e->front_cfg[k->set];
  a.xscratch = x.xs;
  a.Fcol = k->Fcol;
  a.alpha = k->alpha;
  a.la = k->la;
  a.lu = k->lu;
  a.ainv = k->ainv;
  a.supp = k->supp;
  a.rej = k->rej;
  const int w = si.m - 12;
  const SetTables &tb = e->st[k->set];
  a.ft = FieldTables{e->lg[w], e->ex[w], tb.twl, tb.plo, tb.phi, tb.ptlo, tb.pthi};
  a.key = d_k;
  a.status = d_status;
  a.S = x.S;
  a.msg = x.msg;
  a.queue = x.queue;
  a.nops = (int)n;
  a.decode_blocks = e->decode_blocks[k->set];
  a.cus = e->cus;
  a.fft_form = e->fft_form[k->set];
  if (form == 1) {
    HIPCHK(launch_plain_dec_set(k->set, a, *px, s));
    return 0;
  }
  HIPCHK(launch_dec_set(k->set, a, s));
  return 0;
}

}  // namespace tmce

namespace {

int dec_device(tmce_key *k, uint8_t *d_k, int *d_status, const uint8_t *d_ct, size_t n, hipStream_t s) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  if (!k || k->kind != 2) return fail(TMCE_EINVAL, "decapsulation needs a secret-key handle");
  if (n == 0) return 0;
  if (n > (size_t)INT32_MAX / 2) return fail(TMCE_EINVAL, "batch too large");
  const SetInfo &si = set_info(k->set);
  if (!s) s = e->stream;
  const int TT = ((2 * si.t + 31) / 32) * 32;
  const int RW = si.m * TT / 32;
  const int MSGW = si.msg_words;
  if (k->pc && e->form == 1) return fail(TMCE_EINVAL, "the plain form covers the variants without plaintext confirmation");
  if (int r = scratch_wait(e->use_dec, s)) return r;
  if (int r = e->S.ensure(n * RW * sizeof(uint32_t))) return r;
  if (k->pc)
    if (int r = e->ct_synd.ensure(n * si.synd_bytes)) return r;
  if (int r = e->msg_dec.ensure(n * MSGW * sizeof(uint32_t))) return r;
  if (int r = e->dqueue.ensure(1024)) return r;
  DecScratch x{};
  x.S = e->S.as<uint32_t>();
  x.msg = e->msg_dec.as<uint32_t>();
  x.queue = e->dqueue.as<unsigned>();
  x.ct_synd = k->pc ? e->ct_synd.as<uint8_t>() : nullptr;
  if (e->front_mfma[k->set]) {
    const size_t xb = gf2_scratch_bytes(k->set, GF2_FRONT, n);
    if (xb) {
      if (int r = e->xs_dec.ensure(xb)) return r;
      x.xs = e->xs_dec.as<uint32_t>();
    }
  }
  PlainDecScratch px{};
  if (e->form == 1) {
    const int NEW2 = 2 * ((si.n + 63) / 64);
    if (int r = e->p_loc.ensure(n * (si.t + 1) * 2)) return r;
    if (int r = e->p_erow_dec.ensure(n * NEW2 * 4)) return r;
    if (int r = e->p_weight.ensure(n * 4)) return r;
    px = PlainDecScratch{e->p_loc.as<uint16_t>(), e->p_erow_dec.as<uint32_t>(), e->p_weight.as<int>()};
  }
  if (int r = dec_launch(e, k, d_k, d_status, d_ct, n, s, x, e->form, &px)) return r;
  return scratch_done(e->use_dec, s);
}

// operands of the dense syndrome forms, built from the T rows on first use
int ensure_synd_key(Engine *e, tmce_key *k, int method, hipStream_t s) {
  const SetInfo &si = set_info(k->set);
  bool built = false;
  if (method == TMCE_SYND_MFMA && !k->Tb) {
    const size_t nb = gf2_msw_bytes(k->set, GF2_SYND);
    if (hipMalloc((void **)&k->Tb, nb) != hipSuccess) return fail(TMCE_ENOMEM, "hipMalloc Tb");
    HIPCHK(launch_build_msw(k->set, GF2_SYND, k->Trows, (si.k + 31) / 32, k->Tb, s));
    k->bytes += nb;
    built = true;
  }
  if (method == TMCE_SYND_VALU && !k->Tv) {
    const size_t nv = (size_t)((si.mt + 31) / 32) * ((si.k + 31) / 32) * 32;
    if (hipMalloc((void **)&k->Tv, nv * 4) != hipSuccess) return fail(TMCE_ENOMEM, "hipMalloc Tv");
    HIPCHK(launch_build_tv(k->set, k->Trows, k->Tv, s));
    k->bytes += nv * 4;
    built = true;
  }
  (void)e;
  // key material is shared by every stream: a newly built operand is complete before the handle
  // is used again
  if (built) HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int synd_device(tmce_key *k, uint8_t *d_ct, const uint8_t *d_e, size_t n, int method, hipStream_t s) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  if (!k || k->kind != 1) return fail(TMCE_EINVAL, "the syndrome needs a public-key handle");
  if (method < TMCE_SYND_SPARSE || method > TMCE_SYND_VALU) return fail(TMCE_EINVAL, "unknown syndrome method");
  if (n == 0) return 0;
  if (n > (size_t)INT32_MAX / 2) return fail(TMCE_EINVAL, "batch too large");
  if (!s) s = e->stream;
  if (int r = ensure_synd_key(e, k, method, s)) return r;
  if (int r = scratch_wait(e->use_enc, s)) return r;
  uint32_t *xs = nullptr;
  if (method == TMCE_SYND_MFMA) {
    const size_t xb = gf2_scratch_bytes(k->set, GF2_SYND, n);
    if (xb) {
      if (int r = e->xs_synd.ensure(xb)) return r;
      xs = e->xs_synd.as<uint32_t>();
    }
  }
  SyndArgs a{d_e, k->Tb, k->Tv, k->cols, d_ct, xs, (int)n};
  HIPCHK(launch_synd_set(k->set, a, method, e->synd_cfg[k->set], s));
  return scratch_done(e->use_enc, s);
}

// Encapsulation of given error supports (host buffers, pos[i*t .. i*t+t)): the rows the sampler
// would write -- the support and the message prefix [1][e] -- are built here, and the finishing
// kernel (syndrome by column gather, confirmation, session key) runs on them.
int enc_support(tmce_key *k, uint8_t *ct, uint8_t *kk, int *status, const uint16_t *pos, size_t n) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  if (!k || k->kind != 1) return fail(TMCE_EINVAL, "encapsulation needs a public-key handle");
  if (n == 0) return 0;
  std::lock_guard<std::mutex> lk(e->mu);
  const SetInfo &si = set_info(k->set);
  const size_t ctb = key_ct_bytes(k);
  if (k->pad_bad) {  // LM:crypto_kem/6960119/vec/kem_enc.c: outputs cleared, return -1
    std::memset(ct, 0, n * ctb);
    std::memset(kk, 0, n * 32);
    if (status)
      for (size_t i = 0; i < n; ++i) status[i] = -1;
    return 0;
  }
  const size_t IS = (size_t)((si.t + 1) & ~1), MW = (size_t)si.msg_words;
  std::vector<uint16_t> ind(n * IS, 0);
  std::vector<uint32_t> msg(n * MW, 0);
  for (size_t i = 0; i < n; ++i) {
    uint8_t *m8 = reinterpret_cast<uint8_t *>(&msg[i * MW]);
    m8[0] = 1;
    for (int j = 0; j < si.t; ++j) {
      const uint16_t p = pos[i * si.t + j];
      ind[i * IS + j] = p;
      m8[1 + p / 8] |= (uint8_t)(1u << (p % 8));
    }
  }
  hipStream_t s = e->stream;
  if (int r = scratch_wait(e->use_enc, s)) return r;
  if (int r = e->ind.ensure(ind.size() * sizeof(uint16_t))) return r;
  if (int r = e->msg_enc.ensure(msg.size() * sizeof(uint32_t))) return r;
  if (int r = e->ct.ensure(n * ctb)) return r;
  if (int r = e->key.ensure(n * 32)) return r;
  if (int r = e->status.ensure(n * sizeof(int))) return r;
  HIPCHK(hipMemcpyAsync(e->ind.p, ind.data(), ind.size() * sizeof(uint16_t), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(e->msg_enc.p, msg.data(), msg.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  EncArgs a{};
  a.cols = k->cols;
  a.ct = e->ct.as<uint8_t>();
  a.key = e->key.as<uint8_t>();
  a.status = e->status.as<int>();
  a.ind = e->ind.as<uint16_t>();
  a.msg = e->msg_enc.as<uint32_t>();
  a.nops = (int)n;
  a.pc = k->pc ? 1 : 0;
  HIPCHK(launch_enc_finish_set(k->set, a, s));
  if (int r = scratch_done(e->use_enc, s)) return r;
  HIPCHK(hipMemcpyAsync(ct, e->ct.p, n * ctb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(kk, e->key.p, n * 32, hipMemcpyDeviceToHost, s));
  if (status) HIPCHK(hipMemcpyAsync(status, e->status.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

// key cache for the single-operation libmceliece entry points
struct CachedKey {
  std::vector<uint8_t> bytes;
  tmce_key *key = nullptr;
};
std::mutex g_cache_mu;
std::map<std::pair<int, int>, CachedKey> g_cache;  // (set, kind) -> last key

tmce_key *cached_key(int variant, int kind, const uint8_t *bytes, size_t len) {
  auto &c = g_cache[{variant, kind}];
  if (c.key && c.bytes.size() == len && std::memcmp(c.bytes.data(), bytes, len) == 0) return c.key;
  if (c.key) tmce_key_free(c.key);
  c.key = nullptr;
  tmce_key *k = nullptr;
  int r = kind == 1 ? tmce_key_from_pk(variant, bytes, &k) : tmce_key_from_sk(variant, bytes, &k);
  if (r) return nullptr;
  c.key = k;
  c.bytes.assign(bytes, bytes + len);
  return k;
}

// libmceliece's enc: the error vector is drawn on the host from randombytes() exactly as
// libmceliece's gen_e draws it (so any randomness source -- its test driver's included -- gives
// libmceliece's bytes), the rest runs on the GPU
int single_enc(int variant, unsigned char *c, unsigned char *k, const unsigned char *pk) {
  const int set = variant_base(variant);
  std::vector<uint16_t> pos((size_t)set_info(set).t);
  host_gen_e(set, rnd_global, nullptr, pos.data());
  std::lock_guard<std::mutex> lk(g_cache_mu);
  tmce_key *key = cached_key(variant, 1, pk, (size_t)set_info(set).pk_bytes);
  int st = 0;
  // a device failure (no gfx950 device, allocation, launch) returns its TMCE_E* code with the
  // outputs cleared; tmce_last_error() says what failed
  const int r = key ? enc_support(key, c, k, &st, pos.data(), 1) : TMCE_EDEVICE;
  if (r != 0) {
    std::memset(c, 0, (size_t)variant_ct_bytes(variant));
    std::memset(k, 0, 32);
    return r;
  }
  return st;
}

int single_dec(int variant, unsigned char *k, const unsigned char *c, const unsigned char *sk) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  const SetInfo &si = set_info(variant_base(variant));
  uint8_t ctc[256];
  std::memcpy(ctc, c, (size_t)variant_ct_bytes(variant));  // k and c may alias
  tmce_key *key = cached_key(variant, 2, sk, (size_t)si.sk_bytes);
  int st = 0;
  const int r = key ? tmce_dec_batch(key, k, &st, ctc, 1) : TMCE_EDEVICE;
  if (r != 0) {  // device failure: its TMCE_E* code, k cleared
    std::memset(k, 0, 32);
    return r;
  }
  return st;
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

int tmce_init(int device) { return engine_init(device); }
const char *tmce_last_error(void) { return g_err.c_str(); }

void tmce_set_randombytes(tmce_randombytes_fn fn) { g_randombytes = fn; }

int tmce_key_from_pk(int variant, const unsigned char *pk, tmce_key **out) {
  if (!variant_valid(variant) || !pk || !out) return fail(TMCE_EINVAL, "tmce_key_from_pk: bad argument");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  std::lock_guard<std::mutex> lk(e->mu);
  const int set = variant_base(variant);
  auto k = std::make_unique<tmce_key>();
  k->set = set;
  k->variant = variant;
  k->pc = variant_pc(variant);
  k->kind = 1;
  k->pad_bad = pk_padding_bad(set, pk);
  std::vector<uint32_t> cols, trows;
  build_enc_cols(set, pk, cols);
  build_enc_rows(set, pk, trows);
  if (int r = upload(&k->cols, cols)) return r;
  if (int r = upload(&k->Trows, trows)) return r;
  k->bytes = (cols.size() + trows.size()) * 4;
  *out = k.release();
  return 0;
}

int tmce_key_from_sk(int variant, const unsigned char *sk, tmce_key **out) {
  if (!variant_valid(variant) || !sk || !out) return fail(TMCE_EINVAL, "tmce_key_from_sk: bad argument");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  std::lock_guard<std::mutex> lk(e->mu);
  const int set = variant_base(variant);
  auto k = std::make_unique<tmce_key>();
  k->set = set;
  k->variant = variant;
  k->pc = variant_pc(variant);
  k->kind = 2;
  DecKeyHost h;
  build_dec_key(set, sk, h);
  if (int r = upload(&k->F, h.F)) return r;
  const size_t fsw = gf2_msw_bytes(set, GF2_FRONT);
  if (hipMalloc((void **)&k->Fsw, fsw) != hipSuccess) return fail(TMCE_ENOMEM, "hipMalloc Fsw");
  {
    const SetInfo &si = set_info(set);
    const int kwp = (((si.synd_bytes + 3) / 4) + 3) / 4 * 4;
    HIPCHK(launch_build_msw(set, GF2_FRONT, k->F, kwp, k->Fsw, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
  }
  if (int r = upload(&k->Fcol, h.Fcol)) return r;
  if (int r = upload(&k->alpha, h.alpha)) return r;
  if (int r = upload(&k->la, h.la)) return r;
  if (int r = upload(&k->lu, h.lu)) return r;
  if (int r = upload(&k->ainv, h.ainv)) return r;
  if (int r = upload(&k->supp, h.supp)) return r;
  if (int r = upload(&k->rej, h.rej)) return r;
  k->bytes = (h.F.size() + h.Fcol.size() + h.supp.size()) * 4 + fsw +
             (h.alpha.size() + h.la.size() + h.lu.size() + h.ainv.size()) * 2 + h.rej.size();
  *out = k.release();
  return 0;
}

void tmce_key_free(tmce_key *k) {
  if (!k) return;
  if (Engine *e = g_engine) {
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipStreamSynchronize(e->stream);
  }
  (void)hipFree(k->cols);
  (void)hipFree(k->Trows);
  (void)hipFree(k->Tb);
  (void)hipFree(k->Tv);
  (void)hipFree(k->F);
  (void)hipFree(k->Fsw);
  (void)hipFree(k->Fcol);
  (void)hipFree(k->alpha);
  (void)hipFree(k->la);
  (void)hipFree(k->lu);
  (void)hipFree(k->ainv);
  (void)hipFree(k->supp);
  (void)hipFree(k->rej);
  delete k;
}

size_t tmce_key_device_bytes(const tmce_key *k) { return k ? k->bytes : 0; }

int tmce_keypair_seeded(int variant, unsigned char *pk, unsigned char *sk, const unsigned char *seed) {
  if (!variant_valid(variant) || !pk || !sk || !seed) return fail(TMCE_EINVAL, "tmce_keypair_seeded: bad argument");
  SeedChain c;
  c.state[0] = 0x40;
  std::memcpy(c.state + 1, seed, 32);
  host_keypair(variant, pk, sk, rnd_chain, &c);
  return 0;
}

int tmce_enc_batch_device(tmce_key *pk, unsigned char *d_ct, unsigned char *d_k, int *d_status,
                          const unsigned char *d_seeds, size_t n, void *stream) {
  if (!pk || (n && (!d_ct || !d_k || !d_seeds))) return fail(TMCE_EINVAL, "tmce_enc_batch_device: bad argument");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  std::lock_guard<std::mutex> lk(e->mu);
  return enc_device(pk, d_ct, d_k, d_status, d_seeds, n, static_cast<hipStream_t>(stream));
}

int tmce_dec_batch_device(tmce_key *sk, unsigned char *d_k, int *d_status, const unsigned char *d_ct, size_t n,
                          void *stream) {
  if (!sk || (n && (!d_ct || !d_k))) return fail(TMCE_EINVAL, "tmce_dec_batch_device: bad argument");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  std::lock_guard<std::mutex> lk(e->mu);
  return dec_device(sk, d_k, d_status, d_ct, n, static_cast<hipStream_t>(stream));
}

int tmce_enc_batch(tmce_key *pk, unsigned char *ct, unsigned char *k, int *status, const unsigned char *seeds,
                   size_t n) {
  if (!pk || (n && (!ct || !k || !seeds))) return fail(TMCE_EINVAL, "tmce_enc_batch: bad argument");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  if (n == 0) return 0;
  std::lock_guard<std::mutex> lk(e->mu);
  const size_t ctb = key_ct_bytes(pk);
  if (int r = e->seeds.ensure(n * 32)) return r;
  if (int r = e->ct.ensure(n * ctb)) return r;
  if (int r = e->key.ensure(n * 32)) return r;
  if (int r = e->status.ensure(n * sizeof(int))) return r;
  HIPCHK(hipMemcpyAsync(e->seeds.p, seeds, n * 32, hipMemcpyHostToDevice, e->stream));
  if (int r = enc_device(pk, e->ct.as<uint8_t>(), e->key.as<uint8_t>(), e->status.as<int>(), e->seeds.as<uint8_t>(), n,
                         e->stream))
    return r;
  HIPCHK(hipMemcpyAsync(ct, e->ct.p, n * ctb, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipMemcpyAsync(k, e->key.p, n * 32, hipMemcpyDeviceToHost, e->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, e->status.p, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return 0;
}

int tmce_dec_batch(tmce_key *sk, unsigned char *k, int *status, const unsigned char *ct, size_t n) {
  if (!sk || (n && (!ct || !k))) return fail(TMCE_EINVAL, "tmce_dec_batch: bad argument");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  if (n == 0) return 0;
  std::lock_guard<std::mutex> lk(e->mu);
  const size_t ctb = key_ct_bytes(sk);
  if (int r = e->ct.ensure(n * ctb)) return r;
  if (int r = e->key.ensure(n * 32)) return r;
  if (int r = e->status.ensure(n * sizeof(int))) return r;
  HIPCHK(hipMemcpyAsync(e->ct.p, ct, n * ctb, hipMemcpyHostToDevice, e->stream));
  if (int r = dec_device(sk, e->key.as<uint8_t>(), e->status.as<int>(), e->ct.as<uint8_t>(), n, e->stream)) return r;
  HIPCHK(hipMemcpyAsync(k, e->key.p, n * 32, hipMemcpyDeviceToHost, e->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, e->status.p, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return 0;
}

int tmce_syndrome_batch_device(tmce_key *pk, unsigned char *d_ct, const unsigned char *d_e, size_t n, int method,
                               void *stream) {
  if (!pk || (n && (!d_ct || !d_e))) return fail(TMCE_EINVAL, "tmce_syndrome_batch_device: bad argument");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  std::lock_guard<std::mutex> lk(e->mu);
  return synd_device(pk, d_ct, d_e, n, method, static_cast<hipStream_t>(stream));
}

int tmce_syndrome_batch(tmce_key *pk, unsigned char *ct, const unsigned char *ev, size_t n, int method) {
  if (!pk || (n && (!ct || !ev))) return fail(TMCE_EINVAL, "tmce_syndrome_batch: bad argument");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  if (n == 0) return 0;
  if (pk->kind != 1) return fail(TMCE_EINVAL, "the syndrome needs a public-key handle");
  std::lock_guard<std::mutex> lk(e->mu);
  const SetInfo &si = set_info(pk->set);
  if (int r = e->seeds.ensure(n * (si.n / 8))) return r;  // error vectors
  if (int r = e->ct.ensure(n * si.synd_bytes)) return r;
  HIPCHK(hipMemcpyAsync(e->seeds.p, ev, n * (si.n / 8), hipMemcpyHostToDevice, e->stream));
  if (int r = synd_device(pk, e->ct.as<uint8_t>(), e->seeds.as<uint8_t>(), n, method, e->stream)) return r;
  HIPCHK(hipMemcpyAsync(ct, e->ct.p, n * si.synd_bytes, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return 0;
}

int tmce_tune_set(int variant, int param, int value) {
  if (!variant_valid(variant)) return fail(TMCE_EINVAL, "tmce_tune_set: bad set");
  const int set = variant_base(variant);  // launch shapes are per parameter set
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  std::lock_guard<std::mutex> lk(e->mu);
  switch (param) {
    case TMCE_TUNE_SYND_MFMA_CFG:
      if (value < 0 || value >= SYND_MFMA_CFGS) return fail(TMCE_EINVAL, "synd_mfma_cfg out of range");
      e->synd_cfg[set] = value;
      return 0;
    case TMCE_TUNE_GEN_BLOCKS:
      if (value < 0 || value > (1 << 20)) return fail(TMCE_EINVAL, "gen_blocks out of range");
      e->gen_blocks[set] = value;
      return 0;
    case TMCE_TUNE_DECODE_BLOCKS:
      if (value < 0 || value > (1 << 20)) return fail(TMCE_EINVAL, "decode_blocks out of range");
      e->decode_blocks[set] = value;
      return 0;
    case TMCE_TUNE_FRONT_MFMA:
      if (value != 0 && value != 1) return fail(TMCE_EINVAL, "front_mfma must be 0 or 1");
      e->front_mfma[set] = value;
      return 0;
    case TMCE_TUNE_FRONT_CFG:
      if (value < 0 || value >= SYND_MFMA_CFGS) return fail(TMCE_EINVAL, "front_cfg out of range");
      e->front_cfg[set] = value;
      return 0;
    case TMCE_TUNE_GEN_FORM:
      if (value != 0 && value != 1) return fail(TMCE_EINVAL, "gen_form must be 0 or 1");
      e->gen_form[set] = value;
      return 0;
    case TMCE_TUNE_FFT_FORM:
      if (value != 0 && value != 1) return fail(TMCE_EINVAL, "fft_form must be 0 or 1");
      e->fft_form[set] = value;
      return 0;
  }
  return fail(TMCE_EINVAL, "tmce_tune_set: unknown parameter");
}

int tmce_tune_get(int variant, int param) {
  if (!variant_valid(variant)) return fail(TMCE_EINVAL, "tmce_tune_get: bad set");
  const int set = variant_base(variant);
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  std::lock_guard<std::mutex> lk(e->mu);
  switch (param) {
    case TMCE_TUNE_SYND_MFMA_CFG: return e->synd_cfg[set];
    case TMCE_TUNE_GEN_BLOCKS: return e->gen_blocks[set];
    case TMCE_TUNE_DECODE_BLOCKS: return e->decode_blocks[set];
    case TMCE_TUNE_FRONT_MFMA: return e->front_mfma[set];
    case TMCE_TUNE_FRONT_CFG: return e->front_cfg[set];
    case TMCE_TUNE_GEN_FORM: return e->gen_form[set];
    case TMCE_TUNE_FFT_FORM: return e->fft_form[set];
  }
  return fail(TMCE_EINVAL, "tmce_tune_get: unknown parameter");
}

void *tmce_dev_alloc(size_t bytes) {
  if (!engine()) return nullptr;
  void *p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
  return p;
}
void tmce_dev_free(void *p) { (void)hipFree(p); }

void *tmce_host_alloc(size_t bytes) {
  if (!engine()) return nullptr;
  void *h = nullptr, *d = nullptr;
  if (hipHostMalloc(&h, bytes ? bytes : 1, hipHostMallocMapped) != hipSuccess) {
    fail(TMCE_ENOMEM, "hipHostMalloc slab");
    return nullptr;
  }
  if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess || !d) {
    (void)hipHostFree(h);
    fail(TMCE_EDEVICE, "hipHostGetDevicePointer");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_slab_mu);
  g_slabs[reinterpret_cast<uintptr_t>(h)] = Slab{reinterpret_cast<uintptr_t>(h) + (bytes ? bytes : 1),
                                                 static_cast<uint8_t *>(d)};
  return h;
}

int tmce_host_free(void *p) {
  if (!p) return 0;
  {
    std::lock_guard<std::mutex> lk(g_slab_mu);
    auto it = g_slabs.find(reinterpret_cast<uintptr_t>(p));
    if (it == g_slabs.end()) return fail(TMCE_EINVAL, "tmce_host_free: not a slab from tmce_host_alloc");
    g_slabs.erase(it);
  }
  HIPCHK(hipHostFree(p));
  return 0;
}
int tmce_memcpy_h2d(void *d, const void *h, size_t bytes) {
  HIPCHK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
  return 0;
}
int tmce_memcpy_d2h(void *h, const void *d, size_t bytes) {
  HIPCHK(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
  return 0;
}
int tmce_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return 0;
}
int tmce_sync(void) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipDeviceSynchronize());
  return 0;
}

int tmce_stream_create(void **stream) {
  if (!stream) return fail(TMCE_EINVAL, "tmce_stream_create: null argument");
  if (!engine()) return TMCE_ENODEV;
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return 0;
}
int tmce_stream_destroy(void *stream) {
  if (!stream) return 0;
  HIPCHK(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return 0;
}
int tmce_stream_sync(void *stream) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  HIPCHK(hipStreamSynchronize(stream ? static_cast<hipStream_t>(stream) : e->stream));
  return 0;
}
int tmce_event_create(void **event) {
  if (!event) return fail(TMCE_EINVAL, "tmce_event_create: null argument");
  if (!engine()) return TMCE_ENODEV;
  hipEvent_t ev = nullptr;
  HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  *event = ev;
  return 0;
}
int tmce_event_destroy(void *event) {
  if (!event) return 0;
  HIPCHK(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return 0;
}
int tmce_event_record(void *event, void *stream) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  if (!event) return fail(TMCE_EINVAL, "tmce_event_record: null event");
  HIPCHK(hipEventRecord(static_cast<hipEvent_t>(event), stream ? static_cast<hipStream_t>(stream) : e->stream));
  return 0;
}
int tmce_stream_wait_event(void *stream, void *event) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  if (!event) return fail(TMCE_EINVAL, "tmce_stream_wait_event: null event");
  HIPCHK(hipStreamWaitEvent(stream ? static_cast<hipStream_t>(stream) : e->stream, static_cast<hipEvent_t>(event), 0));
  return 0;
}

int tmce_set_form(int form) {
  if (form != 0 && form != 1) return fail(TMCE_EINVAL, "form must be 0 (engine) or 1 (plain)");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  std::lock_guard<std::mutex> lk(e->mu);
  e->form = form;
  return 0;
}

int tmce_timing_enable(int on) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  std::lock_guard<std::mutex> lk(e->mu);
  e->timer.drain();
  if (on) {
    e->timer.reset();
    g_ktimer = &e->timer;
  } else {
    g_ktimer = nullptr;
  }
  return 0;
}

int tmce_timing_read(int max, char *names, long *counts, double *total_ms) {
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  std::lock_guard<std::mutex> lk(e->mu);
  e->timer.drain();
  int i = 0;
  const auto acc = e->timer.snapshot();
  for (auto &kv : acc) {
    if (i >= max) break;
    std::snprintf(names + 48 * i, 48, "%s", kv.first.c_str());
    counts[i] = kv.second.first;
    total_ms[i] = kv.second.second;
    ++i;
  }
  return i;
}

#define TMCE_DEFINE_SET(S, ID)                                                                    \
  void mceliece_kem_##S##_keypair(unsigned char *pk, unsigned char *sk) {                         \
    host_keypair(ID, pk, sk, rnd_global, nullptr);                                                \
  }                                                                                               \
  int mceliece_kem_##S##_enc(unsigned char *c, unsigned char *k, const unsigned char *pk) {       \
    return single_enc(ID, c, k, pk);                                                              \
  }                                                                                               \
  int mceliece_kem_##S##_dec(unsigned char *k, const unsigned char *c, const unsigned char *sk) { \
    return single_dec(ID, k, c, sk);                                                              \
  }                                                                                               \
  const char *mceliece_kem_##S##_implementation(void) { return "tmce-gfx950"; }                   \
  const char *mceliece_kem_##S##_compiler(void) { return "hipcc " __VERSION__; }                  \
  void (*mceliece_dispatch_kem_##S##_keypair(long long))(unsigned char *, unsigned char *) {      \
    return mceliece_kem_##S##_keypair;                                                            \
  }                                                                                               \
  int (*mceliece_dispatch_kem_##S##_enc(long long))(unsigned char *, unsigned char *,             \
                                                    const unsigned char *) {                      \
    return mceliece_kem_##S##_enc;                                                                \
  }                                                                                               \
  int (*mceliece_dispatch_kem_##S##_dec(long long))(unsigned char *, const unsigned char *,       \
                                                    const unsigned char *) {                      \
    return mceliece_kem_##S##_dec;                                                                \
  }                                                                                               \
  const char *mceliece_dispatch_kem_##S##_implementation(long long) { return "tmce-gfx950"; }     \
  const char *mceliece_dispatch_kem_##S##_compiler(long long) { return "hipcc " __VERSION__; }    \
  long long mceliece_numimpl_kem_##S(void) { return 1; }

TMCE_DEFINE_SET(348864, 0)
TMCE_DEFINE_SET(348864f, 0 | VAR_F)
TMCE_DEFINE_SET(348864pc, 0 | VAR_PC)
TMCE_DEFINE_SET(348864pcf, 0 | VAR_PC | VAR_F)
TMCE_DEFINE_SET(460896, 1)
TMCE_DEFINE_SET(460896f, 1 | VAR_F)
TMCE_DEFINE_SET(460896pc, 1 | VAR_PC)
TMCE_DEFINE_SET(460896pcf, 1 | VAR_PC | VAR_F)
TMCE_DEFINE_SET(6688128, 2)
TMCE_DEFINE_SET(6688128f, 2 | VAR_F)
TMCE_DEFINE_SET(6688128pc, 2 | VAR_PC)
TMCE_DEFINE_SET(6688128pcf, 2 | VAR_PC | VAR_F)
TMCE_DEFINE_SET(6960119, 3)
TMCE_DEFINE_SET(6960119f, 3 | VAR_F)
TMCE_DEFINE_SET(6960119pc, 3 | VAR_PC)
TMCE_DEFINE_SET(6960119pcf, 3 | VAR_PC | VAR_F)
TMCE_DEFINE_SET(8192128, 4)
TMCE_DEFINE_SET(8192128f, 4 | VAR_F)
TMCE_DEFINE_SET(8192128pc, 4 | VAR_PC)
TMCE_DEFINE_SET(8192128pcf, 4 | VAR_PC | VAR_F)

}  // extern "C"
