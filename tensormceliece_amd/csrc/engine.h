// engine.h -- host-side internals of libtmce.so shared by the C ABI (api.cpp), the multi-batch
// pipelines (pipeline.cpp) and the serving loop (server.cpp): device-resident key handles, the
// per-device engine context, device scratch buffers and error reporting.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tmce.h"
#include "host_math.h"
#include "kem_kernels.h"

struct tmce_key {
  int set = -1;      // base parameter set (0..4)
  int variant = -1;  // set | VAR_F | VAR_PC as given at creation
  bool pc = false;   // plaintext confirmation: ciphertexts are SYND_BYTES + 32 bytes
  int kind = 0;      // 1 = public (encapsulation), 2 = secret (decapsulation)
  size_t bytes = 0;
  bool pad_bad = false;
  uint32_t *cols = nullptr;
  uint32_t *Trows = nullptr;  // plain form: T as MT rows of ceil(K/32) words
  uint8_t *Tb = nullptr;      // syndrome GEMM operands, built on first use: bit rows in MFMA step-pair pieces ...
  uint32_t *Tv = nullptr;     // ... and 32-row blocks of words (VALU)
  uint32_t *F = nullptr;
  uint8_t *Fsw = nullptr;     // F as bit rows in MFMA step-pair pieces (INT8 MFMA front end)
  uint32_t *Fcol = nullptr;   // syndrome map over all N positions, [N + 1][FCW], column N zero (re-encryption check)
  uint16_t *alpha = nullptr, *la = nullptr, *lu = nullptr;
  uint16_t *ainv = nullptr;   // field element -> support index (0xFFFF off the support)
  uint32_t *supp = nullptr;   // support flags over the field (decoder point order)
  uint8_t *rej = nullptr;
};

namespace tmce {

// records `msg` as tmce_last_error() of the calling thread and returns `code`
int fail(int code, const std::string &msg);

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(TMCE_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) return fail(TMCE_ENOMEM, "hipMalloc scratch");
    cap = bytes;
    return 0;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

struct EventTimer : KernelTimer {
  struct Rec { std::string name; hipEvent_t a, b; };
  std::mutex tmu;  // begin/end/drain may come from several host threads
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<long, double>> acc;
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
  long begin(const char *name, hipStream_t s) override {
    std::lock_guard<std::mutex> lk(tmu);
    Rec r{name, get(), get()};
    (void)hipEventRecord(r.a, s);
    pending.push_back(r);
    return (long)pending.size() - 1;
  }
  void end(long h, hipStream_t s) override {
    std::lock_guard<std::mutex> lk(tmu);
    if (h >= 0 && h < (long)pending.size()) (void)hipEventRecord(pending[(size_t)h].b, s);
  }
  std::map<std::string, std::pair<long, double>> snapshot() {
    std::lock_guard<std::mutex> lk(tmu);
    return acc;
  }
  void reset() {
    std::lock_guard<std::mutex> lk(tmu);
    acc.clear();
  }
  void drain() {
    std::lock_guard<std::mutex> lk(tmu);
    for (auto &r : pending) {
      (void)hipEventSynchronize(r.b);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, r.a, r.b);
      auto &x = acc[r.name];
      x.first += 1;
      x.second += ms;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
};

struct SetTables {
  uint16_t *twl = nullptr, *plo = nullptr, *phi = nullptr, *ptlo = nullptr, *pthi = nullptr;
};

// The last stream that used one set of device scratch, and an event recorded after that use: a
// call on another stream waits for the event first (hipStreamWaitEvent), so calls that share the
// set never overlap on it, while calls with separate sets may run concurrently on two streams.
struct ScratchUse {
  hipEvent_t done = nullptr;
  hipStream_t last = nullptr;
};

struct Engine {
  int device = -1;
  int cus = 256;  // compute units
  hipStream_t stream = nullptr;
  std::mutex mu;
  uint16_t *lg[2] = {nullptr, nullptr}, *ex[2] = {nullptr, nullptr};
  SetTables st[NUM_SETS];
  DevBuf seeds, ct, key, status;  // staging of the host-buffer entry points (engine stream)
  // encapsulation and syndrome scratch
  DevBuf ind, msg_enc, queue, p_erow_enc, p_ens, p_s, xs_synd;
  ScratchUse use_enc;
  // decapsulation scratch
  DevBuf S, msg_dec, dqueue, xs_dec, p_loc, p_erow_dec, p_weight, ct_synd;
  ScratchUse use_dec;
  int form = 0;                                // 0 = engine form, 1 = plain per-stage form
  int synd_cfg[NUM_SETS] = {7, 7, 7, 7, 7};    // MFMA syndrome tile shape per set (tunable)
  int gen_blocks[NUM_SETS] = {0, 0, 0, 0, 0};  // sampler grid per set (0 = built-in default)
  int decode_blocks[NUM_SETS] = {0, 0, 0, 0, 0};
  int front_mfma[NUM_SETS] = {1, 1, 1, 1, 1};  // decapsulation front end: 1 = INT8 MFMA, 0 = VALU
  int front_cfg[NUM_SETS] = {1, 1, 1, 1, 1};   // its MFMA tile shape
  int gen_form[NUM_SETS] = {1, 1, 1, 1, 1};    // sampler: 1 = pipelined lane groups, 0 = one op per lane
  int fft_form[NUM_SETS] = {1, 1, 1, 1, 1};    // decoder's bottom-up FFT: 1 = bitsliced, 0 = log/exp tables
  EventTimer timer;
  // key generation: pinned host staging for the public-key downloads (allocated on first use and
  // kept; kg_mu serialises the calls that use it)
  std::mutex kg_mu;
  uint8_t *kg_stage = nullptr;
};

extern Engine *g_engine;

template <class T>
int upload(T **dst, const std::vector<T> &src) {
  if (hipMalloc((void **)dst, src.size() * sizeof(T)) != hipSuccess) return fail(TMCE_ENOMEM, "hipMalloc table");
  if (hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return fail(TMCE_EDEVICE, "hipMemcpy table");
  return 0;
}

// device address of the host range [p, p + len) when it lies inside one tmce_host_alloc slab, else
// null (the slab registry is process-wide)
uint8_t *slab_device_ptr(const void *p, size_t len);

int engine_init(int device);
// the engine (initialised on device 0 on first use), or null if there is no usable device
Engine *engine();

// device scratch pointers of one decapsulation batch
struct DecScratch {
  uint32_t *S = nullptr;         // syndrome planes [n][RW]
  uint32_t *msg = nullptr;       // session-key messages [n][MSG_WORDS]
  unsigned *queue = nullptr;     // decoder work counters, 1024 bytes
  uint8_t *ct_synd = nullptr;    // pc variants: syndrome part of each ciphertext [n][SYND_BYTES]
  uint32_t *xs = nullptr;        // MFMA front end: repacked ciphertext words (null if not needed)
};
size_t dec_scratch_layout(int set, bool pc, bool front_mfma, size_t n, uint8_t *base, DecScratch *x);
int dec_launch(Engine *e, tmce_key *k, uint8_t *d_k, int *d_status, const uint8_t *d_ct, size_t n, hipStream_t s,
               const DecScratch &x, int form, const PlainDecScratch *px);
// records the no-allocation entry point's scratch address on the current device (engine set-up)
int noalloc_init(int device);

inline bool valid_set(int set) { return set >= 0 && set < NUM_SETS; }
// ciphertext bytes of a key's variant
inline size_t key_ct_bytes(const tmce_key *k) { return (size_t)set_info(k->set).synd_bytes + (k->pc ? 32 : 0); }

}  // namespace tmce
