// kem_kernels.h -- launch interface between the host engine (api.cpp) and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstdint>
#include "params.h"
#include "device_common.h"

namespace tmce {

// ---- multi-batch plan (device memory)
// Sub-batches of one parameter set, possibly under different keys, mapped onto one combined op
// index space: key group g owns combined ops [gstart[g], gstart[g+1]) (PLAN_ALIGN-aligned, so an
// MFMA workgroup tile never spans two keys); its sub-batches sit contiguously from gstart[g] and the
// rest of the group is padding (dead ops, computed and discarded).  Stage kernels given a plan take
// their op count (total) and per-op key material from it; gather/scatter kernels move the sub-batch
// inputs and outputs between the callers' buffers and the flat pipeline buffers.
constexpr int PLAN_MAX_GROUPS = 16;
constexpr int PLAN_ALIGN = 256;
struct KeyDev {
  const uint32_t *cols;   // encapsulation: column-major [I | T]
  const uint32_t *F;      // decapsulation: front-end matrix (VALU form)
  const uint8_t *Fsw;     // decapsulation: front-end matrix (MFMA form)
  const uint32_t *Fcol;   // decapsulation: syndrome map columns [N + 1][FCW], column N zero
  const uint16_t *alpha, *la, *lu;
  const uint16_t *ainv;   // decapsulation: field element -> support index (0xFFFF off the support)
  const uint32_t *supp;   // decapsulation: support flags over the field (decoder point order)
  const uint8_t *rej;
  int pad_bad;            // 6960119 public key with nonzero padding bits
  int reserved;
};
struct SubDev {
  const uint8_t *in;      // seeds [n][32] (enc) or ciphertexts [n][SYND_BYTES] (dec)
  uint8_t *ct;            // enc: ciphertexts out
  uint8_t *key;           // session keys out
  int *status;            // per-op status out (optional)
  int start, n;           // combined index of the first op, op count
};
struct PlanDev {
  int total;              // combined ops (live + padding)
  int ngroups, nsub, reserved;
  int gstart[PLAN_MAX_GROUPS + 1];
  KeyDev keys[PLAN_MAX_GROUPS];
  const SubDev *subs;     // nsub entries sorted by start
};
__device__ __forceinline__ int plan_group(const PlanDev *p, int op) {
  int g = 0;
  while (g + 1 < p->ngroups && op >= p->gstart[g + 1]) ++g;
  return g;
}
// a pointer that is the same in every lane, kept in scalar registers (so address arithmetic on it
// stays scalar); the caller guarantees uniformity
template <class T>
__device__ __forceinline__ T *uniform_ptr(T *p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return reinterpret_cast<T *>(((uint64_t)hi << 32) | lo);
}
// the same pointer typed as global memory: loads through it are global_load (a pointer read from
// memory is otherwise generic and its loads become flat_load)
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *global_ptr(const T *p) {
  return (const __attribute__((address_space(1))) T *)p;
}
typedef unsigned v4u32 __attribute__((ext_vector_type(4)));
typedef unsigned v2u32 __attribute__((ext_vector_type(2)));
// 8-byte global load of element i
__device__ __forceinline__ uint2 global_load8(const uint2 *p, size_t i) {
  const v2u32 v = ((const __attribute__((address_space(1))) v2u32 *)p)[i];
  return make_uint2(v.x, v.y);
}
// 16-byte global load of element i
__device__ __forceinline__ uint4 global_load16(const uint4 *p, size_t i) {
  const v4u32 v = ((const __attribute__((address_space(1))) v4u32 *)p)[i];
  return make_uint4(v.x, v.y, v.z, v.w);
}
// sub-batch holding combined op `op`, or -1 for padding
__device__ __forceinline__ int plan_sub(const PlanDev *p, int op) {
  int lo = 0, hi = p->nsub - 1, r = -1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    if (p->subs[mid].start <= op) { r = mid; lo = mid + 1; } else hi = mid - 1;
  }
  return (r >= 0 && op < p->subs[r].start + p->subs[r].n) ? r : -1;
}
// plan_sub for a lane of a wave whose ops are consecutive: the sub-batch of the wave's first active
// lane's op is found once (uniform, scalar loads) and serves every lane whose op it holds; the
// other lanes (a wave that crosses into the next sub-batch or into padding) search on their own
__device__ __forceinline__ int plan_sub_wave(const PlanDev *p, int op) {
  const int op0 = __builtin_amdgcn_readfirstlane(op);
  const int s0 = __builtin_amdgcn_readfirstlane(plan_sub(p, op0));
  if (s0 >= 0) {
    const int st = __builtin_amdgcn_readfirstlane(p->subs[s0].start);
    const int n = __builtin_amdgcn_readfirstlane(p->subs[s0].n);
    if (op >= st && op < st + n) return s0;
  }
  return plan_sub(p, op);
}
// gather sub-batch inputs into the flat buffer (row_bytes per op; padding rows zero) and scatter
// the flat outputs back (enc: ct_bytes > 0 also scatters ciphertexts)
hipError_t launch_plan_gather(const PlanDev *plan, uint8_t *flat, int row_bytes, int cap, hipStream_t s);
hipError_t launch_plan_scatter(const PlanDev *plan, const uint8_t *ct, int ct_bytes, const uint8_t *key,
                               const int *status, int cap, hipStream_t s);

// Syndrome bits of one decapsulation op (output of the front end, RW words): syndrome j occupies
// bits [j*M, j*M + M) (syndrome-major, so one value is at most two words).
template <class P>
__device__ __forceinline__ uint32_t synd_value(const uint32_t *Srow, int j) {
  const int bit = j * P::M, w = bit >> 5, sh = bit & 31;
  uint64_t v = Srow[w];
  if (sh + P::M > 32) v |= (uint64_t)Srow[w + 1] << 32;
  return (uint32_t)(v >> sh) & ((1u << P::M) - 1u);
}

// Device buffers of one encapsulation batch (all pointers are device pointers).
struct EncArgs {
  const uint8_t *seeds;  // [n][32]
  const uint32_t *cols;  // key: [N][CWP] column-major [I | T]
  uint8_t *ct;           // [n][SYND_BYTES], + 32 bytes per op when pc
  uint8_t *key;          // [n][32]
  int *status;           // [n] 0, or -2 if the sampler gave up (optional)
  uint16_t *ind;         // scratch [n][IND_STRIDE] support of e
  uint32_t *msg;         // scratch [n][MSG_WORDS] session-key message
  uint32_t *attempts;    // optional [n] gen_e attempts
  unsigned *queue;       // scratch work-queue counter
  int nops;              // op count, or with a plan the capacity (the live count is plan->total)
  int gen_blocks;        // persistent grid size of the sampler (0 = default)
  int gen_form;          // sampler: 1 = pipelined lane groups (k_gen_e_pipe), 0 = one op per lane
  int pc;                // plaintext confirmation (pc variants): ciphertext = c || SHAKE256(2 || e, 32)
  const PlanDev *plan;   // multi-batch plan (per-group keys; null: one key, `cols`)
};

// Device-resident tables of one GF(2^m) field and of the additive FFT used by decapsulation.
struct FieldTables {
  const uint16_t *lg;    // [2^m]
  const uint16_t *ex;    // [2^m - 1]
  const uint16_t *twl;   // FFT twist logs, level d at offset 2D - 2(D >> d)
  const uint16_t *plo;   // FFT butterfly multipliers, lane part [KD][64]
  const uint16_t *phi;   // FFT butterfly multipliers, register part (levels concatenated)
  const uint16_t *ptlo;  // (lane)^t, [64]       (monic x^t term when t = D), or null
  const uint16_t *pthi;  // (64 r)^t, [2^m / 64]
};

// Device buffers of one decapsulation batch.
struct DecArgs {
  const uint8_t *ct;     // [n][SYND_BYTES], + 32 bytes per op when pc
  int pc;                // plaintext confirmation (pc variants)
  uint8_t *ct_synd;      // pc: scratch [n][SYND_BYTES], the syndrome part of each ciphertext
  const uint32_t *F;     // key: [R][KWP] front-end matrix (plane-major rows)
  const uint8_t *Fsw;    // key: F as bit rows in MFMA step-pair pieces
  int front_mfma;        // front end on INT8 MFMA (1) or on the VALU (0)
  int front_cfg;         // MFMA tile shape of the front end
  uint32_t *xscratch;    // gf2_scratch_bytes(set, GF2_FRONT, n) bytes (MFMA front end; null if 0)
  const uint32_t *Fcol;  // key: [N + 1][FCW] syndromes of each unit vector (rows as in S, zero-padded), column N zero
  const uint16_t *alpha; // key: [N] support
  const uint16_t *la;    // key: [N] log alpha (0xFFFF for alpha = 0)
  const uint16_t *lu;    // key: [N] log 1/g(alpha)^2 (0xFFFF for 0)
  const uint16_t *ainv;  // key: [2^m] support index of each field element (0xFFFF off the support)
  const uint32_t *supp;  // key: [2^m / 32] support flags: element 64 g + l is bit g % 32 of word l 2^(m-11) + g / 32
  const uint8_t *rej;    // key: [N/8] implicit-rejection string s
  FieldTables ft;
  uint8_t *key;          // [n][32]
  int *status;           // [n] 0, or -1 for a ciphertext with nonzero padding bits (6960119)
  uint32_t *S;           // scratch [n][RW] syndrome bit planes
  uint32_t *msg;         // scratch [n][MSG_WORDS]
  unsigned *queue;       // scratch: the decoder's work counters, 1024 bytes (cleared by the launcher)
  int nops;              // op count, or with a plan the capacity (the live count is plan->total)
  int decode_blocks;     // grid of the decoder (0 = default)
  int cus;               // compute units of the device (0: 256)
  int fft_form;          // bottom-up FFT: 1 = bitsliced GF(2^m) arithmetic, 0 = log/exp tables in LDS
  const PlanDev *plan;   // multi-batch plan (per-group keys; null: one key)
};

// Plain one-launch-per-stage form (baseline).
struct PlainEncArgs {
  const uint8_t *seeds;
  const uint32_t *Trows;  // key: public key T as MT rows of ceil(K/32) words
  uint8_t *ct;
  uint8_t *key;
  int *status;
  uint32_t *erow;  // scratch [n][ceil(N/32)]
  uint32_t *Ens;   // scratch [ceil(K/32)][n]
  uint32_t *S;     // scratch [CW][n]
  uint32_t *msg;   // scratch [n][MSG_WORDS]
  int nops;
};
struct PlainDecScratch {
  uint16_t *loc;   // [n][T+1]
  uint32_t *erow;  // [n][2*ceil(N/64)]
  int *weight;     // [n]
};

// Syndrome alone (PKE encryption of given error vectors).
struct SyndArgs {
  const uint8_t *e;      // [n][N/8] error vectors (libmceliece byte layout)
  const uint8_t *Tb;     // key: T as bit rows in MFMA step-pair pieces (INT8 MFMA form), or null
  const uint32_t *Tv;    // key: T per 32-row block as [word][row] (VALU form), or null
  const uint32_t *cols;  // key: column-major [I | T] (sparse form)
  uint8_t *ct;           // [n][SYND_BYTES]
  uint32_t *xscratch;    // gf2_scratch_bytes(set, GF2_SYND, n) bytes (MFMA form; null if 0)
  int nops;
};
// method: 0 = sparse column gather, 1 = INT8 MFMA GEMM (cfg = tile shape index), 2 = VALU GEMM
hipError_t launch_synd_set(int set, const SyndArgs &a, int method, int cfg, hipStream_t s);
constexpr int SYND_MFMA_CFGS = 8;
// Swizzled 0/1-byte operands of the INT8 MFMA GF(2) products: GF2_SYND = the public key T
// (MT x K), GF2_FRONT = the decapsulation front-end matrix F (R x KBITS).
enum { GF2_SYND = 0, GF2_FRONT = 1 };
size_t gf2_msw_bytes(int set, int which);
// from a row-major bit matrix with rw 32-bit words per row
hipError_t launch_build_msw(int set, int which, const uint32_t *rows, int rw, uint8_t *out, hipStream_t s);
// VALU operand Tv (ceil(MT/32) * ceil(K/32) * 32 words) from the T rows
hipError_t launch_build_tv(int set, const uint32_t *Trows, uint32_t *Tv, hipStream_t s);
// decapsulation front end S = F c on INT8 MFMA (same output layout as the VALU front end)
hipError_t launch_front_mfma_set(int set, const uint8_t *ct, uint32_t *scratch, const uint8_t *Fsw, uint32_t *S,
                                 int nops, int cfg, const PlanDev *plan, hipStream_t s);
// bytes of scratch the MFMA products need for n ops (x rows repacked to aligned words), 0 if none
size_t gf2_scratch_bytes(int set, int which, size_t n);

// Batched key generation, systematic-form step: one candidate (g, alpha) per workgroup.
struct KeygenArgs {
  const uint16_t *g;      // [ncand][T+1] Goppa polynomials (monic)
  const uint16_t *alpha;  // [ncand][N] supports
  const uint16_t *lg;     // field log table, 2^m entries
  const uint16_t *ex;     // field exp table, 2^m - 1 entries
  uint64_t *H;            // scratch [ncand][pk_gen_h_bytes(set) / 8]
  uint8_t *pk;            // [ncand][PK_BYTES]
  int *ok;                // [ncand] 1: pk written, 0: the left mt x mt block is singular
  uint64_t *pivots;       // f variants: [ncand] pivot mask of the column step (null otherwise)
  const int *valid;       // [ncand] 0: the attempt failed before the systematic form (null: all valid)
  int ncand;
};
hipError_t launch_pk_gen_set(int set, const KeygenArgs &a, hipStream_t s);
// Goppa polynomials (monic, [ncand][T+1]) of candidates f ([ncand][T]); valid[c] is cleared when the
// powers of f are dependent; candidates with valid[c] == 0 on entry are skipped
hipError_t launch_genpoly_set(int set, const uint16_t *f, const uint16_t *lg, const uint16_t *ex, uint16_t *g,
                              int *valid, int ncand, hipStream_t s);
// Benes control bits of n keys' support permutations (pi: [n][2^m] u16) into cb: [n][(2m-1) 2^m / 8]
// bytes (LM:crypto_kem/<set>/vec/controlbits.c), one workgroup per key
hipError_t launch_controlbits(int m, const uint16_t *pi, uint8_t *cb, int nkeys, hipStream_t s);
size_t pk_gen_h_bytes(int set);

hipError_t launch_enc_set(int set, const EncArgs &a, hipStream_t s);
// the finishing launch alone, for error supports and message rows [1][e] written by the caller
hipError_t launch_enc_finish_set(int set, const EncArgs &a, hipStream_t s);
hipError_t launch_plain_enc_set(int set, const PlainEncArgs &a, hipStream_t s);
hipError_t launch_plain_dec_set(int set, const DecArgs &a, const PlainDecScratch &x, hipStream_t s);
hipError_t launch_dec_front_only(int set, const DecArgs &a, hipStream_t s);
hipError_t launch_dec_hash_only(int set, const DecArgs &a, hipStream_t s);
hipError_t launch_dec_set(int set, const DecArgs &a, hipStream_t s);
hipError_t launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s);
hipError_t launch_fill_u32(unsigned *p, unsigned v, int n, hipStream_t s);
hipError_t launch_fill_i32(int *p, int v, size_t n, hipStream_t s);

// Optional per-kernel timing: when set, each engine launch is bracketed by events on its stream.
// begin() returns a handle for the matching end(), so launches timed from several host threads
// keep their own event pairs; a thread that captures a graph sets g_ktimer_off (thread-local), so
// no event record lands inside the capture while other threads' launches are still timed.
struct KernelTimer {
  virtual long begin(const char *name, hipStream_t s) = 0;
  virtual void end(long handle, hipStream_t s) = 0;
  virtual ~KernelTimer() = default;
};
extern std::atomic<KernelTimer *> g_ktimer;
extern thread_local bool g_ktimer_off;
#define TMCE_TIMED(name, stream, ...)                                                      \
  do {                                                                                     \
    ::tmce::KernelTimer *kt_ = ::tmce::g_ktimer_off ? nullptr : ::tmce::g_ktimer.load();   \
    const long kh_ = kt_ ? kt_->begin(name, stream) : -1;                                  \
    __VA_ARGS__;                                                                           \
    if (kt_) kt_->end(kh_, stream);                                                        \
  } while (0)

}  // namespace tmce
