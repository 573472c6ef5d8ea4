// noalloc.hip -- the decapsulation entry point that allocates nothing.
//
// Callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc) pass inputs and outputs; every other buffer the decapsulation
// kernels need -- work counters, syndrome planes, session-key messages, the pc variants' syndrome
// copy, the MFMA front end's repacked words -- is carved out of `scratch`, a file-scope device
// array of NOALLOC_SCRATCH_BYTES that each GPU holds for the life of the process.
//
// In host code the array's name is a host-side stand-in, not a device address: the device address
// is looked up once per device with hipGetSymbolAddress at engine set-up (with that device
// current) and only that address is used afterwards.
//
// Every call on a GPU shares that GPU's array: two calls that overlap on one GPU (two streams, or
// two host threads) race on it and must be ordered by the caller.  The arrays keep the last call's
// contents; the only part that must start at zero, the decoder's work counters, is cleared by a
// fill kernel inside each call.
#include <hip/hip_runtime.h>

#include "engine.h"

namespace tmce {
namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)TMCE_NOALLOC_SCRATCH_MB << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
uint8_t *g_scratch_ptr[MAX_DEVICES] = {};  // device address of `scratch` on each device

}  // namespace

int noalloc_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return fail(TMCE_ENODEV, "device index out of range");
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(TMCE_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<uint8_t *>(p);
  return 0;
}

}  // namespace tmce

using namespace tmce;

extern "C" int tmce_dec_batch_device_noalloc(tmce_key *sk, unsigned char *d_k, int *d_status,
                                             const unsigned char *d_ct, size_t n, void *stream) {
  if (!sk || (n && (!d_ct || !d_k))) return fail(TMCE_EINVAL, "tmce_dec_batch_device_noalloc: bad argument");
  if (sk->kind != 2) return fail(TMCE_EINVAL, "decapsulation needs a secret-key handle");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  if (n == 0) return 0;
  if (n > (size_t)TMCE_NOALLOC_MAX_OPS)
    return fail(TMCE_EINVAL, "tmce_dec_batch_device_noalloc: more than TMCE_NOALLOC_MAX_OPS operations");
  std::lock_guard<std::mutex> lk(e->mu);
  uint8_t *base = g_scratch_ptr[e->device];
  if (!base) return fail(TMCE_EDEVICE, "no-allocation scratch not set up on this device");
  const bool mfma = e->front_mfma[sk->set] != 0;
  const size_t need = dec_scratch_layout(sk->set, sk->pc, mfma, n, nullptr, nullptr);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(TMCE_EINVAL, "tmce_dec_batch_device_noalloc: batch needs more than the static scratch holds");
  DecScratch x{};
  dec_scratch_layout(sk->set, sk->pc, mfma, n, base, &x);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : e->stream;
  return dec_launch(e, sk, d_k, d_status, d_ct, n, s, x, 0, nullptr);
}
