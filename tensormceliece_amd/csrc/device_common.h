// device_common.h -- device primitives shared by the encapsulation and decapsulation kernels:
// Keccak-f[1600] / SHAKE256 on one lane, GF(2^m) arithmetic through log/exp tables held in LDS,
// and wave-level helpers for the 64-lane CDNA wavefront.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tmce {

// ------------------------------------------------------------------ Keccak-f[1600]
// FIPS 202 round constants (iota).
__constant__ static const uint64_t kKeccakRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

// 64-bit rotate as two 32-bit funnel shifts (v_alignbit_b32): for 0 < r < 32,
// hi' = (hi:lo) >> (32-r), lo' = (lo:hi) >> (32-r); r >= 32 swaps the halves first.
template <int R>
__device__ __forceinline__ uint64_t rotl64(uint64_t x) {
  if constexpr (R == 0) {
    return x;
  } else {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    if constexpr (R >= 32) { const uint32_t t = lo; lo = hi; hi = t; }
    constexpr int r = R & 31;
    if constexpr (r == 0) {
      return ((uint64_t)hi << 32) | lo;
    } else {
      const uint32_t nhi = __builtin_amdgcn_alignbit(hi, lo, 32 - r);
      const uint32_t nlo = __builtin_amdgcn_alignbit(lo, hi, 32 - r);
      return ((uint64_t)nhi << 32) | nlo;
    }
  }
}

// Round constants held in two VGPRs: lane r (r < 24) has RC[r]; read back with v_readlane, so the
// rolled round loop has no scalar-memory wait per round.
struct KeccakRC {
  uint32_t lo, hi;
  __device__ __forceinline__ KeccakRC() {
    const int l = __lane_id();
    const uint64_t rc = kKeccakRC[l < 24 ? l : 0];
    lo = (uint32_t)rc;
    hi = (uint32_t)(rc >> 32);
  }
  __device__ __forceinline__ uint64_t get(int r) const {
    return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)lo, r) |
           ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, r) << 32);
  }
};

// One lane computes the whole permutation; the state stays in 50 VGPRs.  Lane (x,y) is A[x+5y].
// All lanes of the wave must be active when `rc` is read (it is broadcast with v_readlane).
__device__ __forceinline__ void keccak_f1600(uint64_t (&A)[25], const KeccakRC &rc) {
#pragma unroll 1
  for (int round = 0; round < 24; ++round) {
    uint64_t C0 = A[0] ^ A[5] ^ A[10] ^ A[15] ^ A[20];
    uint64_t C1 = A[1] ^ A[6] ^ A[11] ^ A[16] ^ A[21];
    uint64_t C2 = A[2] ^ A[7] ^ A[12] ^ A[17] ^ A[22];
    uint64_t C3 = A[3] ^ A[8] ^ A[13] ^ A[18] ^ A[23];
    uint64_t C4 = A[4] ^ A[9] ^ A[14] ^ A[19] ^ A[24];
    uint64_t D0 = C4 ^ rotl64<1>(C1), D1 = C0 ^ rotl64<1>(C2), D2 = C1 ^ rotl64<1>(C3);
    uint64_t D3 = C2 ^ rotl64<1>(C4), D4 = C3 ^ rotl64<1>(C0);
    // theta + rho + pi: B[y + 5*((2x+3y)%5)] = rot(A[x+5y] ^ D[x], r[x][y])
    uint64_t B[25];
    B[0]  = rotl64<0>(A[0] ^ D0);
    B[10] = rotl64<1>(A[1] ^ D1);
    B[20] = rotl64<62>(A[2] ^ D2);
    B[5]  = rotl64<28>(A[3] ^ D3);
    B[15] = rotl64<27>(A[4] ^ D4);
    B[16] = rotl64<36>(A[5] ^ D0);
    B[1]  = rotl64<44>(A[6] ^ D1);
    B[11] = rotl64<6>(A[7] ^ D2);
    B[21] = rotl64<55>(A[8] ^ D3);
    B[6]  = rotl64<20>(A[9] ^ D4);
    B[7]  = rotl64<3>(A[10] ^ D0);
    B[17] = rotl64<10>(A[11] ^ D1);
    B[2]  = rotl64<43>(A[12] ^ D2);
    B[12] = rotl64<25>(A[13] ^ D3);
    B[22] = rotl64<39>(A[14] ^ D4);
    B[23] = rotl64<41>(A[15] ^ D0);
    B[8]  = rotl64<45>(A[16] ^ D1);
    B[18] = rotl64<15>(A[17] ^ D2);
    B[3]  = rotl64<21>(A[18] ^ D3);
    B[13] = rotl64<8>(A[19] ^ D4);
    B[14] = rotl64<18>(A[20] ^ D0);
    B[24] = rotl64<2>(A[21] ^ D1);
    B[9]  = rotl64<61>(A[22] ^ D2);
    B[19] = rotl64<56>(A[23] ^ D3);
    B[4]  = rotl64<14>(A[24] ^ D4);
    // chi + iota
#pragma unroll
    for (int y = 0; y < 5; ++y) {
      uint64_t b0 = B[5 * y], b1 = B[5 * y + 1], b2 = B[5 * y + 2], b3 = B[5 * y + 3], b4 = B[5 * y + 4];
      A[5 * y + 0] = b0 ^ (~b1 & b2);
      A[5 * y + 1] = b1 ^ (~b2 & b3);
      A[5 * y + 2] = b2 ^ (~b3 & b4);
      A[5 * y + 3] = b3 ^ (~b4 & b0);
      A[5 * y + 4] = b4 ^ (~b0 & b1);
    }
    A[0] ^= rc.get(round);
  }
}

// Same permutation with the round constant read from constant memory: safe in divergent code
// (the plain per-thread kernels), at the cost of a scalar load per round.
__device__ __forceinline__ void keccak_f1600_c(uint64_t (&A)[25]) {
  struct ConstRC {
    __device__ __forceinline__ uint64_t get(int r) const { return kKeccakRC[r]; }
  };
#pragma unroll 1
  for (int round = 0; round < 24; ++round) {
    uint64_t C0 = A[0] ^ A[5] ^ A[10] ^ A[15] ^ A[20];
    uint64_t C1 = A[1] ^ A[6] ^ A[11] ^ A[16] ^ A[21];
    uint64_t C2 = A[2] ^ A[7] ^ A[12] ^ A[17] ^ A[22];
    uint64_t C3 = A[3] ^ A[8] ^ A[13] ^ A[18] ^ A[23];
    uint64_t C4 = A[4] ^ A[9] ^ A[14] ^ A[19] ^ A[24];
    uint64_t D0 = C4 ^ rotl64<1>(C1), D1 = C0 ^ rotl64<1>(C2), D2 = C1 ^ rotl64<1>(C3);
    uint64_t D3 = C2 ^ rotl64<1>(C4), D4 = C3 ^ rotl64<1>(C0);
    uint64_t B[25];
    B[0] = rotl64<0>(A[0] ^ D0);   B[10] = rotl64<1>(A[1] ^ D1);  B[20] = rotl64<62>(A[2] ^ D2);
    B[5] = rotl64<28>(A[3] ^ D3);  B[15] = rotl64<27>(A[4] ^ D4); B[16] = rotl64<36>(A[5] ^ D0);
    B[1] = rotl64<44>(A[6] ^ D1);  B[11] = rotl64<6>(A[7] ^ D2);  B[21] = rotl64<55>(A[8] ^ D3);
    B[6] = rotl64<20>(A[9] ^ D4);  B[7] = rotl64<3>(A[10] ^ D0);  B[17] = rotl64<10>(A[11] ^ D1);
    B[2] = rotl64<43>(A[12] ^ D2); B[12] = rotl64<25>(A[13] ^ D3); B[22] = rotl64<39>(A[14] ^ D4);
    B[23] = rotl64<41>(A[15] ^ D0); B[8] = rotl64<45>(A[16] ^ D1); B[18] = rotl64<15>(A[17] ^ D2);
    B[3] = rotl64<21>(A[18] ^ D3); B[13] = rotl64<8>(A[19] ^ D4); B[14] = rotl64<18>(A[20] ^ D0);
    B[24] = rotl64<2>(A[21] ^ D1); B[9] = rotl64<61>(A[22] ^ D2); B[19] = rotl64<56>(A[23] ^ D3);
    B[4] = rotl64<14>(A[24] ^ D4);
#pragma unroll
    for (int y = 0; y < 5; ++y) {
      uint64_t b0 = B[5 * y], b1 = B[5 * y + 1], b2 = B[5 * y + 2], b3 = B[5 * y + 3], b4 = B[5 * y + 4];
      A[5 * y + 0] = b0 ^ (~b1 & b2);
      A[5 * y + 1] = b1 ^ (~b2 & b3);
      A[5 * y + 2] = b2 ^ (~b3 & b4);
      A[5 * y + 3] = b3 ^ (~b4 & b0);
      A[5 * y + 4] = b4 ^ (~b0 & b1);
    }
    A[0] ^= ConstRC().get(round);
  }
}

__device__ __forceinline__ void shake256_blocks_32_c(const uint64_t *__restrict__ msg, int nblocks, uint64_t out[4]) {
  uint64_t A[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = 0;
  for (int b = 0; b < nblocks; ++b) {
#pragma unroll
    for (int w = 0; w < 17; ++w) A[w] ^= msg[17 * b + w];
    keccak_f1600_c(A);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) out[i] = A[i];
}

// SHAKE256 over a message already laid out in memory as whole 136-byte blocks with the
// padding applied (0x1F after the message, 0x80 in the last byte); writes 32 bytes.
__device__ __forceinline__ void shake256_blocks_32(const uint64_t *__restrict__ msg, int nblocks,
                                                   uint64_t out[4], const KeccakRC &rc) {
  uint64_t A[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = 0;
  for (int b = 0; b < nblocks; ++b) {
#pragma unroll
    for (int w = 0; w < 17; ++w) A[w] ^= msg[17 * b + w];
    keccak_f1600(A, rc);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) out[i] = A[i];
}

// ------------------------------------------------------------------ GF(2^m) through LDS tables
// lg[a] = discrete log of a (lg[0] unused), ex[i] = g^i for 0 <= i < q, q = 2^m - 1.
struct GfLds {
  const uint16_t *lg;
  const uint16_t *ex;
  uint32_t q;
  __device__ __forceinline__ uint32_t addmod(uint32_t a, uint32_t b) const {
    uint32_t s = a + b;
    return s >= q ? s - q : s;
  }
  __device__ __forceinline__ uint32_t mul(uint32_t a, uint32_t b) const {
    uint32_t r = ex[addmod(lg[a], lg[b])];
    return (a == 0u || b == 0u) ? 0u : r;
  }
  // a * g^lb where lb is a log (lb == 0xFFFF encodes the factor 0)
  __device__ __forceinline__ uint32_t mul_log(uint32_t a, uint32_t lb) const {
    uint32_t r = ex[addmod(lg[a], lb == 0xFFFFu ? 0u : lb)];
    return (a == 0u || lb == 0xFFFFu) ? 0u : r;
  }
};

// ------------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t wave_xor(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v ^= (uint32_t)__shfl_xor((int)v, o);
  return v;
}

// Orders this wave's LDS accesses across lanes (LDS executes one wave's requests in order; the
// fences stop the compiler from moving accesses across this point).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Bitonic sorting network over a register array (all indices compile-time).
template <int NN>
__device__ __forceinline__ void bitonic_sort(uint32_t (&v)[NN]) {
#pragma unroll
  for (int k = 2; k <= NN; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < NN; ++i) {
        const int l = i ^ j;
        if (l > i) {
          uint32_t a = v[i], b = v[l];
          uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
          if ((i & k) == 0) { v[i] = lo; v[l] = hi; }
          else { v[i] = hi; v[l] = lo; }
        }
      }
}

}  // namespace tmce
