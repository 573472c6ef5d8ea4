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

// gfx950 three-input bitwise operation (v_bitop3_b32): bit i of `lut` is the result for the input
// bits (a, b, c) with i = 4a + 2b + c.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ uint32_t chi3(uint32_t a, uint32_t b, uint32_t c) {  // a ^ (~b & c)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0xD2);
}
// (hi:lo) rotated left by R into (oh:ol)
template <int R>
__device__ __forceinline__ void rotl2(uint32_t lo, uint32_t hi, uint32_t &ol, uint32_t &oh) {
  if constexpr (R >= 32) {
    rotl2<R - 32>(hi, lo, ol, oh);
  } else if constexpr (R == 0) {
    ol = lo;
    oh = hi;
  } else {
    oh = __builtin_amdgcn_alignbit(hi, lo, 32 - R);
    ol = __builtin_amdgcn_alignbit(lo, hi, 32 - R);
  }
}

// Keccak-f[1600] on one lane, the state as 25 (lo, hi) register pairs.  Per round: column parities
// C[x] by two 3-input XORs per half, theta as A ^= C[x-1] ^ rotl1(C[x+1]) in one 3-input XOR per
// half, rho + pi as 24 two-word funnel rotations, chi as one 3-input operation per half, iota.
template <class RC>
__device__ __forceinline__ void keccak_rounds(uint64_t (&A)[25], const RC &rc) {
  uint32_t lo[25], hi[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) {
    lo[i] = (uint32_t)A[i];
    hi[i] = (uint32_t)(A[i] >> 32);
  }
#pragma unroll 1
  for (int round = 0; round < 24; ++round) {
    uint32_t cl[5], ch[5], rl[5], rh[5];
#pragma unroll
    for (int x = 0; x < 5; ++x) {
      cl[x] = xor3(xor3(lo[x], lo[x + 5], lo[x + 10]), lo[x + 15], lo[x + 20]);
      ch[x] = xor3(xor3(hi[x], hi[x + 5], hi[x + 10]), hi[x + 15], hi[x + 20]);
    }
#pragma unroll
    for (int x = 0; x < 5; ++x) rotl2<1>(cl[x], ch[x], rl[x], rh[x]);
#pragma unroll
    for (int i = 0; i < 25; ++i) {
      const int x = i % 5;
      lo[i] = xor3(lo[i], cl[(x + 4) % 5], rl[(x + 1) % 5]);
      hi[i] = xor3(hi[i], ch[(x + 4) % 5], rh[(x + 1) % 5]);
    }
    // rho + pi: B[y + 5*((2x+3y)%5)] = rot(A[x+5y], r[x][y])
    uint32_t bl[25], bh[25];
    rotl2<0>(lo[0], hi[0], bl[0], bh[0]);
    rotl2<1>(lo[1], hi[1], bl[10], bh[10]);
    rotl2<62>(lo[2], hi[2], bl[20], bh[20]);
    rotl2<28>(lo[3], hi[3], bl[5], bh[5]);
    rotl2<27>(lo[4], hi[4], bl[15], bh[15]);
    rotl2<36>(lo[5], hi[5], bl[16], bh[16]);
    rotl2<44>(lo[6], hi[6], bl[1], bh[1]);
    rotl2<6>(lo[7], hi[7], bl[11], bh[11]);
    rotl2<55>(lo[8], hi[8], bl[21], bh[21]);
    rotl2<20>(lo[9], hi[9], bl[6], bh[6]);
    rotl2<3>(lo[10], hi[10], bl[7], bh[7]);
    rotl2<10>(lo[11], hi[11], bl[17], bh[17]);
    rotl2<43>(lo[12], hi[12], bl[2], bh[2]);
    rotl2<25>(lo[13], hi[13], bl[12], bh[12]);
    rotl2<39>(lo[14], hi[14], bl[22], bh[22]);
    rotl2<41>(lo[15], hi[15], bl[23], bh[23]);
    rotl2<45>(lo[16], hi[16], bl[8], bh[8]);
    rotl2<15>(lo[17], hi[17], bl[18], bh[18]);
    rotl2<21>(lo[18], hi[18], bl[3], bh[3]);
    rotl2<8>(lo[19], hi[19], bl[13], bh[13]);
    rotl2<18>(lo[20], hi[20], bl[14], bh[14]);
    rotl2<2>(lo[21], hi[21], bl[24], bh[24]);
    rotl2<61>(lo[22], hi[22], bl[9], bh[9]);
    rotl2<56>(lo[23], hi[23], bl[19], bh[19]);
    rotl2<14>(lo[24], hi[24], bl[4], bh[4]);
    // chi
#pragma unroll
    for (int y = 0; y < 5; ++y)
#pragma unroll
      for (int x = 0; x < 5; ++x) {
        lo[5 * y + x] = chi3(bl[5 * y + x], bl[5 * y + (x + 1) % 5], bl[5 * y + (x + 2) % 5]);
        hi[5 * y + x] = chi3(bh[5 * y + x], bh[5 * y + (x + 1) % 5], bh[5 * y + (x + 2) % 5]);
      }
    // iota
    const uint64_t c = rc.get(round);
    lo[0] ^= (uint32_t)c;
    hi[0] ^= (uint32_t)(c >> 32);
  }
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = (uint64_t)lo[i] | ((uint64_t)hi[i] << 32);
}

// All lanes of the wave must be active when `rc` is read (it is broadcast with v_readlane).
__device__ __forceinline__ void keccak_f1600(uint64_t (&A)[25], const KeccakRC &rc) { keccak_rounds(A, rc); }

// Same permutation with the round constant read from constant memory: safe in divergent code
// (the plain per-thread kernels), at the cost of a scalar load per round.
__device__ __forceinline__ void keccak_f1600_c(uint64_t (&A)[25]) {
  struct ConstRC {
    __device__ __forceinline__ uint64_t get(int r) const { return kKeccakRC[r]; }
  };
  keccak_rounds(A, ConstRC());
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
// padding applied (0x1F after the message, 0x80 in the last byte); writes 32 bytes.  With
// PREFETCH the next block's 17 words are loaded before the current block's permutation, so their
// memory latency overlaps it (34 more VGPRs live through the permutation).
template <bool PREFETCH = true>
__device__ __forceinline__ void shake256_blocks_32(const uint64_t *__restrict__ msg, int nblocks,
                                                   uint64_t out[4], const KeccakRC &rc) {
  uint64_t A[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = 0;
  if constexpr (PREFETCH) {
    uint64_t nxt[17];
#pragma unroll
    for (int w = 0; w < 17; ++w) nxt[w] = msg[w];
    for (int b = 0; b < nblocks; ++b) {
#pragma unroll
      for (int w = 0; w < 17; ++w) A[w] ^= nxt[w];
      if (b + 1 < nblocks) {
#pragma unroll
        for (int w = 0; w < 17; ++w) nxt[w] = msg[17 * (b + 1) + w];
      }
      keccak_f1600(A, rc);
    }
  } else {
    for (int b = 0; b < nblocks; ++b) {
#pragma unroll
      for (int w = 0; w < 17; ++w) A[w] ^= msg[17 * b + w];
      keccak_f1600(A, rc);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) out[i] = A[i];
}

// SHAKE256 over the first LEN bytes of `msg` with byte 0 replaced by `domain` (the bytes from LEN
// on are ignored; the padding is applied here); writes 32 bytes.  msg must hold whole 136-byte
// blocks up to the block containing byte LEN.
template <int LEN>
__device__ __forceinline__ void shake256_prefix_32(const uint64_t *__restrict__ msg, uint32_t domain, uint64_t out[4],
                                                   const KeccakRC &rc) {
  constexpr int NB = LEN / 136 + 1;  // blocks after padding
  constexpr int LB = 136 * (NB - 1);  // first byte of the last block
  uint64_t A[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = 0;
  for (int b = 0; b < NB - 1; ++b) {  // whole blocks of message bytes
#pragma unroll
    for (int w = 0; w < 17; ++w) {
      uint64_t v = msg[17 * b + w];
      if (w == 0 && b == 0) v = (v & ~0xFFull) | domain;
      A[w] ^= v;
    }
    keccak_f1600(A, rc);
  }
#pragma unroll
  for (int w = 0; w < 17; ++w) {  // the last block: message tail, 0x1F, zeros, 0x80
    const int b0 = LB + 8 * w;
    uint64_t v = 0;
    if (b0 < LEN) {
      v = msg[17 * (NB - 1) + w];
      if (b0 + 8 > LEN) v &= (1ull << (8 * (LEN - b0))) - 1ull;
    }
    if (b0 == 0) v = (v & ~0xFFull) | domain;
    if (b0 <= LEN && LEN < b0 + 8) v |= 0x1Full << (8 * (LEN - b0));
    if (w == 16) v |= 0x80ull << 56;
    A[w] ^= v;
  }
  keccak_f1600(A, rc);
#pragma unroll
  for (int i = 0; i < 4; ++i) out[i] = A[i];
}

// ------------------------------------------------------------------ Keccak on a lane pair
// One state per pair of adjacent lanes: lane p = 0 holds the low 32 bits of each of the 25 state
// words, lane p = 1 the high 32 bits.  A 64-bit rotation by R gives both lanes the same form
// (R < 32: alignbit(own, partner, 32 - R); R > 32: alignbit(partner, own, 64 - R)), the partner's
// half coming by a DPP swap within the pair.  ~120 instructions per round per lane against 182 for
// a whole state in one lane: lower latency per permutation when a batch is too small to fill the
// chip one op per lane.  Every lane of the wave must be active (round constants by v_readlane,
// DPP swaps).
__device__ __forceinline__ uint32_t pair_swap(uint32_t v) {  // quad_perm [1, 0, 3, 2]
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);
}
template <int R>
__device__ __forceinline__ uint32_t rotl_pair(uint32_t own) {
  if constexpr (R == 0) {
    return own;
  } else if constexpr (R == 32) {
    return pair_swap(own);
  } else if constexpr (R < 32) {
    return __builtin_amdgcn_alignbit(own, pair_swap(own), 32 - R);
  } else {
    return __builtin_amdgcn_alignbit(pair_swap(own), own, 64 - R);
  }
}
__device__ __forceinline__ void keccak_f1600_pair(uint32_t (&s)[25], bool hi, const KeccakRC &rc) {
#pragma unroll 1
  for (int round = 0; round < 24; ++round) {
    uint32_t c[5], r1[5];
#pragma unroll
    for (int x = 0; x < 5; ++x) c[x] = xor3(xor3(s[x], s[x + 5], s[x + 10]), s[x + 15], s[x + 20]);
#pragma unroll
    for (int x = 0; x < 5; ++x) r1[x] = __builtin_amdgcn_alignbit(c[x], pair_swap(c[x]), 31);  // rotl 1
#pragma unroll
    for (int i = 0; i < 25; ++i) s[i] = xor3(s[i], c[(i % 5 + 4) % 5], r1[(i % 5 + 1) % 5]);
    // rho + pi: B[y + 5*((2x+3y)%5)] = rot(A[x+5y], r[x][y])
    uint32_t b[25];
    b[0] = s[0];
    b[10] = rotl_pair<1>(s[1]);
    b[20] = rotl_pair<62>(s[2]);
    b[5] = rotl_pair<28>(s[3]);
    b[15] = rotl_pair<27>(s[4]);
    b[16] = rotl_pair<36>(s[5]);
    b[1] = rotl_pair<44>(s[6]);
    b[11] = rotl_pair<6>(s[7]);
    b[21] = rotl_pair<55>(s[8]);
    b[6] = rotl_pair<20>(s[9]);
    b[7] = rotl_pair<3>(s[10]);
    b[17] = rotl_pair<10>(s[11]);
    b[2] = rotl_pair<43>(s[12]);
    b[12] = rotl_pair<25>(s[13]);
    b[22] = rotl_pair<39>(s[14]);
    b[23] = rotl_pair<41>(s[15]);
    b[8] = rotl_pair<45>(s[16]);
    b[18] = rotl_pair<15>(s[17]);
    b[3] = rotl_pair<21>(s[18]);
    b[13] = rotl_pair<8>(s[19]);
    b[14] = rotl_pair<18>(s[20]);
    b[24] = rotl_pair<2>(s[21]);
    b[9] = rotl_pair<61>(s[22]);
    b[19] = rotl_pair<56>(s[23]);
    b[4] = rotl_pair<14>(s[24]);
#pragma unroll
    for (int y = 0; y < 5; ++y)
#pragma unroll
      for (int x = 0; x < 5; ++x) s[5 * y + x] = chi3(b[5 * y + x], b[5 * y + (x + 1) % 5], b[5 * y + (x + 2) % 5]);
    const uint32_t clo = (uint32_t)__builtin_amdgcn_readlane((int)rc.lo, round);
    const uint32_t chi = (uint32_t)__builtin_amdgcn_readlane((int)rc.hi, round);
    s[0] ^= hi ? chi : clo;
  }
}
// SHAKE256 over whole padded 136-byte blocks (as shake256_blocks_32) on a lane pair: this lane's
// halves of the 17 rate words of each block are the 32-bit words 2 w + hi of the block; the next
// block is loaded before the current one is permuted.  out[i] = this lane's half of output word i.
__device__ __forceinline__ void shake256_blocks_32_pair(const uint32_t *__restrict__ msg, int nblocks, bool hi,
                                                        uint32_t out[4], const KeccakRC &rc) {
  uint32_t s[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) s[i] = 0;
  const uint32_t *m = msg + (hi ? 1 : 0);
  uint32_t nxt[17];
#pragma unroll
  for (int w = 0; w < 17; ++w) nxt[w] = m[2 * w];
  for (int b = 0; b < nblocks; ++b) {
#pragma unroll
    for (int w = 0; w < 17; ++w) s[w] ^= nxt[w];
    if (b + 1 < nblocks) {
#pragma unroll
      for (int w = 0; w < 17; ++w) nxt[w] = m[34 * (b + 1) + 2 * w];
    }
    keccak_f1600_pair(s, hi, rc);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) out[i] = s[i];
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
