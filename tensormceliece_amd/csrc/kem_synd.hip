// kem_synd.hip -- batched syndrome (PKE encryption) c = H e = e[0:MT) + T e[MT:N) over GF(2), three
// ways, for error vectors given as libmceliece's e bytes (N/8 per op):
//
//   k_synd_mfma    the dense GF(2) GEMM on CDNA4 INT8 MFMA (v_mfma_i32_16x16x64_i8): 0/1 bytes,
//                  int32 accumulation (exact: K < 2^31), parity = bit 0.  The public key is held
//                  per key as 0/1 bytes (Tb, [MTP][KP], L2 resident); the workgroup's error bits
//                  sit in LDS and are expanded to bytes on the way into the A operand.
//   k_synd_valu    the bit-packed GEMM on the VALU: one op per lane, T words wave-uniform
//                  (scalar loads), 32-bit AND and 3-input XOR per word pair, parity by popcount.
//   k_synd_sparse  the column gather: c = XOR of the columns of [I | T] at the set bits of e
//                  (column-major key, as in the encapsulation engine).
//
// Semantics: LM:crypto_kem/<set>/vec/encrypt.c syndrome() (:94-132); the reference computes the
// same product as a batched 1-bit / 0-1-byte GEMM (mceliece-gpu-bmma/tc_kernels.cuh:97-123,
// int8_tensor_enc.h:60-102) with the finalize c = e[0:MT) + s (tc_kernels.cuh:126-147).  Bits of
// the last ciphertext byte at positions >= MT are zero (6960119).
#include "kem_kernels.h"

namespace tmce {

typedef int v4i __attribute__((ext_vector_type(4)));

// Bits [MT + 32 w, MT + 32 w + 32) of e (zero beyond N).
template <class P>
__device__ __forceinline__ uint32_t e_ns_word(const uint8_t *__restrict__ erow, int w) {
  if constexpr (P::MT % 32 == 0 && P::N % 32 == 0) {
    const int wi = P::MT / 32 + w;
    return wi < P::N / 32 ? reinterpret_cast<const uint32_t *>(erow)[wi] : 0u;
  } else {
    const int b = P::MT + 32 * w;
    if (b >= P::N) return 0u;
    const int B = b >> 3;
    uint64_t v = 0;
#pragma unroll
    for (int i = 0; i < 5; ++i)
      if (B + i < P::E_BYTES) v |= (uint64_t)erow[B + i] << (8 * i);
    uint32_t r = (uint32_t)(v >> (b & 7));
    if (P::N - b < 32) r &= (1u << (P::N - b)) - 1u;
    return r;
  }
}

// 16 bits -> 16 bytes of 0/1 (byte i = bit i): per nibble, the multiply puts bit j at bits j + 7i
// (i = 0..3, all distinct for j < 4, so no carries) and the mask keeps bit j at byte j.
__device__ __forceinline__ v4i expand16(uint32_t x) {
  v4i r;
  r[0] = (int)(__umul24(x & 0xFu, 0x00204081u) & 0x01010101u);
  r[1] = (int)(__umul24((x >> 4) & 0xFu, 0x00204081u) & 0x01010101u);
  r[2] = (int)(__umul24((x >> 8) & 0xFu, 0x00204081u) & 0x01010101u);
  r[3] = (int)(__umul24((x >> 12) & 0xFu, 0x00204081u) & 0x01010101u);
  return r;
}

// u16 pair of ciphertext bytes [2h, 2h+2) = e's first bytes, masked to the MT bits of H e.
template <class P>
__device__ __forceinline__ uint32_t e_head16(const uint8_t *__restrict__ erow, int h) {
  uint32_t v = reinterpret_cast<const uint16_t *>(erow)[h];  // e rows are 2-byte aligned (N/8 even)
  if constexpr (P::MT % 16 != 0) {
    if (2 * h + 2 > P::MT / 8) {
      const int keep = P::MT - 16 * h;  // bits of this pair below MT
      v &= keep >= 16 ? 0xFFFFu : ((1u << keep) - 1u);
    }
  }
  return v;
}

// ---------------------------------------------------------------------------------- INT8 MFMA
// Generic batched GF(2) matrix product on v_mfma_i32_16x16x64_i8: for every op, output bit r =
// parity(M[r] . x) with x a bit row of the op's input.  0/1 bytes in, int32 accumulation (exact),
// bit 0 out.  G describes the product (rows, depth, where x sits in the input row, output row).
//
// The matrix is stored swizzled per 16x64 MFMA fragment: Msw[((ks * RTP) + rt) * 1024 + 16 l + b] =
// M[16 rt + l % 16][64 ks + 16 (l / 16) + b], so one fragment is 1 KB contiguous and the fragments
// of consecutive row tiles at one k step are adjacent.
//
// Workgroup tile: BM = 16*OPW*WO ops x BN = 16*RTW*WR matrix rows, WO x WR waves, each wave
// OPW x RTW 16x16 int32 tiles.  The k loop runs in super-steps of KSUB MFMA steps (64 deep each)
// over a three-slot LDS ring filled two super-steps ahead by global_load_lds (HBM/L2 -> LDS with no
// registers):
//   B (matrix): BN/16 fragments (1 KB each) per step, 16 bytes per lane;
//   A (x bits): BM x 64 bits per step, one aligned 32-bit word per lane, expanded to 0/1 bytes on
//               the way into the MFMA A operand.
// The copies are issued from inline assembly with explicit vmcnt waits (the compiler, which cannot
// tell the ring slots apart, would otherwise wait for every outstanding copy before any LDS read).
// Before the barrier ending super-step q a wave waits until only the copies it issued for q+2 are
// outstanding, so slot q+1 is complete for every wave after the barrier.
// Output: the D tile gives lane l column l%16 (a matrix row) of rows 4*(l/16)+v (ops); a wave ballot
// of bit 0 per v gives, in 16-bit chunk g, 16 consecutive output bits of op 4g+v, stored as one
// 16-bit word.
template <class G, int OPW, int RTW, int WO, int WR, int KSUB>
struct GemmShape {
  static constexpr int NT = 64 * WO * WR;
  static constexpr int BM = 16 * OPW * WO, BN = 16 * RTW * WR;
  static constexpr int KS = G::KP / 64;
  static constexpr int QS = ceil_div(KS, KSUB);     // super-steps
  static constexpr int NF = BN / 16;                // fragments per step
  static constexpr int FPW = NF / (WO * WR);        // fragments per wave per step
  static constexpr int APT = 2 * BM / NT;           // x words per thread per step
  static constexpr int BSTEP = NF * 1024, ASTEP = BM * 8;
  static constexpr int SLOT = KSUB * (BSTEP + ASTEP);
  static constexpr int NI = KSUB * (FPW + APT);     // copies one wave issues per super-step
  static constexpr size_t LDS = (size_t)3 * SLOT;
  static_assert(FPW >= 1 && FPW * WO * WR == NF, "fragments must split evenly over the waves");
  static_assert(APT >= 1 && APT * NT == 2 * BM, "x words must split evenly over the threads");
  static_assert(ceil_div(G::ROWS, BN) * NF <= G::RTP, "row blocks exceed the padded matrix");
  static_assert(NI < 64, "vmcnt range");
  static_assert(PLAN_ALIGN % BM == 0, "plan key groups must hold whole workgroup tiles");
};

__device__ __forceinline__ uint32_t lds_off(const void *p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}
// LDS DMA: lane l's 16 (4) bytes at g land at LDS address m0 + 16 l (4 l); m0 wave-uniform.
__device__ __forceinline__ void dma16(const void *g, uint32_t m0) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "s"(m0) : "memory", "m0");
}
__device__ __forceinline__ void dma4(const void *g, uint32_t m0) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dword %0, off" ::"v"(g), "s"(m0) : "memory", "m0");
}
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// xw: x as aligned 32-bit words, row stride xstride words, xvalid words per row (reads past it are
// clamped to the last word: the matrix is zero there).  head: the original input rows (G::head16).
template <class G, int OPW, int RTW, int WO, int WR, int KSUB>
__global__ __launch_bounds__(64 * WO * WR) void k_gf2_mfma(const uint32_t *__restrict__ xw, int xstride, int xvalid,
                                                          const uint8_t *__restrict__ head,
                                                          const uint4 *__restrict__ Msw,
                                                          uint8_t *__restrict__ out, int nops,
                                                          const PlanDev *__restrict__ plan) {
  using S = GemmShape<G, OPW, RTW, WO, WR, KSUB>;
  extern __shared__ __align__(16) uint8_t lds[];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63, li = lane & 15, lg = lane >> 4;
  const int wo = wave % WO, wr = wave / WO;
  const int op0 = blockIdx.x * S::BM;
  if (plan) {  // key groups are PLAN_ALIGN-aligned (a multiple of BM): one key per workgroup
    nops = plan->total;
    if (op0 >= nops) return;
    Msw = uniform_ptr(reinterpret_cast<const uint4 *>(G::plan_matrix(plan->keys[plan_group(plan, op0)])));
  }
  const int rt0 = blockIdx.y * S::NF;  // first row tile of the block
  const uint32_t lds0 = lds_off(lds);

  // x word i = tid + NT*p of a step is word 2 ks + i%2 of op i/2 (rows past the batch: last op)
  const uint32_t *xrow[S::APT];
#pragma unroll
  for (int p = 0; p < S::APT; ++p) {
    const int op = op0 + (tid + S::NT * p) / 2;
    xrow[p] = xw + (size_t)(op < nops ? op : nops - 1) * xstride;
  }
  auto issue = [&](int q, int slot) {
#pragma unroll
    for (int u = 0; u < KSUB; ++u) {
      const int ks = min(q * KSUB + u, S::KS - 1);  // a short last super-step repeats a valid step
      const uint32_t sb = lds0 + slot * S::SLOT + u * S::BSTEP;
#pragma unroll
      for (int f = 0; f < S::FPW; ++f) {
        const int fr = wave * S::FPW + f;
        dma16(Msw + ((size_t)ks * G::RTP + rt0 + fr) * 64 + lane, __builtin_amdgcn_readfirstlane(sb + fr * 1024));
      }
      const uint32_t sa = lds0 + slot * S::SLOT + KSUB * S::BSTEP + u * S::ASTEP;
#pragma unroll
      for (int p = 0; p < S::APT; ++p) {
        const int w = min(2 * ks + ((tid + S::NT * p) & 1), xvalid - 1);
        dma4(xrow[p] + w, __builtin_amdgcn_readfirstlane(sa + (wave * 64 + S::NT * p) * 4));
      }
    }
  };

  issue(0, 0);
  if (S::QS > 1) {
    issue(1, 1);
    wait_vmcnt<S::NI>();
  } else {
    wait_vmcnt<0>();
  }
  __syncthreads();

  v4i acc[OPW][RTW];
#pragma unroll
  for (int i = 0; i < OPW; ++i)
#pragma unroll
    for (int j = 0; j < RTW; ++j) acc[i][j] = v4i{0, 0, 0, 0};

  int slot = 0;
#pragma unroll 1
  for (int q = 0; q < S::QS; ++q) {
    const int s1 = slot == 2 ? 0 : slot + 1, s2 = s1 == 2 ? 0 : s1 + 1;
    const bool ahead = q + 2 < S::QS;
    if (ahead) issue(q + 2, s2);
#pragma unroll
    for (int u = 0; u < KSUB; ++u) {
      if (KSUB == 1 || q * KSUB + u < S::KS) {
        const uint8_t *sb = lds + slot * S::SLOT + u * S::BSTEP;
        const uint16_t *sa = reinterpret_cast<const uint16_t *>(lds + slot * S::SLOT + KSUB * S::BSTEP + u * S::ASTEP);
        v4i b[RTW];
#pragma unroll
        for (int j = 0; j < RTW; ++j) {
          const uint4 bv = reinterpret_cast<const uint4 *>(sb)[(RTW * wr + j) * 64 + lane];
          b[j] = v4i{(int)bv.x, (int)bv.y, (int)bv.z, (int)bv.w};
        }
#pragma unroll
        for (int i = 0; i < OPW; ++i) {
          const v4i a = expand16(sa[(16 * OPW * wo + 16 * i + li) * 4 + lg]);
#pragma unroll
          for (int j = 0; j < RTW; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[j], acc[i][j], 0, 0, 0);
        }
      }
    }
    if (ahead) wait_vmcnt<S::NI>();
    else wait_vmcnt<0>();
    __syncthreads();
    slot = s1;
  }

  // parity bits -> 16-bit words in LDS ([op][row tile] per wave; the ring is free after the last
  // barrier), then one op per lane: XOR the head bits and store the wave's words of the op's row
  uint16_t *ep = reinterpret_cast<uint16_t *>(lds) + wave * (16 * OPW) * RTW;
#pragma unroll
  for (int j = 0; j < RTW; ++j)
#pragma unroll
    for (int i = 0; i < OPW; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint64_t m = __ballot(acc[i][j][v] & 1);
        if (li == 0) ep[(16 * i + 4 * lg + v) * RTW + j] = (uint16_t)(m >> (16 * lg));
      }
  wave_lds_sync();
#pragma unroll
  for (int o = lane; o < 16 * OPW; o += 64) {
    const int op = op0 + 16 * OPW * wo + o;
    if (op < nops) {
      const uint8_t *hrow = head + (size_t)op * G::IN_BYTES;
      uint16_t *orow = reinterpret_cast<uint16_t *>(out + (size_t)op * G::OUT_BYTES);
#pragma unroll
      for (int j = 0; j < RTW; ++j) {
        const int h = rt0 + RTW * wr + j;  // 16-bit word index of the output row
        if (16 * h < G::OUT_BITS) orow[h] = (uint16_t)(ep[o * RTW + j] ^ G::head16(hrow, h));
      }
    }
  }
}

// Syndrome: rows = MT, x = e bits [MT, N), out = ciphertext (c = e[0:MT) + s).  When MT and N are
// multiples of 32 the x words are read in place; otherwise (6960119) they are first repacked.
template <int SET>
struct GemmSynd {
  using P = Par<SET>;
  static constexpr int ROWS = P::MT, KB = P::K, KP = round_up(KB, 64);
  static constexpr int RTP = round_up(ceil_div(ROWS, 16), 16);
  static constexpr int IN_BYTES = P::E_BYTES, OUT_BYTES = P::SYND_BYTES, OUT_BITS = 8 * P::SYND_BYTES;
  static constexpr bool DIRECT = P::MT % 32 == 0 && P::N % 32 == 0;
  static constexpr int XOFF_WORDS = P::MT / 32, XSTRIDE = P::E_BYTES / 4, XVALID = P::K / 32;
  __device__ static uint32_t x_word(const uint8_t *row, int w) { return e_ns_word<P>(row, w); }
  __device__ static uint32_t head16(const uint8_t *row, int h) { return e_head16<P>(row, h); }
  __device__ static const void *plan_matrix(const KeyDev &) { return nullptr; }  // no plan form
};

// Decapsulation front end: rows = R (bit b of syndrome j at row j*M + b), x = ciphertext bits
// (KBITS), out = syndrome bit planes [RW] words.
template <int SET>
struct GemmFront {
  using P = Par<SET>;
  static constexpr int ROWS = P::R, KB = P::KBITS, KP = round_up(KB, 64);
  static constexpr int RTP = round_up(ceil_div(ROWS, 16), 16);
  static constexpr int IN_BYTES = P::SYND_BYTES, OUT_BYTES = 4 * P::RW, OUT_BITS = P::R;
  static constexpr bool DIRECT = P::SYND_BYTES % 4 == 0;
  static constexpr int XOFF_WORDS = 0, XSTRIDE = P::SYND_BYTES / 4, XVALID = P::SYND_BYTES / 4;
  __device__ static uint32_t x_word(const uint8_t *row, int w) {  // ciphertext rows are 2-byte aligned
    const uint16_t *h = reinterpret_cast<const uint16_t *>(row);
    uint32_t v = 0;
    if (4 * w < P::SYND_BYTES) v = h[2 * w];
    if (4 * w + 2 < P::SYND_BYTES) v |= (uint32_t)h[2 * w + 1] << 16;
    return v;
  }
  __device__ static uint32_t head16(const uint8_t *, int) { return 0u; }
  __device__ static const void *plan_matrix(const KeyDev &k) { return k.Fsw; }
};

// x rows -> aligned words [n][KP/32] (for the sets where they are not aligned in place)
template <class G>
__global__ void k_repack_x(const uint8_t *__restrict__ x, uint32_t *__restrict__ xw, int nops,
                           const PlanDev *__restrict__ plan) {
  constexpr int W = G::KP / 32;
  if (plan) nops = plan->total;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nops * W) return;
  const int op = (int)(i / W), w = (int)(i % W);
  xw[i] = G::x_word(x + (size_t)op * G::IN_BYTES, w);
}

template <class G>
struct MswDims {
  static constexpr int KS = G::KP / 64;
  static constexpr size_t BYTES = (size_t)KS * G::RTP * 1024;
};

template <class G, int OPW, int RTW, int WO, int WR, int KSUB>
hipError_t launch_gf2_cfg(const uint32_t *xw, int xstride, int xvalid, const uint8_t *x, const uint8_t *Msw,
                          uint8_t *out, int nops, const PlanDev *plan, const char *name, hipStream_t s) {
  using S = GemmShape<G, OPW, RTW, WO, WR, KSUB>;
  static const hipError_t attr = hipFuncSetAttribute(
      reinterpret_cast<const void *>(&k_gf2_mfma<G, OPW, RTW, WO, WR, KSUB>),
      hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::LDS);
  if (attr != hipSuccess) return attr;
  dim3 grid(ceil_div(nops, S::BM), ceil_div(G::ROWS, S::BN));
  TMCE_TIMED(name, s,
             (k_gf2_mfma<G, OPW, RTW, WO, WR, KSUB><<<grid, S::NT, S::LDS, s>>>(
                 xw, xstride, xvalid, x, reinterpret_cast<const uint4 *>(Msw), out, nops, plan)));
  return hipGetLastError();
}

// Tile shapes offered to the tuner: (OPW, RTW, WO, WR, KSUB) -> BM x BN workgroup tiles
//   0: 256 x 128, 8 waves, 1 step per barrier     1: 256 x 128, 8 waves, 2 steps per barrier
//   2: 256 x 256, 8 waves, 2 steps per barrier    3: 256 x 256, 8 waves (128 ops each), 2 steps
//   4: 256 x 128, 4 waves (128 ops each), 2 steps 5: 256 x 256, 8 waves, 1 step per barrier
// x: input rows (IN_BYTES each); scratch: n * KP/8 bytes when the x bits are not word aligned in
// place (G::DIRECT false), else unused.
template <class G>
hipError_t launch_gf2(const uint8_t *x, uint32_t *scratch, const uint8_t *Msw, uint8_t *out, int nops, int cfg,
                      const PlanDev *plan, const char *name, hipStream_t s) {
  const uint32_t *xw;
  int xstride, xvalid;
  if constexpr (G::DIRECT) {
    xw = reinterpret_cast<const uint32_t *>(x) + G::XOFF_WORDS;
    xstride = G::XSTRIDE;
    xvalid = G::XVALID;
  } else {
    if (!scratch) return hipErrorInvalidValue;
    const size_t nw = (size_t)nops * (G::KP / 32);
    TMCE_TIMED("gf2_repack", s, k_repack_x<G><<<(unsigned)((nw + 255) / 256), 256, 0, s>>>(x, scratch, nops, plan));
    xw = scratch;
    xstride = xvalid = G::KP / 32;
  }
  switch (cfg) {
    case 0: return launch_gf2_cfg<G, 4, 4, 4, 2, 1>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 1: return launch_gf2_cfg<G, 4, 4, 4, 2, 2>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 2: return launch_gf2_cfg<G, 4, 8, 4, 2, 2>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 3: return launch_gf2_cfg<G, 8, 4, 2, 4, 2>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 4: return launch_gf2_cfg<G, 8, 4, 2, 2, 2>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 5: return launch_gf2_cfg<G, 4, 8, 4, 2, 1>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------- VALU GEMM
// One op per lane, 256 ops per workgroup (128 for the larger sets); the workgroup's error words sit
// in LDS (lane-private rows, odd stride).  T is stored per block of 32 rows as [word][row] so the 32 row words of one k word
// are contiguous and wave-uniform (scalar loads).  acc_r ^= T[r][w] & e[w], parity = popcount & 1.
template <class P>
struct SyndValuDims {
  static constexpr int KW = ceil_div(P::K, 32);
  static constexpr int ES = KW | 1;
  static constexpr int RB = ceil_div(P::MT, 32);
  static constexpr int OPS = 256 * ES * 4 <= 160 * 1024 ? 256 : 128;  // ops (lanes) per workgroup
};

template <int SET>
__global__ __launch_bounds__(SyndValuDims<Par<SET>>::OPS) void k_synd_valu(const uint8_t *__restrict__ e, const uint32_t *__restrict__ Tv,
                                                   uint8_t *__restrict__ ct, int nops) {
  using P = Par<SET>;
  using D = SyndValuDims<P>;
  constexpr int OPS = D::OPS;
  extern __shared__ uint32_t lds_e[];  // [OPS][ES]
  const int tid = threadIdx.x;
  const int op0 = blockIdx.x * OPS;
  for (int i = tid; i < OPS * D::KW; i += OPS) {
    const int o = i / D::KW, w = i - o * D::KW;
    const int op = op0 + o;
    lds_e[o * D::ES + w] = op < nops ? e_ns_word<P>(e + (size_t)op * P::E_BYTES, w) : 0u;
  }
  __syncthreads();
  const int op = op0 + tid;
  const uint32_t *er = lds_e + tid * D::ES;
  const uint8_t *erow = e + (size_t)(op < nops ? op : 0) * P::E_BYTES;
  uint16_t *crow = reinterpret_cast<uint16_t *>(ct + (size_t)op * P::SYND_BYTES);

#pragma unroll 1
  for (int rb = 0; rb < D::RB; ++rb) {
    const uint32_t *tv = Tv + (size_t)rb * D::KW * 32;
    uint32_t acc[32];
#pragma unroll
    for (int r = 0; r < 32; ++r) acc[r] = 0;
#pragma unroll 1
    for (int w = 0; w + 1 < D::KW; w += 2) {
      const uint32_t e0 = er[w], e1 = er[w + 1];
#pragma unroll
      for (int r = 0; r < 32; ++r) acc[r] ^= (tv[w * 32 + r] & e0) ^ (tv[(w + 1) * 32 + r] & e1);
    }
    if constexpr (D::KW & 1) {
      const uint32_t e0 = er[D::KW - 1];
#pragma unroll
      for (int r = 0; r < 32; ++r) acc[r] ^= tv[(D::KW - 1) * 32 + r] & e0;
    }
    uint32_t word = 0;
#pragma unroll
    for (int r = 0; r < 32; ++r) word |= ((uint32_t)__popc(acc[r]) & 1u) << r;
    if (op < nops) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int h = 2 * rb + q;
        if (2 * h < P::SYND_BYTES) crow[h] = (uint16_t)((word >> (16 * q)) ^ e_head16<P>(erow, h));
      }
    }
  }
}

// ---------------------------------------------------------------------------------- sparse
// One op per lane: walk the set bits of e and XOR the columns of [I | T]
// (column-major, CWP words per column, L2 resident).  Works for any weight of e.
template <int SET>
__global__ __launch_bounds__(256) void k_synd_sparse(const uint8_t *__restrict__ e, const uint4 *__restrict__ cols,
                                                     uint8_t *__restrict__ ct, int nops) {
  using P = Par<SET>;
  constexpr int NW = ceil_div(P::N, 32);
  constexpr int C4 = P::CWP / 4;
  const int op = blockIdx.x * 256 + threadIdx.x;
  if (op >= nops) return;
  const uint8_t *er = e + (size_t)op * P::E_BYTES;
  auto word = [&](int w) -> uint32_t {
    if constexpr (P::E_BYTES % 4 == 0) {
      return reinterpret_cast<const uint32_t *>(er)[w];
    } else {  // rows are 2-byte aligned; the last word may be half
      const uint16_t *eh = reinterpret_cast<const uint16_t *>(er);
      uint32_t v = eh[2 * w];
      if (4 * w + 2 < P::E_BYTES) v |= (uint32_t)eh[2 * w + 1] << 16;
      return v;
    }
  };
  uint4 c[C4];
#pragma unroll
  for (int q = 0; q < C4; ++q) c[q] = make_uint4(0, 0, 0, 0);
  int w = 0;
  uint32_t cur = word(0);
  for (;;) {
    while (cur == 0 && w + 1 < NW) cur = word(++w);
    if (cur == 0) break;
    const int p = 32 * w + __ffs(cur) - 1;
    cur &= cur - 1;
    const uint4 *col = cols + (size_t)p * C4;
#pragma unroll
    for (int q = 0; q < C4; ++q) {
      const uint4 x = col[q];
      c[q].x ^= x.x; c[q].y ^= x.y; c[q].z ^= x.z; c[q].w ^= x.w;
    }
  }
  uint32_t cw[P::CWP];
#pragma unroll
  for (int q = 0; q < C4; ++q) { cw[4 * q] = c[q].x; cw[4 * q + 1] = c[q].y; cw[4 * q + 2] = c[q].z; cw[4 * q + 3] = c[q].w; }
  uint16_t *crow = reinterpret_cast<uint16_t *>(ct + (size_t)op * P::SYND_BYTES);
#pragma unroll
  for (int h = 0; h < P::SYND_BYTES / 2; ++h) crow[h] = (uint16_t)(cw[h / 2] >> (16 * (h & 1)));
}

// ---------------------------------------------------------------------------------- launch
template <int SET>
hipError_t launch_synd(const SyndArgs &a, int method, int cfg, hipStream_t s) {
  using P = Par<SET>;
  if (a.nops <= 0) return hipSuccess;
  switch (method) {
    case 0:
      TMCE_TIMED("synd_sparse", s,
                 k_synd_sparse<SET><<<ceil_div(a.nops, 256), 256, 0, s>>>(a.e, reinterpret_cast<const uint4 *>(a.cols),
                                                                       a.ct, a.nops));
      return hipGetLastError();
    case 1:
      return launch_gf2<GemmSynd<SET>>(a.e, a.xscratch, a.Tb, a.ct, a.nops, cfg, nullptr, "synd_mfma", s);
    case 2: {
      using D = SyndValuDims<P>;
      const size_t lds = (size_t)D::OPS * D::ES * 4;
      static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_synd_valu<SET>),
                                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (attr != hipSuccess) return attr;
      TMCE_TIMED("synd_valu", s,
                 k_synd_valu<SET><<<ceil_div(a.nops, D::OPS), D::OPS, lds, s>>>(a.e, a.Tv, a.ct, a.nops));
      return hipGetLastError();
    }
  }
  return hipErrorInvalidValue;
}

// Key operands of the dense forms, built on the device from a row-major bit matrix (rows of `rw`
// 32-bit words, bits at or past `kb` ignored): Msw as described at k_gf2_mfma; Tv[(rb*KW + w)*32 + i]
// = word w of row 32 rb + i (zero past MT).
__global__ void k_build_msw(const uint32_t *__restrict__ rows, int nrows, int rw, int kb, int rtp, int ks_n,
                            uint8_t *__restrict__ out) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)ks_n * rtp * 1024) return;
  const int b = (int)(idx & 15), l = (int)((idx >> 4) & 63);
  const size_t frag = idx >> 10;
  const int rt = (int)(frag % rtp), ks = (int)(frag / rtp);
  const int r = 16 * rt + (l & 15), k = 64 * ks + 16 * (l >> 4) + b;
  out[idx] = (r < nrows && k < kb) ? (uint8_t)((rows[(size_t)r * rw + k / 32] >> (k % 32)) & 1u) : (uint8_t)0;
}
__global__ void k_build_tv(const uint32_t *__restrict__ Trows, uint32_t *__restrict__ Tv, int mt, int kw, int rbs) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)rbs * kw * 32) return;
  const int i = (int)(idx % 32), w = (int)((idx / 32) % kw), rb = (int)(idx / (32 * (size_t)kw));
  const int r = 32 * rb + i;
  Tv[idx] = r < mt ? Trows[(size_t)r * kw + w] : 0u;
}

template <class G>
hipError_t build_msw(const uint32_t *rows, int rw, uint8_t *out, hipStream_t s) {
  const size_t nb = MswDims<G>::BYTES;
  k_build_msw<<<(unsigned)((nb + 255) / 256), 256, 0, s>>>(rows, G::ROWS, rw, G::KB, G::RTP, MswDims<G>::KS, out);
  return hipGetLastError();
}

size_t gf2_msw_bytes(int set, int which) {
  switch (set) {
#define TMCE_MSW(S) \
  case S: return which == GF2_SYND ? MswDims<GemmSynd<S>>::BYTES : MswDims<GemmFront<S>>::BYTES;
    TMCE_MSW(S348864) TMCE_MSW(S460896) TMCE_MSW(S6688128) TMCE_MSW(S6960119) TMCE_MSW(S8192128)
#undef TMCE_MSW
  }
  return 0;
}

hipError_t launch_build_msw(int set, int which, const uint32_t *rows, int rw, uint8_t *out, hipStream_t s) {
  switch (set) {
#define TMCE_MSW(S)                                                                          \
  case S:                                                                                    \
    return which == GF2_SYND ? build_msw<GemmSynd<S>>(rows, rw, out, s) : build_msw<GemmFront<S>>(rows, rw, out, s);
    TMCE_MSW(S348864) TMCE_MSW(S460896) TMCE_MSW(S6688128) TMCE_MSW(S6960119) TMCE_MSW(S8192128)
#undef TMCE_MSW
  }
  return hipErrorInvalidValue;
}

hipError_t launch_build_tv(int set, const uint32_t *Trows, uint32_t *Tv, hipStream_t s) {
  const SetInfo &si = set_info(set);
  const int kw = (si.k + 31) / 32, rbs = (si.mt + 31) / 32;
  const size_t nv = (size_t)rbs * kw * 32;
  k_build_tv<<<(unsigned)((nv + 255) / 256), 256, 0, s>>>(Trows, Tv, si.mt, kw, rbs);
  return hipGetLastError();
}

template <int SET>
hipError_t front_mfma(const uint8_t *ct, uint32_t *scratch, const uint8_t *Fsw, uint32_t *S, int nops, int cfg,
                      const PlanDev *plan, hipStream_t s) {
  if (nops <= 0) return hipSuccess;
  return launch_gf2<GemmFront<SET>>(ct, scratch, Fsw, reinterpret_cast<uint8_t *>(S), nops, cfg, plan,
                                    "dec_front_mfma", s);
}
hipError_t launch_front_mfma_set(int set, const uint8_t *ct, uint32_t *scratch, const uint8_t *Fsw, uint32_t *S,
                                 int nops, int cfg, const PlanDev *plan, hipStream_t s) {
  switch (set) {
    case S348864: return front_mfma<S348864>(ct, scratch, Fsw, S, nops, cfg, plan, s);
    case S460896: return front_mfma<S460896>(ct, scratch, Fsw, S, nops, cfg, plan, s);
    case S6688128: return front_mfma<S6688128>(ct, scratch, Fsw, S, nops, cfg, plan, s);
    case S6960119: return front_mfma<S6960119>(ct, scratch, Fsw, S, nops, cfg, plan, s);
    case S8192128: return front_mfma<S8192128>(ct, scratch, Fsw, S, nops, cfg, plan, s);
  }
  return hipErrorInvalidValue;
}

size_t gf2_scratch_bytes(int set, int which, size_t n) {
  switch (set) {
#define TMCE_SCR(S)                                                                                   \
  case S:                                                                                             \
    if (which == GF2_SYND) return GemmSynd<S>::DIRECT ? 0 : n * (GemmSynd<S>::KP / 8);                \
    return GemmFront<S>::DIRECT ? 0 : n * (GemmFront<S>::KP / 8);
    TMCE_SCR(S348864) TMCE_SCR(S460896) TMCE_SCR(S6688128) TMCE_SCR(S6960119) TMCE_SCR(S8192128)
#undef TMCE_SCR
  }
  return 0;
}

hipError_t launch_synd_set(int set, const SyndArgs &a, int method, int cfg, hipStream_t s) {
  switch (set) {
    case S348864: return launch_synd<S348864>(a, method, cfg, s);
    case S460896: return launch_synd<S460896>(a, method, cfg, s);
    case S6688128: return launch_synd<S6688128>(a, method, cfg, s);
    case S6960119: return launch_synd<S6960119>(a, method, cfg, s);
    case S8192128: return launch_synd<S8192128>(a, method, cfg, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace tmce
