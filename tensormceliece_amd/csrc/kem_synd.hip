// kem_synd.hip -- batched syndrome (PKE encryption) c = H e = e[0:MT) + T e[MT:N) over GF(2), three
// ways, for error vectors given as libmceliece's e bytes (N/8 per op):
//
//   k_synd_mfma    the dense GF(2) GEMM on CDNA4 INT8 MFMA (v_mfma_i32_16x16x64_i8): 0/1 bytes,
//                  int32 accumulation (exact: K < 2^31), parity = bit 0.  The public key is held
//                  per key as bit rows cut into 128-bit pieces (Tb, L2 resident); the key's and the
//                  errors' bits reach the MFMA operands by one shift per VGPR (bit 0 of each byte).
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

// LDS ring of the MFMA GEMMs.  Ring depth does not change the GEMMs' own speed (3 to 10 slots
// within 3 %); for the decapsulation front end it sets the LDS footprint, which decides how many of
// its workgroups share a CU with the encapsulation sampler's when the two directions run on
// concurrent streams: at 60 KB a CU that holds a sampler workgroup (47 KB) takes one front-end
// workgroup beside it, at 36 KB two, whose waves then take issue slots from the sampler's chain
// (mceliece348864 two-stream step 0.70 -> 0.79 ms).  So the syndrome GEMM uses a fixed 4 slots and
// the front end as many slots as fill GF2_FRONT_LDS bytes, whatever its tile.
#ifndef GF2_NSLOT_SYND
#define GF2_NSLOT_SYND 4
#endif
#ifndef GF2_FRONT_LDS
#define GF2_FRONT_LDS (60 * 1024)
#endif

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

// The A operand of one MFMA step from a 32-bit word w of x: VGPR v = w >> (4 s + v), so byte b of
// VGPR v holds bit 8 b + 4 s + v of w in its bit 0 (s = 0, 1: the two MFMA steps that one word
// feeds).  The bits above bit 0 of each byte are other bits of w: a byte x (signed int8) times a
// 0/1 matrix entry e adds (x - bit0(x)) e, an even number, beyond bit0(x) e to the int32 sum, so
// the sum's parity is that of the 0/1 product (|sum| <= 128 K < 2^31: no overflow).  No masking or
// spreading of the bits is needed, one shift per VGPR.
__device__ __forceinline__ v4i shifted_bytes(uint32_t w, int s) {
  return v4i{(int)(w >> (4 * s)), (int)(w >> (4 * s + 1)), (int)(w >> (4 * s + 2)), (int)(w >> (4 * s + 3))};
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
// parity(M[r] . x) with x a bit row of the op's input.  Both operands are bit rows, turned into
// MFMA bytes whose bit 0 is the bit by one shift per VGPR (shifted_bytes); int32 accumulation
// (exact), bit 0 out.  G describes the product (rows, depth, where x sits in the input row, output
// row).
//
// The k depth is taken in pairs of MFMA steps (128 bits, G::KP a multiple of 128).  Lane l of a wave
// reads, for its 16x16 tiles, the 32-bit word l / 16 of the pair's 128 bits of its op's x row (A
// operand, op row l % 16 of the tile) and of its matrix row (B operand, matrix row l % 16 of the
// tile), and feeds MFMA step s of the pair with shifted_bytes(w, s): byte 4 v + b of both operands
// carries bit 32 (l / 16) + 8 b + 4 s + v of the pair.  An MFMA pairs byte p of A lane l with byte p
// of B lane l' for l / 16 = l' / 16, so the products are those of equal k, in an order that does not
// change the sum.  The matrix is stored per key as Mb[pair][row][4 words] (RTP*16 rows, zero past
// ROWS and past depth KB): the plain bit rows, cut into 16-byte pieces per pair.
//
// Workgroup tile: BM = 16*OPW*WO ops x BN = 16*RTW*WR matrix rows, WO x WR waves, each wave
// OPW x RTW 16x16 int32 tiles.  The k loop runs in super-steps of KPAIR step pairs over an LDS ring
// of NS slots filled NS - 1 super-steps ahead by global_load_lds (HBM/L2 -> LDS with no registers):
// per pair BN x 16 bytes of matrix rows and BM x 16 bytes of x, one dword per lane and copy
// instruction.  The x rows stream from HBM, so the copies need to be issued far ahead of their use.
// The copies are issued from inline assembly with explicit vmcnt waits (the compiler, which cannot
// tell the ring slots apart, would otherwise wait for every outstanding copy before any LDS read).
// Before the barrier ending super-step q a wave waits until only the copies it issued for q+2 ..
// q+NS-1 are outstanding (vector memory loads complete in issue order), so slot q+1 is complete
// for every wave after the barrier.
// Output: the D tile gives lane l column l%16 (a matrix row) of rows 4*(l/16)+v (ops); a wave ballot
// of bit 0 per v gives, in 16-bit chunk g, 16 consecutive output bits of op 4g+v, stored as one
// 16-bit word.
template <class G, int OPW, int RTW, int WO, int WR, int KPAIR, int NS>
struct GemmShape {  // NS > 0: ring slots; NS < 0: as many slots as fill -NS bytes
  static constexpr int NT = 64 * WO * WR;
  static constexpr int BM = 16 * OPW * WO, BN = 16 * RTW * WR;
  static_assert(G::KP % 128 == 0, "k depth in pairs of MFMA steps");
  static constexpr int KS2 = G::KP / 128;           // step pairs
  static constexpr int QS = ceil_div(KS2, KPAIR);   // super-steps
  static constexpr int MPT = 4 * BN / NT;           // matrix words per thread per pair
  static constexpr int APT = 4 * BM / NT;           // x words per thread per pair
  static constexpr int MPAIR = BN * 16, APAIR = BM * 16;
  static constexpr int SLOT = KPAIR * (MPAIR + APAIR);
  static constexpr int NI = KPAIR * (MPT + APT);    // copies one wave issues per super-step
  // ring slots: NS (or -NS bytes' worth, at least 3), capped by the vmcnt range and the CU's LDS
  static constexpr int NSQ = NS > 0 ? NS : (-NS / SLOT > 3 ? -NS / SLOT : 3);
  static constexpr int NSL = NSQ < 2 + 63 / NI ? (NSQ < 160 * 1024 / SLOT ? NSQ : 160 * 1024 / SLOT)
                                               : (2 + 63 / NI < 160 * 1024 / SLOT ? 2 + 63 / NI : 160 * 1024 / SLOT);
  static constexpr size_t LDS = (size_t)NSL * SLOT;
  static_assert(MPT >= 1 && MPT * NT == 4 * BN, "matrix words must split evenly over the threads");
  static_assert(APT >= 1 && APT * NT == 4 * BM, "x words must split evenly over the threads");
  static_assert(ceil_div(G::ROWS, BN) * BN <= 16 * G::RTP, "row blocks exceed the padded matrix");
  static_assert(NSL >= 3 && (NSL - 2) * NI < 64, "vmcnt range");
  static_assert(LDS <= 160 * 1024, "LDS ring exceeds the CU's LDS");
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
// wait until at most k * NI copies are outstanding (k wave-uniform, 0 <= k <= K)
template <int NI, int K>
__device__ __forceinline__ void wait_pending(int k) {
  if constexpr (K > 0) {
    if (k >= K) {
      wait_vmcnt<K * NI>();
      return;
    }
    wait_pending<NI, K - 1>(k);
  } else {
    wait_vmcnt<0>();
  }
}

// xw: x as aligned 32-bit words, row stride xstride words, xvalid words per row (reads past it are
// clamped to the last word: the matrix is zero there).  head: the original input rows (G::head16).
template <class G, int OPW, int RTW, int WO, int WR, int KPAIR, int NS>
__global__ __launch_bounds__(64 * WO * WR) void k_gf2_mfma(const uint32_t *__restrict__ xw, int xstride, int xvalid,
                                                          const uint8_t *__restrict__ head,
                                                          const uint32_t *__restrict__ Mb,
                                                          uint8_t *__restrict__ out, int nops,
                                                          const PlanDev *__restrict__ plan) {
  using S = GemmShape<G, OPW, RTW, WO, WR, KPAIR, NS>;
  extern __shared__ __align__(16) uint8_t lds[];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63, li = lane & 15, lg = lane >> 4;
  const int wo = wave % WO, wr = wave / WO;
  const int op0 = blockIdx.x * S::BM;
  if (plan) {  // key groups are PLAN_ALIGN-aligned (a multiple of BM): one key per workgroup
    nops = plan->total;
    if (op0 >= nops) return;
    Mb = uniform_ptr(reinterpret_cast<const uint32_t *>(G::plan_matrix(plan->keys[plan_group(plan, op0)])));
  }
  const int row0 = blockIdx.y * S::BN;  // first matrix row of the block
  const uint32_t lds0 = lds_off(lds);

  // per pair, copy i = tid + NT*p is word i%4 of x row (op) i/4, or of matrix row i/4; in LDS the
  // pair's pieces are [row][4 words] for the matrix, then [op][4 words] for x
  const uint32_t *xrow[S::APT];
#pragma unroll
  for (int p = 0; p < S::APT; ++p) {
    const int op = op0 + (tid + S::NT * p) / 4;
    xrow[p] = xw + (size_t)(op < nops ? op : nops - 1) * xstride;
  }
  const uint32_t *mrow = Mb + (size_t)row0 * 4 + tid;
  auto issue = [&](int q, int slot) {
#pragma unroll
    for (int u = 0; u < KPAIR; ++u) {
      const int pr = min(q * KPAIR + u, S::KS2 - 1);  // a short last super-step repeats a valid pair
      const uint32_t sb = lds0 + slot * S::SLOT + u * (S::MPAIR + S::APAIR);
#pragma unroll
      for (int p = 0; p < S::MPT; ++p)
        dma4(mrow + (size_t)pr * G::RTP * 64 + S::NT * p, __builtin_amdgcn_readfirstlane(sb + (wave * 64 + S::NT * p) * 4));
      const uint32_t sa = sb + S::MPAIR;
#pragma unroll
      for (int p = 0; p < S::APT; ++p) {
        const int w = min(4 * pr + ((tid + S::NT * p) & 3), xvalid - 1);
        dma4(xrow[p] + w, __builtin_amdgcn_readfirstlane(sa + (wave * 64 + S::NT * p) * 4));
      }
    }
  };

  // prologue: super-steps 0 .. NS-2 into slots 0 .. NS-2, slot 0 complete before the barrier
  constexpr int NS_ = S::NSL;
  const int pre = min(NS_ - 1, S::QS);
  for (int q = 0; q < pre; ++q) issue(q, q);
  wait_pending<S::NI, NS_ - 2>(pre - 1);
  __syncthreads();

  v4i acc[OPW][RTW];
#pragma unroll
  for (int i = 0; i < OPW; ++i)
#pragma unroll
    for (int j = 0; j < RTW; ++j) acc[i][j] = v4i{0, 0, 0, 0};

  int slot = 0;
#pragma unroll 1
  for (int q = 0; q < S::QS; ++q) {
    // slot (q + NS_ - 1) mod NS_ = (q - 1) mod NS_ was read in super-step q - 1, before its barrier
    const int qa = q + NS_ - 1;
    if (qa < S::QS) issue(qa, slot == 0 ? NS_ - 1 : slot - 1);
#pragma unroll
    for (int u = 0; u < KPAIR; ++u) {
      if (KPAIR == 1 || q * KPAIR + u < S::KS2) {
        const uint32_t *sm = reinterpret_cast<const uint32_t *>(lds + slot * S::SLOT + u * (S::MPAIR + S::APAIR));
        const uint32_t *sa = sm + S::MPAIR / 4;
        uint32_t mw[RTW], xwd[OPW];
#pragma unroll
        for (int j = 0; j < RTW; ++j) mw[j] = sm[(16 * (RTW * wr + j) + li) * 4 + lg];
#pragma unroll
        for (int i = 0; i < OPW; ++i) xwd[i] = sa[(16 * OPW * wo + 16 * i + li) * 4 + lg];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          v4i b[RTW];
#pragma unroll
          for (int j = 0; j < RTW; ++j) b[j] = shifted_bytes(mw[j], h);
#pragma unroll
          for (int i = 0; i < OPW; ++i) {
            const v4i a = shifted_bytes(xwd[i], h);
#pragma unroll
            for (int j = 0; j < RTW; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[j], acc[i][j], 0, 0, 0);
          }
        }
      }
    }
    // copies of super-steps q+2 .. min(q+NS-1, QS-1) may stay outstanding
    wait_pending<S::NI, NS_ - 2>(min(qa, S::QS - 1) - (q + 1));
    __syncthreads();
    slot = slot == NS_ - 1 ? 0 : slot + 1;
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
        const int h = row0 / 16 + RTW * wr + j;  // 16-bit word index of the output row
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
  static constexpr int ROWS = P::MT, KB = P::K, KP = round_up(KB, 128);
  static constexpr int RTP = round_up(ceil_div(ROWS, 16), 16);
  static constexpr int IN_BYTES = P::E_BYTES, OUT_BYTES = P::SYND_BYTES, OUT_BITS = 8 * P::SYND_BYTES;
  static constexpr bool DIRECT = P::MT % 32 == 0 && P::N % 32 == 0;
  static constexpr int XOFF_WORDS = P::MT / 32, XSTRIDE = P::E_BYTES / 4, XVALID = P::K / 32;
  static constexpr int NSLOT = GF2_NSLOT_SYND;
  __device__ static uint32_t x_word(const uint8_t *row, int w) { return e_ns_word<P>(row, w); }
  __device__ static uint32_t head16(const uint8_t *row, int h) { return e_head16<P>(row, h); }
  __device__ static const void *plan_matrix(const KeyDev &) { return nullptr; }  // no plan form
};

// Decapsulation front end: rows = R (bit b of syndrome j at row j*M + b), x = ciphertext bits
// (KBITS), out = syndrome bit planes [RW] words.
template <int SET>
struct GemmFront {
  using P = Par<SET>;
  static constexpr int ROWS = P::R, KB = P::KBITS, KP = round_up(KB, 128);
  static constexpr int RTP = round_up(ceil_div(ROWS, 16), 16);
  static constexpr int IN_BYTES = P::SYND_BYTES, OUT_BYTES = 4 * P::RW, OUT_BITS = P::R;
  static constexpr bool DIRECT = P::SYND_BYTES % 4 == 0;
  static constexpr int XOFF_WORDS = 0, XSTRIDE = P::SYND_BYTES / 4, XVALID = P::SYND_BYTES / 4;
  static constexpr int NSLOT = -(GF2_FRONT_LDS);  // slots filling that many bytes
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
struct MswDims {  // Mb[pair][RTP * 16 rows][4 words]
  static constexpr int KS2 = G::KP / 128;
  static constexpr size_t WORDS = (size_t)KS2 * G::RTP * 16 * 4;
  static constexpr size_t BYTES = WORDS * 4;
};

template <class G, int OPW, int RTW, int WO, int WR, int KPAIR, int NS = G::NSLOT>
hipError_t launch_gf2_cfg(const uint32_t *xw, int xstride, int xvalid, const uint8_t *x, const uint8_t *Msw,
                          uint8_t *out, int nops, const PlanDev *plan, const char *name, hipStream_t s) {
  using S = GemmShape<G, OPW, RTW, WO, WR, KPAIR, NS>;
  static const hipError_t attr = hipFuncSetAttribute(
      reinterpret_cast<const void *>(&k_gf2_mfma<G, OPW, RTW, WO, WR, KPAIR, NS>),
      hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::LDS);
  if (attr != hipSuccess) return attr;
  dim3 grid(ceil_div(nops, S::BM), ceil_div(G::ROWS, S::BN));
  TMCE_TIMED(name, s,
             (k_gf2_mfma<G, OPW, RTW, WO, WR, KPAIR, NS><<<grid, S::NT, S::LDS, s>>>(
                 xw, xstride, xvalid, x, reinterpret_cast<const uint32_t *>(Msw), out, nops, plan)));
  return hipGetLastError();
}

// Tile shapes offered to the tuner: (OPW, RTW, WO, WR, KPAIR) -> BM x BN workgroup tiles
//   0: 256 x 128, 8 waves of 32 ops x 128 rows, 1 pair per barrier
//   1: 256 x 128, 8 waves of 64 x 64, 1 pair per barrier      2: 256 x 256, 8 waves of 64 x 128, 1 pair
//   3: 256 x 256, 8 waves of 128 x 64, 1 pair                 4: 256 x 128, 4 waves of 128 x 64, 1 pair
//   5: 256 x 128, 8 waves of 64 x 64, 2 pairs per barrier
//   6: 128 x 64, 4 waves of 64 x 32, 1 pair               7: 128 x 128, 8 waves of 32 x 64, 1 pair
// (6 and 7 give four and two times the workgroups of 1 for small batches: N = 8192 at 256 x 128
// tiles is 192 workgroups for 256 CUs)
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
    case 0: return launch_gf2_cfg<G, 2, 8, 8, 1, 1>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 1: return launch_gf2_cfg<G, 4, 4, 4, 2, 1>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 2: return launch_gf2_cfg<G, 4, 8, 4, 2, 1>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 3: return launch_gf2_cfg<G, 8, 4, 2, 4, 1>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 4: return launch_gf2_cfg<G, 8, 4, 2, 2, 1>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 5: return launch_gf2_cfg<G, 4, 4, 4, 2, 2>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 6: return launch_gf2_cfg<G, 4, 2, 2, 2, 1>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
    case 7: return launch_gf2_cfg<G, 2, 4, 4, 2, 1>(xw, xstride, xvalid, x, Msw, out, nops, plan, name, s);
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
// 32-bit words, bits at or past `kb` ignored): Mb as described at k_gf2_mfma (word w of row r of
// pair pr = bits [128 pr + 32 w, +32) of row r, zero past nrows and past kb);
// Tv[(rb*KW + w)*32 + i] = word w of row 32 rb + i (zero past MT).
__global__ void k_build_msw(const uint32_t *__restrict__ rows, int nrows, int rw, int kb, int rtp, int ks2,
                            uint32_t *__restrict__ out) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)ks2 * rtp * 64) return;
  const int w = (int)(idx & 3), r = (int)((idx >> 2) % (16 * (size_t)rtp)), pr = (int)(idx / (64 * (size_t)rtp));
  const int k0 = 128 * pr + 32 * w;  // first bit of the word
  uint32_t v = 0;
  if (r < nrows && k0 < kb) {
    v = rows[(size_t)r * rw + k0 / 32];
    if (kb - k0 < 32) v &= (1u << (kb - k0)) - 1u;
  }
  out[idx] = v;
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
  const size_t nw = MswDims<G>::WORDS;
  k_build_msw<<<(unsigned)((nw + 255) / 256), 256, 0, s>>>(rows, G::ROWS, rw, G::KB, G::RTP, MswDims<G>::KS2,
                                                            reinterpret_cast<uint32_t *>(out));
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
