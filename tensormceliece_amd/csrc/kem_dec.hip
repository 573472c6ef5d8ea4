// kem_dec.hip -- batched Classic McEliece decapsulation on gfx950.
//
// Engine form (three launches):
//   k_dec_front   one op per lane; the 2t syndromes s_j = sum_i c_i alpha_i^j / g(alpha_i)^2 as a
//                 GF(2) matrix-vector product (AND + parity over 32-bit words) with the key's
//                 front-end matrix staged through LDS (all lanes read the same row: broadcast).
//   k_dec_decode  four ops per wavefront: Berlekamp-Massey on 16-lane rows (DPP reductions and
//                 shifts, log-domain tables in LDS), Gao-Mateer additive FFT of the locator over
//                 all 2^m field elements with the values in registers, root gather at the support,
//                 weight check, re-encryption check (syndrome of the found error recomputed from the
//                 key's log tables), session-key message.
//   k_hash_msg    one op per lane; K = SHAKE256(b || (b ? e : s) || c); 6960119 padding rule; the
//                 pc variants' confirmation check.
//
// Semantics: LM:crypto_kem/<set>/vec/decrypt.c (decrypt :135-190, weight_check :101-116) and
// kem_dec.c; 6960119 ciphertext padding LM:crypto_kem/6960119/vec/kem_dec.c (check_c_padding).
// Any correct Berlekamp-Massey variant yields the same root set (its connection polynomial differs
// by a nonzero scalar), and success requires weight t and an exact syndrome match, so the session
// key depends only on the unique weight-t error (if any) and is identical to libmceliece's.
#include "kem_kernels.h"

namespace tmce {

constexpr int FRONT_WG = 256;
#ifndef DEC_WAVES_M12
#define DEC_WAVES_M12 8
#endif
#ifndef DEC_WAVES_M13
#define DEC_WAVES_M13 12
#endif
constexpr int DEC_MIN_WAVES = 4;  // fewest waves per decoder workgroup a launch uses
#ifndef DEC_REENC_LOADS
#define DEC_REENC_LOADS 16  // re-encryption column loads in flight per pass (8: m = 13 decoders 1-2.5 % slower)
#endif
// decoder workgroups: the per-set field tables in LDS are shared by the workgroup's waves, so the
// wave count per workgroup is what sets the waves per SIMD.  m = 12 (tables 40 KB): 8-wave
// workgroups, two per CU, four waves per SIMD (at most 128 VGPRs); m = 13 (tables 80 KB): one
// 12-wave workgroup per CU, three waves per SIMD (at most 168 VGPRs); the table form of the FFT
// (BS false) keeps 8 waves for m = 13, its multiplier-log table does not leave LDS for 12
#ifndef DEC_OPS8_M13
#define DEC_OPS8_M13 1  // eight-op batches for m = 13 too (0: four-op batches only)
#endif
#ifndef DEC_HASH_PAIR_MAX
#define DEC_HASH_PAIR_MAX 32768  // batches up to this size hash each op on a lane pair (0: never)
#endif
#ifndef DEC_OPS8_MIN
#define DEC_OPS8_MIN 2  // eight-op batches once every resident wave gets at least this many of them
#endif
template <int SET, bool BS, int OPS = 4> struct DecShape {
  static constexpr int WAVES = Par<SET>::M == 12 ? DEC_WAVES_M12 : (BS ? DEC_WAVES_M13 : 8);
  static constexpr int WG = 64 * WAVES;
  static constexpr int WAVES_PER_EU = Par<SET>::M == 12 ? 4 : (BS ? 3 : 2);
  static constexpr int WGS_PER_CU = Par<SET>::M == 12 ? 16 / WAVES : 1;  // resident workgroups per CU
};
// decoder work counters: DEC_QUEUES of them, DEC_QSTRIDE words apart (separate 128-byte lines)
constexpr int DEC_QUEUES = 8, DEC_QSTRIDE = 32;

// ------------------------------------------------------------------ front end
template <int SET>
__global__ __launch_bounds__(FRONT_WG) void k_dec_front(const uint8_t *__restrict__ ct, const uint4 *__restrict__ F,
                                                        uint32_t *__restrict__ S, int nops,
                                                        const PlanDev *__restrict__ plan) {
  using P = Par<SET>;
  constexpr int Q4 = P::KWP / 4;
  __shared__ uint4 rows[32 * Q4];
  if (plan) {  // key groups are PLAN_ALIGN-aligned: one key per workgroup
    nops = plan->total;
    if ((int)(blockIdx.x * FRONT_WG) >= nops) return;
    F = uniform_ptr(reinterpret_cast<const uint4 *>(plan->keys[plan_group(plan, blockIdx.x * FRONT_WG)].F));
  }
  const int tid = threadIdx.x;
  const int op = blockIdx.x * FRONT_WG + tid;
  const bool valid = op < nops;
  uint32_t x[P::KWP];
#pragma unroll
  for (int w = 0; w < P::KWP; ++w) x[w] = 0;
  if (valid) {
    const uint16_t *cp = reinterpret_cast<const uint16_t *>(ct + (size_t)op * P::SYND_BYTES);
#pragma unroll
    for (int h = 0; h < P::SYND_BYTES / 2; ++h) x[h / 2] |= (uint32_t)cp[h] << (16 * (h & 1));
  }
  for (int rw = 0; rw < P::RW; ++rw) {
    __syncthreads();
    for (int i = tid; i < 32 * Q4; i += FRONT_WG) rows[i] = global_load16(F, (size_t)rw * 32 * Q4 + i);
    __syncthreads();
    uint32_t word = 0;
#pragma unroll 4
    for (int r = 0; r < 32; ++r) {
      uint32_t acc = 0;
#pragma unroll
      for (int q = 0; q < Q4; ++q) {
        const uint4 m = rows[r * Q4 + q];
        acc ^= (m.x & x[4 * q]) ^ (m.y & x[4 * q + 1]) ^ (m.z & x[4 * q + 2]) ^ (m.w & x[4 * q + 3]);
      }
      word |= ((uint32_t)__popc(acc) & 1u) << r;
    }
    if (valid) S[(size_t)op * P::RW + rw] = word;
  }
}

// ------------------------------------------------------------------ decoder
// LDS field tables: lg[a] = log a with lg[0] = ZL = 2q (sentinel); ex[i] = g^(i mod q) for
// i < 2q and 0 for 2q <= i <= 4q, so a*b = ex[lg[a] + lg[b]] with no zero test and no reduction.
// The decoder keeps every field element and every log doubled (a byte offset into these u16
// tables, which store doubled values): lookups need no address shift, and XOR commutes with the
// doubling.
__device__ __forceinline__ uint32_t tab(const uint16_t *t, uint32_t byte_off) {
  return *reinterpret_cast<const uint16_t *>(reinterpret_cast<const uint8_t *>(t) + byte_off);
}
// LDS byte address of a pointer into LDS, and a 16-bit LDS read at such an address
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}
__device__ __forceinline__ uint32_t lds_u16(uint32_t addr) {
  return *reinterpret_cast<const __attribute__((address_space(3))) uint16_t *>((uintptr_t)addr);
}
// Ops per wave and batch (OPS): Berlekamp-Massey runs one op per row of LR = 64 / OPS lanes.
// Four ops on 16-lane rows, or with the bitsliced FFT also eight ops on 8-lane rows: the per-step
// fixed work (row reduction, length update, loop control) is then shared by twice the ops and the
// active coefficients are padded to 8 instead of 16, and the FFT and the checks take the batch as
// two halves of four.  Eight-op batches double the work per batch, so the launcher uses them only
// when every resident wave gets at least two (348864 at N = 65536: 0.430 -> 0.421 ms, m = 13
// 4-10 % faster; at N = 8192 they would be 25 % slower).  For m = 13 the eight windows fit LDS
// beside 12 waves with one row width of window padding and the FFT buffers in the rows.
template <int SET, int OPS>
struct DecWave {
  using P = Par<SET>;
  static constexpr int LR = 64 / OPS;
  // syndromes of the wave's ops (doubled logs, 2 ZL for 0) behind SPAD = LR zero entries, so the
  // Berlekamp-Massey window s_{N-i} reads zero for i > N: step N reads i up to LR NA - 1 for NA
  // active slots, and NA slots are entered only once A(N) >= LR (NA - 1) (A, the highest active
  // coefficient, at most N + 1), so i <= N + LR; after BM, rows 4..7 hold the second half's
  // locators and FFT buffers
  static constexpr int SPAD = LR;
  uint16_t s[OPS][SPAD + P::TT];
  // the top-down FFT's two buffers of op q of half h: A (the locator, then every second level's
  // result) and B.  When a syndrome row holds 2 D entries (t = D) both lie in the op's own row, past
  // the padding; otherwise the first half uses f and its row, the second half its row and f.
  static constexpr bool F_IN_ROWS = P::TT >= 2 * P::D;
  uint16_t f[F_IN_ROWS ? 1 : 4][F_IN_ROWS ? 1 : P::D];
  __device__ uint16_t *bufA(int h, int q) { return F_IN_ROWS || h ? s[4 * h + q] + SPAD : f[q]; }
  __device__ uint16_t *bufB(int h, int q) {
    return F_IN_ROWS ? s[4 * h + q] + SPAD + P::D : (h ? f[q] : s[q] + SPAD);
  }
  uint16_t pos[round_up(P::T, 4)];  // error positions of the op being finished, then padding
  uint64_t m64[17 * P::MSG_BLOCKS];  // session-key message of the op being finished
};

// 0 as far as the compiler can tell only at run time: an index offset that keeps reads of constant
// LDS tables inside the loop that uses them (hoisted out, each read would hold a VGPR throughout)
__device__ __forceinline__ int opaque_zero() {
  int z = 0;
  asm volatile("" : "+v"(z));
  return z;
}

// row_ror:n -- lane l of a 16-lane row reads lane (l - n) mod 16 of the same row
constexpr int DPP_ROR1 = 0x121, DPP_ROR2 = 0x122, DPP_ROR4 = 0x124, DPP_ROR8 = 0x128;

// row_ror:n without a previous value (every lane of a rotation has a source lane), so the move
// folds into the VALU instruction that consumes it
template <int CTRL>
__device__ __forceinline__ uint32_t ror(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

// Berlekamp-Massey rows (one op per row of LR lanes, coefficient i in the row's lane i mod LR of
// register slot i / LR).  LR = 16: op r on the 16-lane DPP row r, coefficient lane l16.  LR = 8:
// two ops per 16-lane DPP row, interleaved -- op r on the even (r even) or odd lanes of DPP row
// r / 2, coefficient lane 2 j + (r & 1) for j = i mod 8 -- so that the x-shift of a slot is one
// row_ror:2 (the coefficient-0 lanes take the previous slot's rotated lanes), as row_ror:1 is for
// LR = 16, and the row sum is three row_ror steps.
template <int LR>
__device__ __forceinline__ int bm_row(int lane) {
  return LR == 16 ? lane >> 4 : 2 * (lane >> 4) + (lane & 1);
}
template <int LR>
__device__ __forceinline__ int bm_col(int lane) {
  return LR == 16 ? lane & 15 : (lane & 15) >> 1;
}
template <int LR>
constexpr int bm_lane(int row, int col) {
  return LR == 16 ? 16 * row + col : 16 * (row >> 1) + 2 * col + (row & 1);
}
// x-shift of one slot within the rows (lane col of a row reads col - 1 of the same row)
template <int LR>
__device__ __forceinline__ uint32_t bm_shift(uint32_t v) {
  static_assert(LR == 16 || LR == 8, "rows of 8 or 16 lanes");
  return LR == 16 ? ror<DPP_ROR1>(v) : ror<DPP_ROR2>(v);
}
// XOR over each row of LR lanes
template <int LR>
__device__ __forceinline__ uint32_t row_xor(uint32_t v) {
  if constexpr (LR == 16) v ^= ror<DPP_ROR8>(v);
  v ^= ror<LR == 16 ? DPP_ROR4 : DPP_ROR8>(v);
  v ^= ror<LR == 16 ? DPP_ROR2 : DPP_ROR4>(v);
  v ^= ror<LR == 16 ? DPP_ROR1 : DPP_ROR2>(v);
  return v;
}

// One Berlekamp-Massey step on the wave's rows (one op per row of LR lanes: coefficient i of the
// row's op in lane i mod LR of the row, slot i / LR), for NA active coefficient slots: slots >= NA
// hold zero coefficients of C and B in every row (deg C <= L, deg B <= N + 1 - L), so they are
// skipped.  The window W_i = s_{N-i} of slot s is read from the row's padded syndrome array at
// LDS address wp + 2 LR (NSB - 1 - s) (wp = the lane's window base, advanced 2 bytes per step).
// Straight line: a zero discrepancy turns the update factor into the zero sentinel and the length
// change into a select, so no lane diverges.  lb = log b (doubled), L2 = twice the row's length L.
template <int NA, int NSB, int T, int Q, int LR>
__device__ __forceinline__ void bm_step(int N, uint32_t wp, int lr, uint32_t (&C)[NSB], uint32_t (&Bl)[NSB],
                                        int &L2, uint32_t &lb, const uint16_t *lg_s, const uint16_t *ex_s) {
  constexpr uint32_t ZL2 = 4u * Q, Q2 = 2u * Q;  // doubled sentinel and modulus
  uint32_t lc[NA], pr[NA];
#pragma unroll
  for (int s = 0; s < NA; ++s) lc[s] = tab(lg_s, C[s]);
#pragma unroll
  for (int s = 0; s < NA; ++s) {
    const uint32_t wv = lds_u16(wp + 2 * LR * (NSB - 1 - s));
    pr[s] = tab(ex_s, lc[s] + wv);
  }
  // the products' sum by three-input XORs on 32-bit words (the entries are zero-extended: no
  // 16-bit arithmetic and no re-masking before the row sum)
  uint32_t dp = pr[0];
#pragma unroll
  for (int s = 1; s < NA; s += 2)
    dp = s + 1 < NA ? xor3(dp, pr[s], pr[s + 1]) : __builtin_amdgcn_bitop3_b32(dp, pr[s], 0u, 0x3C);  // dp ^ pr[s]
  const uint32_t d = row_xor<LR>(dp);
  const uint32_t ld = tab(lg_s, d);
  uint32_t lf = ld + Q2 - lb;  // log (d / b): ld - lb mod Q2, the subtraction wrapping when lf < Q2
  lf = min(lf, lf - Q2);
  lf = d != 0 ? lf : ZL2;
  const bool upd = d != 0 && L2 <= N;
#pragma unroll
  for (int s = 0; s < NA; ++s) {
    C[s] ^= tab(ex_s, lf + Bl[s]);
    Bl[s] = upd ? lc[s] : Bl[s];
  }
  L2 = upd ? 2 * (N + 1) - L2 : L2;
  lb = upd ? ld : lb;
  // B <- x B on the slots that can hold a nonzero coefficient afterwards: within a row a shift by
  // one coefficient; the row's coefficient-0 lane takes the previous slot's last coefficient,
  // which the same rotation of that slot brings there
  constexpr int NS = NA < NSB ? NA + 1 : NSB;
  uint32_t R[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) R[s] = bm_shift<LR>(Bl[s]);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const uint32_t v = (lr == 0) ? (s ? R[s > 0 ? s - 1 : 0] : ZL2) : R[s];
    Bl[s] = (s < NSB - 1 || LR * s + lr <= T) ? v : ZL2;
  }
}
// A(N) = max over the wave's rows of max(L, N + 1 - L): step N touches coefficients 0..A(N),
// i.e. A(N) / LR + 1 slots.  A grows by at most one per step (without a length change
// max(L, N + 2 - L) <= A + 1; with one, max(N + 1 - L, L + 1) <= A + 1).
// The rows' lengths are kept doubled (L2 = 2 L, so a step tests L2 <= N with no shift).
template <int LR>
__device__ __forceinline__ int bm_active(int L2, int N) {
  int Lmax = __builtin_amdgcn_readlane(L2, 0), Lmin = Lmax;
#pragma unroll
  for (int r = 1; r < 64 / LR; ++r) {
    const int Lr = __builtin_amdgcn_readlane(L2, bm_lane<LR>(r, 0));
    Lmax = max(Lmax, Lr);
    Lmin = min(Lmin, Lr);
  }
  return max(Lmax >> 1, N + 1 - (Lmin >> 1));
}
// Berlekamp-Massey steps N.. while NA slots cover the active coefficients, then the next count
// (the count never decreases).  Since A grows by at most one per step, NA slots (coefficients
// 0..LR NA - 1) stay enough for LR NA - A(N) more steps: those run without re-reading the lengths.
template <int NA, int NSB, int T, int Q, int LR>
__device__ __forceinline__ void bm_run(int &N, uint32_t wl, int lr, uint32_t (&C)[NSB], uint32_t (&Bl)[NSB],
                                       int &L2, uint32_t &lb, const uint16_t *lg_s, const uint16_t *ex_s) {
#pragma unroll 1
  while (N < 2 * T) {
    int rem = 2 * T - N;
    if constexpr (NA < NSB) {
      const int A = bm_active<LR>(L2, N);
      if (A > LR * NA - 1) break;
      rem = min(rem, LR * NA - A);
    }
    // the window's address advances 2 bytes per step, kept whole in one register (split into a
    // scalar step part and a lane part, every slot's read would need an address add)
    uint32_t wp = wl + 2 * N;
    rem = __builtin_amdgcn_readfirstlane(rem);
#pragma unroll 1
    for (int k = 0; k < rem; ++k) {
      bm_step<NA, NSB, T, Q, LR>(N, wp, lr, C, Bl, L2, lb, lg_s, ex_s);
      ++N;
      wp += 2;
      asm volatile("" : "+v"(wp));
    }
  }
  if constexpr (NA < NSB) bm_run<NA + 1, NSB, T, Q, LR>(N, wl, lr, C, Bl, L2, lb, lg_s, ex_s);
}

// ---- bitsliced GF(2^m) arithmetic (the decoder's bottom-up FFT, fft_form 1)
// A bitsliced vector holds 32 field elements in m words: bit i of word b is the coefficient of x^b
// of element i (polynomial basis, the field polynomials x^12 + x^3 + 1 and x^13 + x^4 + x^3 + x + 1
// of LM:crypto_kem/<set>/vec/gf.c).
__device__ __forceinline__ uint32_t and_xor(uint32_t acc, uint32_t a, uint32_t b) {  // acc ^ (a & b)
  return __builtin_amdgcn_bitop3_b32(acc, a, b, 0x78);
}
__device__ __forceinline__ uint32_t select_bits(uint32_t m, uint32_t a, uint32_t b) {  // m ? a : b, per bit
  return __builtin_amdgcn_bitop3_b32(m, a, b, 0xCA);
}
// r = x + a * b elementwise (32 products): schoolbook over the m x m coefficient pairs, one
// three-input AND-XOR each, the low words' sums starting from x (so the addition costs nothing),
// then reduction of the 2m - 1 product words by the field polynomial
template <int M>
__device__ __forceinline__ void bs_mul_add(const uint32_t (&x)[M], const uint32_t (&a)[M], const uint32_t (&b)[M],
                                           uint32_t (&r)[M]) {
  uint32_t p[2 * M - 1];
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < M; ++j) {
      if (i == 0 || j == M - 1)  // first term of word i + j
        p[i + j] = i + j < M ? and_xor(x[i + j], a[i], b[j]) : a[i] & b[j];
      else
        p[i + j] = and_xor(p[i + j], a[i], b[j]);
    }
  // reduction: the high words h_k (k >= M, after what folds into them from above) land on
  // r_j = p_j + sum of h_{j + M - e} over the exponents e of the field polynomial's low terms,
  // summed three at a time
  if constexpr (M == 12) {  // x^12 = x^3 + 1: h_12 += h_21, h_13 += h_22; r_j = p_j + h_{j+12} + h_{j+9}
    p[12] ^= p[21];
    p[13] ^= p[22];
#pragma unroll
    for (int j = 0; j < 12; ++j) {
      if (j < 3) r[j] = p[j] ^ p[j + 12];
      else if (j <= 10) r[j] = xor3(p[j], p[j + 12], p[j + 9]);
      else r[j] = p[j] ^ p[j + 9];
    }
  } else {  // x^13 = x^4 + x^3 + x + 1: r_j = p_j + h_{j+13} + h_{j+12} + h_{j+10} + h_{j+9}
    p[13] = xor3(p[13], p[23], p[22]);
    p[14] = xor3(p[14], p[24], p[23]);
    p[15] ^= p[24];
#pragma unroll
    for (int j = 0; j < 13; ++j) {
      uint32_t t[4];
      int nt = 0;
      if (j + 13 <= 24) t[nt++] = p[j + 13];
      if (j + 12 >= 13 && j + 12 <= 24) t[nt++] = p[j + 12];
      if (j + 10 >= 13 && j + 10 <= 24) t[nt++] = p[j + 10];
      if (j + 9 >= 13 && j + 9 <= 24) t[nt++] = p[j + 9];
      uint32_t v = p[j];
      int u = 0;
      for (; u + 2 <= nt; u += 2) v = xor3(v, t[u], t[u + 1]);
      if (u < nt) v ^= t[u];
      r[j] = v;
    }
  }
}
// Exchange bit j of the position inside a word with the word index: x and y are the words with
// index bit 0 and 1; afterwards x holds the entries whose old position bit j was 0 and y those
// whose bit was 1, each at its old position with bit j replaced by the old word index bit.
__device__ __forceinline__ void swap_word_bit(int j, uint32_t &x, uint32_t &y) {
  uint32_t nx, ny;
  if (j == 3) {  // bytes: x = [x0 y0 x2 y2], y = [x1 y1 x3 y3]
    nx = __builtin_amdgcn_perm(y, x, 0x06020400u);
    ny = __builtin_amdgcn_perm(y, x, 0x07030501u);
  } else if (j == 4) {  // half words: x = [x01 y01], y = [x23 y23]
    nx = __builtin_amdgcn_perm(y, x, 0x05040100u);
    ny = __builtin_amdgcn_perm(y, x, 0x07060302u);
  } else {
    const int hj = 1 << j;
    const uint32_t mj = j == 0 ? 0xAAAAAAAAu : j == 1 ? 0xCCCCCCCCu : 0xF0F0F0F0u;  // position bit j set
    nx = select_bits(mj, y << hj, x);
    ny = select_bits(mj, y, x >> hj);
  }
  x = nx;
  y = ny;
}
// the multiplier words of one bottom-up level in this lane: bit q of word b is bit b of
// plo[d][lane] ^ phi[k(q)]; `ph` holds the phi part (bit planes, lane-uniform), plo2 is plo doubled
// (the planes of one multiplier set start 16-byte aligned, M rounded up to 4 words: read as 16-byte
// LDS loads, whose 16-bit offset field holds the table's address, with no address add)
template <int M>
__device__ __forceinline__ void bs_multipliers(const uint32_t *ph, uint32_t plo2, uint32_t (&c)[M]) {
  uint32_t p[(M + 3) & ~3];
#pragma unroll
  for (int q = 0; q < (M + 3) / 4; ++q) {
    const uint4 v = reinterpret_cast<const uint4 *>(ph)[q];
    p[4 * q] = v.x;
    p[4 * q + 1] = v.y;
    p[4 * q + 2] = v.z;
    p[4 * q + 3] = v.w;
  }
#pragma unroll
  for (int b = 0; b < M; ++b) c[b] = p[b] ^ (uint32_t)(-(int)((plo2 >> (b + 1)) & 1u));
}

// One wavefront decodes four ops at a time: Berlekamp-Massey runs on the four 16-lane rows in
// parallel (coefficient i of op `row` lives in lane 16*row + i%16, slot i/16; row reductions and
// the x-shifts of B and of the syndrome window are DPP row operations), then the wave finishes the
// ops one by one: additive FFT with the 2^m values in registers (point p = 64*r + lane, so every
// butterfly pairs two registers of one lane), root flags, weight and re-encryption checks, message.
// m = 12: at most 128 VGPRs, so two 8-wave workgroups share a CU (LDS holds two); m = 13 keeps the
// FFT's 2^13 values in registers (two waves per SIMD).
// BS (fft_form 1): the bottom-up butterflies run on bitsliced values -- 32 points of a lane per
// word and bit plane, the GF(2^m) products by AND-XOR on the VALU -- instead of log/exp table
// gathers in LDS (fft_form 0), whose random 16-bit reads are mostly LDS bank conflicts.
// LDS table fill in a workgroup of wg threads: thread tid writes entries tid, tid + wg, ... below
// N, dst(i, src(i)), in passes of K whose loads are issued together (src is called with indices
// clamped to N - 1, so that it reads in bounds without a branch)
template <int N, int K = 8, class Src, class Dst>
__device__ __forceinline__ void wg_fill(int tid, int wg, Src &&src, Dst &&dst) {
  for (int i0 = tid; i0 < N; i0 += K * wg) {
    uint32_t v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = src(min(i0 + k * wg, N - 1));
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (i0 + k * wg < N) dst(i0 + k * wg, v[k]);
  }
}

template <int SET, bool BS, int OPS>
__global__ __launch_bounds__((DecShape<SET, BS, OPS>::WG)) __attribute__((amdgpu_waves_per_eu(DecShape<SET, BS, OPS>::WAVES_PER_EU)))
void k_dec_decode(DecArgs a) {
  constexpr int DEC_WAVES = DecShape<SET, BS, OPS>::WAVES;  // at most; a launch may use fewer waves
  using P = Par<SET>;
  constexpr int M = P::M, T = P::T, Dn = P::D, KD = P::KD, Q = P::Q;
  constexpr int NPTS = 1 << M;
  constexpr int NR = NPTS / 64;  // registers per lane holding field values (== D)
  constexpr uint32_t ZL = 2u * Q;
  constexpr int LR = DecWave<SET, OPS>::LR;  // lanes per BM row
  static_assert(OPS == 4 || (OPS == 8 && BS && (P::M == 12 || DEC_OPS8_M13)), "batch shapes");
  constexpr int NSB = ceil_div(T + 1, LR);  // BM coefficient slots per lane
  constexpr int LOG2D = Dn == 64 ? 6 : 7;
  static_assert((1 << LOG2D) == Dn, "D is 64 or 128");
  static_assert(M - KD == 6 && NR == Dn, "FFT layout assumes 2^(m-KD) = 64 points per constant");
  constexpr int NPHI = (1 << (M - 6)) - (1 << (M - 6 - KD));  // sum over levels of 2^(m-7-d)
  constexpr int NG = NR / 32;                     // bitsliced words per plane and lane (2 or 4)
  constexpr int MP = (M + 3) & ~3;                // plane words per multiplier set, 16-byte rows
  // the workgroup's LDS in one object, the field tables first: their bases then fit the 16-bit
  // offset field of ds_read (lg at 0, ex at 2^(m+1) bytes), so a lookup at a register index costs
  // no address add (placed by the compiler after the per-wave buffers, lg sat at 66 KB and every
  // Berlekamp-Massey step spent one add per active slot on its base)
  struct Lds {
    uint16_t lg[NPTS];
    uint16_t ex[4 * Q + 2];
    uint16_t phi[NPHI];
    uint16_t pthi[NR];
    uint16_t twl[2 * Dn];
    uint16_t plo[(KD + 1) * 64];  // butterfly multipliers' lane parts [KD][64], then (lane)^t
    uint16_t lpt[BS ? 1 : NPHI * 64];  // logs of the butterfly multipliers [phi index][lane]
    // bitsliced form: the multipliers' register parts as bit planes, [level + KD g][b] (g: which
    // 32-point half of a 64-point butterfly span), and the monic term's register part [group][b]
    alignas(16) uint32_t phib[BS ? 2 * KD * MP : 1];
    alignas(16) uint32_t pthib[BS ? NG * MP : 1];
    alignas(16) DecWave<SET, OPS> Wv[DEC_WAVES];
  };
  __shared__ Lds L;
  static_assert(2 * NPTS < 65536, "the power table's base fits a ds_read offset");
  uint16_t *const lg_s = L.lg, *const ex_s = L.ex, *const phi_s = L.phi, *const pthi_s = L.pthi;
  uint16_t *const twl_s = L.twl, *const plo_s = L.plo, *const lpt_s = L.lpt;
  uint32_t *const phib_s = L.phib, *const pthib_s = L.pthib;
  DecWave<SET, OPS> *const Wv = L.Wv;

  const int tid = threadIdx.x, wg = __builtin_amdgcn_readfirstlane((int)blockDim.x);
  // tables doubled (byte offsets of 16-bit entries); the entries that are not copies -- log 0, the
  // zero padding past the doubled power table, the monic term's parts when t < D -- written after
  // the copies, by the threads that copied them (same thread, program order)
  wg_fill<NPTS>(tid, wg, [&](int i) { return 2u * a.ft.lg[i]; }, [&](int i, uint32_t x) { lg_s[i] = (uint16_t)x; });
  wg_fill<2 * Q>(tid, wg, [&](int i) { return 2u * a.ft.ex[i < Q ? i : i - Q]; },
                 [&](int i, uint32_t x) { ex_s[i] = (uint16_t)x; });
  wg_fill<NPHI>(tid, wg, [&](int i) { return 2u * a.ft.phi[i]; }, [&](int i, uint32_t x) { phi_s[i] = (uint16_t)x; });
  wg_fill<2 * Dn>(tid, wg, [&](int i) { return 2u * a.ft.twl[i]; }, [&](int i, uint32_t x) { twl_s[i] = (uint16_t)x; });
  wg_fill<KD * 64>(tid, wg, [&](int i) { return 2u * a.ft.plo[i]; }, [&](int i, uint32_t x) { plo_s[i] = (uint16_t)x; });
  if constexpr (P::TOP_TERM) {
    wg_fill<NR>(tid, wg, [&](int i) { return 2u * a.ft.pthi[i]; }, [&](int i, uint32_t x) { pthi_s[i] = (uint16_t)x; });
    wg_fill<64>(tid, wg, [&](int i) { return 2u * a.ft.ptlo[i]; },
                [&](int i, uint32_t x) { plo_s[KD * 64 + i] = (uint16_t)x; });
  } else {
    for (int i = tid; i < NR; i += wg) pthi_s[i] = 0;
    for (int i = tid; i < 64; i += wg) plo_s[KD * 64 + i] = 0;
  }
  for (int i = 2 * Q + tid; i < 4 * Q + 2; i += wg) ex_s[i] = 0;
  if (tid == 0) lg_s[0] = 2 * ZL;
  __syncthreads();
  // the bottom-up multiplier of phi index k in lane l is plo[d][l] ^ phi[k], where level d holds the
  // indices [2^(m-6) - 2^(m-6-d), 2^(m-6) - 2^(m-7-d)); its log is looked up once here, so the
  // butterflies read it lane-contiguous (no bank conflict) instead of by a table gather
  if constexpr (!BS) {
    for (int i = tid; i < NPHI * 64; i += wg) {
      const int k = i >> 6, l = i & 63;
      int d = 0;
      while (d + 1 < KD && k >= (1 << (M - 6)) - (1 << (M - 7 - d))) ++d;
      lpt_s[i] = (uint16_t)tab(lg_s, plo_s[64 * d + l] ^ phi_s[k]);
    }
  } else {
    // level d pairs points 64 R + l and 64 (R + h) + l, h = 2^(m-7-d), with the multiplier
    // plo[d][l] ^ phi[off_d + (R mod h)], off_d = 2^(m-6) - 2^(m-6-d).  Bit q of plane word
    // (d + KD g, b) is bit b of phi[off_d + (q mod h)] for h < 32 (pairs inside a 32-point word),
    // of phi[off_d + 32 g + q] for h >= 32 (pairs of words, g = word index mod h/32)
    for (int i = tid; i < 2 * KD * MP; i += wg) {
      const int slot = i / MP, b = i % MP, d = slot % KD, g = slot / KD;
      const int h = 1 << (M - 7 - d), off = (1 << (M - 6)) - (1 << (M - 6 - d));
      uint32_t w = 0;
      if (b < M && (h < 32 || g < h / 32))
        for (int q = 0; q < 32; ++q) {
          const int j = h < 32 ? (q & (h - 1)) : 32 * g + q;
          w |= ((uint32_t)(phi_s[off + j] >> (b + 1)) & 1u) << q;  // phi_s is doubled
        }
      phib_s[i] = w;
    }
    for (int i = tid; i < NG * MP; i += wg) {
      const int G = i / MP, b = i % MP;
      uint32_t w = 0;
      if (b < M)
        for (int q = 0; q < 32; ++q) w |= ((uint32_t)(pthi_s[32 * G + q] >> (b + 1)) & 1u) << q;
      pthib_s[i] = w;
    }
  }
  __syncthreads();
  const int nops = a.plan ? a.plan->total : a.nops;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane0 = tid & 63;  // wave: uniform (scalar)
  constexpr int SPAD = DecWave<SET, OPS>::SPAD;
  DecWave<SET, OPS> &w = Wv[wave];
  for (int i = lane0; i < OPS * SPAD; i += 64) w.s[i / SPAD][i % SPAD] = 2 * ZL;  // window padding (zero)
  // m = 12, t = D: the monic term x^t at this lane's points, (lane)^t + (64 R)^t, as bit planes in
  // the FFT's layout -- the same for every op, so it is built once and kept in registers (2 x 12)
  constexpr bool TT_HELD = BS && P::TOP_TERM && NG == 2;
  uint32_t tth[TT_HELD ? NG : 1][TT_HELD ? M : 1];
  if constexpr (TT_HELD) {
    const uint32_t ptlo2 = plo_s[64 * KD + lane0];
#pragma unroll
    for (int G = 0; G < NG; ++G)
#pragma unroll
      for (int b = 0; b < M; ++b) tth[G][b] = pthib_s[G * MP + b] ^ (uint32_t)(-(int)((ptlo2 >> (b + 1)) & 1u));
  }

  // the wave takes batches of OPS ops from work counters, one per XCD (workgroups are dispatched
  // to the 8 XCDs round-robin; counter x hands out batches x, x + 8, x + 16, ...), and from the
  // other counters once its own is exhausted: workgroups that become resident late -- the GPU
  // shared with other kernels -- take fewer ops, and each counter sees an eighth of the requests
  const int nbat = (nops + OPS - 1) / OPS;
  int xq = blockIdx.x & (DEC_QUEUES - 1), tried = 0;
  for (;;) {
    // the lane index re-derived in every batch through an opaque zero: values computed from it
    // (per-lane addresses, bounds, masks) are then recomputed per batch instead of being hoisted
    // out of the loop and kept live (or spilled) for the whole kernel
    const int lane = lane0 + opaque_zero(), row = lane >> 4, l16 = lane & 15;
    int bidx = nbat;
    while (tried < DEC_QUEUES) {
      unsigned got = 0;
      if (lane == 0) got = atomicAdd(a.queue + DEC_QSTRIDE * xq, 1u);
      bidx = xq + DEC_QUEUES * __builtin_amdgcn_readfirstlane((int)got);
      if (bidx < nbat) break;
      xq = (xq + 1) & (DEC_QUEUES - 1);
      ++tried;
    }
    if (bidx >= nbat) break;
    const int base = OPS * bidx;
    // ---- 1. syndromes of the batch's ops (syndrome-major bits from the front end)
#pragma unroll 1
    for (int q = 0; q < OPS; ++q) {
      const int op = base + q;
      const uint32_t *Srow = a.S + (size_t)(op < nops ? op : 0) * P::RW;
      for (int j = lane; j < P::TT; j += 64) w.s[q][SPAD + j] = lg_s[op < nops ? synd_value<P>(Srow, j) : 0u];
    }
    wave_lds_sync();

    // ---- 2. Berlekamp-Massey, one op per row; C in normal form, B and the syndrome window
    //         W_i = s_{N-i} in log form
    {
      const int brow = bm_row<LR>(lane), lr = bm_col<LR>(lane);
      // the lane's window base: slot s reads s_{N - LR s - lr} at wl + 2 N + 2 LR (NSB - 1 - s)
      const uint32_t wl = (uint32_t)((int)lds_addr(w.s[brow]) + 2 * (SPAD - lr - LR * (NSB - 1)));
      uint32_t C[NSB], Bl[NSB];
#pragma unroll
      for (int s = 0; s < NSB; ++s) {
        const int i = LR * s + lr;
        C[s] = (i == 0) ? 2u : 0u;  // doubled values and logs
        Bl[s] = (i == 1) ? 0u : 2 * ZL;
      }
      int L2 = 0, N = 0;  // twice the length L
      uint32_t lb = 0;  // log b, b = 1
      bm_run<1, NSB, T, Q, LR>(N, wl, lr, C, Bl, L2, lb, lg_s, ex_s);
      // locator sigma_j = C_{t-j} (sigma_t = C_0 = 1 is added from the power table when t = D)
      // into buffer A of the op (f or its now dead syndrome row, past the zero padding that the
      // next batch's windows read)
      wave_lds_sync();
      uint16_t *lrow = w.bufA(brow >> 2, brow & 3);
      for (int j = lr; j < Dn; j += LR) lrow[j] = 0;
      wave_lds_sync();
#pragma unroll
      for (int s = 0; s < NSB; ++s) {
        const int i = LR * s + lr;
        if (i <= T && T - i < Dn) lrow[T - i] = (uint16_t)C[s];
      }
      wave_lds_sync();
    }

#pragma unroll 1
    for (int hf = 0; hf < OPS / 4; ++hf) {
    // ---- 3. additive FFT, top-down on the D coefficients of four locators at once (row r: op r of
    //         the half, in its buffers A and B; the rows' zero padding is never written): twist,
    //         Taylor expansion at x^2 + x, even/odd split
    {
      uint16_t *fr = w.bufA(hf, row);
      uint16_t *tr = w.bufB(hf, row);
      static_assert(P::TT >= Dn, "syndrome array doubles as the FFT buffer");
#pragma unroll 1
      for (int d = 0; d < KD; ++d) {
        const int lgL = LOG2D - d, Ld = 1 << lgL;
        const uint16_t *tw = twl_s + (2 * Dn - 2 * Ld) + opaque_zero();
        for (int c = l16; c < Dn; c += 16) fr[c] = tab(ex_s, tab(lg_s, fr[c]) + tw[c & (Ld - 1)]);
        wave_lds_sync();
        for (int lq = lgL - 2; lq >= 0; --lq) {  // block length 4 q with q = 2^lq
          const int q4 = 1 << lq;
          for (int u = l16; u < Dn / 4; u += 16) {
            const int bs = (u >> lq) << (lq + 2), i = u & (q4 - 1);
            fr[bs + 2 * q4 + i] ^= fr[bs + 3 * q4 + i];
          }
          wave_lds_sync();
          for (int u = l16; u < Dn / 4; u += 16) {
            const int bs = (u >> lq) << (lq + 2), i = u & (q4 - 1);
            fr[bs + q4 + i] ^= fr[bs + 2 * q4 + i];
          }
          wave_lds_sync();
        }
        for (int c = l16; c < Dn; c += 16) {
          const int bs = c & ~(Ld - 1), i = c & (Ld - 1);
          tr[bs + (i >> 1) + (i & 1) * (Ld >> 1)] = fr[c];
        }
        wave_lds_sync();
        uint16_t *x = fr;
        fr = tr;
        tr = x;
      }
    }

    // ---- 4. finish the half's four ops one after the other
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
      const int op = base + 4 * hf + q;
      if (op >= nops) break;
      const int lane = lane0 + opaque_zero();  // per op, as above: no per-lane masks hoisted out
      // key material of the op (one key, or its key group under a plan)
      const uint16_t *k_ainv = a.ainv;
      const uint32_t *k_supp = a.supp;
      const uint32_t *k_fcol = a.Fcol;
      const uint8_t *k_rej = a.rej;  // (loads below go through global_ptr: these may come from memory)
      if (a.plan) {  // the wave's ops share a key group (groups are PLAN_ALIGN-aligned)
        const KeyDev &kd = a.plan->keys[plan_group(a.plan, __builtin_amdgcn_readfirstlane(op))];
        k_ainv = uniform_ptr(kd.ainv);
        k_supp = uniform_ptr(kd.supp);
        k_fcol = uniform_ptr(kd.Fcol);
        k_rej = uniform_ptr(kd.rej);
      }
      // support flags of this lane's field points (point 64 (32 w + r) + lane is bit r of word w),
      // fetched ahead
      constexpr int ZPL = (1 << M) / 32 / 64;  // flag words per lane: 2 (m = 12), 4 (m = 13)
      static_assert(ZPL == 2 || ZPL == 4, "field of 2^12 or 2^13 points");
      uint32_t sw[ZPL];
      if constexpr (ZPL == 2) {
        const uint2 v = global_load8(reinterpret_cast<const uint2 *>(k_supp), lane);
        sw[0] = v.x;
        sw[1] = v.y;
      } else {
        const uint4 v = global_load16(reinterpret_cast<const uint4 *>(k_supp), lane);
        sw[0] = v.x;
        sw[1] = v.y;
        sw[2] = v.z;
        sw[3] = v.w;
      }
      // the ciphertext (for the session-key message), fetched now, stored to LDS after the message
      // buffer's previous reader is done
      constexpr int CH16 = P::SYND_BYTES / 2, CPL = ceil_div(CH16, 64);
      uint32_t ctw[CPL];
      {
        const uint16_t *ct16 = reinterpret_cast<const uint16_t *>(a.ct + (size_t)op * P::SYND_BYTES);
#pragma unroll
        for (int c = 0; c < CPL; ++c) ctw[c] = 64 * c + lane < CH16 ? (uint32_t)ct16[64 * c + lane] : 0u;
      }
      uint8_t *m8 = reinterpret_cast<uint8_t *>(w.m64);
      uint32_t *m32 = reinterpret_cast<uint32_t *>(w.m64);
      // [b][e] starts as zero (the bits of e are set at the roots below; b is written last)
      constexpr int EW = (1 + P::N / 8 + 3) / 4;
      for (int k = lane; k < EW; k += 64) m32[k] = 0;
      // the top-down result: after KD buffer swaps it is in the first buffer (KD even) or the second
      const uint16_t *fq = (KD & 1) ? w.bufB(hf, q) : w.bufA(hf, q);
      // ciphertext bytes into the message buffer (before the butterflies need the registers)
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        const int h = 64 * c + lane;
        if (h < CH16) {
          m8[1 + P::N / 8 + 2 * h] = (uint8_t)ctw[c];
          m8[2 + P::N / 8 + 2 * h] = (uint8_t)(ctw[c] >> 8);
        }
      }
      static_assert(NG == ZPL, "one zero-flag word per 32-point group");
      uint32_t zm[ZPL];
      if constexpr (!BS) {
        // bottom-up butterflies in registers: V2[r] packs point 64r+lane (low half) and point
        // 64(r+NR/2)+lane (high half); the two butterflies of a packed pair share their multiplier
        constexpr int NH = NR / 2;
        const int zo = opaque_zero();
        const uint16_t *lpq = lpt_s + zo + lane, *ptq = pthi_s + zo, *plq = plo_s + zo;
        uint32_t V2[NH];
  #pragma unroll
        for (int r = 0; r < NH; ++r) V2[r] = (uint32_t)fq[r] | ((uint32_t)fq[r + NH] << 16);
  #pragma unroll
        for (int d = KD - 1; d >= 1; --d) {  // deepest level first
          constexpr int dummy = 0;
          (void)dummy;
          const int step = 1 << (M - 7 - d);
          const int phoff = (1 << (M - 6)) - (1 << (M - 6 - d));
  #pragma unroll
          for (int j = 0; j < step; ++j) {
            const uint32_t lp = lpq[64 * (phoff + j)];
  #pragma unroll
            for (int r = j; r < NH; r += 2 * step) {
              const uint32_t y2 = V2[r + step];
              const uint32_t plo_ = tab(ex_s, tab(lg_s, y2 & 0xFFFFu) + lp);
              const uint32_t phi_ = tab(ex_s, tab(lg_s, y2 >> 16) + lp);
              const uint32_t x2 = V2[r] ^ (plo_ | (phi_ << 16));
              V2[r] = x2;
              V2[r + step] = x2 ^ y2;
            }
          }
        }
        {  // level 0 (half = 2^(m-1)): the pair is (low, high) of one register
  #pragma unroll
          for (int r = 0; r < NH; ++r) {
            const uint32_t lp = lpq[64 * r];
            const uint32_t x = V2[r] & 0xFFFFu, y = V2[r] >> 16;
            const uint32_t xn = x ^ tab(ex_s, tab(lg_s, y) + lp);
            V2[r] = xn | ((xn ^ y) << 16);
          }
        }
        // zero flags of sigma (adding the monic x^t term where it lies outside the FFT), transposed:
        // bit r of zf[w] in lane l flags the point 64 g + l of group g = 32 w + r, where the low
        // halves of V2 are groups 0..NH-1 and the high halves groups NH..2NH-1; masked by the
        // support flags in the same layout, then counted (the weight of e is the sum over the wave)
        static_assert(NH % 32 == 0 && 2 * NH == 32 * ZPL, "mask words per lane");
  #pragma unroll
        for (int z = 0; z < ZPL; ++z) zm[z] = 0;
        {
          const uint32_t ptlo = plq[64 * KD + lane];
          const uint32_t ptlo2 = ptlo | (ptlo << 16);
  #pragma unroll
          for (int r = 0; r < NH; ++r) {
            uint32_t v = V2[r];
            if (P::TOP_TERM) v = xor3(v, ptlo2, (uint32_t)ptq[r] | ((uint32_t)ptq[r + NH] << 16));
            // x - 1 has bit 31 set exactly for x = 0 (field values are below 2^15)
            zm[r / 32] |= (((v & 0xFFFFu) - 1u) >> 31) << (r % 32);
            zm[(r + NH) / 32] |= (((v >> 16) - 1u) >> 31) << (r % 32);
          }
        }
      } else {
        // bitsliced bottom-up on the points 64 R + lane, R = 0 .. NR-1 (bits r0 .. r6 of R), kept
        // in NG words per bit plane; the level with span h = 2^k pairs R with R + h, i.e. it flips
        // r_k.  Each level's r_k is bit 0 of the word index, so every butterfly pairs two whole
        // words: the first level (r0) starts there, and before each later level r_k moves in from
        // bit k - 1 of the position inside the word, which takes the previous level's bit in
        // exchange (one bit swap between the two words of a pair).  The last swap restores the
        // standard layout: Vb[G][b] bit i = bit b of the value at R = 32 G + i.
        const int zo = opaque_zero();
        const uint16_t *plq = plo_s + zo;
        const uint32_t *phq = phib_s + zo;
        uint32_t Vb[NG][M];
        // the D sub-polynomial constants (lane-uniform, doubled) into bit planes in the starting
        // layout -- word G = (r6 when NG = 4, r0), bit i = (R >> 1) mod 32: lane l reads constant
        // R = 64 hh + 2 (l mod 32) + l / 32 and one ballot per bit gives the words 2 hh and 2 hh + 1
#pragma unroll
        for (int hh = 0; hh < NG / 2; ++hh) {
          const uint32_t v = fq[64 * hh + (((lane & 31) << 1) | (lane >> 5))];
#pragma unroll
          for (int b = 0; b < M; ++b) {
            const uint64_t bal = __ballot((v >> (b + 1)) & 1u);
            Vb[2 * hh][b] = (uint32_t)bal;
            Vb[2 * hh + 1][b] = (uint32_t)(bal >> 32);
          }
        }
#pragma unroll
        for (int d = KD - 1; d >= 0; --d) {  // deepest level first: h = 1, 2, ..., 2^(m-7)
          const int k = M - 7 - d, h = 1 << k;
          const uint32_t plo2 = plq[64 * d + lane];
          if (k >= 1 && k <= 5) {  // r_k (at position bit k - 1) <-> r_(k-1) (word index bit 0)
#pragma unroll
            for (int G = 0; G < NG; G += 2)
#pragma unroll
              for (int b = 0; b < M; ++b) swap_word_bit(k - 1, Vb[G][b], Vb[G + 1][b]);
          }
          // the multiplier of the pair (R, R + h) depends on R mod h, which is the position bits
          // below k (h < 32), the position (h = 32), or bit 0 of the word index and the position
          // (h = 64, pairs (G, G + 2))
          const int gap = h < 64 ? 1 : 2;
#pragma unroll
          for (int G = 0; G < NG; ++G) {
            if (G & gap) continue;
            uint32_t C[M], Z[M];
            bs_multipliers<M>(phq + (d + KD * (gap == 2 ? G : 0)) * MP, plo2, C);
            bs_mul_add<M>(Vb[G], Vb[G + gap], C, Z);  // x' = x + c y
#pragma unroll
            for (int b = 0; b < M; ++b) {
              Vb[G][b] = Z[b];
              Vb[G + gap][b] ^= Z[b];  // y' = x' + y
            }
          }
        }
        if constexpr (TT_HELD) {
          // (added inside the zero-flag OR below)
        } else if constexpr (P::TOP_TERM) {  // + x^t = (lane)^t + (64 R)^t (t a power of two)
          const uint32_t ptlo2 = plq[64 * KD + lane];
          const uint4 *ptq = reinterpret_cast<const uint4 *>(pthib_s + zo);
          uint32_t pt[NG][MP];
#pragma unroll
          for (int G = 0; G < NG; ++G)
#pragma unroll
            for (int q = 0; q < MP / 4; ++q) {
              const uint4 v = ptq[G * (MP / 4) + q];
              pt[G][4 * q] = v.x;
              pt[G][4 * q + 1] = v.y;
              pt[G][4 * q + 2] = v.z;
              pt[G][4 * q + 3] = v.w;
            }
#pragma unroll
          for (int b = 0; b < M; ++b) {
            const uint32_t msk = (uint32_t)(-(int)((ptlo2 >> (b + 1)) & 1u));
#pragma unroll
            for (int G = 0; G < NG; ++G) Vb[G][b] = xor3(Vb[G][b], pt[G][b], msk);
          }
        }
#pragma unroll
        for (int G = 0; G < NG; ++G) {  // zero flags: no bit set in any plane
          uint32_t o;
          if constexpr (TT_HELD) {  // o | (v ^ x^t) per plane, one three-input operation each
            o = Vb[G][0] ^ tth[G][0];
#pragma unroll
            for (int b = 1; b < M; ++b) o = __builtin_amdgcn_bitop3_b32(o, Vb[G][b], tth[G][b], 0xF6);
          } else {
            o = Vb[G][0];
#pragma unroll
            for (int b = 1; b < M; ++b) o |= Vb[G][b];
          }
          zm[G] = ~o;
        }
      }
      uint32_t cnt = 0;
#pragma unroll
      for (int z = 0; z < ZPL; ++z) {
        zm[z] &= sw[z];
        cnt += (uint32_t)__popc(zm[z]);
      }
      // roots at the support: an exclusive prefix of the per-lane root counts (one count bit per
      // ballot) places each lane's roots in the position list (as field elements)
      int pre = 0, wt = 0;
      for (int k = 0; __any(cnt != 0u); ++k) {  // at most 8 passes (cnt <= 128)
        const uint64_t bk = __ballot(cnt & 1u);
        pre += (int)mbcnt64(bk) << k;
        wt += __popcll(bk) << k;
        cnt >>= 1;
      }
      {
        int idx = pre;
#pragma unroll
        for (int z = 0; z < ZPL; ++z) {
          uint32_t bits = zm[z];
          while (bits) {
            const int bi = __builtin_ctz(bits);
            bits &= bits - 1u;
            if (idx < T) w.pos[idx] = (uint16_t)(64 * (32 * z + bi) + lane);
            ++idx;
          }
        }
      }
      wave_lds_sync();
      // re-encryption: the syndromes of e (the XOR of the syndrome-map columns at its t positions,
      // L2-resident, rows as in S) must equal the ciphertext's, word for word
      bool ok = (wt == T);
      if (ok) {
        // the columns are gathered CPI per load instruction: each span of SPN lanes (16 or 32)
        // loads one column, lane li of the span its 16-byte word li (FQ words per column)
        constexpr int FQ = P::FCW / 4;
        constexpr int SPN = FQ <= 16 ? 16 : 32;
        constexpr int CPI = 64 / SPN;
        constexpr int NJ = ceil_div(T, CPI);  // load instructions per op
        static_assert(FQ <= SPN, "a column fits its span");
        static_assert(CPI * NJ <= round_up(T, 4), "the spans' positions fit the padded list");
        constexpr int KP = ceil_div(T, 64);
        // support index of each root (lane k holds root k), and its bit of e
        uint32_t pos_r[KP];
        const auto inv = global_ptr(k_ainv);
#pragma unroll
        for (int kk = 0; kk < KP; ++kk) {
          const int k = 64 * kk + lane;
          pos_r[kk] = k < T ? (uint32_t)inv[w.pos[k]] : 0u;
        }
#pragma unroll
        for (int kk = 0; kk < KP; ++kk) {
          const int k = 64 * kk + lane;
          if (k < T) {
            atomicOr(&m32[(8u + pos_r[kk]) >> 5], 1u << ((8u + pos_r[kk]) & 31u));
            w.pos[k] = (uint16_t)pos_r[kk];  // support indices replace the points (read above)
          } else if (k < round_up(T, CPI)) {
            w.pos[k] = (uint16_t)P::N;  // the last load's spans past t read the key's zero column N
          }
        }
        wave_lds_sync();
        const int g = lane / SPN, li = lane % SPN;
        const int lq = li < FQ ? li : FQ - 1;  // lanes past the column re-read its last word
        const uint4 *fc = reinterpret_cast<const uint4 *>(k_fcol);
        uint4 acc = make_uint4(0, 0, 0, 0);
        const uint16_t *pg = w.pos + g;  // span g reads positions g, g + CPI, ... (< round_up(T, CPI))
        // DEC_REENC_LOADS load instructions in flight per pass; a rolled outer loop (the compiler
        // would otherwise unroll every pass and hold all loads in registers at once)
#pragma unroll 1
        for (int j0 = 0; j0 < NJ; j0 += DEC_REENC_LOADS) {
          uint32_t pk[DEC_REENC_LOADS];
#pragma unroll
          for (int jj = 0; jj < DEC_REENC_LOADS; ++jj)
            pk[jj] = pg[CPI * (NJ % DEC_REENC_LOADS == 0 ? j0 + jj : min(j0 + jj, NJ - 1))];
#pragma unroll
          for (int jj = 0; jj < DEC_REENC_LOADS; ++jj) {
            if (NJ % DEC_REENC_LOADS != 0 && j0 + jj >= NJ) break;
            const uint4 v = global_load16(fc, (size_t)pk[jj] * FQ + lq);
            acc.x ^= v.x;
            acc.y ^= v.y;
            acc.z ^= v.z;
            acc.w ^= v.w;
          }
        }
        // XOR of the spans' partial sums
#pragma unroll
        for (int sh = SPN; sh < 64; sh *= 2) {
          acc.x ^= (uint32_t)__shfl_xor((int)acc.x, sh);
          acc.y ^= (uint32_t)__shfl_xor((int)acc.y, sh);
          acc.z ^= (uint32_t)__shfl_xor((int)acc.z, sh);
          acc.w ^= (uint32_t)__shfl_xor((int)acc.w, sh);
        }
        bool mism = false;
        if (lane < FQ) {  // words 4 lane .. 4 lane + 3 of the op's syndrome row (RW words, even)
          const uint2 *srow = reinterpret_cast<const uint2 *>(a.S + (size_t)op * P::RW);
          const uint2 s0 = global_load8(srow, 2 * lane);
          mism = s0.x != acc.x || s0.y != acc.y;
          if (4 * lane + 2 < P::RW) {
            const uint2 s1 = global_load8(srow, 2 * lane + 1);
            mism = mism || s1.x != acc.z || s1.y != acc.w;
          }
        }
        ok = !__any(mism);
      }
      // session-key message [b][b ? e : s][c][SHAKE padding] assembled in LDS, stored as 8-byte words
      {
        constexpr int EB = P::N / 8, MB = 136 * P::MSG_BLOCKS;
        if (lane == 0) m8[0] = ok ? 1 : 0;
        if (!ok) {  // implicit rejection: the key's string s replaces e (ok is uniform in the wave)
          const auto rj = global_ptr(k_rej);
          for (int b = lane; b < EB; b += 64) m8[1 + b] = rj[b];
        }
        for (int b = P::MSG_LEN + lane; b < MB; b += 64) {
          uint32_t v = 0;
          if (b == P::MSG_LEN) v |= 0x1Fu;
          if (b == MB - 1) v |= 0x80u;
          m8[b] = (uint8_t)v;
        }
        wave_lds_sync();
        uint64_t *mrow = reinterpret_cast<uint64_t *>(a.msg + (size_t)op * P::MSG_WORDS);
        for (int k = lane; k < MB / 8; k += 64) mrow[k] = w.m64[k];
      }
      wave_lds_sync();
    }
    }  // halves
  }
}

// The pc variants (LM:crypto_kem/<set>pc/vec/kem_dec.c:43-59): a decoded op (b = 1) whose
// SHAKE256(2 || e, 32) differs from the ciphertext's last 32 bytes is rejected as well (b = 0, s in
// place of e), and the message carries the whole ciphertext: its last 32 bytes are appended here.
template <int SET, bool PC>
__global__ __launch_bounds__(256) void k_hash_msg(uint32_t *__restrict__ msg, const uint8_t *__restrict__ ct,
                                                  uint8_t *__restrict__ key, int *__restrict__ status, int nops,
                                                  const PlanDev *__restrict__ plan, const uint8_t *__restrict__ rej) {
  using P = Par<SET>;
  constexpr int CTB = PC ? P::CT_BYTES_PC : P::SYND_BYTES;
  // every lane stays active through the hash (round constants via v_readlane)
  const KeccakRC rc;
  if (plan) nops = plan->total;
  if ((int)(blockIdx.x * blockDim.x) >= nops) return;  // whole workgroup past the batch (uniform)
  const int gop = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = gop < nops;
  const int op = live ? gop : nops - 1;
  uint32_t *mrow = msg + (size_t)op * P::MSG_WORDS;
  if constexpr (PC) {
    const uint16_t *tag16 = reinterpret_cast<const uint16_t *>(ct + (size_t)op * CTB + P::SYND_BYTES);
    uint32_t tw[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) tw[i] = (uint32_t)tag16[2 * i] | ((uint32_t)tag16[2 * i + 1] << 16);
    uint64_t cf[4];
    shake256_prefix_32<P::CONF_LEN>(reinterpret_cast<const uint64_t *>(mrow), 2u, cf, rc);
    bool same = true;
#pragma unroll
    for (int i = 0; i < 4; ++i) same = same && cf[i] == ((uint64_t)tw[2 * i] | ((uint64_t)tw[2 * i + 1] << 32));
    if (live && (mrow[0] & 0xFFu) == 1u && !same) {  // confirmation failed: implicit rejection
      if (plan) rej = plan->keys[plan_group(plan, blockIdx.x * blockDim.x)].rej;
      uint8_t *m8 = reinterpret_cast<uint8_t *>(mrow);
      m8[0] = 0;
      for (int i = 0; i < P::N / 8; ++i) m8[1 + i] = rej[i];
    }
    // the tag at byte MSG_LEN, then SHAKE padding to the end of block MSG_BLOCKS_PC
    constexpr int KB = P::MSG_LEN / 4, RS = 8 * (P::MSG_LEN % 4);
    constexpr int PADW = P::MSG_LEN_PC / 4, PADS = 8 * (P::MSG_LEN_PC % 4);
    const uint32_t keep = RS ? (mrow[KB] & ((1u << RS) - 1u)) : 0u;
#pragma unroll
    for (int k = KB; k < 34 * P::MSG_BLOCKS_PC; ++k) {
      const int j = k - KB;
      uint32_t v = 0;
      if (j < 8) v |= tw[j] << RS;
      if (RS != 0 && j >= 1 && j - 1 < 8) v |= tw[j - 1] >> ((32 - RS) & 31);
      if (k == KB) v |= keep;
      if (k == PADW) v |= 0x1Fu << PADS;
      if (k == 34 * P::MSG_BLOCKS_PC - 1) v |= 0x80u << 24;
      if (live) mrow[k] = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // this lane's message stores complete
  }
  uint64_t out[4];
  shake256_blocks_32(reinterpret_cast<const uint64_t *>(mrow), PC ? P::MSG_BLOCKS_PC : P::MSG_BLOCKS, out, rc);
  int st = 0;
  if (SET == S6960119) {
    const uint8_t last = ct[(size_t)op * CTB + P::SYND_BYTES - 1];
    if (last >> (P::MT % 8)) {
      st = -1;
#pragma unroll
      for (int i = 0; i < 4; ++i) out[i] = ~0ull;
    }
  }
  if (!live) return;
  uint64_t *kp = reinterpret_cast<uint64_t *>(key + 32 * (size_t)op);
#pragma unroll
  for (int i = 0; i < 4; ++i) kp[i] = out[i];
  if (status) status[op] = st;
}

// k_hash_msg (the variants without plaintext confirmation) with each op's SHAKE256 on a lane pair
// (the state's low halves in the even lane, high halves in the odd one): ~1.5x lower latency per
// permutation for twice the lanes, used for batches too small to give every SIMD a wave one op per
// lane.  128 ops per 256-thread workgroup.
template <int SET>
__global__ __launch_bounds__(256) void k_hash_msg_pair(const uint32_t *__restrict__ msg, const uint8_t *__restrict__ ct,
                                                       uint8_t *__restrict__ key, int *__restrict__ status, int nops,
                                                       const PlanDev *__restrict__ plan) {
  using P = Par<SET>;
  const KeccakRC rc;  // every lane stays active through the hash
  if (plan) nops = plan->total;
  if ((int)(blockIdx.x * (blockDim.x / 2)) >= nops) return;  // whole workgroup past the batch (uniform)
  const int gop = (int)(blockIdx.x * (blockDim.x / 2) + (threadIdx.x >> 1));
  const bool hi = (threadIdx.x & 1) != 0;
  const bool live = gop < nops;
  const int op = live ? gop : nops - 1;
  uint32_t out[4];
  shake256_blocks_32_pair(msg + (size_t)op * P::MSG_WORDS, P::MSG_BLOCKS, hi, out, rc);
  int st = 0;
  if (SET == S6960119) {
    const uint8_t last = ct[(size_t)op * P::SYND_BYTES + P::SYND_BYTES - 1];
    if (last >> (P::MT % 8)) {
      st = -1;
#pragma unroll
      for (int i = 0; i < 4; ++i) out[i] = ~0u;
    }
  }
  if (!live) return;
  uint32_t *kp = reinterpret_cast<uint32_t *>(key + 32 * (size_t)op) + (hi ? 1 : 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) kp[2 * i] = out[i];
  if (status && !hi) status[op] = st;
}

// the syndrome part (first SYND_BYTES) of each pc ciphertext, packed densely for the front end and
// the decoder; one thread per 16-bit half
template <int SET>
__global__ void k_ct_synd(const uint8_t *__restrict__ ct, uint8_t *__restrict__ out, int nops) {
  using P = Par<SET>;
  constexpr int H = P::SYND_BYTES / 2;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nops * H) return;
  const long op = i / H, h = i - op * H;
  reinterpret_cast<uint16_t *>(out)[i] = reinterpret_cast<const uint16_t *>(ct + op * P::CT_BYTES_PC)[h];
}

template <int SET>
hipError_t launch_dec(const DecArgs &args, hipStream_t s) {
  using P = Par<SET>;
  if (args.nops <= 0) return hipSuccess;
  DecArgs a = args;  // the front end and decoder read the syndrome part of the ciphertexts
  if (a.pc) {
    if (!a.ct_synd) return hipErrorInvalidValue;
    const long halves = (long)a.nops * (P::SYND_BYTES / 2);
    k_ct_synd<SET><<<(unsigned)((halves + 255) / 256), 256, 0, s>>>(args.ct, a.ct_synd, a.nops);
    a.ct = a.ct_synd;
  }
  if (a.front_mfma) {
    const hipError_t e = launch_front_mfma_set(SET, a.ct, a.xscratch, a.Fsw, a.S, a.nops, a.front_cfg, a.plan, s);
    if (e != hipSuccess) return e;
  } else {
    TMCE_TIMED("dec_front", s,
               k_dec_front<SET><<<ceil_div(a.nops, FRONT_WG), FRONT_WG, 0, s>>>(
                   a.ct, reinterpret_cast<const uint4 *>(a.F), a.S, a.nops, a.plan));
  }
  // waves per workgroup: the full shape when the batch gives every CU a workgroup of them, fewer
  // (down to DEC_MIN_WAVES) when it does not, so that a small batch still reaches every CU; the
  // workgroup is the unit that holds the tables in LDS, so only its wave count shrinks
  const int max_waves4 = a.fft_form ? DecShape<SET, true>::WAVES : DecShape<SET, false>::WAVES;
  const int cus = a.cus > 0 ? a.cus : 256;
  // eight-op batches (bitsliced FFT; m = 12, and m = 13 when built with them) when every resident
  // wave gets at least DEC_OPS8_MIN of them
  constexpr bool OPS8 = P::M == 12 || DEC_OPS8_M13;
  const int max_waves8 = DecShape<SET, true, 8>::WAVES;
  const int resident = DecShape<SET, true, 8>::WGS_PER_CU * cus * max_waves8;
  const bool ops8 = OPS8 && a.fft_form && ceil_div(a.nops, 8) >= DEC_OPS8_MIN * resident;
  const int max_waves = ops8 ? max_waves8 : max_waves4;
  const int nbat = ceil_div(a.nops, ops8 ? 8 : 4);
  const int waves = nbat >= cus * max_waves ? max_waves : std::max(DEC_MIN_WAVES, std::min(max_waves, ceil_div(nbat, cus)));
  int grid = a.decode_blocks > 0 ? a.decode_blocks : DecShape<SET, true>::WGS_PER_CU * cus;
  const int maxgrid = ceil_div(nbat, waves);
  if (grid > maxgrid) grid = maxgrid;
  {
    const hipError_t e = launch_fill_u32(a.queue, 0u, DEC_QUEUES * DEC_QSTRIDE, s);  // a kernel, not a memset
    if (e != hipSuccess) return e;
  }
  if (!a.fft_form)
    TMCE_TIMED("dec_decode", s, k_dec_decode<SET, false, 4><<<grid, 64 * waves, 0, s>>>(a));
  else if constexpr (OPS8) {
    if (ops8)
      TMCE_TIMED("dec_decode", s, k_dec_decode<SET, true, 8><<<grid, 64 * waves, 0, s>>>(a));
    else
      TMCE_TIMED("dec_decode", s, k_dec_decode<SET, true, 4><<<grid, 64 * waves, 0, s>>>(a));
  } else
    TMCE_TIMED("dec_decode", s, k_dec_decode<SET, true, 4><<<grid, 64 * waves, 0, s>>>(a));
  if (a.pc)
    TMCE_TIMED("dec_hash", s,
               k_hash_msg<SET, true><<<ceil_div(a.nops, 256), 256, 0, s>>>(a.msg, args.ct, a.key, a.status, a.nops,
                                                                           a.plan, a.rej));
  else if (a.nops <= DEC_HASH_PAIR_MAX)  // lane-pair hash: 128 ops per workgroup
    TMCE_TIMED("dec_hash", s,
               k_hash_msg_pair<SET><<<ceil_div(a.nops, 128), 256, 0, s>>>(a.msg, a.ct, a.key, a.status, a.nops,
                                                                         a.plan));
  else
    TMCE_TIMED("dec_hash", s,
               k_hash_msg<SET, false><<<ceil_div(a.nops, 256), 256, 0, s>>>(a.msg, a.ct, a.key, a.status, a.nops,
                                                                            a.plan, a.rej));
  return hipGetLastError();
}

template <int SET>
hipError_t front_only(const DecArgs &a, hipStream_t s) {
  k_dec_front<SET><<<ceil_div(a.nops, FRONT_WG), FRONT_WG, 0, s>>>(a.ct, reinterpret_cast<const uint4 *>(a.F), a.S,
                                                                    a.nops, nullptr);
  return hipGetLastError();
}
template <int SET>
hipError_t hash_only(const DecArgs &a, hipStream_t s) {
  k_hash_msg<SET, false><<<ceil_div(a.nops, 256), 256, 0, s>>>(a.msg, a.ct, a.key, a.status, a.nops, nullptr, a.rej);
  return hipGetLastError();
}
hipError_t launch_dec_front_only(int set, const DecArgs &a, hipStream_t s) {
  switch (set) {
    case S348864: return front_only<S348864>(a, s);
    case S460896: return front_only<S460896>(a, s);
    case S6688128: return front_only<S6688128>(a, s);
    case S6960119: return front_only<S6960119>(a, s);
    case S8192128: return front_only<S8192128>(a, s);
  }
  return hipErrorInvalidValue;
}
hipError_t launch_dec_hash_only(int set, const DecArgs &a, hipStream_t s) {
  switch (set) {
    case S348864: return hash_only<S348864>(a, s);
    case S460896: return hash_only<S460896>(a, s);
    case S6688128: return hash_only<S6688128>(a, s);
    case S6960119: return hash_only<S6960119>(a, s);
    case S8192128: return hash_only<S8192128>(a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_dec_set(int set, const DecArgs &a, hipStream_t s) {
  switch (set) {
    case S348864: return launch_dec<S348864>(a, s);
    case S460896: return launch_dec<S460896>(a, s);
    case S6688128: return launch_dec<S6688128>(a, s);
    case S6960119: return launch_dec<S6960119>(a, s);
    case S8192128: return launch_dec<S8192128>(a, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace tmce
