// kem_dec.hip -- batched Classic McEliece decapsulation on gfx950.
//
// Engine form (three launches):
//   k_dec_front   one op per lane; the 2t syndromes s_j = sum_i c_i alpha_i^j / g(alpha_i)^2 as a
//                 GF(2) matrix-vector product (AND + parity over 32-bit words) with the key's
//                 front-end matrix staged through LDS (all lanes read the same row: broadcast).
//   k_dec_decode  one op per wavefront: Berlekamp-Massey with the coefficients spread over the
//                 64 lanes, Gao-Mateer additive FFT of the locator over all 2^m field elements in
//                 LDS, root gather at the support, weight check, re-encryption check (syndrome of
//                 the found error recomputed from the key's log tables), session-key message.
//   k_hash_msg    one op per lane; K = SHAKE256(b || (b ? e : s) || c); 6960119 padding rule.
//
// Semantics: LM:crypto_kem/<set>/vec/decrypt.c (decrypt :135-190, weight_check :101-116) and
// kem_dec.c; 6960119 ciphertext padding LM:crypto_kem/6960119/vec/kem_dec.c (check_c_padding).
// Any correct Berlekamp-Massey variant yields the same root set (its connection polynomial differs
// by a nonzero scalar), and success requires weight t and an exact syndrome match, so the session
// key depends only on the unique weight-t error (if any) and is identical to libmceliece's.
#include "kem_kernels.h"

namespace tmce {

constexpr int FRONT_WG = 256;
constexpr int DEC_WG = 256;
constexpr int DEC_WAVES = DEC_WG / 64;

// ------------------------------------------------------------------ front end
template <int SET>
__global__ __launch_bounds__(FRONT_WG) void k_dec_front(const uint8_t *__restrict__ ct, const uint4 *__restrict__ F,
                                                        uint32_t *__restrict__ S, int nops) {
  using P = Par<SET>;
  constexpr int Q4 = P::KWP / 4;
  __shared__ uint4 rows[32 * Q4];
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
    for (int i = tid; i < 32 * Q4; i += FRONT_WG) rows[i] = F[(size_t)rw * 32 * Q4 + i];
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

// ------------------------------------------------------------------ decoder (one op per wave)
template <int SET>
struct DecLds {
  using P = Par<SET>;
  uint16_t s[P::TT];
  uint16_t f[P::D];
  uint16_t tmp[P::D];
  uint16_t pos[P::T];
  uint32_t ebits[ceil_div(P::N, 64) * 2];
  uint16_t V[1 << P::M];
};

template <int SET>
__device__ __forceinline__ uint32_t msg_byte(int b, bool ok, const uint8_t *ebytes, const uint8_t *rej,
                                             const uint8_t *ctrow) {
  using P = Par<SET>;
  constexpr int EB = P::N / 8;
  uint32_t v = 0;
  if (b == 0) v = ok ? 1u : 0u;
  else if (b <= EB) v = ok ? ebytes[b - 1] : rej[b - 1];
  else if (b <= EB + P::SYND_BYTES) v = ctrow[b - 1 - EB];
  if (b == P::MSG_LEN) v |= 0x1Fu;
  if (b == 136 * P::MSG_BLOCKS - 1) v |= 0x80u;
  return v;
}

template <int SET>
__global__ __launch_bounds__(DEC_WG) void k_dec_decode(DecArgs a) {
  using P = Par<SET>;
  constexpr int M = P::M, T = P::T, Dn = P::D, KD = P::KD;
  constexpr int NPTS = 1 << M;
  __shared__ uint16_t lg_s[NPTS];
  __shared__ uint16_t ex_s[NPTS];
  __shared__ DecLds<SET> W[DEC_WAVES];

  const int tid = threadIdx.x;
  for (int i = tid; i < NPTS; i += DEC_WG) {
    lg_s[i] = a.ft.lg[i];
    ex_s[i] = i < P::Q ? a.ft.ex[i] : 0;
  }
  __syncthreads();
  const GfLds gf{lg_s, ex_s, (uint32_t)P::Q};
  const int wave = tid >> 6, l = tid & 63;
  DecLds<SET> &w = W[wave];

  for (int op = blockIdx.x * DEC_WAVES + wave; op < a.nops; op += gridDim.x * DEC_WAVES) {
    // ---- 1. syndromes from the bit planes
    const uint32_t *Srow = a.S + (size_t)op * P::RW;
    for (int j = l; j < P::TT; j += 64) {
      uint32_t v = 0;
#pragma unroll
      for (int b = 0; b < M; ++b) v |= ((Srow[b * (P::TT / 32) + (j >> 5)] >> (j & 31)) & 1u) << b;
      w.s[j] = (uint16_t)v;
    }
    wave_lds_sync();

    // ---- 2. Berlekamp-Massey (coefficient i = slot*64 + lane)
    constexpr int NS = P::BM_SLOTS;
    uint32_t C[NS], B[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int i = s * 64 + l;
      C[s] = (i == 0) ? 1u : 0u;
      B[s] = (i == 1) ? 1u : 0u;
    }
    int L = 0;
    uint32_t bb = 1;
    for (int N = 0; N < 2 * T; ++N) {
      uint32_t dp = 0;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int i = s * 64 + l;
        if (i <= T && i <= N) dp ^= gf.mul(C[s], w.s[N - i]);
      }
      const uint32_t d = wave_xor(dp);
      if (d != 0) {
        const uint32_t lf = gf.addmod(lg_s[d], P::Q - lg_s[bb]);
        const bool upd = (2 * L <= N);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          const uint32_t tc = C[s];
          C[s] ^= gf.mul_log(B[s], lf);
          if (upd) B[s] = tc;
        }
        if (upd) { L = N + 1 - L; bb = d; }
      }
      // B <- x * B (coefficient t+1 and above dropped)
      uint32_t nb[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const uint32_t up = (uint32_t)__shfl_up((int)B[s], 1);
        const uint32_t carry = (s > 0) ? (uint32_t)__shfl((int)B[s > 0 ? s - 1 : 0], 63) : 0u;
        nb[s] = (l == 0) ? carry : up;
      }
#pragma unroll
      for (int s = 0; s < NS; ++s) B[s] = (s * 64 + l <= T) ? nb[s] : 0u;
    }

    // ---- 3. locator coefficients sigma_j = C_{t-j}, j < D (sigma_t = C_0 = 1 via the power table)
    for (int j = l; j < Dn; j += 64) w.f[j] = 0;
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int i = s * 64 + l;
      const int j = T - i;
      if (i <= T && j < Dn) w.f[j] = (uint16_t)C[s];
    }
    wave_lds_sync();

    // ---- 4. additive FFT, top-down: twist, Taylor expansion at x^2+x, split even/odd
#pragma unroll 1
    for (int d = 0; d < KD; ++d) {
      const int Ld = Dn >> d;
      const uint16_t *tw = a.ft.twl + (2 * Dn - 2 * Ld);
      for (int c = l; c < Dn; c += 64) w.f[c] = (uint16_t)gf.mul_log(w.f[c], tw[c & (Ld - 1)]);
      wave_lds_sync();
      for (int len = Ld; len >= 4; len >>= 1) {
        const int q4 = len >> 2;
        for (int u = l; u < Dn / 4; u += 64) {
          const int base = (u / q4) * len, i = u % q4;
          w.f[base + 2 * q4 + i] ^= w.f[base + 3 * q4 + i];
        }
        wave_lds_sync();
        for (int u = l; u < Dn / 4; u += 64) {
          const int base = (u / q4) * len, i = u % q4;
          w.f[base + q4 + i] ^= w.f[base + 2 * q4 + i];
        }
        wave_lds_sync();
      }
      for (int c = l; c < Dn; c += 64) {
        const int base = c & ~(Ld - 1), i = c & (Ld - 1);
        w.tmp[base + (i >> 1) + (i & 1) * (Ld >> 1)] = w.f[c];
      }
      wave_lds_sync();
      for (int c = l; c < Dn; c += 64) w.f[c] = w.tmp[c];
      wave_lds_sync();
    }
    // ---- 5. bottom-up butterflies over all 2^m points
    for (int c = l; c < NPTS; c += 64) w.V[c] = w.f[c >> (M - KD)];
    wave_lds_sync();
#pragma unroll 1
    for (int d = KD - 1; d >= 0; --d) {
      const int hb = M - d - 1;  // log2(half)
      const uint16_t *pl = a.ft.plg + (NPTS - (NPTS >> d));
      for (int u = l; u < NPTS / 2; u += 64) {
        const int c = u & ((1 << hb) - 1);
        const int base = (u >> hb) << (hb + 1);
        uint32_t x = w.V[base + c];
        const uint32_t y = w.V[base + (1 << hb) + c];
        x ^= gf.mul_log(y, pl[c]);
        w.V[base + c] = (uint16_t)x;
        w.V[base + (1 << hb) + c] = (uint16_t)(x ^ y);
      }
      wave_lds_sync();
    }

    // ---- 6. roots at the support: e_i = [sigma(alpha_i) == 0]
    int wt = 0;
    for (int it = 0; it < ceil_div(P::N, 64); ++it) {
      const int i = it * 64 + l;
      const bool valid = i < P::N;
      const uint32_t av = valid ? a.alpha[i] : 0u;
      uint32_t val = w.V[av];
      if (P::TOP_TERM) val ^= a.ft.powt[av];
      const bool ei = valid && (val == 0);
      const uint64_t mask = __ballot(ei);
      if (ei) {
        const int idx = wt + (int)mbcnt64(mask);
        if (idx < T) w.pos[idx] = (uint16_t)i;
      }
      if (l == 0) { w.ebits[2 * it] = (uint32_t)mask; w.ebits[2 * it + 1] = (uint32_t)(mask >> 32); }
      wt += __popcll(mask);
    }
    wave_lds_sync();

    // ---- 7. re-encryption check: syndrome of e equals the ciphertext syndrome
    bool ok = (wt == T);
    if (ok) {
      uint32_t acc[ceil_div(2 * T, 64)];
#pragma unroll
      for (int k = 0; k < ceil_div(2 * T, 64); ++k) acc[k] = 0;
      for (int k = 0; k < T; ++k) {
        const int p = w.pos[k];
        const uint32_t la = a.la[p], lu = a.lu[p];
#pragma unroll
        for (int kk = 0; kk < ceil_div(2 * T, 64); ++kk) {
          const uint32_t j = (uint32_t)(kk * 64 + l);
          uint32_t v;
          if (lu == 0xFFFFu) v = 0;
          else if (la == 0xFFFFu) v = (j == 0) ? ex_s[lu] : 0u;
          else {
            uint32_t e = j * la + lu;
            e = (e & P::Q) + (e >> M);
            e = (e & P::Q) + (e >> M);
            e = e >= (uint32_t)P::Q ? e - P::Q : e;
            v = ex_s[e];
          }
          acc[kk] ^= v;
        }
      }
      bool mism = false;
#pragma unroll
      for (int kk = 0; kk < ceil_div(2 * T, 64); ++kk) {
        const int j = kk * 64 + l;
        if (j < 2 * T && acc[kk] != w.s[j]) mism = true;
      }
      ok = !__any(mism);
    }

    // ---- 8. session-key message [b][b ? e : s][c][padding]
    const uint8_t *ebytes = reinterpret_cast<const uint8_t *>(w.ebits);
    const uint8_t *ctrow = a.ct + (size_t)op * P::SYND_BYTES;
    uint32_t *mrow = a.msg + (size_t)op * P::MSG_WORDS;
    for (int k = l; k < P::MSG_WORDS; k += 64) {
      uint32_t v = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) v |= msg_byte<SET>(4 * k + q, ok, ebytes, a.rej, ctrow) << (8 * q);
      mrow[k] = v;
    }
    wave_lds_sync();
  }
}

template <int SET>
__global__ __launch_bounds__(256) void k_hash_msg(const uint32_t *__restrict__ msg, const uint8_t *__restrict__ ct,
                                                  uint8_t *__restrict__ key, int *__restrict__ status, int nops) {
  using P = Par<SET>;
  const int op = blockIdx.x * blockDim.x + threadIdx.x;
  if (op >= nops) return;
  uint64_t out[4];
  shake256_blocks_32(reinterpret_cast<const uint64_t *>(msg + (size_t)op * P::MSG_WORDS), P::MSG_BLOCKS, out);
  int st = 0;
  if (SET == S6960119) {
    const uint8_t last = ct[(size_t)op * P::SYND_BYTES + P::SYND_BYTES - 1];
    if (last >> (P::MT % 8)) {
      st = -1;
#pragma unroll
      for (int i = 0; i < 4; ++i) out[i] = ~0ull;
    }
  }
  uint64_t *kp = reinterpret_cast<uint64_t *>(key + 32 * (size_t)op);
#pragma unroll
  for (int i = 0; i < 4; ++i) kp[i] = out[i];
  if (status) status[op] = st;
}

template <int SET>
hipError_t launch_dec(const DecArgs &a, hipStream_t s) {
  using P = Par<SET>;
  if (a.nops <= 0) return hipSuccess;
  TMCE_TIMED("dec_front", s,
             k_dec_front<SET><<<ceil_div(a.nops, FRONT_WG), FRONT_WG, 0, s>>>(
                 a.ct, reinterpret_cast<const uint4 *>(a.F), a.S, a.nops));
  int grid = a.decode_blocks > 0 ? a.decode_blocks : 1024;
  const int maxgrid = ceil_div(a.nops, DEC_WAVES);
  if (grid > maxgrid) grid = maxgrid;
  TMCE_TIMED("dec_decode", s, k_dec_decode<SET><<<grid, DEC_WG, 0, s>>>(a));
  TMCE_TIMED("dec_hash", s,
             k_hash_msg<SET><<<ceil_div(a.nops, 256), 256, 0, s>>>(a.msg, a.ct, a.key, a.status, a.nops));
  return hipGetLastError();
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
