// kem_plain.hip -- the plain one-launch-per-stage GPU form of batched encapsulation and
// decapsulation, the baseline the engine form (kem_enc.hip / kem_dec.hip) is measured against.
//
// It mirrors the stage structure of the reference pipelines (mceliece-gpu-params/main.cu:213-262
// encapsulation: gen_e + pack, syndrome GEMM, finalize; :326-370 decapsulation: front end, BM,
// root search, error extraction, re-encryption, tail), with straightforward mappings: one op per
// thread for the per-op serial stages, a bit-packed GF(2) GEMM for the syndrome, intermediate
// results through HBM between launches.  Results are byte-identical to the engine form.
#include "kem_kernels.h"

namespace tmce {

constexpr int PG_WG = 64;  // per-thread stages with lane-private LDS arrays

// ------------------------------------------------------------------ encapsulation
// gen_e per thread, libmceliece's structure (LM:crypto_kem/<set>/vec/encrypt.c:41-92): draws,
// range filter, sort the t indices, reject on equal neighbours.  Writes e (row-major bytes) and
// e's last K bits as the GEMM operand E[K/32][op] (k-major 32-bit words).
template <int SET>
__global__ __launch_bounds__(PG_WG) void k_plain_gen_e(const uint8_t *__restrict__ seeds, uint32_t *__restrict__ erow,
                                                       uint32_t *__restrict__ Ens, int nops) {
  using P = Par<SET>;
  constexpr uint32_t GFMASK = (1u << P::M) - 1u;
  constexpr int NEW = ceil_div(P::N, 32);
  constexpr int KWE = ceil_div(P::K, 32);
  __shared__ uint16_t lst[P::TP * PG_WG];
  const int tid = threadIdx.x;
  const int op = blockIdx.x * PG_WG + tid;
  if (op >= nops) return;
  uint32_t st[8];
  const uint32_t *sp = reinterpret_cast<const uint32_t *>(seeds + 32 * (size_t)op);
  for (int i = 0; i < 8; ++i) st[i] = sp[i];
  for (int attempt = 0; attempt < P::MAX_ATTEMPTS; ++attempt) {
    uint64_t A[25];
#pragma unroll
    for (int i = 0; i < 25; ++i) A[i] = 0;
    const uint64_t s0 = st[0] | ((uint64_t)st[1] << 32), s1 = st[2] | ((uint64_t)st[3] << 32);
    const uint64_t s2 = st[4] | ((uint64_t)st[5] << 32), s3 = st[6] | ((uint64_t)st[7] << 32);
    A[0] = 0x40ull | (s0 << 8);
    A[1] = (s0 >> 56) | (s1 << 8);
    A[2] = (s1 >> 56) | (s2 << 8);
    A[3] = (s2 >> 56) | (s3 << 8);
    A[4] = (s3 >> 56) | (0x1Full << 8);
    A[16] = 0x80ull << 56;
    int count = 0;
    for (int blk = 0; blk < P::DRAW_BLOCKS; ++blk) {
      keccak_f1600_c(A);
      if (blk == 0)
        for (int i = 0; i < 4; ++i) { st[2 * i] = (uint32_t)A[i]; st[2 * i + 1] = (uint32_t)(A[i] >> 32); }
      for (int w = 0; w < 17; ++w)
        for (int q = 0; q < 4; ++q) {
          const int idx = blk * 68 + w * 4 + q;
          const uint32_t v = (uint32_t)(A[w] >> (16 * q)) & GFMASK;
          if (idx < P::NDRAW && v < (uint32_t)P::N && count < P::T) lst[(count++) * PG_WG + tid] = (uint16_t)v;
        }
    }
    if (count < P::T) continue;
    for (int i = P::T; i < P::TP; ++i) lst[i * PG_WG + tid] = 0xFFFF;
    // bitonic sort of the lane's list in LDS
    for (int k = 2; k <= P::TP; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1)
        for (int i = 0; i < P::TP; ++i) {
          const int l = i ^ j;
          if (l > i) {
            const uint16_t a = lst[i * PG_WG + tid], b = lst[l * PG_WG + tid];
            const bool up = (i & k) == 0;
            if ((a > b) == up) { lst[i * PG_WG + tid] = b; lst[l * PG_WG + tid] = a; }
          }
        }
    bool dup = false;
    for (int i = 1; i < P::T; ++i) dup |= lst[(i - 1) * PG_WG + tid] == lst[i * PG_WG + tid];
    if (dup) continue;
    // e words by walking the sorted support
    uint32_t *er = erow + (size_t)op * NEW;
    int ptr = 0;
    for (int w = 0; w < NEW; ++w) {
      uint32_t word = 0;
      while (ptr < P::T && lst[ptr * PG_WG + tid] < 32u * (w + 1)) word |= 1u << (lst[(ptr++) * PG_WG + tid] & 31u);
      er[w] = word;
    }
    // E[kw][op] = bits MT + 32 kw ... of e
    for (int kw = 0; kw < KWE; ++kw) {
      const int bit = P::MT + 32 * kw, w0 = bit >> 5, sh = bit & 31;
      uint32_t v = er[w0] >> sh;
      if (sh && w0 + 1 < NEW) v |= er[w0 + 1] << (32 - sh);
      if (32 * kw + 32 > P::K) v &= (1u << (P::K - 32 * kw)) - 1u;
      Ens[(size_t)kw * nops + op] = v;
    }
    return;
  }
  // abandoned (never expected): e = 0 marks it for the finalize stage
  for (int w = 0; w < NEW; ++w) erow[(size_t)op * NEW + w] = 0;
  for (int kw = 0; kw < KWE; ++kw) Ens[(size_t)kw * nops + op] = 0;
}

// S = T . E over GF(2): T row-major bits (pk rows, MT x K), E k-major words; one op per lane,
// 32 rows x 32 words of T staged in LDS per step; S[MT/32 words][op].
template <int SET>
__global__ __launch_bounds__(256) void k_plain_syndrome(const uint32_t *__restrict__ Trows, const uint32_t *__restrict__ Ens,
                                                        uint32_t *__restrict__ S, int nops) {
  using P = Par<SET>;
  constexpr int KWE = ceil_div(P::K, 32);
  constexpr int KC = 32;
  __shared__ uint32_t tile[32][KC];
  const int tid = threadIdx.x;
  const int op = blockIdx.x * 256 + tid;
  const bool live = op < nops;
  for (int rb = 0; rb < P::CW; ++rb) {
    uint32_t acc[32];
#pragma unroll
    for (int r = 0; r < 32; ++r) acc[r] = 0;
    for (int kb = 0; kb < KWE; kb += KC) {
      uint32_t x[KC];
#pragma unroll
      for (int w = 0; w < KC; ++w) x[w] = (live && kb + w < KWE) ? Ens[(size_t)(kb + w) * nops + op] : 0u;
      __syncthreads();
      for (int i = tid; i < 32 * KC; i += 256) {
        const int r = rb * 32 + i / KC, w = kb + i % KC;
        tile[i / KC][i % KC] = (r < P::MT && w < KWE) ? Trows[(size_t)r * KWE + w] : 0u;
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 32; ++r)
#pragma unroll
        for (int w = 0; w < KC; ++w) acc[r] ^= tile[r][w] & x[w];
    }
    uint32_t word = 0;
#pragma unroll
    for (int r = 0; r < 32; ++r) word |= ((uint32_t)__popc(acc[r]) & 1u) << r;
    if (live) S[(size_t)rb * nops + op] = word;
  }
}

// c = e[0:MT] + S, K = SHAKE256(1 || e || c), per thread.
template <int SET>
__global__ __launch_bounds__(256) void k_plain_finalize(const uint32_t *__restrict__ erow, const uint32_t *__restrict__ S,
                                                        uint32_t *__restrict__ msg, uint8_t *__restrict__ ct,
                                                        uint8_t *__restrict__ key, int *__restrict__ status, int nops) {
  using P = Par<SET>;
  constexpr int NEW = ceil_div(P::N, 32);
  const int op = blockIdx.x * blockDim.x + threadIdx.x;
  if (op >= nops) return;
  const uint32_t *er = erow + (size_t)op * NEW;
  const uint8_t *eb = reinterpret_cast<const uint8_t *>(er);
  bool zero = true;
  for (int w = 0; w < NEW; ++w) zero &= er[w] == 0;
  uint8_t *cp = ct + (size_t)op * P::SYND_BYTES;
  uint8_t *mb = reinterpret_cast<uint8_t *>(msg + (size_t)op * P::MSG_WORDS);
  for (int b = 0; b < P::SYND_BYTES; ++b) {
    const uint32_t sw = S[(size_t)(b / 4) * nops + op] >> (8 * (b % 4));
    uint32_t v = (eb[b] ^ sw) & 0xFFu;
    if (8 * b + 8 > P::MT) v &= (1u << (P::MT - 8 * b)) - 1u;  // bits >= MT are zero
    cp[b] = (uint8_t)v;
  }
  mb[0] = 1;
  for (int b = 0; b < P::N / 8; ++b) mb[1 + b] = eb[b];
  for (int b = 0; b < P::SYND_BYTES; ++b) mb[1 + P::N / 8 + b] = cp[b];
  for (int b = P::MSG_LEN; b < 136 * P::MSG_BLOCKS; ++b) mb[b] = 0;
  mb[P::MSG_LEN] ^= 0x1F;
  mb[136 * P::MSG_BLOCKS - 1] ^= 0x80;
  uint64_t out[4];
  shake256_blocks_32_c(reinterpret_cast<const uint64_t *>(mb), P::MSG_BLOCKS, out);
  uint64_t *kp = reinterpret_cast<uint64_t *>(key + 32 * (size_t)op);
  for (int i = 0; i < 4; ++i) kp[i] = zero ? 0ull : out[i];
  if (zero) for (int b = 0; b < P::SYND_BYTES; ++b) cp[b] = 0;
  if (status) status[op] = zero ? -2 : 0;
}

// ------------------------------------------------------------------ decapsulation
// Berlekamp-Massey per thread (LM:crypto_kem/<set>/vec/bm.c semantics; the division form whose
// connection polynomials are scalar multiples of the library's), polynomials in lane-private
// LDS, GF(2^m) through LDS log/exp tables.  Output: locator sigma_j = C_{t-j}, [op][T+1].
template <int SET>
__global__ __launch_bounds__(PG_WG) void k_plain_bm(const uint32_t *__restrict__ Sbits, const uint16_t *__restrict__ lg,
                                                    const uint16_t *__restrict__ ex, uint16_t *__restrict__ loc, int nops) {
  using P = Par<SET>;
  constexpr int T = P::T, NPTS = 1 << P::M, Q = P::Q;
  constexpr int TS = T + 1;
  __shared__ uint16_t lg_s[NPTS];
  __shared__ uint16_t ex_s[NPTS];
  __shared__ uint16_t Cs[TS * PG_WG], Bs[TS * PG_WG], Ss[2 * T * PG_WG];
  const int tid = threadIdx.x;
  for (int i = tid; i < NPTS; i += PG_WG) { lg_s[i] = lg[i]; ex_s[i] = i < Q ? ex[i] : 0; }
  __syncthreads();
  const GfLds gf{lg_s, ex_s, (uint32_t)Q};
  const int op = blockIdx.x * PG_WG + tid;
  if (op >= nops) return;
  const uint32_t *Srow = Sbits + (size_t)op * P::RW;
  for (int j = 0; j < 2 * T; ++j) {
    uint32_t v = 0;
    v = synd_value<P>(Srow, j);
    Ss[j * PG_WG + tid] = (uint16_t)v;
  }
  for (int i = 0; i < TS; ++i) { Cs[i * PG_WG + tid] = i == 0; Bs[i * PG_WG + tid] = i == 1; }
  int L = 0;
  uint32_t b = 1;
  for (int N = 0; N < 2 * T; ++N) {
    uint32_t d = 0;
    for (int i = 0; i <= (N < T ? N : T); ++i) d ^= gf.mul(Cs[i * PG_WG + tid], Ss[(N - i) * PG_WG + tid]);
    const bool upd = d != 0 && 2 * L <= N;
    if (d) {
      const uint32_t lf = gf.addmod(lg_s[d], Q - lg_s[b]);
      for (int i = 0; i < TS; ++i) {
        const uint32_t c = Cs[i * PG_WG + tid], bi = Bs[i * PG_WG + tid];
        Cs[i * PG_WG + tid] = (uint16_t)(c ^ gf.mul_log(bi, lf));
        if (upd) Bs[i * PG_WG + tid] = (uint16_t)c;
      }
      if (upd) { L = N + 1 - L; b = d; }
    }
    for (int i = T; i >= 1; --i) Bs[i * PG_WG + tid] = Bs[(i - 1) * PG_WG + tid];
    Bs[tid] = 0;
  }
  for (int j = 0; j <= T; ++j) loc[(size_t)op * TS + j] = Cs[(T - j) * PG_WG + tid];
}

// Root search, one op per wave: the additive FFT of the locator over all 2^m field elements
// with the values in LDS, then e_i = [sigma(alpha_i) == 0]; writes e rows and the weight.
template <int SET>
__global__ __launch_bounds__(256) void k_plain_roots(const uint16_t *__restrict__ loc, DecArgs a,
                                                     uint32_t *__restrict__ erow, int *__restrict__ weight) {
  using P = Par<SET>;
  constexpr int M = P::M, T = P::T, Dn = P::D, KD = P::KD, NPTS = 1 << M, Q = P::Q;
  constexpr int NEW2 = ceil_div(P::N, 64) * 2;
  __shared__ uint16_t lg_s[NPTS];
  __shared__ uint16_t ex_s[NPTS];
  __shared__ uint16_t Vs[4][NPTS];
  __shared__ uint16_t fs[4][Dn], ts[4][Dn];
  const int tid = threadIdx.x;
  for (int i = tid; i < NPTS; i += 256) { lg_s[i] = a.ft.lg[i]; ex_s[i] = i < Q ? a.ft.ex[i] : 0; }
  __syncthreads();
  const GfLds gf{lg_s, ex_s, (uint32_t)Q};
  const int wave = tid >> 6, lane = tid & 63;
  const int op = blockIdx.x * 4 + wave;
  if (op >= a.nops) return;
  uint16_t *V = Vs[wave], *f = fs[wave], *tmp = ts[wave];
  for (int j = lane; j < Dn; j += 64) f[j] = j <= T ? loc[(size_t)op * (T + 1) + j] : 0;
  wave_lds_sync();
  for (int d = 0; d < KD; ++d) {
    const int Ld = Dn >> d;
    const uint16_t *tw = a.ft.twl + (2 * Dn - 2 * Ld);
    for (int c = lane; c < Dn; c += 64) f[c] = (uint16_t)gf.mul_log(f[c], tw[c & (Ld - 1)]);
    wave_lds_sync();
    for (int len = Ld; len >= 4; len >>= 1) {
      const int q4 = len >> 2;
      for (int u = lane; u < Dn / 4; u += 64) f[(u / q4) * len + 2 * q4 + u % q4] ^= f[(u / q4) * len + 3 * q4 + u % q4];
      wave_lds_sync();
      for (int u = lane; u < Dn / 4; u += 64) f[(u / q4) * len + q4 + u % q4] ^= f[(u / q4) * len + 2 * q4 + u % q4];
      wave_lds_sync();
    }
    for (int c = lane; c < Dn; c += 64) {
      const int bs = c & ~(Ld - 1), i = c & (Ld - 1);
      tmp[bs + (i >> 1) + (i & 1) * (Ld >> 1)] = f[c];
    }
    wave_lds_sync();
    for (int c = lane; c < Dn; c += 64) f[c] = tmp[c];
    wave_lds_sync();
  }
  for (int c = lane; c < NPTS; c += 64) V[c] = f[c >> (M - KD)];
  wave_lds_sync();
  for (int d = KD - 1; d >= 0; --d) {
    const int hb = M - d - 1;
    const int phoff = (1 << (M - 6)) - (1 << (M - 6 - d));
    for (int u = lane; u < NPTS / 2; u += 64) {
      const int c = u & ((1 << hb) - 1), base = (u >> hb) << (hb + 1);
      const uint32_t p = a.ft.plo[64 * d + (c & 63)] ^ a.ft.phi[phoff + (c >> 6)];
      uint32_t x = V[base + c];
      const uint32_t y = V[base + (1 << hb) + c];
      x ^= gf.mul(y, p);
      V[base + c] = (uint16_t)x;
      V[base + (1 << hb) + c] = (uint16_t)(x ^ y);
    }
    wave_lds_sync();
  }
  int wt = 0;
  uint32_t *er = erow + (size_t)op * NEW2;
  for (int it = 0; it < ceil_div(P::N, 64); ++it) {
    const int i = it * 64 + lane;
    const bool valid = i < P::N;
    const uint32_t av = valid ? a.alpha[i] : 0u;
    uint32_t val = V[av];
    if (P::TOP_TERM) val ^= a.ft.ptlo[av & 63] ^ a.ft.pthi[av >> 6];
    const uint64_t mask = __ballot(valid && val == 0);
    if (lane == 0) { er[2 * it] = (uint32_t)mask; er[2 * it + 1] = (uint32_t)(mask >> 32); }
    wt += __popcll(mask);
  }
  if (lane == 0) weight[op] = wt;
}

// Re-encryption check and session-key message, one op per wave: success iff weight t and the
// syndromes of e equal the ciphertext's; message [b][b ? e : s][c][padding].
template <int SET>
__global__ __launch_bounds__(256) void k_plain_reenc(const uint32_t *__restrict__ erow, const int *__restrict__ weight,
                                                     DecArgs a) {
  using P = Par<SET>;
  constexpr int M = P::M, T = P::T, NPTS = 1 << M, Q = P::Q;
  constexpr int NEW2 = ceil_div(P::N, 64) * 2;
  constexpr int JS = ceil_div(2 * T, 64);
  __shared__ uint16_t ex_s[NPTS];
  const int tid = threadIdx.x;
  for (int i = tid; i < NPTS; i += 256) ex_s[i] = i < Q ? a.ft.ex[i] : 0;
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int op = blockIdx.x * 4 + wave;
  if (op >= a.nops) return;
  const uint32_t *er = erow + (size_t)op * NEW2;
  bool ok = weight[op] == T;
  if (ok) {
    uint32_t acc[JS] = {0};
    for (int w = 0; w < NEW2; ++w) {
      uint32_t bits = er[w];
      while (bits) {
        const int p = 32 * w + __ffs(bits) - 1;
        bits &= bits - 1;
        const uint32_t la = a.la[p], lu = a.lu[p];
        for (int k = 0; k < JS; ++k) {
          const uint32_t j = (uint32_t)(k * 64 + lane);
          uint32_t v = 0;
          if (lu != 0xFFFFu) {
            if (la == 0xFFFFu) v = j == 0 ? ex_s[lu] : 0u;
            else v = ex_s[(j * la + lu) % Q];
          }
          acc[k] ^= v;
        }
      }
    }
    // syndromes of the ciphertext from the front end's bit planes
    bool mism = false;
    const uint32_t *Srow = a.S + (size_t)op * P::RW;
    for (int k = 0; k < JS; ++k) {
      const int j = k * 64 + lane;
      if (j >= 2 * T) continue;
      uint32_t s = 0;
      s = synd_value<P>(Srow, j);
      mism |= acc[k] != s;
    }
    ok = !__any(mism);
  }
  const uint8_t *ebytes = reinterpret_cast<const uint8_t *>(er);
  const uint8_t *ctrow = a.ct + (size_t)op * P::SYND_BYTES;
  uint32_t *mrow = a.msg + (size_t)op * P::MSG_WORDS;
  constexpr int EB = P::N / 8;
  for (int k = lane; k < P::MSG_WORDS; k += 64) {
    uint32_t v = 0;
    for (int q = 0; q < 4; ++q) {
      const int b = 4 * k + q;
      uint32_t x = 0;
      if (b == 0) x = ok ? 1u : 0u;
      else if (b <= EB) x = ok ? ebytes[b - 1] : a.rej[b - 1];
      else if (b <= EB + P::SYND_BYTES) x = ctrow[b - 1 - EB];
      if (b == P::MSG_LEN) x |= 0x1Fu;
      if (b == 136 * P::MSG_BLOCKS - 1) x |= 0x80u;
      v |= x << (8 * q);
    }
    mrow[k] = v;
  }
}

// ------------------------------------------------------------------ launchers
template <int SET>
hipError_t launch_plain_enc(const PlainEncArgs &a, hipStream_t s) {
  if (a.nops <= 0) return hipSuccess;
  TMCE_TIMED("plain_gen_e", s, k_plain_gen_e<SET><<<ceil_div(a.nops, PG_WG), PG_WG, 0, s>>>(a.seeds, a.erow, a.Ens, a.nops));
  TMCE_TIMED("plain_syndrome", s,
             k_plain_syndrome<SET><<<ceil_div(a.nops, 256), 256, 0, s>>>(a.Trows, a.Ens, a.S, a.nops));
  TMCE_TIMED("plain_finalize", s,
             k_plain_finalize<SET><<<ceil_div(a.nops, 256), 256, 0, s>>>(a.erow, a.S, a.msg, a.ct, a.key, a.status,
                                                                        a.nops));
  return hipGetLastError();
}

template <int SET>
hipError_t launch_plain_dec(const DecArgs &a, const PlainDecScratch &x, hipStream_t s) {
  if (a.nops <= 0) return hipSuccess;
  hipError_t e = hipSuccess;
  TMCE_TIMED("plain_front", s, e = launch_dec_front_only(SET, a, s));
  if (e != hipSuccess) return e;
  TMCE_TIMED("plain_bm", s,
             k_plain_bm<SET><<<ceil_div(a.nops, PG_WG), PG_WG, 0, s>>>(a.S, a.ft.lg, a.ft.ex, x.loc, a.nops));
  TMCE_TIMED("plain_roots", s, k_plain_roots<SET><<<ceil_div(a.nops, 4), 256, 0, s>>>(x.loc, a, x.erow, x.weight));
  TMCE_TIMED("plain_reenc", s, k_plain_reenc<SET><<<ceil_div(a.nops, 4), 256, 0, s>>>(x.erow, x.weight, a));
  TMCE_TIMED("plain_hash", s, e = launch_dec_hash_only(SET, a, s));
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t launch_plain_enc_set(int set, const PlainEncArgs &a, hipStream_t s) {
  switch (set) {
    case S348864: return launch_plain_enc<S348864>(a, s);
    case S460896: return launch_plain_enc<S460896>(a, s);
    case S6688128: return launch_plain_enc<S6688128>(a, s);
    case S6960119: return launch_plain_enc<S6960119>(a, s);
    case S8192128: return launch_plain_enc<S8192128>(a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_plain_dec_set(int set, const DecArgs &a, const PlainDecScratch &x, hipStream_t s) {
  switch (set) {
    case S348864: return launch_plain_dec<S348864>(a, x, s);
    case S460896: return launch_plain_dec<S460896>(a, x, s);
    case S6688128: return launch_plain_dec<S6688128>(a, x, s);
    case S6960119: return launch_plain_dec<S6960119>(a, x, s);
    case S8192128: return launch_plain_dec<S8192128>(a, x, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace tmce
