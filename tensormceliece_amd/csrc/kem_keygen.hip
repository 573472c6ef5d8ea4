// kem_keygen.hip -- batched key generation on gfx950: the systematic-form step.
//
// One workgroup per candidate key (Goppa polynomial g, support alpha, from the host part of the
// generator).  It builds the parity-check matrix H in HBM -- row i*m + b, column j: bit b of
// alpha_j^i / g(alpha_j) -- with the column bits of 64 columns turned into row words by wave
// ballots, reduces H to [I | T] by Gauss-Jordan elimination, and writes the public key T.  Row
// swaps are kept as an index permutation in LDS; each elimination step gives every wave whole
// rows and every lane one or two of the row's words, so rows without the pivot bit are skipped
// uniformly.  The reduced form of a full-rank matrix is unique, so the bytes do not depend on the
// pivot order, and the step fails exactly when the left mt x mt block is singular -- the outcome
// of LM:crypto_kem/<set>/vec/pk_gen.c for the same candidate.
//
// The f variants (LM:crypto_kem/<set>f/vec/pk_gen.c, mov_columns :87-153 called at row mt-32
// :251-255) add a column step when the elimination reaches column c0 = mt - 32: the 32 rows still
// without a pivot are zero left of c0, and the pivot columns pv[0..32) of their 32 x 64 window
// (columns c0 .. c0+63) in row-echelon form -- pv[i] the lowest column where any row not yet used
// is nonzero; these depend only on the rows' span -- give the pivot mask and the column swaps
// c0+i <-> c0+pv[i] (i = 0..31 in order) applied to every row of H.  The elimination then finishes
// on the permuted matrix; the host applies the same swaps to the support permutation.
#include "kem_kernels.h"

namespace tmce {

constexpr int KG_WG = 1024;

// minimum over the wave, in every lane: DPP within each row of 16 lanes (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror), then the four rows' values through scalar
// registers -- no LDS round trips, unlike a butterfly of lane shuffles
__device__ __forceinline__ int wave_min(int v) {
  v = min(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false));
  const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
  const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
  return min(min(a, b), min(c, d));
}
constexpr int KG_WAVES = KG_WG / 64;

template <int SET, bool F>
__global__ __launch_bounds__(KG_WG) void k_pk_gen(const uint16_t *__restrict__ g_all,
                                                 const uint16_t *__restrict__ alpha_all,
                                                 const uint16_t *__restrict__ lg, const uint16_t *__restrict__ ex,
                                                 uint64_t *__restrict__ H_all, uint8_t *__restrict__ pk_all,
                                                 int *__restrict__ ok_all, uint64_t *__restrict__ piv_all,
                                                 const int *__restrict__ valid) {
  using P = Par<SET>;
  constexpr int M = P::M, T = P::T, N = P::N, MT = P::MT, Q = P::Q;
  constexpr int W = ceil_div(N, 64);  // 64-bit words per row of H
  static_assert(W <= 128, "two row words per lane");
  // field tables while H is built; afterwards the same LDS stages the panel's pivot rows
  constexpr int CW2 = 32;  // words of a row per phase-2 pass
  constexpr int TABW = ceil_div(2 * ((1 << M) + 2 * Q), 8), STAGEW = 64 * CW2 + 256 * CW2;
  __shared__ uint64_t tab_lds[TABW > STAGEW ? TABW : STAGEW];
  uint16_t *lg_s = reinterpret_cast<uint16_t *>(tab_lds);
  uint16_t *ex_s = lg_s + (1 << M);  // ex_s[i] = generator^(i mod q): a sum of two logs needs no reduction
  uint64_t *srow = tab_lds;          // [pivot bit j][CW2 words] of the current panel
  uint64_t *tbl = tab_lds + 64 * CW2;  // [nibble g][16 combinations][CW2 words]
  __shared__ uint16_t rowix[MT];     // logical row -> physical row of H
  __shared__ uint16_t lix[MT];       // physical row -> logical row
  __shared__ uint64_t pw[MT];        // the panel's word of every physical row
  __shared__ uint64_t mk[MT];        // pivot rows (by bit) XORed into each physical row in this panel
  __shared__ int red[KG_WAVES];
  __shared__ int pvs[32];            // f variants: pivot columns of the window (relative to mt - 32)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cand = blockIdx.x;
  if (valid && !valid[cand]) {  // the attempt failed before the systematic form (uniform)
    if (tid == 0) ok_all[cand] = 0;
    return;
  }
  const uint16_t *g = g_all + (size_t)cand * (T + 1);
  const uint16_t *alpha = alpha_all + (size_t)cand * N;
  uint64_t *H = H_all + (size_t)cand * MT * W;
  for (int i = tid; i < (1 << M); i += KG_WG) lg_s[i] = lg[i];
  for (int i = tid; i < 2 * Q; i += KG_WG) ex_s[i] = ex[i < Q ? i : i - Q];
  for (int i = tid; i < MT; i += KG_WG) rowix[i] = lix[i] = (uint16_t)i;
  __syncthreads();
  auto mul = [&](uint32_t a, uint32_t b) -> uint32_t { return (a && b) ? (uint32_t)ex_s[lg_s[a] + lg_s[b]] : 0u; };

  // ---- H, 64 columns per wave pass: lane l takes column 64 wi + l (zero past n)
  for (int wi = wave; wi < W; wi += KG_WAVES) {
    const int j = 64 * wi + lane;
    const uint32_t a = j < N ? alpha[j] : 0u;
    uint32_t v = g[T];
    for (int k = T - 1; k >= 0; --k) v = mul(v, a) ^ g[k];
    v = j < N ? (uint32_t)ex_s[Q - lg_s[v]] : 0u;  // 1 / g(alpha_j) (g has no root in the field)
    for (int i = 0; i < T; ++i) {
      uint64_t out = 0;
#pragma unroll
      for (int b = 0; b < M; ++b) {
        const uint64_t bal = __ballot((v >> b) & 1u);
        out = lane == b ? bal : out;
      }
      if (lane < M) H[(size_t)(i * M + lane) * W + wi] = out;
      v = mul(v, a);
    }
  }
  __syncthreads();

  // ---- Gauss-Jordan elimination to [I | T], one panel at a time.  A panel is a run of pivot
  // columns inside one 64-bit word cw (it ends at the word's end, at mt, and for the f variants at
  // mt - 32).  Phase 1 eliminates the panel's columns on the word cw of every row, in LDS, and
  // records in mk[r] which of the panel's pivot rows (by column bit) row r has absorbed, as
  // combinations of the pivot rows' contents at the panel's start.  Phase 2 applies those
  // combinations to the words right of cw, 64 words at a time: each row is read and written once
  // per panel instead of once per column.
  for (int c = 0; c < MT;) {
    if constexpr (F) {
      if (c == MT - 32) {  // the f variants' column step (uniform branch)
        constexpr int C0 = MT - 32, W0 = C0 >> 6, SH = C0 & 63;
        auto window = [&](int prow) -> uint64_t {
          const uint64_t *row = H + (size_t)prow * W;
          return SH ? (row[W0] >> SH) | (row[W0 + 1] << (64 - SH)) : row[W0];
        };
        if (wave == 0) {  // echelon pivots of the 32 rows' windows, lane i holding row i
          uint64_t b = lane < 32 ? window(rowix[C0 + lane]) : 0ull;
          bool full = true;
          for (int i = 0; i < 32; ++i) {
            const int z = (lane >= i && lane < 32 && b) ? __builtin_ctzll(b) : 64;
            int m = z;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o));
            if (m == 64) { full = false; break; }  // uniform
            const uint64_t who = __ballot(z == m);
            const int wl = __builtin_ctzll(who);  // a row with the lowest leading column
            const uint64_t bw = __shfl(b, wl), bi = __shfl(b, i);
            if (lane == i) b = bw;
            else if (lane == wl) b = bi;
            if (lane > i && lane < 32 && ((b >> m) & 1ull)) b ^= bw;
            if (lane == 0) pvs[i] = m;
          }
          if (lane == 0) red[0] = full ? 1 : 0;
        }
        __syncthreads();
        if (red[0] == 0) {  // rank < 32 (uniform over the workgroup)
          if (tid == 0) ok_all[cand] = 0;
          return;
        }
        uint64_t mask = 0;
        int pv[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) { pv[i] = pvs[i]; mask |= 1ull << pv[i]; }
        for (int r = tid; r < MT; r += KG_WG) {  // the swaps on every physical row
          uint64_t t = window(r);
#pragma unroll
          for (int i = 0; i < 32; ++i) {
            const uint64_t d = ((t >> i) ^ (t >> pv[i])) & 1ull;
            t ^= (d << i) | (d << pv[i]);
          }
          uint64_t *row = H + (size_t)r * W;
          if (SH) {
            row[W0] = (row[W0] & ((1ull << SH) - 1ull)) | (t << SH);
            row[W0 + 1] = (row[W0 + 1] & ~((1ull << SH) - 1ull)) | (t >> (64 - SH));
          } else {
            row[W0] = t;
          }
        }
        if (tid == 0) piv_all[cand] = mask;
        __syncthreads();
      }
    }
    const int cw = c >> 6;
    int ce = 64 * (cw + 1) < MT ? 64 * (cw + 1) : MT;
    if (F && c < MT - 32 && ce > MT - 32) ce = MT - 32;
    for (int r = tid; r < MT; r += KG_WG) {
      pw[r] = H[(size_t)r * W + cw];
      mk[r] = 0;
    }
    __syncthreads();
    // phase 1: one pass over each thread's physical rows both applies the previous step's pivot
    // row and looks for the next pivot (the rows' logical indices in lix; the two rows of the
    // previous swap, which tid 0 writes meanwhile, by their new indices), so a step costs two
    // workgroup barriers
    int ppr = -1, ppc = -1, pbest = -1;  // previous step: pivot row, row it displaces, pivot's old logical index
    uint64_t pbm = 0, ppw = 0, pmk = 0;  // previous step: column bit, pivot row word and mask
    for (int cc = c;; ++cc) {
      const bool last = cc == ce;
      const uint64_t bm = last ? 0ull : 1ull << (cc & 63);
      if (tid == 0 && ppr >= 0 && pbest != cc - 1) {  // the previous step's row swap
        rowix[cc - 1] = (uint16_t)ppr;
        rowix[pbest] = (uint16_t)ppc;
        lix[ppr] = (uint16_t)(cc - 1);
        lix[ppc] = (uint16_t)pbest;
      }
      int best = MT;  // first logical row at or below cc with the column bit
      for (int r = tid; r < MT; r += KG_WG) {
        uint64_t v = pw[r];
        if (ppr >= 0 && r != ppr && (v & pbm)) {
          v ^= ppw;
          pw[r] = v;
          mk[r] ^= pmk;
        }
        if (!last) {
          const int l = r == ppr ? cc - 1 : r == ppc ? pbest : (int)lix[r];
          if (l >= cc && (v & bm) && l < best) best = l;
        }
      }
      if (last) break;
      best = wave_min(best);
      if (lane == 0) red[wave] = best;
      __syncthreads();
      best = red[0];
#pragma unroll
      for (int w = 1; w < KG_WAVES; ++w) best = min(best, red[w]);
      if (best == MT) {  // the left mt x mt block is singular (uniform over the workgroup)
        if (tid == 0) ok_all[cand] = 0;
        return;
      }
      ppr = rowix[best];
      ppc = rowix[cc];
      pbest = best;
      pbm = bm;
      ppw = pw[ppr];
      pmk = mk[ppr] ^ bm;
      __syncthreads();  // red, rowix and the pivot row read by every thread
    }
    __syncthreads();  // the last step's updates and swap
    for (int r = tid; r < MT; r += KG_WG) H[(size_t)r * W + cw] = pw[r];
    // phase 2: the pivot of column 64 cw + j sits at logical row 64 cw + j; its words right of cw
    // are still as at the panel's start (phase 1 wrote word cw only)
    // Four Russians, 32 words at a time: the panel's pivot rows are staged, then for each group
    // of four pivot bits a table of the 16 XOR combinations of those rows' words; a row's
    // combination is then 16 table reads, one per nibble of its mask, all independent.  A wave
    // takes two rows per pass (one per half-wave, lane = word) and loads the row words of its next
    // PG passes while it works on the current ones.
    constexpr int PG = 2;
    const int hh = lane >> 5, hl = lane & 31;
    for (int w0 = cw + 1; w0 < W; w0 += CW2) {
      for (int i = tid; i < 64 * CW2; i += KG_WG) {
        const int j = i / CW2, col = 64 * cw + j, w = w0 + i % CW2;
        srow[i] = (col >= c && col < ce && w < W) ? H[(size_t)rowix[col] * W + w] : 0ull;
      }
      __syncthreads();
      for (int i = tid; i < 256 * CW2; i += KG_WG) {  // tbl[g][e][x]: XOR of srow[4 g + b][x], bits b of e
        const int x = i % CW2, e = (i / CW2) & 15, g4 = 4 * (i / (16 * CW2));
        uint64_t v = 0;
#pragma unroll
        for (int bb = 0; bb < 4; ++bb)
          if ((e >> bb) & 1) v ^= srow[(g4 + bb) * CW2 + x];
        tbl[i] = v;
      }
      __syncthreads();
      const bool wok = w0 + hl < W;
      // pass k: rows 2 (KG_WAVES k + wave) + hh
      auto load = [&](int k, uint64_t &m, uint64_t &h) {
        const int r = 2 * (KG_WAVES * k + wave) + hh;
        m = r < MT ? mk[r] : 0ull;
        h = (m && wok) ? H[(size_t)r * W + w0 + hl] : 0ull;
      };
      constexpr int NPASS = ceil_div(MT, 2 * KG_WAVES);
      uint64_t mc[PG], hc[PG];
#pragma unroll
      for (int q = 0; q < PG; ++q) load(q, mc[q], hc[q]);
      for (int k0 = 0; k0 < NPASS; k0 += PG) {
        uint64_t mn[PG], hn[PG];
#pragma unroll
        for (int q = 0; q < PG; ++q) load(k0 + PG + q, mn[q], hn[q]);  // rows past mt load nothing
#pragma unroll
        for (int q = 0; q < PG; ++q) {
          const uint64_t m = mc[q];
          uint64_t acc = 0;
#pragma unroll
          for (int g = 0; g < 16; ++g) acc ^= tbl[(16 * g + (int)((m >> (4 * g)) & 15u)) * CW2 + hl];
          const int r = 2 * (KG_WAVES * (k0 + q) + wave) + hh;
          if (m && wok) H[(size_t)r * W + w0 + hl] = hc[q] ^ acc;
        }
#pragma unroll
        for (int q = 0; q < PG; ++q) {
          mc[q] = mn[q];
          hc[q] = hn[q];
        }
      }
      __syncthreads();  // srow and tbl are rewritten for the next words
    }
    c = ce;
  }

  // ---- pk row r = columns mt .. n-1 of logical row r (bits past n are zero in H)
  uint8_t *pk = pk_all + (size_t)cand * P::PK_BYTES;
  for (int idx = tid; idx < MT * P::PK_ROW_BYTES; idx += KG_WG) {
    const int r = idx / P::PK_ROW_BYTES, jb = idx - r * P::PK_ROW_BYTES;
    const uint64_t *row = H + (size_t)rowix[r] * W;
    const int bit = MT + 8 * jb, wb = bit >> 6, sh = bit & 63;
    uint64_t v = row[wb] >> sh;
    if (sh > 56 && wb + 1 < W) v |= row[wb + 1] << (64 - sh);
    pk[idx] = (uint8_t)v;
  }
  if (tid == 0) ok_all[cand] = 1;
}

// ------------------------------------------------------------------ Goppa polynomial
// g = the minimal polynomial of f over GF(2^m) in GF((2^m)^t) = GF(2^m)[y] / F(y)
// (LM:crypto_kem/<set>/vec/sk_gen.c genpoly_gen, gf_2mt_mul.c for F): the powers f^0 .. f^t as the
// columns of a t x (t + 1) matrix, Gauss-Jordan on its first t columns, g_i = the last column
// (the solution is unique when the system is nonsingular, so the elimination order is free).  One
// workgroup per candidate; products through log/exp tables in LDS.  y^t = F_low(y): the taps of F
// below t with their coefficients (all 1, except z = 2 at y^0 for 348864).
template <int SET> struct ExtTaps;
template <> struct ExtTaps<S348864> { static constexpr int NT = 3; static constexpr int tap[3] = {3, 1, 0}; static constexpr int coef[3] = {1, 1, 2}; };
template <> struct ExtTaps<S460896> { static constexpr int NT = 4; static constexpr int tap[4] = {10, 9, 6, 0}; static constexpr int coef[4] = {1, 1, 1, 1}; };
template <> struct ExtTaps<S6960119> { static constexpr int NT = 2; static constexpr int tap[2] = {8, 0}; static constexpr int coef[2] = {1, 1}; };
template <> struct ExtTaps<S6688128> { static constexpr int NT = 4; static constexpr int tap[4] = {7, 2, 1, 0}; static constexpr int coef[4] = {1, 1, 1, 1}; };
template <> struct ExtTaps<S8192128> { static constexpr int NT = 4; static constexpr int tap[4] = {7, 2, 1, 0}; static constexpr int coef[4] = {1, 1, 1, 1}; };

constexpr int GP_WG = 1024;
// threads per product coefficient: the largest power of two S with S (2t - 1) <= GP_WG
constexpr int gp_split(int t) {
  int s = 1;
  while (2 * s * (2 * t - 1) <= GP_WG) s *= 2;
  return s;
}

template <int SET>
__global__ __launch_bounds__(GP_WG) void k_genpoly(const uint16_t *__restrict__ f_all, const uint16_t *__restrict__ lg,
                                                   const uint16_t *__restrict__ ex, uint16_t *__restrict__ g_all,
                                                   int *__restrict__ valid) {
  using P = Par<SET>;
  using X = ExtTaps<SET>;
  constexpr int M = P::M, T = P::T, Q = P::Q;
  constexpr uint32_t ZL = 0xFFFFu;  // log of 0
  constexpr int C = T + 1;          // matrix columns (powers 0 .. t)
  static_assert(T >= 2 * 10, "taps stay below t/2");
  __shared__ uint16_t lg_s[1 << M];
  __shared__ uint16_t ex_s[2 * Q];  // ex_s[i] = generator^(i mod q)
  __shared__ uint16_t mat[T * C];   // mat[i C + j] = coefficient i of f^j
  __shared__ uint16_t lf[T], la[T], fac[T];
  __shared__ uint16_t prod[2 * T];
  __shared__ int red[GP_WG / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, cand = blockIdx.x;
  if (!valid[cand]) return;  // the attempt failed already (support sort keys repeat): uniform
  for (int i = tid; i < (1 << M); i += GP_WG) lg_s[i] = lg[i];
  for (int i = tid; i < 2 * Q; i += GP_WG) ex_s[i] = ex[i < Q ? i : i - Q];
  __syncthreads();
  auto lgz = [&](uint32_t a) -> uint32_t { return a ? (uint32_t)lg_s[a] : ZL; };
  auto mul_l = [&](uint32_t la_, uint32_t lb_) -> uint32_t {  // product of two logs
    return (la_ == ZL || lb_ == ZL) ? 0u : (uint32_t)ex_s[la_ + lb_];
  };
  const uint16_t *f = f_all + (size_t)cand * T;
  for (int i = tid; i < T; i += GP_WG) {
    const uint32_t fi = f[i];
    lf[i] = (uint16_t)lgz(fi);
    mat[i * C + 0] = i == 0 ? 1 : 0;
    mat[i * C + 1] = (uint16_t)fi;
  }
  __syncthreads();
  // ---- f^j = f^(j-1) f mod F(y), j = 2 .. t
  for (int j = 2; j <= T; ++j) {
    for (int i = tid; i < T; i += GP_WG) la[i] = (uint16_t)lgz(mat[i * C + j - 1]);
    __syncthreads();
    {  // coefficient k of the product by S adjacent lanes, each over every S-th term
      constexpr int S = gp_split(T);
      static_assert(S <= 64 && S * (2 * T - 1) <= GP_WG, "one pass, lanes of one wave");
      const int k = tid / S, part = tid % S;
      uint32_t acc = 0;
      if (k < 2 * T - 1) {
        const int lo = k - T + 1 > 0 ? k - T + 1 : 0, hi = k < T - 1 ? k : T - 1;
        for (int i = lo + part; i <= hi; i += S) acc ^= mul_l(la[i], lf[k - i]);
      }
#pragma unroll
      for (int o = 1; o < S; o <<= 1) acc ^= (uint32_t)__shfl_xor((int)acc, o);
      if (k < 2 * T - 1 && part == 0) prod[k] = (uint16_t)acc;
    }
    __syncthreads();
    // y^i for i >= t: the reduction from the top, in closed form.  Index q in [t, 2t-2] receives
    // tap contributions only from indices >= 2t - 10, which receive none, so
    // c_q = p_q + sum_taps coef c_(q + t - tap) with c = p on the right; then the output below t
    // is p_k + sum_taps coef c_(k + t - tap).
    for (int q = T + tid; q < 2 * T - 1; q += GP_WG) {
      uint32_t v = prod[q];
#pragma unroll
      for (int u = 0; u < X::NT; ++u) {
        const int src = q + T - X::tap[u];
        if (src <= 2 * T - 2) {
          const uint32_t sv = prod[src];
          v ^= X::coef[u] == 1 ? sv : mul_l(lgz(sv), lg_s[X::coef[u]]);
        }
      }
      la[q - T] = (uint16_t)v;  // c_q, parked in la (free now) at q - t <= t - 2
    }
    __syncthreads();
    for (int k = tid; k < T; k += GP_WG) {
      uint32_t v = prod[k];
#pragma unroll
      for (int u = 0; u < X::NT; ++u) {
        const int src = k + T - X::tap[u];
        if (src >= T && src <= 2 * T - 2) {
          const uint32_t cv = la[src - T];
          v ^= X::coef[u] == 1 ? cv : mul_l(lgz(cv), lg_s[X::coef[u]]);
        }
      }
      mat[k * C + j] = (uint16_t)v;
    }
    __syncthreads();
  }
  // ---- Gauss-Jordan on columns 0 .. t-1; column t is the right-hand side
  for (int c = 0; c < T; ++c) {
    int best = T;
    for (int r = c + tid; r < T; r += GP_WG)
      if (mat[r * C + c]) {
        best = r;
        break;
      }
    best = wave_min(best);
    if (lane == 0) red[wave] = best;
    __syncthreads();
    best = red[0];
#pragma unroll
    for (int w = 1; w < GP_WG / 64; ++w) best = min(best, red[w]);
    if (best == T) {  // 1, f, ..., f^(t-1) are linearly dependent: the attempt fails (uniform)
      if (tid == 0) valid[cand] = 0;
      return;
    }
    if (best != c)
      for (int j = c + tid; j < C; j += GP_WG) {
        const uint16_t t0 = mat[c * C + j];
        mat[c * C + j] = mat[best * C + j];
        mat[best * C + j] = t0;
      }
    __syncthreads();
    const uint32_t linv = (uint32_t)(Q - lg_s[mat[c * C + c]]);  // log of the pivot's inverse (< 2q)
    __syncthreads();
    for (int j = c + tid; j < C; j += GP_WG) mat[c * C + j] = (uint16_t)mul_l(lgz(mat[c * C + j]), linv);
    for (int r = tid; r < T; r += GP_WG) fac[r] = r == c ? 0 : mat[r * C + c];
    __syncthreads();
    for (int e = tid; e < T * (C - c); e += GP_WG) {
      const int r = e / (C - c), j = c + e % (C - c);
      const uint32_t fr = fac[r];
      if (fr) mat[r * C + j] ^= (uint16_t)mul_l(lgz(fr), lgz(mat[c * C + j]));
    }
    __syncthreads();
  }
  uint16_t *g = g_all + (size_t)cand * (T + 1);
  for (int i = tid; i < T; i += GP_WG) g[i] = mat[i * C + T];
  if (tid == 0) g[T] = 1;
}

hipError_t launch_genpoly_set(int set, const uint16_t *f, const uint16_t *lg, const uint16_t *ex, uint16_t *g,
                              int *valid, int ncand, hipStream_t s) {
  if (ncand <= 0) return hipSuccess;
  switch (set) {
    case S348864: k_genpoly<S348864><<<ncand, GP_WG, 0, s>>>(f, lg, ex, g, valid); break;
    case S460896: k_genpoly<S460896><<<ncand, GP_WG, 0, s>>>(f, lg, ex, g, valid); break;
    case S6688128: k_genpoly<S6688128><<<ncand, GP_WG, 0, s>>>(f, lg, ex, g, valid); break;
    case S6960119: k_genpoly<S6960119><<<ncand, GP_WG, 0, s>>>(f, lg, ex, g, valid); break;
    case S8192128: k_genpoly<S8192128><<<ncand, GP_WG, 0, s>>>(f, lg, ex, g, valid); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------ Benes control bits
// The recursion of LM:crypto_kem/<set>/vec/controlbits.c (cbrecursion), level by level: at level d
// (w = m - d) the 2^d subproblems are the segments of 2^w entries of the level's permutation array,
// each a permutation of its own 0 .. 2^w - 1, and every step of the recursion body runs on all
// segments at once (elementwise, or as a sort inside each segment, done as the scatter that the
// sort's keys make it equal to, below).  Subproblem s of level d (its choices
// child 0 / child 1 from the top are the bits of s, most significant first) writes its first-layer
// bits at d n/2 + rev_d(s) + 2^d j and its last-layer bits at (2m - 2 - d) n/2 + rev_d(s) + 2^d k,
// rev_d = d-bit reversal, which is where the recursion's position and step arithmetic puts them;
// its children are its two halves, segments 2s and 2s + 1 of level d + 1.  One workgroup per key.
// Every "sort" of the recursion sorts words whose high field (bits sh and up) runs over a
// permutation of 0 .. 2^w - 1 inside each segment of 2^w entries, so the sorted segment holds word x
// at position (its high field): the sort is a scatter into a second array, copied back.  Wave v of
// the workgroup owns the entries [CH v, CH (v + 1)) (CH = 2^m / 16): the elementwise steps touch
// only their own entry or its pair partner, and a scatter inside segments no wider than a chunk
// stays inside the wave's chunk, so those need only the wave's own LDS ordering; a scatter across
// chunks runs between workgroup barriers.
constexpr int CB_WG = 1024;

template <int MM>
__device__ __forceinline__ void cb_sort(uint32_t *v, uint32_t *tmp, int w, int sh, int tid) {
  constexpr int NN = 1 << MM, CH = NN / (CB_WG / 64);
  const int lane = tid & 63, c0 = (tid >> 6) * CH, lm = (1 << w) - 1;
  const bool wide = (1 << w) > CH;
  // wide: every wave is done with tmp (its previous copy-back) before any wave writes it
  if (wide) __syncthreads();
  else wave_lds_sync();
  for (int x = c0 + lane; x < c0 + CH; x += 64) {
    const uint32_t a = v[x];
    tmp[(x & ~lm) | ((a >> sh) & lm)] = a;
  }
  if (wide) __syncthreads();
  else wave_lds_sync();
  for (int x = c0 + lane; x < c0 + CH; x += 64) v[x] = tmp[x];
  wave_lds_sync();
}

template <int MM>
__global__ __launch_bounds__(CB_WG) void k_controlbits(const uint16_t *__restrict__ pi_all, uint8_t *__restrict__ cb_all) {
  constexpr int NN = 1 << MM, HALF = NN / 2, CH = NN / (CB_WG / 64);
  constexpr int CBW = (2 * MM - 1) * HALF / 32;  // control-bit words: (2m - 1) layers of n/2 bits
  __shared__ uint32_t A[NN], B[NN], C[NN], cbw[CBW];
  __shared__ uint16_t PI[NN];
  const int tid = threadIdx.x, lane = tid & 63, c0 = (tid >> 6) * CH, c1 = c0 + CH;
  const uint16_t *pi = pi_all + (size_t)blockIdx.x * NN;
  for (int x = tid; x < NN; x += CB_WG) PI[x] = pi[x];
  for (int i = tid; i < CBW; i += CB_WG) cbw[i] = 0;
  __syncthreads();
  auto flip = [&](int pos, uint32_t b) {
    if (b & 1u) atomicXor(&cbw[pos >> 5], 1u << (pos & 31));
  };
  // x: the wave's own entries; p: the wave's own pairs (2p, 2p + 1)
#define CB_OWN(x) for (int x = c0 + lane; x < c1; x += 64)
#define CB_OWN_PAIRS(p) for (int p = (c0 >> 1) + lane; p < (c1 >> 1); p += 64)
  for (int d = 0; d < MM - 1; ++d) {
    const int w = MM - d, ns = 1 << w, lm = ns - 1;
    CB_OWN(x) A[x] = ((uint32_t)(PI[x] ^ 1u) << 16) | PI[x ^ 1];
    cb_sort<MM>(A, C, w, 16, tid);
    CB_OWN(x) {
      const uint32_t px = A[x] & 0xffffu, lx = (uint32_t)(x & lm);
      B[x] = (px << 16) | (px < lx ? px : lx);
      A[x] = (A[x] << 16) | lx;
    }
    cb_sort<MM>(A, C, w, 16, tid);
    CB_OWN(x) A[x] = (A[x] << 16) + (B[x] >> 16);
    cb_sort<MM>(A, C, w, 16, tid);
    if (w <= 10) {
      CB_OWN(x) B[x] = ((A[x] & 0xffffu) << 10) | (B[x] & 0x3ffu);
      for (int i = 1; i < w - 1; ++i) {
        CB_OWN(x) A[x] = ((B[x] & ~0x3ffu) << 6) | (uint32_t)(x & lm);
        cb_sort<MM>(A, C, w, 16, tid);
        CB_OWN(x) A[x] = (A[x] << 20) | B[x];
        cb_sort<MM>(A, C, w, 20, tid);
        CB_OWN(x) {
          const uint32_t u = (A[x] & 0xffc00u) | (B[x] & 0x3ffu), v = A[x] & 0xfffffu;
          B[x] = u < v ? u : v;
        }
      }
      CB_OWN(x) B[x] &= 0x3ffu;
    } else {
      CB_OWN(x) B[x] = (A[x] << 16) | (B[x] & 0xffffu);
      for (int i = 1; i < w - 1; ++i) {
        CB_OWN(x) A[x] = (B[x] & ~0xffffu) | (uint32_t)(x & lm);
        cb_sort<MM>(A, C, w, 16, tid);
        CB_OWN(x) A[x] = (A[x] << 16) | (B[x] & 0xffffu);
        if (i < w - 2) {
          CB_OWN(x) B[x] = (A[x] & ~0xffffu) | (B[x] >> 16);
          cb_sort<MM>(B, C, w, 16, tid);
          CB_OWN(x) B[x] = (B[x] << 16) | (A[x] & 0xffffu);
        }
        cb_sort<MM>(A, C, w, 16, tid);
        CB_OWN(x) {
          const uint32_t u = (B[x] & ~0xffffu) | (A[x] & 0xffffu);
          B[x] = B[x] < u ? B[x] : u;
        }
      }
      CB_OWN(x) B[x] &= 0xffffu;
    }
    CB_OWN(x) A[x] = ((uint32_t)PI[x] << 16) + (uint32_t)(x & lm);
    cb_sort<MM>(A, C, w, 16, tid);
    const int rb = d ? 32 - d : 0;  // rev_d(seg) = bit reversal of the segment index in d bits
    CB_OWN_PAIRS(p) {  // first layer: pair (2j, 2j + 1) of segment seg
      const int x = 2 * p, seg = x >> w, j = (x & lm) >> 1;
      const int rv = d ? (int)(__builtin_bitreverse32((uint32_t)seg) >> rb) : 0;
      const uint32_t fj = B[x] & 1u, Fz = (uint32_t)(x & lm) + fj;
      flip(d * HALF + rv + (j << d), fj);
      B[x] = (A[x] << 16) | Fz;
      B[x + 1] = (A[x + 1] << 16) | (Fz ^ 1u);
    }
    cb_sort<MM>(B, C, w, 16, tid);
    CB_OWN_PAIRS(p) {  // last layer
      const int y = 2 * p, seg = y >> w, k = (y & lm) >> 1;
      const int rv = d ? (int)(__builtin_bitreverse32((uint32_t)seg) >> rb) : 0;
      const uint32_t lk = B[y] & 1u, Ly = (uint32_t)(y & lm) + lk;
      flip((2 * MM - 2 - d) * HALF + rv + (k << d), lk);
      A[y] = (Ly << 16) | (B[y] & 0xffffu);
      A[y + 1] = ((Ly ^ 1u) << 16) | (B[y + 1] & 0xffffu);
    }
    cb_sort<MM>(A, C, w, 16, tid);
    // every wave has read this level's PI (the scatters above of a segment wider than a chunk hold
    // workgroup barriers; a segment within a chunk is read and written by its own wave only)
    CB_OWN_PAIRS(p) {  // the children's permutations: the segment's halves
      const int y = 2 * p, seg = y >> w, j = (y & lm) >> 1, base = seg << w;
      PI[base + j] = (uint16_t)((A[y] & 0xffffu) >> 1);
      PI[base + (ns >> 1) + j] = (uint16_t)((A[y + 1] & 0xffffu) >> 1);
    }
    __syncthreads();
  }
#undef CB_OWN
#undef CB_OWN_PAIRS
  for (int p = tid; p < HALF; p += CB_WG) {  // w = 1: the middle layer, one bit per segment of 2
    const int rv = (int)(__builtin_bitreverse32((uint32_t)p) >> (32 - (MM - 1)));
    flip((MM - 1) * HALF + rv, PI[2 * p]);
  }
  __syncthreads();
  uint32_t *out = reinterpret_cast<uint32_t *>(cb_all + (size_t)blockIdx.x * (CBW * 4));
  for (int i = tid; i < CBW; i += CB_WG) out[i] = cbw[i];
}

hipError_t launch_controlbits(int m, const uint16_t *pi, uint8_t *cb, int nkeys, hipStream_t s) {
  if (nkeys <= 0) return hipSuccess;
  if (m == 12) k_controlbits<12><<<nkeys, CB_WG, 0, s>>>(pi, cb);
  else if (m == 13) k_controlbits<13><<<nkeys, CB_WG, 0, s>>>(pi, cb);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int SET>
hipError_t pk_gen(const KeygenArgs &a, hipStream_t s) {
  if (a.ncand <= 0) return hipSuccess;
  if (a.pivots)
    k_pk_gen<SET, true><<<a.ncand, KG_WG, 0, s>>>(a.g, a.alpha, a.lg, a.ex, a.H, a.pk, a.ok, a.pivots, a.valid);
  else
    k_pk_gen<SET, false><<<a.ncand, KG_WG, 0, s>>>(a.g, a.alpha, a.lg, a.ex, a.H, a.pk, a.ok, nullptr, a.valid);
  return hipGetLastError();
}

hipError_t launch_pk_gen_set(int set, const KeygenArgs &a, hipStream_t s) {
  switch (set) {
    case S348864: return pk_gen<S348864>(a, s);
    case S460896: return pk_gen<S460896>(a, s);
    case S6688128: return pk_gen<S6688128>(a, s);
    case S6960119: return pk_gen<S6960119>(a, s);
    case S8192128: return pk_gen<S8192128>(a, s);
  }
  return hipErrorInvalidValue;
}

size_t pk_gen_h_bytes(int set) {
  const SetInfo &si = set_info(set);
  return (size_t)si.mt * ((si.n + 63) / 64) * 8;
}

}  // namespace tmce
