// host_math.cpp -- host-side field arithmetic, key precomputation and key generation.
#include "host_math.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>

#include "params.h"

namespace tmce {

// ------------------------------------------------------------------ GF(2^m)
// m=12: z^12+z^3+1, m=13: z^13+z^4+z^3+z+1 (LM:crypto_kem/<set>/vec/gf.c)
static uint16_t mul_poly(int m, uint16_t a, uint16_t b) {
  const uint32_t poly = m == 12 ? 0x1009u : 0x201Bu;
  uint32_t r = 0;
  for (int i = 0; i < m; ++i)
    if ((b >> i) & 1) r ^= (uint32_t)a << i;
  for (int i = 2 * m - 2; i >= m; --i)
    if ((r >> i) & 1) r ^= poly << (i - m);
  return (uint16_t)r;
}

HostField::HostField(int m_) : m(m_), q((1u << m_) - 1) {
  lg.assign(1u << m, 0);
  ex.assign(2 * q, 0);
  for (uint16_t g = 2;; ++g) {  // smallest generator of the multiplicative group
    uint16_t x = 1;
    uint32_t ord = 0;
    do { x = mul_poly(m, x, g); ++ord; } while (x != 1 && ord <= q);
    if (ord != q) continue;
    x = 1;
    for (uint32_t i = 0; i < 2 * q; ++i) {
      ex[i] = x;
      if (i < q) lg[x] = (uint16_t)i;
      x = mul_poly(m, x, g);
    }
    break;
  }
}

uint16_t HostField::pw(uint16_t a, uint32_t e) const {
  if (e == 0) return 1;
  if (!a) return 0;
  return ex[(uint64_t)lg[a] * e % q];
}

const HostField &host_field(int m) {
  static HostField f12(12), f13(13);
  if (m == 12) return f12;
  if (m == 13) return f13;
  throw std::invalid_argument("unsupported field");
}

uint16_t bitrev(uint16_t a, int m) {
  uint16_t r = 0;
  for (int i = 0; i < m; ++i) r |= (uint16_t)(((a >> i) & 1u) << (m - 1 - i));
  return r;
}

// ------------------------------------------------------------------ Benes control bits
void perm_from_controlbits(int m, const uint8_t *cb, std::vector<uint16_t> &pi) {
  const size_t n = size_t(1) << m;
  pi.resize(n);
  for (size_t i = 0; i < n; ++i) pi[i] = (uint16_t)i;
  size_t bit = 0;
  for (int layer = 0; layer < 2 * m - 1; ++layer) {
    const int s = layer < m ? layer : 2 * m - 2 - layer;
    const size_t stride = size_t(1) << s;
    for (size_t blk = 0; blk < n; blk += 2 * stride)
      for (size_t j = 0; j < stride; ++j, ++bit)
        if ((cb[bit >> 3] >> (bit & 7)) & 1) std::swap(pi[blk + j], pi[blk + j + stride]);
  }
}

// Control bits of the Benes network for a permutation pi of n = 2^w indices (the bits libmceliece
// stores in sk, LM:crypto_kem/<set>/vec/controlbits.c; the formulas of Bernstein, "Verified fast
// formulas for control bits for permutation networks"), computed here directly from their
// definition with index arrays:
//   sigma(y) = pi(pi^-1(y ^ 1) ^ 1), whose cycles have length at most n/2;
//   c(x)     = the smallest element of x's cycle under sigma;
//   first layer: bit f_j = c(2j) mod 2, and F swaps the pair (2j, 2j + 1) when f_j = 1;
//   last layer:  bit l_k = F(pi(2k)) mod 2, and L swaps the pair (2k, 2k + 1) when l_k = 1;
//   middle:      M = F o pi o L, and the two half networks route q_b(j) = M(2j + b) / 2 (b = 0, 1).
// The first layer's bits sit at pos + j step, the last layer's at pos + (2w - 2)(n/2) step + k step,
// and the two halves continue at pos + (n/2) step and that + step, with twice the step.
namespace {
struct CbWork {
  std::vector<uint16_t> inv, sig, cmin, q[16];
  std::vector<uint8_t> seen;
};

void cb_level(uint8_t *out, long long pos, long long step, const uint16_t *pi, int w, CbWork &wk) {
  auto flip = [&](long long p, uint32_t b) { out[p >> 3] ^= (uint8_t)((b & 1u) << (p & 7)); };
  if (w == 1) {
    flip(pos, pi[0]);
    return;
  }
  const size_t n = size_t(1) << w, half = n / 2;
  wk.inv.resize(n);
  wk.sig.resize(n);
  wk.cmin.resize(n);
  wk.seen.assign(n, 0);
  uint16_t *inv = wk.inv.data(), *sig = wk.sig.data(), *cmin = wk.cmin.data();
  for (size_t x = 0; x < n; ++x) inv[pi[x]] = (uint16_t)x;
  for (size_t y = 0; y < n; ++y) sig[y] = pi[inv[y ^ 1] ^ 1];
  for (size_t x = 0; x < n; ++x) {  // cycle minima: one walk per cycle
    if (wk.seen[x]) continue;
    uint16_t mn = (uint16_t)x;
    for (size_t y = sig[x]; y != x; y = sig[y]) mn = std::min<uint16_t>(mn, (uint16_t)y);
    cmin[x] = mn;
    wk.seen[x] = 1;
    for (size_t y = sig[x]; y != x; y = sig[y]) {
      cmin[y] = mn;
      wk.seen[y] = 1;
    }
  }
  std::vector<uint16_t> &q = wk.q[w];
  q.resize(n);
  // f and l bits; F and L as XOR masks on the low index bit
  auto F = [&](uint32_t x) { return x ^ (cmin[x & ~1u] & 1u); };
  const long long last = pos + (long long)(2 * w - 2) * (long long)half * step;
  for (size_t j = 0; j < half; ++j) flip(pos + (long long)j * step, cmin[2 * j]);
  for (size_t k = 0; k < half; ++k) {
    const uint32_t lk = F(pi[2 * k]) & 1u;
    flip(last + (long long)k * step, lk);
    // M(2k + b) = F(pi(L(2k + b))) = F(pi(2k + (b ^ lk)))
    q[k] = (uint16_t)(F(pi[2 * k + lk]) >> 1);
    q[half + k] = (uint16_t)(F(pi[2 * k + (lk ^ 1u)]) >> 1);
  }
  const long long next = pos + (long long)half * step;
  cb_level(out, next, 2 * step, q.data(), w - 1, wk);
  cb_level(out, next + step, 2 * step, q.data() + half, w - 1, wk);
}
}  // namespace

void controlbits_from_perm(int m, const std::vector<uint16_t> &pi, uint8_t *cb) {
  const size_t n = size_t(1) << m;
  std::memset(cb, 0, ((2 * m - 1) * n / 2 + 7) / 8);
  static thread_local CbWork wk;
  cb_level(cb, 0, 1, pi.data(), m, wk);
}

// ------------------------------------------------------------------ additive FFT tables
// Gao-Mateer recursion on the standard basis (1, z, ..., z^(m-1)) so that output index c is the
// field element c.  Level d: twist by the last basis element b (coefficient i times b^i), Taylor
// expansion at x^2+x, recurse on span(gamma_i^2 + gamma_i) with gamma_i = basis_i / b; the
// bottom-up butterfly at level d multiplies by p(c) = sum_i c_i gamma_i.
FftTables build_fft_tables(int m, int D) {
  const HostField &F = host_field(m);
  int KD = 0;
  while ((1 << KD) < D) ++KD;
  FftTables t;
  t.twl.assign(2 * D, 0);
  t.plo.assign((size_t)KD * 64, 0);
  size_t phi_total = 0;
  for (int d = 0; d < KD; ++d) phi_total += size_t(1) << (m - 7 - d < 0 ? 0 : m - 7 - d);
  t.phi.assign(phi_total, 0);
  size_t phi_off = 0;
  std::vector<uint16_t> basis(m);
  for (int i = 0; i < m; ++i) basis[i] = (uint16_t)(1u << i);
  for (int d = 0; d < KD; ++d) {
    const int Ld = D >> d;
    const uint16_t b = basis.back();
    for (int i = 0; i < Ld; ++i) t.twl[2 * D - 2 * Ld + i] = F.lg[F.pw(b, (uint32_t)i)];
    const uint16_t binv = F.inv(b);
    std::vector<uint16_t> gam(basis.size() - 1);
    for (size_t i = 0; i + 1 < basis.size(); ++i) gam[i] = F.mul(basis[i], binv);
    auto pval = [&](size_t c) {
      uint16_t p = 0;
      for (size_t i = 0; i < gam.size(); ++i)
        if ((c >> i) & 1) p ^= gam[i];
      return p;
    };
    for (size_t l = 0; l < 64; ++l) t.plo[(size_t)d * 64 + l] = pval(l);
    const size_t nhi = size_t(1) << (m - 7 - d);
    for (size_t j = 0; j < nhi; ++j) t.phi[phi_off + j] = pval(j << 6);
    phi_off += nhi;
    for (size_t i = 0; i < gam.size(); ++i) gam[i] = (uint16_t)(F.mul(gam[i], gam[i]) ^ gam[i]);
    basis = gam;
  }
  return t;
}

// ------------------------------------------------------------------ per-key precomputation
template <int SET>
static void build_dec_key_t(const uint8_t *sk, DecKeyHost &o) {
  using P = Par<SET>;
  const HostField &F = host_field(P::M);
  std::vector<uint16_t> pi;
  perm_from_controlbits(P::M, sk + 40 + P::IRR_BYTES, pi);
  std::vector<uint16_t> g(P::T + 1);
  for (int i = 0; i < P::T; ++i) g[i] = (uint16_t)((sk[40 + 2 * i] | (sk[41 + 2 * i] << 8)) & ((1 << P::M) - 1));
  g[P::T] = 1;
  o.alpha.resize(P::N);
  o.la.resize(P::N);
  o.lu.resize(P::N);
  std::vector<uint16_t> u(P::N);
  for (int i = 0; i < P::N; ++i) {
    const uint16_t a = bitrev(pi[i], P::M);
    uint16_t v = g[P::T];
    for (int k = P::T - 1; k >= 0; --k) v = (uint16_t)(F.mul(v, a) ^ g[k]);
    u[i] = F.inv(F.mul(v, v));
    o.alpha[i] = a;
    o.la[i] = F.logv(a);
    o.lu[i] = F.logv(u[i]);
  }
  o.ainv.assign(1u << P::M, 0xFFFFu);
  o.supp.assign((1u << P::M) / 32, 0u);
  for (int i = 0; i < P::N; ++i) {
    o.ainv[o.alpha[i]] = (uint16_t)i;
    // decoder point order: field element a = 64 g + l is bit g % 32 of word l * (2^m / 2048) + g / 32
    const uint32_t av = o.alpha[i], g = av >> 6, l = av & 63;
    o.supp[l * ((1u << P::M) / 2048) + g / 32] |= 1u << (g % 32);
  }
  // front-end matrix: row j*M + b, column i (i < KBITS): bit b of alpha_i^j u_i (syndrome-major)
  // and the same map over all N positions, column-major (the decoder's re-encryption check)
  o.F.assign((size_t)P::R * P::KWP, 0);
  // columns padded to 16-byte rows, and a zero column N (the re-encryption gather's padding reads it)
  o.Fcol.assign((size_t)(P::N + 1) * P::FCW, 0);
  for (int i = 0; i < P::N; ++i) {
    uint16_t v = u[i];
    for (int j = 0; j < 2 * P::T; ++j) {
      for (int b = 0; b < P::M; ++b)
        if ((v >> b) & 1) {
          const int r = j * P::M + b;
          if (i < P::KBITS) o.F[(size_t)r * P::KWP + i / 32] |= 1u << (i % 32);
          o.Fcol[(size_t)i * P::FCW + r / 32] |= 1u << (r % 32);
        }
      v = F.mul(v, o.alpha[i]);
    }
  }
  o.rej.assign(sk + 40 + P::IRR_BYTES + P::COND_BYTES, sk + 40 + P::IRR_BYTES + P::COND_BYTES + P::N / 8);
}

void build_dec_key(int set, const uint8_t *sk, DecKeyHost &out) {
  switch (set) {
    case S348864: return build_dec_key_t<S348864>(sk, out);
    case S460896: return build_dec_key_t<S460896>(sk, out);
    case S6688128: return build_dec_key_t<S6688128>(sk, out);
    case S6960119: return build_dec_key_t<S6960119>(sk, out);
    case S8192128: return build_dec_key_t<S8192128>(sk, out);
  }
  throw std::invalid_argument("set");
}

template <int SET>
static void build_enc_cols_t(const uint8_t *pk, std::vector<uint32_t> &cols) {
  using P = Par<SET>;
  cols.assign((size_t)P::N * P::CWP, 0);
  for (int p = 0; p < P::MT; ++p) cols[(size_t)p * P::CWP + p / 32] = 1u << (p % 32);
  for (int r = 0; r < P::MT; ++r) {
    const uint8_t *row = pk + (size_t)r * P::PK_ROW_BYTES;
    for (int j = 0; j < P::K; ++j)
      if ((row[j >> 3] >> (j & 7)) & 1) cols[(size_t)(P::MT + j) * P::CWP + r / 32] |= 1u << (r % 32);
  }
}

void build_enc_cols(int set, const uint8_t *pk, std::vector<uint32_t> &cols) {
  switch (set) {
    case S348864: return build_enc_cols_t<S348864>(pk, cols);
    case S460896: return build_enc_cols_t<S460896>(pk, cols);
    case S6688128: return build_enc_cols_t<S6688128>(pk, cols);
    case S6960119: return build_enc_cols_t<S6960119>(pk, cols);
    case S8192128: return build_enc_cols_t<S8192128>(pk, cols);
  }
  throw std::invalid_argument("set");
}

void build_enc_rows(int set, const uint8_t *pk, std::vector<uint32_t> &rows) {
  const SetInfo &si = set_info(set);
  const int kw = (si.k + 31) / 32;
  rows.assign((size_t)si.mt * kw, 0);
  for (int r = 0; r < si.mt; ++r)
    for (int j = 0; j < si.k; ++j)
      if ((pk[(size_t)r * si.pk_row_bytes + j / 8] >> (j % 8)) & 1) rows[(size_t)r * kw + j / 32] |= 1u << (j % 32);
}

// LM:crypto_kem/6960119/vec/kem_enc.c check_pk_padding
bool pk_padding_bad(int set, const uint8_t *pk) {
  if (set != S6960119) return false;
  using P = Par<S6960119>;
  uint8_t b = 0;
  for (int i = 0; i < P::MT; ++i) b |= pk[(size_t)i * P::PK_ROW_BYTES + P::PK_ROW_BYTES - 1];
  return (b >> (P::K % 8)) != 0;
}

// ------------------------------------------------------------------ SHAKE256 (host)
// Keccak-f[1600] (FIPS 202): theta, rho and pi fused into one pass over the 25 lanes (lane i of
// the state goes to position PI_DST[i], rotated by ROT[i]), then chi and iota; the loops over
// constant tables unroll into straight-line code with constant rotations.
static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "state lanes are read and written as bytes");
static inline uint64_t rol(uint64_t x, int r) { return (x << r) | (x >> ((64 - r) & 63)); }
static void keccakf(uint64_t s[25]) {
  static constexpr uint64_t RC[24] = {
      0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
      0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
      0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
      0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
      0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
      0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
  // lane x + 5 y: rotation offset, and its position y + 5 ((2 x + 3 y) mod 5) after pi
  static constexpr int ROT[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
  static constexpr int PI_DST[25] = {0, 10, 20, 5, 15, 16, 1, 11, 21, 6, 7, 17, 2, 12, 22, 23, 8, 18, 3, 13, 14, 24, 9, 19, 4};
  uint64_t a[25], b[25];
  std::memcpy(a, s, sizeof a);
  for (int r = 0; r < 24; ++r) {
    uint64_t c[5], d[5];
#pragma unroll
    for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
#pragma unroll
    for (int x = 0; x < 5; ++x) d[x] = c[(x + 4) % 5] ^ rol(c[(x + 1) % 5], 1);
#pragma unroll
    for (int i = 0; i < 25; ++i) b[PI_DST[i]] = rol(a[i] ^ d[i % 5], ROT[i]);
#pragma unroll
    for (int i = 0; i < 25; ++i) a[i] = b[i] ^ (~b[(i / 5) * 5 + (i + 1) % 5] & b[(i / 5) * 5 + (i + 2) % 5]);
    a[0] ^= RC[r];
  }
  std::memcpy(s, a, sizeof a);
}

void shake256(uint8_t *out, long long outlen, const uint8_t *in, long long inlen) {
  uint64_t s[25] = {0};
  auto xor_block = [&](const uint8_t *p, size_t len) {
    uint8_t *sb = reinterpret_cast<uint8_t *>(s);
    for (size_t i = 0; i < len; ++i) sb[i] ^= p[i];
  };
  while (inlen >= 136) {
    xor_block(in, 136);
    keccakf(s);
    in += 136;
    inlen -= 136;
  }
  xor_block(in, (size_t)inlen);
  reinterpret_cast<uint8_t *>(s)[inlen] ^= 0x1F;
  reinterpret_cast<uint8_t *>(s)[135] ^= 0x80;
  while (outlen > 0) {
    keccakf(s);
    const long long take = outlen < 136 ? outlen : 136;
    std::memcpy(out, s, (size_t)take);
    out += take;
    outlen -= take;
  }
}

// ------------------------------------------------------------------ key generation
// GF((2^m)^t) multiplication modulo the set's irreducible F(y) (LM:crypto_kem/<set>/vec/gf_2mt_mul.c)
template <int SET>
static void ext_mul(const HostField &F, const uint16_t *a, const uint16_t *b, uint16_t *out) {
  using P = Par<SET>;
  constexpr int T = P::T;
  uint16_t prod[2 * T - 1] = {0};
  for (int i = 0; i < T; ++i)
    if (a[i])
      for (int j = 0; j < T; ++j) prod[i + j] ^= F.mul(a[i], b[j]);
  for (int i = 2 * T - 2; i >= T; --i) {
    const uint16_t c = prod[i];
    if constexpr (SET == S348864) { prod[i - T + 3] ^= c; prod[i - T + 1] ^= c; prod[i - T] ^= F.mul(c, 2); }
    else if constexpr (SET == S460896) { prod[i - T + 10] ^= c; prod[i - T + 9] ^= c; prod[i - T + 6] ^= c; prod[i - T] ^= c; }
    else if constexpr (SET == S6960119) { prod[i - T + 8] ^= c; prod[i - T] ^= c; }
    else { prod[i - T + 7] ^= c; prod[i - T + 2] ^= c; prod[i - T + 1] ^= c; prod[i - T] ^= c; }
  }
  std::memcpy(out, prod, sizeof(uint16_t) * T);
}

// Goppa polynomial = minimal polynomial of f over GF(2^m) (LM:.../sk_gen.c genpoly_gen);
// false when 1, f, ..., f^(t-1) are linearly dependent.
template <int SET>
static bool min_poly(const HostField &F, const uint16_t *f, uint16_t *g) {
  using P = Par<SET>;
  constexpr int T = P::T;
  std::vector<uint16_t> cols((size_t)(T + 1) * T, 0);  // column j = coefficients of f^j
  cols[0] = 1;
  std::memcpy(&cols[T], f, sizeof(uint16_t) * T);
  for (int j = 2; j <= T; ++j) ext_mul<SET>(F, &cols[(size_t)(j - 1) * T], f, &cols[(size_t)j * T]);
  // Gauss-Jordan on rows = coefficient index, columns 0..T-1 = powers, column T = rhs
  auto at = [&](int row, int col) -> uint16_t & { return cols[(size_t)col * T + row]; };
  for (int c = 0; c < T; ++c) {
    int piv = c;
    while (piv < T && !at(piv, c)) ++piv;
    if (piv == T) return false;
    if (piv != c)
      for (int j = 0; j <= T; ++j) std::swap(at(piv, j), at(c, j));
    const uint16_t inv = F.inv(at(c, c));
    for (int j = c; j <= T; ++j) at(c, j) = F.mul(at(c, j), inv);
    for (int r = 0; r < T; ++r) {
      if (r == c || !at(r, c)) continue;
      const uint16_t fac = at(r, c);
      for (int j = c; j <= T; ++j) at(r, j) ^= F.mul(fac, at(c, j));
    }
  }
  for (int i = 0; i < T; ++i) g[i] = at(i, T);
  return true;
}

// The f variants' column step (LM:crypto_kem/<set>f/vec/pk_gen.c mov_columns :87-153): at column
// c0 = mt - 32 the rows c0..mt-1 are zero left of c0; the pivot columns pv[] of their 32 x 64
// window (columns c0..c0+63) in row-echelon form -- which depend only on the rows' span -- give the
// pivot mask and the swaps c0+i <-> c0+pv[i], i = 0..31 in order, of the matrix's columns.
// Returns false when the window has rank < 32.
template <int SET>
static bool f_column_step(std::vector<uint64_t> &H, uint64_t *pivots, int pv[32]) {
  using P = Par<SET>;
  constexpr int W = ceil_div(P::N, 64), C0 = P::MT - 32, W0 = C0 / 64, SH = C0 % 64;
  auto window = [&](int r) -> uint64_t {
    const uint64_t *row = &H[(size_t)r * W];
    return SH ? (row[W0] >> SH) | (row[W0 + 1] << (64 - SH)) : row[W0];
  };
  uint64_t b[32];
  for (int i = 0; i < 32; ++i) b[i] = window(C0 + i);
  *pivots = 0;
  for (int i = 0; i < 32; ++i) {
    int best = 64, who = -1;
    for (int j = i; j < 32; ++j)
      if (b[j] && __builtin_ctzll(b[j]) < best) { best = __builtin_ctzll(b[j]); who = j; }
    if (who < 0) return false;
    std::swap(b[i], b[who]);
    for (int j = i + 1; j < 32; ++j)
      if ((b[j] >> best) & 1) b[j] ^= b[i];
    pv[i] = best;
    *pivots |= 1ull << best;
  }
  for (int r = 0; r < P::MT; ++r) {
    uint64_t t = window(r);
    for (int i = 0; i < 32; ++i) {
      const uint64_t d = ((t >> i) ^ (t >> pv[i])) & 1;
      t ^= (d << i) | (d << pv[i]);
    }
    uint64_t *row = &H[(size_t)r * W];
    if (SH) {
      row[W0] = (row[W0] & ((1ull << SH) - 1)) | (t << SH);
      row[W0 + 1] = (row[W0 + 1] & ~((1ull << SH) - 1)) | (t >> (64 - SH));
    } else {
      row[W0] = t;
    }
  }
  return true;
}

// Systematic form of the parity-check matrix; false when the left mt x mt block is singular.  With
// `pivots` (the f variants) the column step above runs at column mt - 32 and sets *pivots.
template <int SET>
static bool systematic_pk(const HostField &F, const uint16_t *g, const uint16_t *alpha, uint8_t *pk,
                          uint64_t *pivots = nullptr) {
  using P = Par<SET>;
  constexpr int W = ceil_div(P::N, 64);
  std::vector<uint64_t> H((size_t)P::MT * W, 0);
  for (int j = 0; j < P::N; ++j) {
    uint16_t v = g[P::T];
    for (int k = P::T - 1; k >= 0; --k) v = (uint16_t)(F.mul(v, alpha[j]) ^ g[k]);
    v = F.inv(v);
    for (int i = 0; i < P::T; ++i) {
      for (int b = 0; b < P::M; ++b)
        if ((v >> b) & 1) H[(size_t)(i * P::M + b) * W + j / 64] |= 1ull << (j % 64);
      v = F.mul(v, alpha[j]);
    }
  }
  for (int c = 0; c < P::MT; ++c) {
    if (pivots && c == P::MT - 32) {
      int pv[32];
      if (!f_column_step<SET>(H, pivots, pv)) return false;
    }
    const int cw = c / 64;
    const uint64_t cm = 1ull << (c % 64);
    int piv = c;
    while (piv < P::MT && !(H[(size_t)piv * W + cw] & cm)) ++piv;
    if (piv == P::MT) return false;
    uint64_t *rc = &H[(size_t)c * W];
    if (piv != c) std::swap_ranges(rc + cw, rc + W, &H[(size_t)piv * W + cw]);
    for (int r = 0; r < P::MT; ++r) {
      uint64_t *rr = &H[(size_t)r * W];
      if (r != c && (rr[cw] & cm))
        for (int w = cw; w < W; ++w) rr[w] ^= rc[w];
    }
  }
  std::memset(pk, 0, (size_t)P::PK_BYTES);
  for (int r = 0; r < P::MT; ++r)
    for (int j = 0; j < P::K; ++j) {
      const int col = P::MT + j;
      if ((H[(size_t)r * W + col / 64] >> (col % 64)) & 1) pk[(size_t)r * P::PK_ROW_BYTES + j / 8] |= (uint8_t)(1u << (j % 8));
    }
  return true;
}

// LM:crypto_kem/<set>/vec/kem_keypair.c, split for the batched key generator: one attempt per
// seed of the chain (seed expansion by SHAKE256, Goppa polynomial, random support by sorting
// 32-bit keys) up to the systematic-form step, and the secret key once that step has succeeded
// (Benes control bits, rejection string).
// The support permutation from 2^m random 32-bit keys (little-endian at pp): pi lists the indices
// in ascending key order, as the sort of (key, index) pairs in LM:crypto_kem/<set>/vec/pk_gen.c
// does; false when two keys are equal (the attempt fails).  LSD radix sort, four 8-bit digits.
static bool support_perm(const uint8_t *pp, int nq, uint16_t *pi) {
  thread_local std::vector<uint64_t> a, t;
  a.resize(nq);
  t.resize(nq);
  for (int i = 0; i < nq; ++i) {
    uint32_t k;
    std::memcpy(&k, pp + 4 * i, 4);
    a[i] = ((uint64_t)k << 16) | (uint64_t)i;
  }
  for (int sh = 16; sh < 48; sh += 8) {
    uint32_t cnt[256] = {0};
    for (int i = 0; i < nq; ++i) ++cnt[(a[i] >> sh) & 255];
    uint32_t sum = 0;
    for (int d = 0; d < 256; ++d) {
      const uint32_t c = cnt[d];
      cnt[d] = sum;
      sum += c;
    }
    for (int i = 0; i < nq; ++i) t[cnt[(a[i] >> sh) & 255]++] = a[i];
    a.swap(t);
  }
  for (int i = 1; i < nq; ++i)
    if ((a[i - 1] >> 16) == (a[i] >> 16)) return false;
  for (int i = 0; i < nq; ++i) pi[i] = (uint16_t)a[i];
  return true;
}

template <int SET>
static void candidate_t(uint8_t seed33[33], KeygenCandidate &c) {
  using P = Par<SET>;
  const HostField &F = host_field(P::M);
  constexpr int NQ = 1 << P::M;
  const size_t rlen = P::N / 8 + 4 * (size_t)NQ + 2 * P::T + 32;
  std::vector<uint8_t> r(rlen);
  std::vector<uint16_t> f(P::T);
  c.g.assign(P::T + 1, 0);
  c.alpha.assign(P::N, 0);
  c.pi.assign(NQ, 0);
  for (;;) {
    shake256(r.data(), (long long)rlen, seed33, 33);
    std::memcpy(c.sk_seed, seed33 + 1, 32);
    std::memcpy(seed33 + 1, &r[rlen - 32], 32);  // the next attempt's seed, whatever happens below
    const uint8_t *fp = &r[rlen - 32 - 2 * P::T];
    for (int i = 0; i < P::T; ++i) f[i] = (uint16_t)((fp[2 * i] | (fp[2 * i + 1] << 8)) & ((1 << P::M) - 1));
    if (!min_poly<SET>(F, f.data(), c.g.data())) continue;
    c.g[P::T] = 1;
    if (!support_perm(&r[P::N / 8], NQ, c.pi.data())) continue;
    for (int i = 0; i < P::N; ++i) c.alpha[i] = bitrev(c.pi[i], P::M);
    c.s.assign(r.begin(), r.begin() + P::N / 8);
    return;
  }
}

template <int SET>
static bool step_t(uint8_t seed33[33], KeygenCandidate &c) {
  using P = Par<SET>;
  constexpr int NQ = 1 << P::M;
  const size_t rlen = P::N / 8 + 4 * (size_t)NQ + 2 * P::T + 32;
  thread_local std::vector<uint8_t> r;
  r.resize(rlen);
  c.f.assign(P::T, 0);
  c.g.assign(P::T + 1, 0);
  c.alpha.assign(P::N, 0);
  c.pi.assign(NQ, 0);
  shake256(r.data(), (long long)rlen, seed33, 33);
  std::memcpy(c.sk_seed, seed33 + 1, 32);
  std::memcpy(seed33 + 1, &r[rlen - 32], 32);  // the next attempt's seed, whatever happens below
  const uint8_t *fp = &r[rlen - 32 - 2 * P::T];
  for (int i = 0; i < P::T; ++i) c.f[i] = (uint16_t)((fp[2 * i] | (fp[2 * i + 1] << 8)) & ((1 << P::M) - 1));
  if (!support_perm(&r[P::N / 8], NQ, c.pi.data())) return false;
  for (int i = 0; i < P::N; ++i) c.alpha[i] = bitrev(c.pi[i], P::M);
  c.s.assign(r.begin(), r.begin() + P::N / 8);
  return true;
}

bool keygen_step(int set, uint8_t seed33[33], KeygenCandidate &c) {
  switch (set) {
    case S348864: return step_t<S348864>(seed33, c);
    case S460896: return step_t<S460896>(seed33, c);
    case S6688128: return step_t<S6688128>(seed33, c);
    case S6960119: return step_t<S6960119>(seed33, c);
    case S8192128: return step_t<S8192128>(seed33, c);
  }
  throw std::invalid_argument("set");
}

template <int SET>
static void apply_pivots_t(KeygenCandidate &c, uint64_t pivots) {
  using P = Par<SET>;
  constexpr int C0 = P::MT - 32;
  int i = 0;
  for (int b = 0; b < 64; ++b)
    if ((pivots >> b) & 1) {
      if (b != i) {
        std::swap(c.pi[C0 + i], c.pi[C0 + b]);
        std::swap(c.alpha[C0 + i], c.alpha[C0 + b]);
      }
      ++i;
    }
  c.pivots = pivots;
}

// sk = seed || pivots || g || control bits || s; cb == null: the control bits are computed here
template <int SET>
static void finish_t(const KeygenCandidate &c, const uint8_t *cb, uint8_t *sk) {
  using P = Par<SET>;
  std::memcpy(sk, c.sk_seed, 32);
  for (int i = 0; i < 8; ++i) sk[32 + i] = (uint8_t)(c.pivots >> (8 * i));  // little-endian
  uint8_t *skp = sk + 40;
  for (int i = 0; i < P::T; ++i) { skp[2 * i] = (uint8_t)c.g[i]; skp[2 * i + 1] = (uint8_t)(c.g[i] >> 8); }
  if (cb) std::memcpy(skp + P::IRR_BYTES, cb, P::COND_BYTES);
  else controlbits_from_perm(P::M, c.pi, skp + P::IRR_BYTES);
  std::memcpy(skp + P::IRR_BYTES + P::COND_BYTES, c.s.data(), P::N / 8);
}

template <int SET>
static void keypair_t(uint8_t *pk, uint8_t *sk, rand_fn rnd, void *ctx, bool f) {
  using P = Par<SET>;
  const HostField &F = host_field(P::M);
  uint8_t seed[33];
  seed[0] = 64;
  rnd(seed + 1, 32, ctx);
  KeygenCandidate c;
  uint64_t pivots = 0;
  for (;;) {
    candidate_t<SET>(seed, c);
    if (systematic_pk<SET>(F, c.g.data(), c.alpha.data(), pk, f ? &pivots : nullptr)) break;
  }
  if (f) apply_pivots_t<SET>(c, pivots);
  finish_t<SET>(c, nullptr, sk);
}

void keygen_candidate(int set, uint8_t seed33[33], KeygenCandidate &c) {
  switch (set) {
    case S348864: return candidate_t<S348864>(seed33, c);
    case S460896: return candidate_t<S460896>(seed33, c);
    case S6688128: return candidate_t<S6688128>(seed33, c);
    case S6960119: return candidate_t<S6960119>(seed33, c);
    case S8192128: return candidate_t<S8192128>(seed33, c);
  }
  throw std::invalid_argument("set");
}

void keygen_finish_with_cb(int set, const KeygenCandidate &c, const uint8_t *cb, uint8_t *sk) {
  switch (set) {
    case S348864: return finish_t<S348864>(c, cb, sk);
    case S460896: return finish_t<S460896>(c, cb, sk);
    case S6688128: return finish_t<S6688128>(c, cb, sk);
    case S6960119: return finish_t<S6960119>(c, cb, sk);
    case S8192128: return finish_t<S8192128>(c, cb, sk);
  }
  throw std::invalid_argument("set");
}

void keygen_finish(int set, const KeygenCandidate &c, uint8_t *sk) { keygen_finish_with_cb(set, c, nullptr, sk); }

void host_gen_e(int set, rand_fn rnd, void *ctx, uint16_t *pos) {
  const SetInfo &si = set_info(set);
  const int t = si.t, ndraw = set == S8192128 ? t : 2 * t;
  const uint16_t mask = (uint16_t)((1u << si.m) - 1u);
  std::vector<uint8_t> buf(2 * (size_t)ndraw);
  for (;;) {
    rnd(buf.data(), (long long)buf.size(), ctx);
    int count = 0;
    for (int i = 0; i < ndraw && count < t; ++i) {
      const uint16_t v = (uint16_t)((buf[2 * i] | (buf[2 * i + 1] << 8)) & mask);
      if (v < si.n) pos[count++] = v;
    }
    if (count < t) continue;
    std::sort(pos, pos + t);
    if (std::adjacent_find(pos, pos + t) == pos + t) return;
  }
}

void keygen_apply_pivots(int set, KeygenCandidate &c, uint64_t pivots) {
  switch (set) {
    case S348864: return apply_pivots_t<S348864>(c, pivots);
    case S460896: return apply_pivots_t<S460896>(c, pivots);
    case S6688128: return apply_pivots_t<S6688128>(c, pivots);
    case S6960119: return apply_pivots_t<S6960119>(c, pivots);
    case S8192128: return apply_pivots_t<S8192128>(c, pivots);
  }
  throw std::invalid_argument("set");
}

void host_keypair(int variant, uint8_t *pk, uint8_t *sk, rand_fn rnd, void *ctx) {
  const bool f = variant_f(variant);
  switch (variant_base(variant)) {
    case S348864: return keypair_t<S348864>(pk, sk, rnd, ctx, f);
    case S460896: return keypair_t<S460896>(pk, sk, rnd, ctx, f);
    case S6688128: return keypair_t<S6688128>(pk, sk, rnd, ctx, f);
    case S6960119: return keypair_t<S6960119>(pk, sk, rnd, ctx, f);
    case S8192128: return keypair_t<S8192128>(pk, sk, rnd, ctx, f);
  }
  throw std::invalid_argument("set");
}

}  // namespace tmce
