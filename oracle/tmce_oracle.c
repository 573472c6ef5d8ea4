/*
 * tmce_oracle.c -- CPU oracle for the Classic McEliece KEM, libmceliece-20250507 semantics.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline
 * as the checker; never linked into or called by the shipped library.
 *
 * This is a straightforward restatement of the algorithm (the "ref" formulation of Classic
 * McEliece: support + Goppa polynomial, syndrome by direct evaluation, Berlekamp-Massey, root
 * search by Horner evaluation at every support element), written to reproduce libmceliece's
 * byte-level behaviour.  Citations refer to files inside deps/libmceliece-20250507.tar.gz of the
 * reference checkout ("LM:").  Pinned by the libmceliece KAT checksums (oracle_kat_checksum).
 */
#define _GNU_SOURCE
#include "tmce_oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <sched.h>

/* ------------------------------------------------------------------ parameters */
/* LM:crypto_kem/<set>/vec/params.h, gf_params.h; api sizes LM:api */
#define P(name, m, n, t)                                                                       \
  { name, m, n, t, (m) * (t), (n) - (m) * (t), ((n) - (m) * (t) + 7) / 8, ((m) * (t) + 7) / 8, \
    2 * (t), (1 << ((m) - 4)) * (2 * (m) - 1), (long)(m) * (t) * (((n) - (m) * (t) + 7) / 8),  \
    40L + 2 * (t) + (1 << ((m) - 4)) * (2 * (m) - 1) + (n) / 8 }

static const orc_params PARAMS[5] = {
  P("348864", 12, 3488, 64),
  P("460896", 13, 4608, 96),
  P("6688128", 13, 6688, 128),
  P("6960119", 13, 6960, 119),
  P("8192128", 13, 8192, 128),
};

/* variant ids: base set (0..4) | ORC_F (semi-systematic key generation) | ORC_PC (plaintext
   confirmation); LM:crypto_kem/<set>{,f,pc,pcf} */
static int valid_variant(int v) { return v >= 0 && (v & 7) < 5 && (v & ~(7 | ORC_F | ORC_PC)) == 0; }
const orc_params *oracle_params(int set) { return valid_variant(set) ? &PARAMS[set & 7] : NULL; }
int oracle_ct_bytes(int set) { return valid_variant(set) ? PARAMS[set & 7].synd_bytes + ((set & ORC_PC) ? 32 : 0) : -1; }

/* ------------------------------------------------------------------ GF(2^m) */
/* field polynomials: m=12: z^12+z^3+1; m=13: z^13+z^4+z^3+z+1 (LM:crypto_kem/<set>/vec/gf.c) */
static uint16_t gf_mul_slow(int m, uint16_t a, uint16_t b) {
  uint32_t r = 0;
  for (int i = 0; i < m; i++)
    if ((b >> i) & 1) r ^= (uint32_t)a << i;
  uint32_t poly = (m == 12) ? ((1u << 12) | (1u << 3) | 1u)
                            : ((1u << 13) | (1u << 4) | (1u << 3) | (1u << 1) | 1u);
  for (int i = 2 * m - 2; i >= m; i--)
    if ((r >> i) & 1) r ^= poly << (i - m);
  return (uint16_t)r;
}

typedef struct {
  int ready;
  uint16_t exp[2 * 8192];
  uint16_t log[8192];
} gf_tab;
static gf_tab TAB[2]; /* [0]: m=12, [1]: m=13 */
static pthread_once_t tab_once = PTHREAD_ONCE_INIT;

static void tab_init(void) {
  for (int w = 0; w < 2; w++) {
    int m = 12 + w, q = (1 << m) - 1;
    /* find a primitive element by search */
    for (uint16_t g = 2;; g++) {
      uint16_t x = 1;
      int ord = 0;
      do {
        x = gf_mul_slow(m, x, g);
        ord++;
      } while (x != 1 && ord <= q);
      if (ord == q) {
        x = 1;
        for (int i = 0; i < 2 * q; i++) {
          TAB[w].exp[i] = x;
          if (i < q) TAB[w].log[x] = (uint16_t)i;
          x = gf_mul_slow(m, x, g);
        }
        TAB[w].log[0] = 0;
        break;
      }
    }
    TAB[w].ready = 1;
  }
}

static inline const gf_tab *tab(int m) {
  pthread_once(&tab_once, tab_init);
  return &TAB[m - 12];
}

static inline uint16_t gmul(const gf_tab *T, uint16_t a, uint16_t b) {
  if (!a || !b) return 0;
  return T->exp[T->log[a] + T->log[b]];
}
static inline uint16_t ginv(const gf_tab *T, int m, uint16_t a) {
  if (!a) return 0; /* libmceliece gf_inv(0) = 0^(2^m-2) = 0 */
  int q = (1 << m) - 1;
  return T->exp[(q - T->log[a]) % q];
}

uint16_t oracle_gf_mul(int m, uint16_t a, uint16_t b) { return gmul(tab(m), a, b); }
uint16_t oracle_gf_inv(int m, uint16_t a) { return ginv(tab(m), m, a); }

static inline uint16_t load_gf(const uint8_t *p, int m) {
  return (uint16_t)((p[0] | (p[1] << 8)) & ((1 << m) - 1));
}
static inline uint16_t bitrev_m(uint16_t a, int m) {
  uint16_t r = 0;
  for (int i = 0; i < m; i++) r |= (uint16_t)(((a >> i) & 1) << (m - 1 - i));
  return r;
}

/* ------------------------------------------------------------------ Keccak / SHAKE256 */
static const uint64_t KRC[24] = {
  0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
  0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
  0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
  0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
  0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
  0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

static inline uint64_t rol64(uint64_t x, int r) { return r ? (x << r) | (x >> (64 - r)) : x; }

static void keccak_f(uint64_t A[25]) {
  /* rho offsets r[x][y] for lane (x,y) at index x+5y */
  static const int R[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43,
                            25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
  for (int round = 0; round < 24; round++) {
    uint64_t C[5], D[5], B[25];
    for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
    for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rol64(C[(x + 1) % 5], 1);
    for (int i = 0; i < 25; i++) A[i] ^= D[i % 5];
    /* rho + pi: B[y, 2x+3y] = rot(A[x,y], r[x,y]) */
    for (int x = 0; x < 5; x++)
      for (int y = 0; y < 5; y++) B[y + 5 * ((2 * x + 3 * y) % 5)] = rol64(A[x + 5 * y], R[x + 5 * y]);
    for (int x = 0; x < 5; x++)
      for (int y = 0; y < 5; y++)
        A[x + 5 * y] = B[x + 5 * y] ^ ((~B[(x + 1) % 5 + 5 * y]) & B[(x + 2) % 5 + 5 * y]);
    A[0] ^= KRC[round];
  }
}

void oracle_shake256(uint8_t *out, long long outlen, const uint8_t *in, long long inlen) {
  const int rate = 136;
  uint64_t A[25];
  uint8_t blk[136];
  memset(A, 0, sizeof A);
  while (inlen >= rate) {
    for (int i = 0; i < rate / 8; i++) {
      uint64_t w = 0;
      for (int b = 0; b < 8; b++) w |= (uint64_t)in[8 * i + b] << (8 * b);
      A[i] ^= w;
    }
    keccak_f(A);
    in += rate;
    inlen -= rate;
  }
  memset(blk, 0, sizeof blk);
  memcpy(blk, in, (size_t)inlen);
  blk[inlen] ^= 0x1F;
  blk[rate - 1] ^= 0x80;
  for (int i = 0; i < rate / 8; i++) {
    uint64_t w = 0;
    for (int b = 0; b < 8; b++) w |= (uint64_t)blk[8 * i + b] << (8 * b);
    A[i] ^= w;
  }
  while (outlen > 0) {
    keccak_f(A);
    long long take = outlen < rate ? outlen : rate;
    for (long long i = 0; i < take; i++) out[i] = (uint8_t)(A[i / 8] >> (8 * (i % 8)));
    out += take;
    outlen -= take;
  }
}

/* ------------------------------------------------------------------ randomness */
/* seeded adapter (the per-op chain the reference harness uses in its GPU gen_e,
   mceliece-gpu-params/ref/348864.cuh:3434-3486): state = 0x40||seed, out = SHAKE256(state,len),
   then state[1..32] = out[0..31]. */
enum { RNG_SEEDED = 0, RNG_KNOWN = 1 };
static __thread int rng_mode = RNG_SEEDED;
static __thread uint8_t rng_state[33] = {0x40};

void oracle_rng_seeded(const uint8_t seed[32]) {
  rng_mode = RNG_SEEDED;
  rng_state[0] = 0x40;
  memcpy(rng_state + 1, seed, 32);
}

/* knownrandombytes of LM:command/mceliece-test.c:257-357 (ChaCha20 key-erasing RNG) */
static inline uint32_t ld32le(const uint8_t *x) {
  return (uint32_t)x[0] | ((uint32_t)x[1] << 8) | ((uint32_t)x[2] << 16) | ((uint32_t)x[3] << 24);
}
static inline void st32le(uint8_t *x, uint32_t u) {
  for (int i = 0; i < 4; i++) x[i] = (uint8_t)(u >> (8 * i));
}
static inline uint32_t rotl32(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }
static const uint8_t SIGMA[16] = {'e', 'x', 'p', 'a', 'n', 'd', ' ', '3', '2', '-', 'b', 'y', 't', 'e', ' ', 'k'};

static void chacha_block(uint8_t out[64], const uint8_t in[16], const uint8_t k[32]) {
  uint32_t x[16], y[16];
  for (int i = 0; i < 4; i++) {
    x[i] = ld32le(SIGMA + 4 * i);
    x[12 + i] = ld32le(in + 4 * i);
  }
  for (int i = 0; i < 8; i++) x[4 + i] = ld32le(k + 4 * i);
  memcpy(y, x, sizeof x);
#define QR(a, b, c, d)                                                                             \
  a += b; d = rotl32(d ^ a, 16); c += d; b = rotl32(b ^ c, 12);                                    \
  a += b; d = rotl32(d ^ a, 8);  c += d; b = rotl32(b ^ c, 7);
  for (int r = 0; r < 10; r++) {
    for (int j = 0; j < 4; j++) { QR(x[j], x[j + 4], x[j + 8], x[j + 12]) }
    for (int j = 0; j < 4; j++) { QR(x[j], x[((j + 1) & 3) + 4], x[((j + 2) & 3) + 8], x[((j + 3) & 3) + 12]) }
  }
#undef QR
  for (int i = 0; i < 16; i++) st32le(out + 4 * i, x[i] + y[i]);
}

static void chacha20_stream(uint8_t *c, uint64_t len, const uint8_t nonce[8], const uint8_t k[32]) {
  uint8_t z[16] = {0}, blk[64];
  memcpy(z + 8, nonce, 8);
  while (len > 0) {
    chacha_block(blk, z, k);
    uint64_t take = len < 64 ? len : 64;
    memcpy(c, blk, take);
    c += take;
    len -= take;
    uint32_t u = 1;
    for (int i = 0; i < 8; i++) {
      u += z[i];
      z[i] = (uint8_t)u;
      u >>= 8;
    }
  }
}

#define KRB_OUT 736
static __thread uint8_t krb_key[32];
static __thread uint8_t krb_buf[KRB_OUT];
static __thread long krb_pos = KRB_OUT;

void oracle_rng_known_reset(void) {
  rng_mode = RNG_KNOWN;
  memset(krb_key, 0, sizeof krb_key);
  memset(krb_buf, 0, sizeof krb_buf);
  krb_pos = KRB_OUT;
}

static void known_randombytes(uint8_t *x, long long len) {
  static const uint8_t nonce[8] = {0};
  while (len > 0) {
    if (krb_pos == KRB_OUT) {
      uint8_t tmp[32 + KRB_OUT];
      chacha20_stream(tmp, sizeof tmp, nonce, krb_key);
      memcpy(krb_key, tmp, 32);
      memcpy(krb_buf, tmp + 32, KRB_OUT);
      krb_pos = 0;
    }
    *x++ = krb_buf[krb_pos];
    krb_buf[krb_pos++] = 0;
    len--;
  }
}

void oracle_randombytes(uint8_t *out, long long len) {
  if (rng_mode == RNG_KNOWN) {
    known_randombytes(out, len);
    return;
  }
  oracle_shake256(out, len, rng_state, 33);
  long long c = len < 32 ? len : 32;
  memcpy(rng_state + 1, out, (size_t)c);
}

/* ------------------------------------------------------------------ polynomials over GF(2^m) */
static uint16_t poly_eval(const gf_tab *T, const uint16_t *f, int deg, uint16_t x) {
  uint16_t r = f[deg];
  for (int i = deg - 1; i >= 0; i--) r = gmul(T, r, x) ^ f[i];
  return r;
}

/* minimal polynomial of f in GF((2^m)^t) = GF(2^m)[y]/F(y); LM:crypto_kem/<set>/vec/sk_gen.c
   (genpoly_gen) and gf_2mt_mul.c (F).  Returns 0 and irr[0..t) on success, -1 if singular. */
static void ext_mul(int set, const gf_tab *T, uint16_t *out, const uint16_t *a, const uint16_t *b) {
  const int t = PARAMS[set & 7].t;
  uint16_t prod[256];
  memset(prod, 0, sizeof(prod[0]) * (2 * t - 1));
  for (int i = 0; i < t; i++)
    for (int j = 0; j < t; j++) prod[i + j] ^= gmul(T, a[i], b[j]);
  for (int i = 2 * t - 2; i >= t; i--) {
    uint16_t c = prod[i];
    switch (set & 7) {
      case ORC_348864: /* y^64 = y^3 + y + z */
        prod[i - t + 3] ^= c; prod[i - t + 1] ^= c; prod[i - t] ^= gmul(T, c, 2); break;
      case ORC_460896: /* y^96 = y^10 + y^9 + y^6 + 1 */
        prod[i - t + 10] ^= c; prod[i - t + 9] ^= c; prod[i - t + 6] ^= c; prod[i - t] ^= c; break;
      case ORC_6960119: /* y^119 = y^8 + 1 */
        prod[i - t + 8] ^= c; prod[i - t] ^= c; break;
      default: /* 6688128, 8192128: y^128 = y^7 + y^2 + y + 1 */
        prod[i - t + 7] ^= c; prod[i - t + 2] ^= c; prod[i - t + 1] ^= c; prod[i - t] ^= c; break;
    }
  }
  memcpy(out, prod, sizeof(uint16_t) * t);
}

static int genpoly(int set, uint16_t *irr, const uint16_t *f) {
  const int t = PARAMS[set & 7].t, m = PARAMS[set & 7].m;
  const gf_tab *T = tab(m);
  /* pw[j] = f^j, j = 0..t; solve sum_{j<t} x_j f^j = f^t (coefficient-wise linear system) */
  uint16_t *pw = (uint16_t *)calloc((size_t)(t + 1) * t, sizeof(uint16_t));
  pw[0] = 1;
  memcpy(pw + t, f, sizeof(uint16_t) * t);
  for (int j = 2; j <= t; j++) ext_mul(set, T, pw + (size_t)j * t, pw + (size_t)(j - 1) * t, f);
  /* augmented matrix M[row=coef i][col=j], col t = rhs */
  uint16_t *M = (uint16_t *)malloc(sizeof(uint16_t) * t * (t + 1));
  for (int i = 0; i < t; i++)
    for (int j = 0; j <= t; j++) M[i * (t + 1) + j] = pw[(size_t)j * t + i];
  int ok = 0;
  for (int c = 0; c < t; c++) {
    int piv = -1;
    for (int r = c; r < t; r++)
      if (M[r * (t + 1) + c]) { piv = r; break; }
    if (piv < 0) { ok = -1; break; }
    if (piv != c)
      for (int j = 0; j <= t; j++) {
        uint16_t tmp = M[c * (t + 1) + j];
        M[c * (t + 1) + j] = M[piv * (t + 1) + j];
        M[piv * (t + 1) + j] = tmp;
      }
    uint16_t inv = ginv(T, m, M[c * (t + 1) + c]);
    for (int j = 0; j <= t; j++) M[c * (t + 1) + j] = gmul(T, M[c * (t + 1) + j], inv);
    for (int r = 0; r < t; r++) {
      if (r == c) continue;
      uint16_t fac = M[r * (t + 1) + c];
      if (!fac) continue;
      for (int j = 0; j <= t; j++) M[r * (t + 1) + j] ^= gmul(T, fac, M[c * (t + 1) + j]);
    }
  }
  if (!ok)
    for (int i = 0; i < t; i++) irr[i] = M[i * (t + 1) + t];
  free(M);
  free(pw);
  return ok;
}

/* ------------------------------------------------------------------ control bits */
/* Benes network control bits of a permutation of 2^w elements: the recursion of
   LM:crypto_kem/<set>/vec/controlbits.c (cbrecursion, Nassimi-Sahni / "controlbits-20200923"),
   restated with 32-bit unsigned keys (all keys are non-negative and distinct, so any correct
   sort gives the same order). */
static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}
static void sort_u32(uint32_t *x, long n) { qsort(x, (size_t)n, sizeof(uint32_t), cmp_u32); }
static inline uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

static void put_bit(uint8_t *out, long long pos, uint32_t b) { out[pos >> 3] ^= (uint8_t)((b & 1) << (pos & 7)); }

static void cb_rec(uint8_t *out, long long pos, long long step, const int16_t *pi, int w, long n) {
  if (w == 1) {
    put_bit(out, pos, (uint32_t)pi[0]);
    return;
  }
  uint32_t *A = (uint32_t *)malloc(sizeof(uint32_t) * n);
  uint32_t *B = (uint32_t *)malloc(sizeof(uint32_t) * n);
  for (long x = 0; x < n; x++) A[x] = (((uint32_t)pi[x] ^ 1u) << 16) | (uint32_t)(uint16_t)pi[x ^ 1];
  sort_u32(A, n);
  for (long x = 0; x < n; x++) {
    uint32_t px = A[x] & 0xffff;
    B[x] = (px << 16) | umin(px, (uint32_t)x);
  }
  for (long x = 0; x < n; x++) A[x] = (A[x] << 16) | (uint32_t)x;
  sort_u32(A, n);
  for (long x = 0; x < n; x++) A[x] = (A[x] << 16) + (B[x] >> 16);
  sort_u32(A, n);
  if (w <= 10) {
    for (long x = 0; x < n; x++) B[x] = ((A[x] & 0xffff) << 10) | (B[x] & 0x3ff);
    for (int i = 1; i < w - 1; i++) {
      for (long x = 0; x < n; x++) A[x] = ((B[x] & ~0x3ffu) << 6) | (uint32_t)x;
      sort_u32(A, n);
      for (long x = 0; x < n; x++) A[x] = (A[x] << 20) | B[x];
      sort_u32(A, n);
      for (long x = 0; x < n; x++) {
        uint32_t ppcp = A[x] & 0xfffff;
        uint32_t ppc = (A[x] & 0xffc00) | (B[x] & 0x3ff);
        B[x] = umin(ppc, ppcp);
      }
    }
    for (long x = 0; x < n; x++) B[x] &= 0x3ff;
  } else {
    for (long x = 0; x < n; x++) B[x] = (A[x] << 16) | (B[x] & 0xffff);
    for (int i = 1; i < w - 1; i++) {
      for (long x = 0; x < n; x++) A[x] = (B[x] & ~0xffffu) | (uint32_t)x;
      sort_u32(A, n);
      for (long x = 0; x < n; x++) A[x] = (A[x] << 16) | (B[x] & 0xffff);
      if (i < w - 2) {
        for (long x = 0; x < n; x++) B[x] = (A[x] & ~0xffffu) | (B[x] >> 16);
        sort_u32(B, n);
        for (long x = 0; x < n; x++) B[x] = (B[x] << 16) | (A[x] & 0xffff);
      }
      sort_u32(A, n);
      for (long x = 0; x < n; x++) {
        uint32_t cp = (B[x] & ~0xffffu) | (A[x] & 0xffff);
        B[x] = umin(B[x], cp);
      }
    }
    for (long x = 0; x < n; x++) B[x] &= 0xffff;
  }
  for (long x = 0; x < n; x++) A[x] = ((uint32_t)(uint16_t)pi[x] << 16) + (uint32_t)x;
  sort_u32(A, n);
  for (long j = 0; j < n / 2; j++) {
    long z = 2 * j;
    uint32_t fj = B[z] & 1;
    uint32_t Fz = (uint32_t)z + fj, Fz1 = Fz ^ 1;
    put_bit(out, pos, fj);
    pos += step;
    B[z] = (A[z] << 16) | Fz;
    B[z + 1] = (A[z + 1] << 16) | Fz1;
  }
  sort_u32(B, n);
  pos += (2 * w - 3) * step * (n / 2);
  for (long k = 0; k < n / 2; k++) {
    long y = 2 * k;
    uint32_t lk = B[y] & 1;
    uint32_t Ly = (uint32_t)y + lk, Ly1 = Ly ^ 1;
    put_bit(out, pos, lk);
    pos += step;
    A[y] = (Ly << 16) | (B[y] & 0xffff);
    A[y + 1] = (Ly1 << 16) | (B[y + 1] & 0xffff);
  }
  sort_u32(A, n);
  pos -= (2 * w - 2) * step * (n / 2);
  int16_t *q = (int16_t *)malloc(sizeof(int16_t) * n);
  for (long j = 0; j < n / 2; j++) {
    q[j] = (int16_t)((A[2 * j] & 0xffff) >> 1);
    q[j + n / 2] = (int16_t)((A[2 * j + 1] & 0xffff) >> 1);
  }
  free(A);
  free(B);
  cb_rec(out, pos, step * 2, q, w - 1, n / 2);
  cb_rec(out, pos + step, step * 2, q + n / 2, w - 1, n / 2);
  free(q);
}

void oracle_controlbits(int m, const int16_t *pi, uint8_t *cb) {
  long n = 1L << m;
  memset(cb, 0, (size_t)(((2 * m - 1) * n / 2 + 7) / 8));
  cb_rec(cb, 0, 1, pi, m, n);
}

/* Benes layers in the order LM:crypto_kem/<set>/vec/controlbits.c verifies them
   (stride 2^0 .. 2^(m-1), then 2^(m-2) .. 2^0; n/2 control bits per layer) */
void oracle_perm_from_controlbits(int m, const uint8_t *cb, int16_t *pi) {
  long n = 1L << m;
  for (long i = 0; i < n; i++) pi[i] = (int16_t)i;
  const uint8_t *ptr = cb;
  for (int L = 0; L < 2 * m - 1; L++) {
    int s = L < m ? L : 2 * m - 2 - L;
    long stride = 1L << s, idx = 0;
    for (long i = 0; i < n; i += 2 * stride)
      for (long j = 0; j < stride; j++) {
        if ((ptr[idx >> 3] >> (idx & 7)) & 1) {
          int16_t tmp = pi[i + j];
          pi[i + j] = pi[i + j + stride];
          pi[i + j + stride] = tmp;
        }
        idx++;
      }
    ptr += n >> 4;
  }
}

/* ------------------------------------------------------------------ public key */
static inline int mat_bit(const uint64_t *mat, int W, int r, int c) { return (int)((mat[(size_t)r * W + c / 64] >> (c % 64)) & 1); }
static inline void mat_flip(uint64_t *mat, int W, int r, int c) { mat[(size_t)r * W + c / 64] ^= 1ULL << (c % 64); }

/* The "f" variants' column step (LM:crypto_kem/<set>f/vec/pk_gen.c mov_columns :87-153, called at
   row mt-32 :251-255): when the elimination reaches column c0 = mt-32, rows c0..mt-1 are zero in
   columns < c0.  The pivot columns of their 32 x 64 window (columns c0..c0+63) in row-echelon
   form -- pv[i] = lowest column where any of the still-unused rows is nonzero -- give `pivots`
   (bit pv[i] set) and the column swaps (c0+i <-> c0+pv[i], i = 0..31 in order) applied to the
   matrix and to the support permutation pi.  -1 if the window has rank < 32. */
static int f_move_columns(uint64_t *mat, int W, int mt, int16_t *pi, uint64_t *pivots) {
  const int c0 = mt - 32;
  uint64_t buf[32];
  int pv[32];
  for (int i = 0; i < 32; i++) {
    buf[i] = 0;
    for (int c = 0; c < 64; c++) buf[i] |= (uint64_t)mat_bit(mat, W, c0 + i, c0 + c) << c;
  }
  *pivots = 0;
  for (int i = 0; i < 32; i++) {
    int best = 64, who = -1;
    for (int j = i; j < 32; j++)
      if (buf[j] && __builtin_ctzll(buf[j]) < best) { best = __builtin_ctzll(buf[j]); who = j; }
    if (who < 0) return -1;
    uint64_t tmp = buf[i]; buf[i] = buf[who]; buf[who] = tmp;
    for (int j = i + 1; j < 32; j++)
      if ((buf[j] >> best) & 1) buf[j] ^= buf[i];
    pv[i] = best;
    *pivots |= 1ULL << best;
  }
  for (int i = 0; i < 32; i++) {
    const int a = c0 + i, b = c0 + pv[i];
    if (a == b) continue;
    int16_t tp = pi[a]; pi[a] = pi[b]; pi[b] = tp;
    for (int r = 0; r < mt; r++)
      if (mat_bit(mat, W, r, a) != mat_bit(mat, W, r, b)) { mat_flip(mat, W, r, a); mat_flip(mat, W, r, b); }
  }
  return 0;
}

/* H~ row i*m+b, column j = bit b of alpha_j^i / g(alpha_j); systematic form [I | T];
   pk row r = T row r, column mt+c at bit c (LM:crypto_kem/<set>/vec/pk_gen.c).  With pi and
   pivots given (the "f" variants) the column step above runs at column mt-32; pi is the support
   permutation (alpha_j = bitrev(pi[j])) and is permuted with the columns. */
static int pk_from_support(int set, uint8_t *pk, const uint16_t *g /* t+1, monic */, const uint16_t *alpha,
                           int16_t *pi, uint64_t *pivots) {
  const orc_params *p = &PARAMS[set & 7];
  const int m = p->m, n = p->n, t = p->t, mt = p->mt;
  const gf_tab *T = tab(m);
  const int W = (n + 63) / 64;
  uint64_t *mat = (uint64_t *)calloc((size_t)mt * W, sizeof(uint64_t));
  uint16_t *v = (uint16_t *)malloc(sizeof(uint16_t) * n);
  for (int j = 0; j < n; j++) v[j] = ginv(T, m, poly_eval(T, g, t, alpha[j]));
  for (int i = 0; i < t; i++) {
    for (int j = 0; j < n; j++)
      for (int b = 0; b < m; b++)
        if ((v[j] >> b) & 1) mat[(size_t)(i * m + b) * W + j / 64] |= 1ULL << (j % 64);
    for (int j = 0; j < n; j++) v[j] = gmul(T, v[j], alpha[j]);
  }
  free(v);
  int ok = 0;
  for (int c = 0; c < mt; c++) {
    if (pi && c == mt - 32 && f_move_columns(mat, W, mt, pi, pivots)) { ok = -1; break; }
    int piv = -1;
    for (int r = c; r < mt; r++)
      if ((mat[(size_t)r * W + c / 64] >> (c % 64)) & 1) { piv = r; break; }
    if (piv < 0) { ok = -1; break; }
    if (piv != c)
      for (int w = 0; w < W; w++) {
        uint64_t tmp = mat[(size_t)c * W + w];
        mat[(size_t)c * W + w] = mat[(size_t)piv * W + w];
        mat[(size_t)piv * W + w] = tmp;
      }
    for (int r = 0; r < mt; r++)
      if (r != c && ((mat[(size_t)r * W + c / 64] >> (c % 64)) & 1))
        for (int w = c / 64; w < W; w++) mat[(size_t)r * W + w] ^= mat[(size_t)c * W + w];
  }
  if (!ok) {
    memset(pk, 0, (size_t)p->pk_bytes);
    for (int r = 0; r < mt; r++)
      for (int c = 0; c < p->k; c++) {
        int col = mt + c;
        if ((mat[(size_t)r * W + col / 64] >> (col % 64)) & 1)
          pk[(size_t)r * p->pk_row_bytes + c / 8] |= (uint8_t)(1 << (c % 8));
      }
  }
  free(mat);
  return ok;
}

typedef struct { uint32_t key; uint16_t idx; } perm_ent;
static int cmp_perm(const void *a, const void *b) {
  const perm_ent *x = (const perm_ent *)a, *y = (const perm_ent *)b;
  if (x->key != y->key) return (x->key > y->key) - (x->key < y->key);
  return (x->idx > y->idx) - (x->idx < y->idx);
}

/* ------------------------------------------------------------------ keypair */
/* LM:crypto_kem/<set>/vec/kem_keypair.c */
void oracle_keypair(int set, uint8_t *pk, uint8_t *sk) {
  const orc_params *p = &PARAMS[set & 7];
  const int m = p->m, n = p->n, t = p->t, q = 1 << m;
  uint8_t seed[33];
  seed[0] = 64;
  oracle_randombytes(seed + 1, 32);
  const long rlen = n / 8 + 4L * q + 2L * t + 32;
  uint8_t *r = (uint8_t *)malloc((size_t)rlen);
  uint16_t *f = (uint16_t *)malloc(sizeof(uint16_t) * t);
  uint16_t *irr = (uint16_t *)malloc(sizeof(uint16_t) * (t + 1));
  perm_ent *ent = (perm_ent *)malloc(sizeof(perm_ent) * q);
  int16_t *pi = (int16_t *)malloc(sizeof(int16_t) * q);
  uint16_t *alpha = (uint16_t *)malloc(sizeof(uint16_t) * n);
  for (;;) {
    oracle_shake256(r, rlen, seed, 33);
    memcpy(sk, seed + 1, 32);
    memcpy(seed + 1, r + rlen - 32, 32);
    const uint8_t *fp = r + rlen - 32 - 2 * t;
    for (int i = 0; i < t; i++) f[i] = load_gf(fp + 2 * i, m);
    if (genpoly(set, irr, f)) continue;
    irr[t] = 1;
    const uint8_t *pp = r + n / 8;
    for (int i = 0; i < q; i++) {
      ent[i].key = ld32le(pp + 4 * i);
      ent[i].idx = (uint16_t)i;
    }
    qsort(ent, (size_t)q, sizeof(perm_ent), cmp_perm);
    int dup = 0;
    for (int i = 1; i < q; i++)
      if (ent[i - 1].key == ent[i].key) dup = 1;
    if (dup) continue;
    for (int i = 0; i < q; i++) pi[i] = (int16_t)ent[i].idx;
    for (int i = 0; i < n; i++) alpha[i] = bitrev_m((uint16_t)pi[i], m);
    uint64_t pivots = 0xFFFFFFFFULL; /* LM:crypto_kem/<set>/vec/kem_keypair.c stores 0xFFFFFFFF */
    if (pk_from_support(set, pk, irr, alpha, (set & ORC_F) ? pi : NULL, &pivots)) continue;
    uint8_t *skp = sk + 40;
    for (int i = 0; i < t; i++) {
      skp[2 * i] = (uint8_t)irr[i];
      skp[2 * i + 1] = (uint8_t)(irr[i] >> 8);
    }
    skp += 2 * t;
    oracle_controlbits(m, pi, skp);
    skp += p->cond_bytes;
    memcpy(skp, r, (size_t)(n / 8));
    /* sk[32..40): the 64-bit pivot mask, little-endian (LM:crypto_kem/<set>f/vec/kem_keypair.c:87) */
    for (int i = 0; i < 8; i++) sk[32 + i] = (uint8_t)(pivots >> (8 * i));
    break;
  }
  free(r); free(f); free(irr); free(ent); free(pi); free(alpha);
}

void oracle_keypair_seeded(int set, uint8_t *pk, uint8_t *sk, const uint8_t seed[32]) {
  oracle_rng_seeded(seed);
  oracle_keypair(set, pk, sk);
}

/* ------------------------------------------------------------------ support from sk */
void oracle_support(int set, const uint8_t *sk, uint16_t *alpha) {
  const orc_params *p = &PARAMS[set & 7];
  const int m = p->m;
  int16_t *pi = (int16_t *)malloc(sizeof(int16_t) << m);
  oracle_perm_from_controlbits(m, sk + 40 + p->irr_bytes, pi);
  for (int i = 0; i < p->n; i++) alpha[i] = bitrev_m((uint16_t)pi[i], m);
  free(pi);
}

int oracle_pk_from_sk(int set, const uint8_t *sk, uint8_t *pk) {
  const orc_params *p = &PARAMS[set & 7];
  uint16_t *alpha = (uint16_t *)malloc(sizeof(uint16_t) * p->n);
  uint16_t *g = (uint16_t *)malloc(sizeof(uint16_t) * (p->t + 1));
  oracle_support(set, sk, alpha);
  for (int i = 0; i < p->t; i++) g[i] = load_gf(sk + 40 + 2 * i, p->m);
  g[p->t] = 1;
  int r = pk_from_support(set, pk, g, alpha, NULL, NULL);
  free(alpha);
  free(g);
  return r;
}

/* ------------------------------------------------------------------ encapsulation */
/* LM:crypto_kem/<set>/vec/encrypt.c gen_e (the 8192128 variant draws t values, no range filter) */
void oracle_gen_e(int set, uint8_t *e, int *attempts) {
  const orc_params *p = &PARAMS[set & 7];
  const int m = p->m, n = p->n, t = p->t;
  const int ndraw = ((set & 7) == ORC_8192128) ? t : 2 * t;
  uint8_t buf[512];
  uint16_t ind[128];
  int tries = 0;
  for (;;) {
    tries++;
    oracle_randombytes(buf, 2 * ndraw);
    int count = 0;
    for (int i = 0; i < ndraw && count < t; i++) {
      uint16_t v = load_gf(buf + 2 * i, m);
      if (v < n) ind[count++] = v;
    }
    if (count < t) continue;
    /* sort (insertion) and reject on a repeated index */
    for (int i = 1; i < t; i++) {
      uint16_t v = ind[i];
      int j = i - 1;
      while (j >= 0 && ind[j] > v) { ind[j + 1] = ind[j]; j--; }
      ind[j + 1] = v;
    }
    int coll = 0;
    for (int i = 1; i < t; i++)
      if (ind[i - 1] == ind[i]) coll = 1;
    if (!coll) break;
  }
  memset(e, 0, (size_t)(n / 8));
  for (int i = 0; i < t; i++) e[ind[i] >> 3] |= (uint8_t)(1 << (ind[i] & 7));
  if (attempts) *attempts = tries;
}

/* c = e[0:mt] xor T e[mt:n]; bits >= mt of the last byte are zero */
static void syndrome_pk(int set, uint8_t *c, const uint8_t *pk, const uint8_t *e) {
  const orc_params *p = &PARAMS[set & 7];
  memset(c, 0, (size_t)p->synd_bytes);
  for (int i = 0; i < p->mt; i++)
    if ((e[i >> 3] >> (i & 7)) & 1) c[i >> 3] |= (uint8_t)(1 << (i & 7));
  for (int r = 0; r < p->mt; r++) {
    const uint8_t *row = pk + (size_t)r * p->pk_row_bytes;
    uint32_t par = 0;
    for (int j = 0; j < p->k; j++) {
      int col = p->mt + j;
      par ^= ((row[j >> 3] >> (j & 7)) & (e[col >> 3] >> (col & 7))) & 1;
    }
    c[r >> 3] ^= (uint8_t)(par << (r & 7));
  }
}

/* LM:crypto_kem/<set>/vec/kem_enc.c; the pc variants (LM:crypto_kem/<set>pc/vec/kem_enc.c:32-37)
   append C = SHAKE256(2 || e, 32) to the ciphertext and hash the longer ciphertext into K */
int oracle_enc(int set, uint8_t *ct, uint8_t *key, const uint8_t *pk) {
  const orc_params *p = &PARAMS[set & 7];
  const int n = p->n, clen = oracle_ct_bytes(set);
  uint8_t e[1 + 1024];
  uint8_t pre[1 + 1024 + 240];
  int ret = 0;
  if ((set & 7) == ORC_6960119) {
    /* LM:crypto_kem/6960119/vec/kem_enc.c check_pk_padding */
    uint8_t b = 0;
    for (int i = 0; i < p->mt; i++) b |= pk[(size_t)i * p->pk_row_bytes + p->pk_row_bytes - 1];
    b >>= (p->k % 8);
    ret = b ? -1 : 0;
  }
  oracle_gen_e(set, e + 1, NULL);
  syndrome_pk(set, ct, pk, e + 1);
  if (set & ORC_PC) {
    e[0] = 2;
    oracle_shake256(ct + p->synd_bytes, 32, e, 1 + n / 8);
  }
  pre[0] = 1;
  memcpy(pre + 1, e + 1, (size_t)(n / 8));
  memcpy(pre + 1 + n / 8, ct, (size_t)clen);
  oracle_shake256(key, 32, pre, 1 + n / 8 + clen);
  if (ret) {
    memset(ct, 0, (size_t)clen);
    memset(key, 0, 32);
  }
  return ret;
}

int oracle_enc_seeded(int set, uint8_t *ct, uint8_t *key, const uint8_t *pk, const uint8_t seed[32]) {
  oracle_rng_seeded(seed);
  return oracle_enc(set, ct, key, pk);
}

/* ------------------------------------------------------------------ decapsulation */
typedef struct {
  uint16_t *alpha; /* n */
  uint16_t *u;     /* 1/g(alpha)^2, n */
  uint16_t g[129];
} dec_key;

static void dec_key_load(int set, const uint8_t *sk, dec_key *K) {
  const orc_params *p = &PARAMS[set & 7];
  const gf_tab *T = tab(p->m);
  K->alpha = (uint16_t *)malloc(sizeof(uint16_t) * p->n);
  K->u = (uint16_t *)malloc(sizeof(uint16_t) * p->n);
  oracle_support(set, sk, K->alpha);
  for (int i = 0; i < p->t; i++) K->g[i] = load_gf(sk + 40 + 2 * i, p->m);
  K->g[p->t] = 1;
  for (int i = 0; i < p->n; i++) {
    uint16_t gv = poly_eval(T, K->g, p->t, K->alpha[i]);
    K->u[i] = ginv(T, p->m, gmul(T, gv, gv));
  }
}
static void dec_key_free(dec_key *K) { free(K->alpha); free(K->u); }

/* s[j] = sum_{i<n} r_i alpha_i^j / g(alpha_i)^2, j < 2t */
static void synd_bits(int set, const dec_key *K, const uint8_t *r, int rbits, uint16_t *s) {
  const orc_params *p = &PARAMS[set & 7];
  const gf_tab *T = tab(p->m);
  memset(s, 0, sizeof(uint16_t) * 2 * p->t);
  for (int i = 0; i < rbits; i++) {
    if (!((r[i >> 3] >> (i & 7)) & 1)) continue;
    uint16_t v = K->u[i];
    for (int j = 0; j < 2 * p->t; j++) {
      s[j] ^= v;
      v = gmul(T, v, K->alpha[i]);
    }
  }
}

/* Berlekamp-Massey with division (the formulation of the Classic McEliece "ref" code; the vec
   code's inversion-free variant keeps every intermediate polynomial a nonzero scalar multiple of
   this one, so both give the same root set).  locator[i] = C[t-i]. */
void oracle_bm(int set, const uint16_t *s, uint16_t *locator) {
  const orc_params *p = &PARAMS[set & 7];
  const int t = p->t, m = p->m;
  const gf_tab *T = tab(m);
  uint16_t C[129], B[129], Tm[129];
  memset(C, 0, sizeof C);
  memset(B, 0, sizeof B);
  C[0] = 1;
  B[1] = 1;
  uint16_t b = 1;
  int L = 0;
  for (int N = 0; N < 2 * t; N++) {
    uint16_t d = 0;
    for (int i = 0; i <= (N < t ? N : t); i++) d ^= gmul(T, C[i], s[N - i]);
    int upd_len = (d != 0) && (N >= 2 * L);
    memcpy(Tm, C, sizeof C);
    if (d) {
      uint16_t f = gmul(T, d, ginv(T, m, b));
      for (int i = 0; i <= t; i++) C[i] ^= gmul(T, f, B[i]);
    }
    if (upd_len) {
      L = N + 1 - L;
      memcpy(B, Tm, sizeof B);
      b = d;
    }
    for (int i = t; i >= 1; i--) B[i] = B[i - 1];
    B[0] = 0;
  }
  for (int i = 0; i <= t; i++) locator[i] = C[t - i];
}

void oracle_synd_ct(int set, const uint8_t *sk, const uint8_t *ct, uint16_t *s) {
  const orc_params *p = &PARAMS[set & 7];
  dec_key K;
  dec_key_load(set, sk, &K);
  synd_bits(set, &K, ct, 8 * p->synd_bytes, s);
  dec_key_free(&K);
}

/* LM:crypto_kem/<set>/vec/decrypt.c: r = (ct bits incl. any padding bits of the last byte, 0...);
   e_i = [locator(alpha_i) == 0]; success iff weight(e) == t and synd(e) == synd(r). */
static int decrypt_with(int set, const dec_key *K, const uint8_t *ct, uint8_t *e) {
  const orc_params *p = &PARAMS[set & 7];
  const int t = p->t, n = p->n;
  const gf_tab *T = tab(p->m);
  uint16_t s[256], s2[256], loc[129];
  synd_bits(set, K, ct, 8 * p->synd_bytes, s);
  oracle_bm(set, s, loc);
  memset(e, 0, (size_t)(n / 8));
  int w = 0;
  for (int i = 0; i < n; i++)
    if (poly_eval(T, loc, t, K->alpha[i]) == 0) {
      e[i >> 3] |= (uint8_t)(1 << (i & 7));
      w++;
    }
  synd_bits(set, K, e, n, s2);
  int ok = (w == t);
  for (int j = 0; j < 2 * t; j++)
    if (s[j] != s2[j]) ok = 0;
  return ok ? 0 : 1;
}

int oracle_decrypt(int set, const uint8_t *sk, const uint8_t *ct, uint8_t *e) {
  dec_key K;
  dec_key_load(set, sk, &K);
  int r = decrypt_with(set, &K, ct, e);
  dec_key_free(&K);
  return r;
}

/* LM:crypto_kem/<set>/vec/kem_dec.c (+ 6960119 check_c_padding); the pc variants
   (LM:crypto_kem/<set>pc/vec/kem_dec.c:43-59) also fail when SHAKE256(2 || e, 32) differs from
   the ciphertext's last 32 bytes, and hash the whole ciphertext */
static int dec_with(int set, const dec_key *K, uint8_t *key, const uint8_t *ct, const uint8_t *sk) {
  const orc_params *p = &PARAMS[set & 7];
  const int n = p->n, clen = oracle_ct_bytes(set);
  uint8_t e[1 + 1024], pre[1 + 1024 + 240], ctc[240];
  memcpy(ctc, ct, (size_t)clen); /* ct and key may alias */
  int pad_bad = 0;
  if ((set & 7) == ORC_6960119) pad_bad = (ctc[p->synd_bytes - 1] >> (p->mt % 8)) != 0;
  int fail = decrypt_with(set, K, ctc, e + 1);
  if (set & ORC_PC) {
    uint8_t conf[32];
    e[0] = 2;
    oracle_shake256(conf, 32, e, 1 + n / 8);
    if (memcmp(conf, ctc + p->synd_bytes, 32) != 0) fail = 1;
  }
  const uint8_t *s = sk + 40 + p->irr_bytes + p->cond_bytes;
  pre[0] = fail ? 0 : 1;
  memcpy(pre + 1, fail ? s : e + 1, (size_t)(n / 8));
  memcpy(pre + 1 + n / 8, ctc, (size_t)clen);
  oracle_shake256(key, 32, pre, 1 + n / 8 + clen);
  if (pad_bad) {
    memset(key, 0xFF, 32);
    return -1;
  }
  return 0;
}

int oracle_dec(int set, uint8_t *key, const uint8_t *ct, const uint8_t *sk) {
  const orc_params *p = &PARAMS[set & 7];
  uint8_t *skc = (uint8_t *)malloc((size_t)p->sk_bytes); /* sk and key may alias */
  memcpy(skc, sk, (size_t)p->sk_bytes);
  dec_key K;
  dec_key_load(set, skc, &K);
  int r = dec_with(set, &K, key, ct, skc);
  dec_key_free(&K);
  free(skc);
  return r;
}

/* ------------------------------------------------------------------ batches (threads) */
typedef struct {
  int set, is_dec;
  uint8_t *ct, *key;
  int *ret;
  const uint8_t *pk, *sk, *seeds, *cts;
  long lo, hi;
} job_t;

static void *job_run(void *arg) {
  job_t *j = (job_t *)arg;
  const orc_params *p = &PARAMS[j->set & 7];
  const long clen = oracle_ct_bytes(j->set);
  if (j->is_dec == 2) { /* syndrome of given error vectors (seeds holds e) */
    for (long i = j->lo; i < j->hi; i++)
      syndrome_pk(j->set, j->ct + (size_t)i * p->synd_bytes, j->pk, j->seeds + (size_t)i * (p->n / 8));
  } else if (j->is_dec) {
    dec_key K;
    dec_key_load(j->set, j->sk, &K);
    for (long i = j->lo; i < j->hi; i++) {
      int r = dec_with(j->set, &K, j->key + 32 * i, j->cts + (size_t)i * clen, j->sk);
      if (j->ret) j->ret[i] = r;
    }
    dec_key_free(&K);
  } else {
    for (long i = j->lo; i < j->hi; i++) {
      int r = oracle_enc_seeded(j->set, j->ct + (size_t)i * clen, j->key + 32 * i, j->pk,
                                j->seeds + 32 * i);
      if (j->ret) j->ret[i] = r;
    }
  }
  return NULL;
}

static int online_cores(void) {
  cpu_set_t cs;
  if (sched_getaffinity(0, sizeof cs, &cs) == 0) return CPU_COUNT(&cs);
  long c = sysconf(_SC_NPROCESSORS_ONLN);
  return c > 0 ? (int)c : 1;
}

static int run_jobs(job_t *proto, long n, int nthreads) {
  if (nthreads <= 0) nthreads = online_cores();
  if (nthreads > n) nthreads = (int)(n > 0 ? n : 1);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  job_t *jobs = (job_t *)malloc(sizeof(job_t) * nthreads);
  for (int i = 0; i < nthreads; i++) {
    jobs[i] = *proto;
    jobs[i].lo = n * i / nthreads;
    jobs[i].hi = n * (i + 1) / nthreads;
    pthread_create(&th[i], NULL, job_run, &jobs[i]);
  }
  for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
  free(th);
  free(jobs);
  return nthreads;
}

int oracle_enc_batch(int set, uint8_t *ct, uint8_t *key, int *ret, const uint8_t *pk,
                     const uint8_t *seeds, long n, int nthreads) {
  job_t j = {set, 0, ct, key, ret, pk, NULL, seeds, NULL, 0, 0};
  return run_jobs(&j, n, nthreads);
}

int oracle_syndrome_batch(int set, uint8_t *ct, const uint8_t *pk, const uint8_t *e, long n, int nthreads) {
  job_t j = {set, 2, ct, NULL, NULL, pk, NULL, e, NULL, 0, 0};
  return run_jobs(&j, n, nthreads);
}

int oracle_dec_batch(int set, uint8_t *key, int *ret, const uint8_t *ct, const uint8_t *sk,
                     long n, int nthreads) {
  job_t j = {set, 1, NULL, key, ret, NULL, sk, NULL, ct, 0, 0};
  return run_jobs(&j, n, nthreads);
}

/* ------------------------------------------------------------------ word-level CPU port
   The CPU baseline that is timed beside the GPU: the same results as oracle_enc_seeded and
   oracle_dec (checked against them in tests/test_oracle_fast.py), computed the way a CPU
   implementation would, with the per-key work done once per batch:
   - encapsulation: seeded gen_e with a bitmap collision test; the syndrome as the XOR of the t
     columns of [I | T] at the support of e (64-bit words; T's columns transposed once per key);
   - decapsulation: the 2t syndromes of r as the XOR of per-position columns u_i alpha_i^j
     (u_i = 1/g(alpha_i)^2, 16-bit field elements, built once per key; LM:crypto_kem/<set>/vec/
     decrypt.c computes the same s by its transposed FFT); Berlekamp-Massey with division
     bounded by the current length; the error locator evaluated at every field element by the
     Gao-Mateer additive FFT (the method of LM:crypto_kem/<set>/vec/fft.c) and read at the
     support; the re-encryption check by the same columns.
   Threads: one per core by default, each on a contiguous slice. */
typedef struct {
  int set;
  /* encapsulation */
  const uint8_t *pk;
  uint64_t *tcol; /* column j of T: cw words */
  int cw, pk_bad;
  /* decapsulation */
  const uint8_t *sk;
  dec_key K;
  uint64_t *scol; /* position i: the 2t syndromes of the unit vector e_i, 16 bits each, sw words */
  int sw;
  uint16_t *pow_t; /* x^t for every field element (t = D only) */
} fast_key;

void *oracle_fast_key(int set, const uint8_t *pk, const uint8_t *sk) {
  const orc_params *p = &PARAMS[set & 7];
  const gf_tab *T = tab(p->m);
  fast_key *F = (fast_key *)calloc(1, sizeof(fast_key));
  F->set = set;
  if (pk) {
    F->pk = pk;
    F->cw = (p->mt + 63) / 64;
    F->tcol = (uint64_t *)calloc((size_t)p->k * F->cw, 8);
    for (int r = 0; r < p->mt; r++) {
      const uint8_t *row = pk + (size_t)r * p->pk_row_bytes;
      for (int j = 0; j < p->k; j++)
        if ((row[j >> 3] >> (j & 7)) & 1) F->tcol[(size_t)j * F->cw + (r >> 6)] |= 1ULL << (r & 63);
    }
    if ((set & 7) == ORC_6960119) { /* LM:crypto_kem/6960119/vec/kem_enc.c check_pk_padding */
      uint8_t b = 0;
      for (int i = 0; i < p->mt; i++) b |= pk[(size_t)i * p->pk_row_bytes + p->pk_row_bytes - 1];
      F->pk_bad = (b >> (p->k % 8)) != 0;
    }
  }
  if (sk) {
    F->sk = sk;
    dec_key_load(set, sk, &F->K);
    const int tt = 2 * p->t;
    F->sw = (tt + 3) / 4; /* 2t = 238 for 6960119: the last word holds two */
    F->scol = (uint64_t *)calloc((size_t)p->n * F->sw, 8);
    for (int i = 0; i < p->n; i++) {
      uint16_t *c = (uint16_t *)(F->scol + (size_t)i * F->sw);
      uint16_t v = F->K.u[i];
      for (int j = 0; j < tt; j++) {
        c[j] = v;
        v = gmul(T, v, F->K.alpha[i]);
      }
    }
    if (p->t == 64 || p->t == 128) {
      F->pow_t = (uint16_t *)malloc(sizeof(uint16_t) << p->m);
      for (int x = 0; x < (1 << p->m); x++) {
        uint16_t v = (uint16_t)x;
        for (int k = 1; k < p->t; k <<= 1) v = gmul(T, v, v);
        F->pow_t[x] = v;
      }
    }
  }
  return F;
}

void oracle_fast_key_free(void *h) {
  fast_key *F = (fast_key *)h;
  if (!F) return;
  free(F->tcol);
  if (F->sk) dec_key_free(&F->K);
  free(F->scol);
  free(F->pow_t);
  free(F);
}

/* seeded gen_e (the chain of oracle_rng_seeded) returning the sorted-free index list */
static void fast_gen_e(int set, const uint8_t seed[32], uint16_t *ind) {
  const orc_params *p = &PARAMS[set & 7];
  const int m = p->m, n = p->n, t = p->t;
  const int ndraw = ((set & 7) == ORC_8192128) ? t : 2 * t;
  uint8_t st[33], buf[512];
  uint64_t bm[128];
  st[0] = 0x40;
  memcpy(st + 1, seed, 32);
  for (;;) {
    oracle_shake256(buf, 2 * ndraw, st, 33);
    memcpy(st + 1, buf, 32);
    int count = 0;
    for (int i = 0; i < ndraw && count < t; i++) {
      uint16_t v = load_gf(buf + 2 * i, m);
      if (v < n) ind[count++] = v;
    }
    if (count < t) continue;
    memset(bm, 0, sizeof(uint64_t) * (size_t)((n + 63) / 64));
    int coll = 0;
    for (int i = 0; i < t; i++) {
      const uint64_t b = 1ULL << (ind[i] & 63);
      coll |= (bm[ind[i] >> 6] & b) != 0;
      bm[ind[i] >> 6] |= b;
    }
    if (!coll) return;
  }
}

static int fast_enc_one(const fast_key *F, uint8_t *ct, uint8_t *key, const uint8_t seed[32]) {
  const int set = F->set;
  const orc_params *p = &PARAMS[set & 7];
  const int n = p->n, clen = oracle_ct_bytes(set);
  if (F->pk_bad) {
    memset(ct, 0, (size_t)clen);
    memset(key, 0, 32);
    return -1;
  }
  uint16_t ind[128];
  fast_gen_e(set, seed, ind);
  uint8_t pre[1 + 1024 + 240];
  uint8_t *e = pre + 1;
  memset(pre, 0, (size_t)(1 + n / 8));
  uint64_t c[32];
  memset(c, 0, sizeof c);
  for (int i = 0; i < p->t; i++) {
    const int v = ind[i];
    e[v >> 3] |= (uint8_t)(1 << (v & 7));
    if (v < p->mt) {
      c[v >> 6] ^= 1ULL << (v & 63);
    } else {
      const uint64_t *col = F->tcol + (size_t)(v - p->mt) * F->cw;
      for (int w = 0; w < F->cw; w++) c[w] ^= col[w];
    }
  }
  memcpy(ct, c, (size_t)p->synd_bytes); /* little-endian words = libmceliece's bit order */
  if (set & ORC_PC) {
    pre[0] = 2;
    oracle_shake256(ct + p->synd_bytes, 32, pre, 1 + n / 8);
  }
  pre[0] = 1;
  memcpy(pre + 1 + n / 8, ct, (size_t)clen);
  oracle_shake256(key, 32, pre, 1 + n / 8 + clen);
  return 0;
}

/* Gao-Mateer additive FFT: f[0..L) (L a power of two, destroyed) evaluated at the 2^k points
   sum_j bit_j(i) b[j], out[i] for i < 2^k */
static void fast_afft(const gf_tab *T, int m, uint16_t *f, int L, const uint16_t *b, int k, uint16_t *out) {
  const int NP = 1 << k;
  if (L == 1) {
    for (int i = 0; i < NP; i++) out[i] = f[0];
    return;
  }
  /* g(x) = f(b[k-1] x) */
  const uint16_t bk = b[k - 1];
  uint16_t pw = 1;
  for (int i = 1; i < L; i++) {
    pw = gmul(T, pw, bk);
    f[i] = gmul(T, f[i], pw);
  }
  /* Taylor expansion at x^2 + x: blocks of 4q coefficients [a0 a1 a2 a3] -> [a0, a1+a2+a3, a2+a3, a3] */
  for (int q = L / 4; q >= 1; q >>= 1)
    for (int bs = 0; bs < L; bs += 4 * q)
      for (int i = 0; i < q; i++) {
        f[bs + 2 * q + i] ^= f[bs + 3 * q + i];
        f[bs + q + i] ^= f[bs + 2 * q + i];
      }
  uint16_t g0[128], g1[128], gam[13], del[13], pre[13];
  for (int i = 0; i < L / 2; i++) {
    g0[i] = f[2 * i];
    g1[i] = f[2 * i + 1];
  }
  const uint16_t binv = ginv(T, m, bk);
  for (int j = 0; j < k - 1; j++) {
    gam[j] = gmul(T, b[j], binv);
    del[j] = gmul(T, gam[j], gam[j]) ^ gam[j];
    pre[j] = (j ? pre[j - 1] : 0) ^ gam[j];
  }
  const int half = NP / 2;
  fast_afft(T, m, g0, L / 2, del, k - 1, out);
  fast_afft(T, m, g1, L / 2, del, k - 1, out + half);
  uint16_t G = 0;
  for (int i = 0; i < half; i++) {
    if (i) G ^= pre[__builtin_ctz((unsigned)i)];
    const uint16_t v = out[half + i];
    const uint16_t x = out[i] ^ gmul(T, G, v);
    out[i] = x;
    out[half + i] = x ^ v;
  }
}

/* BM with division (as oracle_bm); deg C <= L (Massey), so the discrepancy sums i <= min(L, N, t),
   and after step N deg C <= N + 1.  B is kept as logs (ZL for a zero coefficient) and the syndromes
   as logs, so each product is one table read: xe[la + lb] with xe zero beyond 2q.
   locator[i] = C[t-i] */
static void fast_bm(const gf_tab *T, int t, int m, const uint16_t *s, uint16_t *locator) {
  const int q = (1 << m) - 1, ZL = 2 * q;
  static __thread uint16_t xe[2][4 * 8191 + 2];
  static __thread int xe_ready[2];
  uint16_t *ex = xe[m - 12];
  if (!xe_ready[m - 12]) {
    for (int i = 0; i <= 4 * q + 1; i++) ex[i] = i < 2 * q ? T->exp[i % q] : 0;
    xe_ready[m - 12] = 1;
  }
  uint16_t C[130], Bl[130], Tm[130], sl[256];
  for (int j = 0; j < 2 * t; j++) sl[j] = s[j] ? T->log[s[j]] : (uint16_t)ZL;
  memset(C, 0, sizeof C);
  for (int i = 0; i <= t; i++) Bl[i] = (uint16_t)ZL;
  C[0] = 1;
  Bl[1] = 0;
  int lb = 0; /* log b, b = 1 */
  int L = 0;
  for (int N = 0; N < 2 * t; N++) {
    uint16_t d = 0;
    int top = L < N ? L : N;
    if (top > t) top = t;
    for (int i = 0; i <= top; i++)
      if (C[i]) d ^= ex[T->log[C[i]] + sl[N - i]];
    const int upd_len = (d != 0) && (N >= 2 * L);
    const int hi = N + 1 < t ? N + 1 : t;
    if (upd_len)
      for (int i = 0; i <= hi; i++) Tm[i] = C[i] ? T->log[C[i]] : (uint16_t)ZL;
    if (d) {
      const int ld = T->log[d];
      const int lf = ld >= lb ? ld - lb : ld + q - lb;
      for (int i = 0; i <= hi; i++) C[i] ^= ex[lf + Bl[i]];
      if (upd_len) {
        L = N + 1 - L;
        for (int i = 0; i <= hi; i++) Bl[i] = Tm[i];
        for (int i = hi + 1; i <= t; i++) Bl[i] = (uint16_t)ZL;
        lb = ld;
      }
    }
    for (int i = (hi + 1 < t ? hi + 1 : t); i >= 1; i--) Bl[i] = Bl[i - 1];
    Bl[0] = (uint16_t)ZL;
  }
  for (int i = 0; i <= t; i++) locator[i] = C[t - i];
}

static int fast_decrypt(const fast_key *F, const uint8_t *ct, uint8_t *e) {
  const int set = F->set;
  const orc_params *p = &PARAMS[set & 7];
  const int t = p->t, n = p->n, m = p->m, sw = F->sw;
  const gf_tab *T = tab(m);
  uint64_t s[64], s2[64];
  memset(s, 0, sizeof(uint64_t) * (size_t)sw);
  const int rbits = 8 * p->synd_bytes;
  for (int w = 0; w < (rbits + 63) / 64; w++) {
    uint64_t x = 0;
    const int nb = p->synd_bytes - 8 * w; /* >= 1 inside the loop */
    memcpy(&x, ct + 8 * w, (size_t)(nb >= 8 ? 8 : (nb > 0 ? nb : 0)));
    while (x) {
      const int i = 64 * w + __builtin_ctzll(x);
      x &= x - 1;
      const uint64_t *col = F->scol + (size_t)i * sw;
      for (int k = 0; k < sw; k++) s[k] ^= col[k];
    }
  }
  uint16_t loc[130];
  fast_bm(T, t, m, (const uint16_t *)s, loc);
  const int D = t <= 64 ? 64 : 128;
  uint16_t f[128], basis[13];
  static __thread uint16_t vals[8192];
  for (int i = 0; i < D; i++) f[i] = i <= t ? loc[i] : 0;
  for (int j = 0; j < m; j++) basis[j] = (uint16_t)(1u << j);
  fast_afft(T, m, f, D, basis, m, vals);
  if (t == D)
    for (int x = 0; x < (1 << m); x++) vals[x] ^= F->pow_t[x]; /* loc[t] = C[0] = 1 */
  memset(e, 0, (size_t)(n / 8));
  memset(s2, 0, sizeof(uint64_t) * (size_t)sw);
  int wt = 0;
  for (int i = 0; i < n; i++)
    if (vals[F->K.alpha[i]] == 0) {
      e[i >> 3] |= (uint8_t)(1 << (i & 7));
      const uint64_t *col = F->scol + (size_t)i * sw;
      for (int k = 0; k < sw; k++) s2[k] ^= col[k];
      wt++;
    }
  int ok = wt == t;
  for (int k = 0; k < sw; k++) ok &= s[k] == s2[k];
  return ok ? 0 : 1;
}

static int fast_dec_one(const fast_key *F, uint8_t *key, const uint8_t *ct) {
  const int set = F->set;
  const orc_params *p = &PARAMS[set & 7];
  const int n = p->n, clen = oracle_ct_bytes(set);
  uint8_t e[1 + 1024], pre[1 + 1024 + 240], ctc[240];
  memcpy(ctc, ct, (size_t)clen);
  int pad_bad = 0;
  if ((set & 7) == ORC_6960119) pad_bad = (ctc[p->synd_bytes - 1] >> (p->mt % 8)) != 0;
  int fail = fast_decrypt(F, ctc, e + 1);
  if (set & ORC_PC) {
    uint8_t conf[32];
    e[0] = 2;
    oracle_shake256(conf, 32, e, 1 + n / 8);
    if (memcmp(conf, ctc + p->synd_bytes, 32) != 0) fail = 1;
  }
  const uint8_t *sv = F->sk + 40 + p->irr_bytes + p->cond_bytes;
  pre[0] = fail ? 0 : 1;
  memcpy(pre + 1, fail ? sv : e + 1, (size_t)(n / 8));
  memcpy(pre + 1 + n / 8, ctc, (size_t)clen);
  oracle_shake256(key, 32, pre, 1 + n / 8 + clen);
  if (pad_bad) {
    memset(key, 0xFF, 32);
    return -1;
  }
  return 0;
}

typedef struct {
  const fast_key *F;
  int is_dec;
  uint8_t *ct, *key;
  int *ret;
  const uint8_t *in;
  long lo, hi;
} fast_job;

static void *fast_job_run(void *arg) {
  fast_job *j = (fast_job *)arg;
  const long clen = oracle_ct_bytes(j->F->set);
  for (long i = j->lo; i < j->hi; i++) {
    const int r = j->is_dec ? fast_dec_one(j->F, j->key + 32 * i, j->in + (size_t)i * clen)
                            : fast_enc_one(j->F, j->ct + (size_t)i * clen, j->key + 32 * i, j->in + 32 * i);
    if (j->ret) j->ret[i] = r;
  }
  return NULL;
}

static int fast_run(fast_job *proto, long n, int nthreads) {
  if (nthreads <= 0) nthreads = online_cores();
  if (nthreads > n) nthreads = (int)(n > 0 ? n : 1);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  fast_job *jobs = (fast_job *)malloc(sizeof(fast_job) * nthreads);
  for (int i = 0; i < nthreads; i++) {
    jobs[i] = *proto;
    jobs[i].lo = n * i / nthreads;
    jobs[i].hi = n * (i + 1) / nthreads;
    pthread_create(&th[i], NULL, fast_job_run, &jobs[i]);
  }
  for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
  free(th);
  free(jobs);
  return nthreads;
}

int oracle_fast_enc_batch(void *h, uint8_t *ct, uint8_t *key, int *ret, const uint8_t *seeds, long n, int nthreads) {
  fast_job j = {(const fast_key *)h, 0, ct, key, ret, seeds, 0, 0};
  return fast_run(&j, n, nthreads);
}

int oracle_fast_dec_batch(void *h, uint8_t *key, int *ret, const uint8_t *ct, long n, int nthreads) {
  fast_job j = {(const fast_key *)h, 1, NULL, key, ret, ct, 0, 0};
  return fast_run(&j, n, nthreads);
}

/* evaluation of f[0..deg] at every field element by the FFT above (test hook) */
void oracle_fast_eval_all(int m, const uint16_t *f, int deg, uint16_t *vals) {
  const gf_tab *T = tab(m);
  uint16_t g[128], basis[13];
  int L = 1;
  while (L < deg + 1) L <<= 1;
  for (int i = 0; i < L; i++) g[i] = i <= deg ? f[i] : 0;
  for (int j = 0; j < m; j++) basis[j] = (uint16_t)(1u << j);
  fast_afft(T, m, g, L, basis, m, vals);
}

/* ------------------------------------------------------------------ KAT checksum */
/* LM:command/mceliece-test.c: salsa20 core (:101-128), testvector/myrandom (:168-196),
   checksum (:374-388) and the KEM loop of test_kem_<set>_impl (:4634-4731). */
static void salsa_core(uint8_t out[64], const uint8_t in[16], const uint8_t k[32]) {
  uint32_t x[16], y[16], w[16], t[4];
  for (int i = 0; i < 4; i++) {
    x[5 * i] = ld32le(SIGMA + 4 * i);
    x[1 + i] = ld32le(k + 4 * i);
    x[6 + i] = ld32le(in + 4 * i);
    x[11 + i] = ld32le(k + 16 + 4 * i);
  }
  memcpy(y, x, sizeof x);
  for (int r = 0; r < 20; r++) {
    for (int j = 0; j < 4; j++) {
      for (int m = 0; m < 4; m++) t[m] = x[(5 * j + 4 * m) % 16];
      t[1] ^= rotl32(t[0] + t[3], 7);
      t[2] ^= rotl32(t[1] + t[0], 9);
      t[3] ^= rotl32(t[2] + t[1], 13);
      t[0] ^= rotl32(t[3] + t[2], 18);
      for (int m = 0; m < 4; m++) w[4 * j + (j + m) % 4] = t[m];
    }
    memcpy(x, w, sizeof x);
  }
  for (int i = 0; i < 16; i++) st32le(out + 4 * i, x[i] + y[i]);
}

typedef struct {
  uint8_t state[64];
  uint8_t tvn[8];
} kat_ctx;

static void kat_checksum(kat_ctx *c, const uint8_t *x, long long len) {
  uint8_t blk[16];
  while (len >= 16) {
    salsa_core(c->state, x, c->state);
    x += 16;
    len -= 16;
  }
  memset(blk, 0, 16);
  memcpy(blk, x, (size_t)len);
  blk[len] = 1;
  c->state[0] ^= 1;
  salsa_core(c->state, blk, c->state);
}

static uint64_t kat_myrandom(kat_ctx *c) {
  static const uint8_t tvk[33] = "generate inputs for test vectors";
  uint8_t z[16] = {0}, out[64];
  memcpy(z, c->tvn, 8);
  salsa_core(out, z, tvk);
  for (int i = 0; i < 8; i++)
    if (++c->tvn[i]) break;
  uint64_t r = 0;
  for (int i = 7; i >= 0; i--) r = (r << 8) | out[i];
  return r;
}

/* the KEM loop of the test driver over three libmceliece-style functions (randomness: the
   driver's knownrandombytes, which a library under test must read through oracle_randombytes) */
static void kat_core(int set, int loops, char *hex_out, orc_keypair_fn kp, orc_enc_fn enc, orc_dec_fn dec) {
  const orc_params *p = &PARAMS[set & 7];
  kat_ctx c;
  memset(&c, 0, sizeof c);
  oracle_rng_known_reset();
  uint8_t *pk = (uint8_t *)malloc((size_t)p->pk_bytes);
  uint8_t *sk = (uint8_t *)malloc((size_t)p->sk_bytes);
  uint8_t ct[240], k[32], t[32];
  const long clen = oracle_ct_bytes(set);
  for (int loop = 0; loop < loops; loop++) {
    kp(pk, sk);
    kat_checksum(&c, pk, p->pk_bytes);
    kat_checksum(&c, sk, p->sk_bytes);
    enc(ct, k, pk);
    kat_checksum(&c, ct, clen);
    kat_checksum(&c, k, 32);
    dec(t, ct, sk);
    kat_checksum(&c, t, 32);
    for (int rep = 0; rep < 3; rep++) {
      /* c[myrandom() % clen] += 1 + (myrandom() % 255): as compiled by gcc, the right operand
         (the increment draw) is evaluated before the index draw */
      uint64_t r2 = kat_myrandom(&c);
      uint64_t r1 = kat_myrandom(&c);
      ct[r1 % (uint64_t)clen] = (uint8_t)(ct[r1 % (uint64_t)clen] + 1 + (r2 % 255));
      if (dec(t, ct, sk) == 0)
        kat_checksum(&c, t, 32);
      else
        kat_checksum(&c, ct, clen);
    }
  }
  for (int i = 0; i < 32; i++) sprintf(hex_out + 2 * i, "%02x", c.state[i]);
  hex_out[64] = 0;
  free(pk);
  free(sk);
}

static __thread int kat_set;
static void own_keypair(uint8_t *pk, uint8_t *sk) { oracle_keypair(kat_set, pk, sk); }
static int own_enc(uint8_t *ct, uint8_t *k, const uint8_t *pk) { return oracle_enc(kat_set, ct, k, pk); }
static int own_dec(uint8_t *k, const uint8_t *ct, const uint8_t *sk) { return oracle_dec(kat_set, k, ct, sk); }

void oracle_kat_checksum(int set, int loops, char *hex_out) {
  kat_set = set;
  kat_core(set, loops, hex_out, own_keypair, own_enc, own_dec);
}

void oracle_kat_checksum_fns(int set, int loops, char *hex_out, orc_keypair_fn kp, orc_enc_fn enc,
                             orc_dec_fn dec) {
  kat_core(set, loops, hex_out, kp, enc, dec);
}

/* testvector(): salsa20 stream under nonce tvn, key "generate inputs for test vectors"
   (LM:command/mceliece-test.c:175-180) */
static void kat_testvector(kat_ctx *c, uint8_t *x, long long len) {
  static const uint8_t tvk[33] = "generate inputs for test vectors";
  uint8_t z[16] = {0}, blk[64];
  memcpy(z, c->tvn, 8);
  while (len > 0) {
    salsa_core(blk, z, tvk);
    long long take = len < 64 ? len : 64;
    memcpy(x, blk, (size_t)take);
    x += take;
    len -= take;
    uint32_t u = 1;
    for (int i = 8; i < 16; i++) {
      u += z[i];
      z[i] = (uint8_t)u;
      u >>= 8;
    }
  }
  for (int i = 0; i < 8; i++)
    if (++c->tvn[i]) break;
}

/* the xof_shake256 checksum of LM:command/mceliece-test.c:502-552 (pins SHAKE256 and the
   checksum/testvector machinery used by oracle_kat_checksum) */
void oracle_kat_shake_checksum(int loops, int maxtest, char *hex_out) {
  kat_ctx c;
  memset(&c, 0, sizeof c);
  uint8_t *m = (uint8_t *)malloc((size_t)maxtest + 1), *h = (uint8_t *)malloc((size_t)maxtest + 1);
  for (int loop = 0; loop < loops; loop++) {
    long long mlen = (long long)(kat_myrandom(&c) % (uint64_t)(maxtest + 1));
    long long hlen = (long long)(kat_myrandom(&c) % (uint64_t)(maxtest + 1));
    kat_testvector(&c, m, mlen);
    oracle_shake256(h, hlen, m, mlen);
    kat_checksum(&c, h, hlen);
  }
  for (int i = 0; i < 32; i++) sprintf(hex_out + 2 * i, "%02x", c.state[i]);
  hex_out[64] = 0;
  free(m);
  free(h);
}
