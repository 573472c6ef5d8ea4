// host_math.h -- host-side GF(2^m) arithmetic and per-key / per-field precomputation for the
// device engine, plus the host key generator behind the crypto_kem keypair entry points.
#pragma once
#include <cstdint>
#include <vector>

namespace tmce {

struct HostField {
  int m = 0;
  uint32_t q = 0;
  std::vector<uint16_t> lg, ex;  // ex has 2q entries (no reduction needed for a sum of two logs)
  explicit HostField(int m_);
  uint16_t mul(uint16_t a, uint16_t b) const { return (a && b) ? ex[lg[a] + lg[b]] : 0; }
  uint16_t inv(uint16_t a) const { return a ? ex[(q - lg[a]) % q] : 0; }
  uint16_t pw(uint16_t a, uint32_t e) const;
  uint16_t logv(uint16_t a) const { return a ? lg[a] : 0xFFFF; }
};
const HostField &host_field(int m);

uint16_t bitrev(uint16_t a, int m);

// Permutation of 2^m elements encoded by Benes control bits laid out as in libmceliece's
// secret key (2m-1 layers of 2^(m-1) bits; strides 1,2,..,2^(m-1),..,2,1).
void perm_from_controlbits(int m, const uint8_t *cb, std::vector<uint16_t> &pi);
// Control bits of a permutation (Nassimi-Sahni recursion, libmceliece controlbits.c).
void controlbits_from_perm(int m, const std::vector<uint16_t> &pi, uint8_t *cb);

// Additive-FFT tables for evaluating a D-coefficient polynomial at all 2^m points.
// The bottom-up butterfly multiplier of level d at point index c is p_d(c) = sum_i c_i gamma_i,
// linear in the bits of c; with c = 64*j + lane it splits as plo[d][lane] ^ phi[d][j].
struct FftTables {
  std::vector<uint16_t> twl;   // 2D entries: twist logs, level d at 2D - 2(D >> d)
  std::vector<uint16_t> plo;   // KD x 64 raw multipliers (lane part)
  std::vector<uint16_t> phi;   // register part, level d at offset PHI_OFF(d) = sum_{d'<d} 2^(m-7-d')
};
FftTables build_fft_tables(int m, int D);

// Per-key data of decapsulation, derived from sk.
struct DecKeyHost {
  std::vector<uint16_t> alpha, la, lu;
  std::vector<uint16_t> ainv;  // 2^m entries: i for the field element alpha_i, 0xFFFF off the support
  std::vector<uint32_t> supp;  // 2^m bits: field element 64 g + l is in the support (bit g % 32 of word l * 2^(m-11) + g / 32)
  std::vector<uint32_t> F;     // R x KWP words
  std::vector<uint32_t> Fcol;  // (N + 1) x FCW words: column i = syndromes of unit vector i (row j*M + b), column N zero
  std::vector<uint8_t> rej;  // N/8 bytes
};
void build_dec_key(int set, const uint8_t *sk, DecKeyHost &out);

// Per-key data of encapsulation, derived from pk: column-major [I | T], N x CWP words.
void build_enc_cols(int set, const uint8_t *pk, std::vector<uint32_t> &cols);
bool pk_padding_bad(int set, const uint8_t *pk);
// The public key T as MT rows of ceil(K/32) little-endian words (plain form's GEMM operand).
void build_enc_rows(int set, const uint8_t *pk, std::vector<uint32_t> &rows);

// Host key generation with libmceliece's keypair semantics; rand(out,len) supplies randomness.
// `variant` is a set id, | VAR_F for the semi-systematic form of the f variants (VAR_PC does not
// change key generation).
typedef void (*rand_fn)(uint8_t *out, long long len, void *ctx);
void host_keypair(int variant, uint8_t *pk, uint8_t *sk, rand_fn rnd, void *ctx);

// The same key generation in two host steps around the systematic-form step (which the batched
// generator runs on the GPU).  keygen_candidate advances seed33 (64 || 32-byte state) through
// attempts until one has an irreducible Goppa polynomial and distinct support sort keys; if the
// systematic form of that candidate fails, calling it again continues the chain (libmceliece's
// loop).  keygen_finish writes sk once the systematic form has succeeded.
struct KeygenCandidate {
  uint8_t sk_seed[32];          // the attempt's seed (sk bytes 0..31)
  std::vector<uint16_t> g;      // Goppa polynomial, t+1 coefficients (monic)
  std::vector<uint16_t> alpha;  // support, n field elements
  std::vector<uint16_t> pi;     // the permutation of 2^m elements behind the support
  std::vector<uint8_t> s;       // implicit-rejection string, n/8 bytes
  uint64_t pivots = 0xFFFFFFFFull;  // sk bytes 32..39 (the f variants' pivot mask)
  std::vector<uint16_t> f;      // keygen_step: the element of GF((2^m)^t) whose minimal polynomial is g
};
void keygen_candidate(int set, uint8_t seed33[33], KeygenCandidate &c);
// One attempt of the chain without the Goppa polynomial (computed elsewhere from c.f): seed
// expansion, f, support by sorting; advances seed33.  False when the support's sort keys repeat
// (the attempt fails whatever g is).
bool keygen_step(int set, uint8_t seed33[33], KeygenCandidate &c);
// the f variants' column step as the systematic form reported it (pivot mask over columns
// mt-32 .. mt+31): swaps positions mt-32+i and mt-32+pv[i] of the support permutation, i = 0..31
// in order, pv = the mask's set bits in increasing order
void keygen_apply_pivots(int set, KeygenCandidate &c, uint64_t pivots);
void keygen_finish(int set, const KeygenCandidate &c, uint8_t *sk);
// the same with the Benes control bits of c.pi computed elsewhere (cb: COND_BYTES bytes)
void keygen_finish_with_cb(int set, const KeygenCandidate &c, const uint8_t *cb, uint8_t *sk);

void shake256(uint8_t *out, long long outlen, const uint8_t *in, long long inlen);

// libmceliece's gen_e (LM:crypto_kem/<set>/vec/encrypt.c gen_e :41-92) drawing from `rnd` exactly
// as it draws from randombytes(): per attempt 4t bytes (2t for 8192128) read as little-endian
// m-bit values; the first t below n (8192128: all t); rejected when a value repeats.  pos[0..t) =
// the error positions, ascending.
void host_gen_e(int set, rand_fn rnd, void *ctx, uint16_t *pos);

}  // namespace tmce
