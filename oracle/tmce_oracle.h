/*
 * tmce_oracle.h -- CPU oracle for the Classic McEliece KEM (libmceliece-20250507 semantics).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or calls this code;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as the
 * checker (or as the CPU timing baseline).  It is a plain-C restatement of the algorithm that
 * libmceliece-20250507 (deps/libmceliece-20250507.tar.gz in the reference checkout) implements
 * for the five parameter sets and their "f" (semi-systematic key generation) and "pc" (plaintext
 * confirmation) variants, written from the specification and the library's behaviour, favouring
 * obviously-correct code over speed.
 *
 * Parity pins (see tests/test_oracle.py):
 *   - whole KEM: the KAT checksums of libmceliece's own test driver
 *     (LM:command/mceliece-test.c:4567-4570 etc.), replayed through oracle_kat_checksum();
 *   - SHAKE256: FIPS-202 (hashlib.shake_256);
 *   - GF(2^m): brute-force polynomial multiplication modulo the field polynomial;
 *   - support/control bits: the Benes network applied to the identity reproduces the permutation.
 */
#ifndef TMCE_ORACLE_H
#define TMCE_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* parameter-set ids; a variant id is a set id | ORC_F and/or | ORC_PC (348864f = 0 | ORC_F, ...).
   Every function below that takes `set` accepts a variant id. */
enum { ORC_348864 = 0, ORC_460896 = 1, ORC_6688128 = 2, ORC_6960119 = 3, ORC_8192128 = 4 };
enum { ORC_F = 8, ORC_PC = 16 };

typedef struct {
  const char *name;
  int m, n, t;
  int mt, k;            /* PK_NROWS, PK_NCOLS */
  int pk_row_bytes;
  int synd_bytes;       /* = ciphertext bytes of the non-pc variants */
  int irr_bytes;
  int cond_bytes;
  long pk_bytes, sk_bytes;
} orc_params;

const orc_params *oracle_params(int set); /* the base set's parameters; NULL for a bad id */
int oracle_ct_bytes(int set);              /* SYND_BYTES, + 32 for a pc variant */

/* randomness source used by keypair/enc (thread-local) */
void oracle_rng_seeded(const uint8_t seed[32]); /* state = 0x40||seed; out = SHAKE256(state,len); state[1..32] = out[0..31] */
void oracle_rng_known_reset(void);              /* libmceliece test driver's "knownrandombytes" (ChaCha20 crypto_rng) */
void oracle_randombytes(uint8_t *out, long long len);

/* the three KEM operations; byte layouts as libmceliece */
void oracle_keypair(int set, uint8_t *pk, uint8_t *sk);
int oracle_enc(int set, uint8_t *ct, uint8_t *key, const uint8_t *pk);
int oracle_dec(int set, uint8_t *key, const uint8_t *ct, const uint8_t *sk);

/* seeded conveniences: keypair/enc with the seeded adapter (state reset from seed first) */
void oracle_keypair_seeded(int set, uint8_t *pk, uint8_t *sk, const uint8_t seed[32]);
int oracle_enc_seeded(int set, uint8_t *ct, uint8_t *key, const uint8_t *pk, const uint8_t seed[32]);

/* batch helpers used by tests and by the CPU baseline (nthreads <= 0: all online cores) */
int oracle_enc_batch(int set, uint8_t *ct, uint8_t *key, int *ret, const uint8_t *pk,
                     const uint8_t *seeds, long n, int nthreads);
int oracle_dec_batch(int set, uint8_t *key, int *ret, const uint8_t *ct, const uint8_t *sk,
                     long n, int nthreads);
/* c_i = e_i[0:mt) + T e_i[mt:n) (LM:crypto_kem/<set>/vec/encrypt.c syndrome), e: n/8 bytes each */
int oracle_syndrome_batch(int set, uint8_t *ct, const uint8_t *pk, const uint8_t *e, long n, int nthreads);

/* stage-level functions (for stage parity tests) */
void oracle_shake256(uint8_t *out, long long outlen, const uint8_t *in, long long inlen);

/* word-level CPU port (the timed CPU baseline; same results as oracle_enc_seeded / oracle_dec):
   a key handle from pk and/or sk (kept by pointer, not copied), batches on nthreads threads
   (0 = every core) */
void *oracle_fast_key(int set, const uint8_t *pk, const uint8_t *sk);
void oracle_fast_key_free(void *h);
int oracle_fast_enc_batch(void *h, uint8_t *ct, uint8_t *key, int *ret, const uint8_t *seeds, long n, int nthreads);
int oracle_fast_dec_batch(void *h, uint8_t *key, int *ret, const uint8_t *ct, long n, int nthreads);
void oracle_fast_eval_all(int m, const uint16_t *f, int deg, uint16_t *vals); /* deg < 128 */
uint16_t oracle_gf_mul(int m, uint16_t a, uint16_t b);
uint16_t oracle_gf_inv(int m, uint16_t a);
void oracle_gen_e(int set, uint8_t *e, int *attempts); /* uses the current rng */
void oracle_support(int set, const uint8_t *sk, uint16_t *alpha); /* alpha[0..n) */
void oracle_perm_from_controlbits(int m, const uint8_t *cb, int16_t *pi);
void oracle_controlbits(int m, const int16_t *pi, uint8_t *cb);
/* s[0..2t) = sum_i r_i alpha_i^j / g(alpha_i)^2 with r = ct bits (zero-extended to n) */
void oracle_synd_ct(int set, const uint8_t *sk, const uint8_t *ct, uint16_t *s);
/* Berlekamp-Massey: locator[0..t] (reversed connection polynomial, roots = error positions) */
void oracle_bm(int set, const uint16_t *s, uint16_t *locator);
/* e bits (n/8 bytes); returns 0 on success, 1 on failure (libmceliece decrypt()) */
int oracle_decrypt(int set, const uint8_t *sk, const uint8_t *ct, uint8_t *e);
/* pk from sk (systematic form from the Goppa polynomial and support stored in sk); 0 = ok */
int oracle_pk_from_sk(int set, const uint8_t *sk, uint8_t *pk);

/* libmceliece test-driver KAT checksum (loops = 8 or 64); writes 64 hex chars + NUL */
void oracle_kat_checksum(int set, int loops, char *hex_out);
/* the same driver loop over another implementation's libmceliece-style entry points (e.g. a
   library under test); the driver's knownrandombytes is served by oracle_randombytes, which the
   implementation must use as its randombytes() */
typedef void (*orc_keypair_fn)(uint8_t *pk, uint8_t *sk);
typedef int (*orc_enc_fn)(uint8_t *ct, uint8_t *key, const uint8_t *pk);
typedef int (*orc_dec_fn)(uint8_t *key, const uint8_t *ct, const uint8_t *sk);
void oracle_kat_checksum_fns(int set, int loops, char *hex_out, orc_keypair_fn kp, orc_enc_fn enc,
                             orc_dec_fn dec);
/* libmceliece test-driver xof_shake256 checksum (64/128 or 512/16384) */
void oracle_kat_shake_checksum(int loops, int maxtest, char *hex_out);

#ifdef __cplusplus
}
#endif
#endif
