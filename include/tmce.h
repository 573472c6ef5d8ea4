/*
 * tmce.h -- C ABI of libtmce.so, the MI355X (gfx950) batched Classic McEliece KEM engine.
 *
 * Two layers:
 *  1. The libmceliece-20250507 crypto_kem interface for the five parameter sets and their f, pc
 *     and pcf variants (LM:api, 20 in all), same names,
 *     argument meaning, sizes and return conventions, so a program linked against libmceliece
 *     can link against this library instead:
 *       mceliece_kem_<set>_keypair / _enc / _dec / _implementation / _compiler
 *     (LM:doc/api.md:1-115; symbol names from LM:hdef; reference callers
 *     mceliece-gpu-params/main.cu:17-21,159-167 and mceliece-gpu-bmma/main.cu:60,139).
 *     Randomness comes from randombytes() exactly as in libmceliece: keypair draws 32 bytes;
 *     enc draws 4t bytes per attempt of gen_e (2t for 8192128; LM:crypto_kem/<set>/vec/
 *     encrypt.c:41-92) on the host and runs the rest on the GPU, so any randomness source --
 *     libmceliece's test driver's knownrandombytes included -- gives libmceliece's bytes.  Install
 *     a source with tmce_set_randombytes() (default: getrandom()).
 *  2. Batched extensions (no counterpart in libmceliece; the reference harness has no exported
 *     API -- its per-stage kernels are launched from mceliece-gpu-params/main.cu:213-262
 *     (encapsulation) and :326-370 (decapsulation)):
 *       device-resident key handles, enc/dec over n operations from host or device buffers.
 *     Encapsulation is seeded per operation: op i uses the chain
 *       state = 0x40 || seed_i;  each randombytes(buf,len) returns SHAKE256(state, len) and
 *       then sets state[1..32] = buf[0..31]
 *     (the reference's GPU gen_e, mceliece-gpu-params/ref/348864.cuh:3434-3486), so its output
 *     is byte-identical to libmceliece enc() run with that randombytes source.
 *
 * Errors: functions return 0 on success; negative TMCE_E* codes for API misuse or device
 * failure (tmce_last_error() describes the last one).  Per-operation KEM results go to the
 * status array with libmceliece's meaning (0, or -1 for the 6960119 padding rules).
 * Thread safety: calls may come from several host threads; they are serialised per device.
 */
#ifndef TMCE_H
#define TMCE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- sizes (LM:api) */
/* the five parameter sets and their variants: f = semi-systematic key generation (same sizes),
   pc = plaintext confirmation (32 more ciphertext bytes) */
#define mceliece_kem_348864_PUBLICKEYBYTES 261120
#define mceliece_kem_348864_SECRETKEYBYTES 6492
#define mceliece_kem_348864_CIPHERTEXTBYTES 96
#define mceliece_kem_348864_BYTES 32
#define mceliece_kem_348864f_PUBLICKEYBYTES 261120
#define mceliece_kem_348864f_SECRETKEYBYTES 6492
#define mceliece_kem_348864f_CIPHERTEXTBYTES 96
#define mceliece_kem_348864f_BYTES 32
#define mceliece_kem_348864pc_PUBLICKEYBYTES 261120
#define mceliece_kem_348864pc_SECRETKEYBYTES 6492
#define mceliece_kem_348864pc_CIPHERTEXTBYTES 128
#define mceliece_kem_348864pc_BYTES 32
#define mceliece_kem_348864pcf_PUBLICKEYBYTES 261120
#define mceliece_kem_348864pcf_SECRETKEYBYTES 6492
#define mceliece_kem_348864pcf_CIPHERTEXTBYTES 128
#define mceliece_kem_348864pcf_BYTES 32
#define mceliece_kem_460896_PUBLICKEYBYTES 524160
#define mceliece_kem_460896_SECRETKEYBYTES 13608
#define mceliece_kem_460896_CIPHERTEXTBYTES 156
#define mceliece_kem_460896_BYTES 32
#define mceliece_kem_460896f_PUBLICKEYBYTES 524160
#define mceliece_kem_460896f_SECRETKEYBYTES 13608
#define mceliece_kem_460896f_CIPHERTEXTBYTES 156
#define mceliece_kem_460896f_BYTES 32
#define mceliece_kem_460896pc_PUBLICKEYBYTES 524160
#define mceliece_kem_460896pc_SECRETKEYBYTES 13608
#define mceliece_kem_460896pc_CIPHERTEXTBYTES 188
#define mceliece_kem_460896pc_BYTES 32
#define mceliece_kem_460896pcf_PUBLICKEYBYTES 524160
#define mceliece_kem_460896pcf_SECRETKEYBYTES 13608
#define mceliece_kem_460896pcf_CIPHERTEXTBYTES 188
#define mceliece_kem_460896pcf_BYTES 32
#define mceliece_kem_6688128_PUBLICKEYBYTES 1044992
#define mceliece_kem_6688128_SECRETKEYBYTES 13932
#define mceliece_kem_6688128_CIPHERTEXTBYTES 208
#define mceliece_kem_6688128_BYTES 32
#define mceliece_kem_6688128f_PUBLICKEYBYTES 1044992
#define mceliece_kem_6688128f_SECRETKEYBYTES 13932
#define mceliece_kem_6688128f_CIPHERTEXTBYTES 208
#define mceliece_kem_6688128f_BYTES 32
#define mceliece_kem_6688128pc_PUBLICKEYBYTES 1044992
#define mceliece_kem_6688128pc_SECRETKEYBYTES 13932
#define mceliece_kem_6688128pc_CIPHERTEXTBYTES 240
#define mceliece_kem_6688128pc_BYTES 32
#define mceliece_kem_6688128pcf_PUBLICKEYBYTES 1044992
#define mceliece_kem_6688128pcf_SECRETKEYBYTES 13932
#define mceliece_kem_6688128pcf_CIPHERTEXTBYTES 240
#define mceliece_kem_6688128pcf_BYTES 32
#define mceliece_kem_6960119_PUBLICKEYBYTES 1047319
#define mceliece_kem_6960119_SECRETKEYBYTES 13948
#define mceliece_kem_6960119_CIPHERTEXTBYTES 194
#define mceliece_kem_6960119_BYTES 32
#define mceliece_kem_6960119f_PUBLICKEYBYTES 1047319
#define mceliece_kem_6960119f_SECRETKEYBYTES 13948
#define mceliece_kem_6960119f_CIPHERTEXTBYTES 194
#define mceliece_kem_6960119f_BYTES 32
#define mceliece_kem_6960119pc_PUBLICKEYBYTES 1047319
#define mceliece_kem_6960119pc_SECRETKEYBYTES 13948
#define mceliece_kem_6960119pc_CIPHERTEXTBYTES 226
#define mceliece_kem_6960119pc_BYTES 32
#define mceliece_kem_6960119pcf_PUBLICKEYBYTES 1047319
#define mceliece_kem_6960119pcf_SECRETKEYBYTES 13948
#define mceliece_kem_6960119pcf_CIPHERTEXTBYTES 226
#define mceliece_kem_6960119pcf_BYTES 32
#define mceliece_kem_8192128_PUBLICKEYBYTES 1357824
#define mceliece_kem_8192128_SECRETKEYBYTES 14120
#define mceliece_kem_8192128_CIPHERTEXTBYTES 208
#define mceliece_kem_8192128_BYTES 32
#define mceliece_kem_8192128f_PUBLICKEYBYTES 1357824
#define mceliece_kem_8192128f_SECRETKEYBYTES 14120
#define mceliece_kem_8192128f_CIPHERTEXTBYTES 208
#define mceliece_kem_8192128f_BYTES 32
#define mceliece_kem_8192128pc_PUBLICKEYBYTES 1357824
#define mceliece_kem_8192128pc_SECRETKEYBYTES 14120
#define mceliece_kem_8192128pc_CIPHERTEXTBYTES 240
#define mceliece_kem_8192128pc_BYTES 32
#define mceliece_kem_8192128pcf_PUBLICKEYBYTES 1357824
#define mceliece_kem_8192128pcf_SECRETKEYBYTES 14120
#define mceliece_kem_8192128pcf_CIPHERTEXTBYTES 240
#define mceliece_kem_8192128pcf_BYTES 32

/* ---------------------------------------------------------------- layer 1: libmceliece API */
#define TMCE_DECLARE_SET(S)                                                                     \
  void mceliece_kem_##S##_keypair(unsigned char *pk, unsigned char *sk);                       \
  int mceliece_kem_##S##_enc(unsigned char *c, unsigned char *k, const unsigned char *pk);     \
  int mceliece_kem_##S##_dec(unsigned char *k, const unsigned char *c, const unsigned char *sk); \
  const char *mceliece_kem_##S##_implementation(void);                                          \
  const char *mceliece_kem_##S##_compiler(void);                                                \
  /* implementation dispatch, as libmceliece's scripts-build/dispatch generates it: one         \
     implementation here; impl < 0 or out of range selects it as well */                         \
  void (*mceliece_dispatch_kem_##S##_keypair(long long impl))(unsigned char *, unsigned char *);  \
  int (*mceliece_dispatch_kem_##S##_enc(long long impl))(unsigned char *, unsigned char *,        \
                                                          const unsigned char *);                \
  int (*mceliece_dispatch_kem_##S##_dec(long long impl))(unsigned char *, const unsigned char *,  \
                                                          const unsigned char *);                \
  const char *mceliece_dispatch_kem_##S##_implementation(long long impl);                       \
  const char *mceliece_dispatch_kem_##S##_compiler(long long impl);                             \
  long long mceliece_numimpl_kem_##S(void);

/* LM:doc/api.md:39-116, one block per parameter set and variant.
   enc/dec return what libmceliece returns (0, or -1 under the 6960119 padding rules).  A device
   failure, which libmceliece cannot have (no gfx950 device, an allocation or launch failing),
   returns its negative TMCE_E* code instead, with the outputs cleared (ct and k for enc, k for
   dec) and tmce_last_error() set; the process is never aborted. */
TMCE_DECLARE_SET(348864)
TMCE_DECLARE_SET(348864f)
TMCE_DECLARE_SET(348864pc)
TMCE_DECLARE_SET(348864pcf)
TMCE_DECLARE_SET(460896)
TMCE_DECLARE_SET(460896f)
TMCE_DECLARE_SET(460896pc)
TMCE_DECLARE_SET(460896pcf)
TMCE_DECLARE_SET(6688128)
TMCE_DECLARE_SET(6688128f)
TMCE_DECLARE_SET(6688128pc)
TMCE_DECLARE_SET(6688128pcf)
TMCE_DECLARE_SET(6960119)
TMCE_DECLARE_SET(6960119f)
TMCE_DECLARE_SET(6960119pc)
TMCE_DECLARE_SET(6960119pcf)
TMCE_DECLARE_SET(8192128)
TMCE_DECLARE_SET(8192128f)
TMCE_DECLARE_SET(8192128pc)
TMCE_DECLARE_SET(8192128pcf)

/* the short names of LM:hdef:118-140 (mceliece<set>_keypair = mceliece_kem_<set>_keypair, ...) */
#define mceliece348864_keypair mceliece_kem_348864_keypair
#define mceliece348864_enc mceliece_kem_348864_enc
#define mceliece348864_dec mceliece_kem_348864_dec
#define mceliece348864_SECRETKEYBYTES mceliece_kem_348864_SECRETKEYBYTES
#define mceliece348864_PUBLICKEYBYTES mceliece_kem_348864_PUBLICKEYBYTES
#define mceliece348864_CIPHERTEXTBYTES mceliece_kem_348864_CIPHERTEXTBYTES
#define mceliece348864_BYTES mceliece_kem_348864_BYTES
#define mceliece348864f_keypair mceliece_kem_348864f_keypair
#define mceliece348864f_enc mceliece_kem_348864f_enc
#define mceliece348864f_dec mceliece_kem_348864f_dec
#define mceliece348864f_SECRETKEYBYTES mceliece_kem_348864f_SECRETKEYBYTES
#define mceliece348864f_PUBLICKEYBYTES mceliece_kem_348864f_PUBLICKEYBYTES
#define mceliece348864f_CIPHERTEXTBYTES mceliece_kem_348864f_CIPHERTEXTBYTES
#define mceliece348864f_BYTES mceliece_kem_348864f_BYTES
#define mceliece348864pc_keypair mceliece_kem_348864pc_keypair
#define mceliece348864pc_enc mceliece_kem_348864pc_enc
#define mceliece348864pc_dec mceliece_kem_348864pc_dec
#define mceliece348864pc_SECRETKEYBYTES mceliece_kem_348864pc_SECRETKEYBYTES
#define mceliece348864pc_PUBLICKEYBYTES mceliece_kem_348864pc_PUBLICKEYBYTES
#define mceliece348864pc_CIPHERTEXTBYTES mceliece_kem_348864pc_CIPHERTEXTBYTES
#define mceliece348864pc_BYTES mceliece_kem_348864pc_BYTES
#define mceliece348864pcf_keypair mceliece_kem_348864pcf_keypair
#define mceliece348864pcf_enc mceliece_kem_348864pcf_enc
#define mceliece348864pcf_dec mceliece_kem_348864pcf_dec
#define mceliece348864pcf_SECRETKEYBYTES mceliece_kem_348864pcf_SECRETKEYBYTES
#define mceliece348864pcf_PUBLICKEYBYTES mceliece_kem_348864pcf_PUBLICKEYBYTES
#define mceliece348864pcf_CIPHERTEXTBYTES mceliece_kem_348864pcf_CIPHERTEXTBYTES
#define mceliece348864pcf_BYTES mceliece_kem_348864pcf_BYTES
#define mceliece460896_keypair mceliece_kem_460896_keypair
#define mceliece460896_enc mceliece_kem_460896_enc
#define mceliece460896_dec mceliece_kem_460896_dec
#define mceliece460896_SECRETKEYBYTES mceliece_kem_460896_SECRETKEYBYTES
#define mceliece460896_PUBLICKEYBYTES mceliece_kem_460896_PUBLICKEYBYTES
#define mceliece460896_CIPHERTEXTBYTES mceliece_kem_460896_CIPHERTEXTBYTES
#define mceliece460896_BYTES mceliece_kem_460896_BYTES
#define mceliece460896f_keypair mceliece_kem_460896f_keypair
#define mceliece460896f_enc mceliece_kem_460896f_enc
#define mceliece460896f_dec mceliece_kem_460896f_dec
#define mceliece460896f_SECRETKEYBYTES mceliece_kem_460896f_SECRETKEYBYTES
#define mceliece460896f_PUBLICKEYBYTES mceliece_kem_460896f_PUBLICKEYBYTES
#define mceliece460896f_CIPHERTEXTBYTES mceliece_kem_460896f_CIPHERTEXTBYTES
#define mceliece460896f_BYTES mceliece_kem_460896f_BYTES
#define mceliece460896pc_keypair mceliece_kem_460896pc_keypair
#define mceliece460896pc_enc mceliece_kem_460896pc_enc
#define mceliece460896pc_dec mceliece_kem_460896pc_dec
#define mceliece460896pc_SECRETKEYBYTES mceliece_kem_460896pc_SECRETKEYBYTES
#define mceliece460896pc_PUBLICKEYBYTES mceliece_kem_460896pc_PUBLICKEYBYTES
#define mceliece460896pc_CIPHERTEXTBYTES mceliece_kem_460896pc_CIPHERTEXTBYTES
#define mceliece460896pc_BYTES mceliece_kem_460896pc_BYTES
#define mceliece460896pcf_keypair mceliece_kem_460896pcf_keypair
#define mceliece460896pcf_enc mceliece_kem_460896pcf_enc
#define mceliece460896pcf_dec mceliece_kem_460896pcf_dec
#define mceliece460896pcf_SECRETKEYBYTES mceliece_kem_460896pcf_SECRETKEYBYTES
#define mceliece460896pcf_PUBLICKEYBYTES mceliece_kem_460896pcf_PUBLICKEYBYTES
#define mceliece460896pcf_CIPHERTEXTBYTES mceliece_kem_460896pcf_CIPHERTEXTBYTES
#define mceliece460896pcf_BYTES mceliece_kem_460896pcf_BYTES
#define mceliece6688128_keypair mceliece_kem_6688128_keypair
#define mceliece6688128_enc mceliece_kem_6688128_enc
#define mceliece6688128_dec mceliece_kem_6688128_dec
#define mceliece6688128_SECRETKEYBYTES mceliece_kem_6688128_SECRETKEYBYTES
#define mceliece6688128_PUBLICKEYBYTES mceliece_kem_6688128_PUBLICKEYBYTES
#define mceliece6688128_CIPHERTEXTBYTES mceliece_kem_6688128_CIPHERTEXTBYTES
#define mceliece6688128_BYTES mceliece_kem_6688128_BYTES
#define mceliece6688128f_keypair mceliece_kem_6688128f_keypair
#define mceliece6688128f_enc mceliece_kem_6688128f_enc
#define mceliece6688128f_dec mceliece_kem_6688128f_dec
#define mceliece6688128f_SECRETKEYBYTES mceliece_kem_6688128f_SECRETKEYBYTES
#define mceliece6688128f_PUBLICKEYBYTES mceliece_kem_6688128f_PUBLICKEYBYTES
#define mceliece6688128f_CIPHERTEXTBYTES mceliece_kem_6688128f_CIPHERTEXTBYTES
#define mceliece6688128f_BYTES mceliece_kem_6688128f_BYTES
#define mceliece6688128pc_keypair mceliece_kem_6688128pc_keypair
#define mceliece6688128pc_enc mceliece_kem_6688128pc_enc
#define mceliece6688128pc_dec mceliece_kem_6688128pc_dec
#define mceliece6688128pc_SECRETKEYBYTES mceliece_kem_6688128pc_SECRETKEYBYTES
#define mceliece6688128pc_PUBLICKEYBYTES mceliece_kem_6688128pc_PUBLICKEYBYTES
#define mceliece6688128pc_CIPHERTEXTBYTES mceliece_kem_6688128pc_CIPHERTEXTBYTES
#define mceliece6688128pc_BYTES mceliece_kem_6688128pc_BYTES
#define mceliece6688128pcf_keypair mceliece_kem_6688128pcf_keypair
#define mceliece6688128pcf_enc mceliece_kem_6688128pcf_enc
#define mceliece6688128pcf_dec mceliece_kem_6688128pcf_dec
#define mceliece6688128pcf_SECRETKEYBYTES mceliece_kem_6688128pcf_SECRETKEYBYTES
#define mceliece6688128pcf_PUBLICKEYBYTES mceliece_kem_6688128pcf_PUBLICKEYBYTES
#define mceliece6688128pcf_CIPHERTEXTBYTES mceliece_kem_6688128pcf_CIPHERTEXTBYTES
#define mceliece6688128pcf_BYTES mceliece_kem_6688128pcf_BYTES
#define mceliece6960119_keypair mceliece_kem_6960119_keypair
#define mceliece6960119_enc mceliece_kem_6960119_enc
#define mceliece6960119_dec mceliece_kem_6960119_dec
#define mceliece6960119_SECRETKEYBYTES mceliece_kem_6960119_SECRETKEYBYTES
#define mceliece6960119_PUBLICKEYBYTES mceliece_kem_6960119_PUBLICKEYBYTES
#define mceliece6960119_CIPHERTEXTBYTES mceliece_kem_6960119_CIPHERTEXTBYTES
#define mceliece6960119_BYTES mceliece_kem_6960119_BYTES
#define mceliece6960119f_keypair mceliece_kem_6960119f_keypair
#define mceliece6960119f_enc mceliece_kem_6960119f_enc
#define mceliece6960119f_dec mceliece_kem_6960119f_dec
#define mceliece6960119f_SECRETKEYBYTES mceliece_kem_6960119f_SECRETKEYBYTES
#define mceliece6960119f_PUBLICKEYBYTES mceliece_kem_6960119f_PUBLICKEYBYTES
#define mceliece6960119f_CIPHERTEXTBYTES mceliece_kem_6960119f_CIPHERTEXTBYTES
#define mceliece6960119f_BYTES mceliece_kem_6960119f_BYTES
#define mceliece6960119pc_keypair mceliece_kem_6960119pc_keypair
#define mceliece6960119pc_enc mceliece_kem_6960119pc_enc
#define mceliece6960119pc_dec mceliece_kem_6960119pc_dec
#define mceliece6960119pc_SECRETKEYBYTES mceliece_kem_6960119pc_SECRETKEYBYTES
#define mceliece6960119pc_PUBLICKEYBYTES mceliece_kem_6960119pc_PUBLICKEYBYTES
#define mceliece6960119pc_CIPHERTEXTBYTES mceliece_kem_6960119pc_CIPHERTEXTBYTES
#define mceliece6960119pc_BYTES mceliece_kem_6960119pc_BYTES
#define mceliece6960119pcf_keypair mceliece_kem_6960119pcf_keypair
#define mceliece6960119pcf_enc mceliece_kem_6960119pcf_enc
#define mceliece6960119pcf_dec mceliece_kem_6960119pcf_dec
#define mceliece6960119pcf_SECRETKEYBYTES mceliece_kem_6960119pcf_SECRETKEYBYTES
#define mceliece6960119pcf_PUBLICKEYBYTES mceliece_kem_6960119pcf_PUBLICKEYBYTES
#define mceliece6960119pcf_CIPHERTEXTBYTES mceliece_kem_6960119pcf_CIPHERTEXTBYTES
#define mceliece6960119pcf_BYTES mceliece_kem_6960119pcf_BYTES
#define mceliece8192128_keypair mceliece_kem_8192128_keypair
#define mceliece8192128_enc mceliece_kem_8192128_enc
#define mceliece8192128_dec mceliece_kem_8192128_dec
#define mceliece8192128_SECRETKEYBYTES mceliece_kem_8192128_SECRETKEYBYTES
#define mceliece8192128_PUBLICKEYBYTES mceliece_kem_8192128_PUBLICKEYBYTES
#define mceliece8192128_CIPHERTEXTBYTES mceliece_kem_8192128_CIPHERTEXTBYTES
#define mceliece8192128_BYTES mceliece_kem_8192128_BYTES
#define mceliece8192128f_keypair mceliece_kem_8192128f_keypair
#define mceliece8192128f_enc mceliece_kem_8192128f_enc
#define mceliece8192128f_dec mceliece_kem_8192128f_dec
#define mceliece8192128f_SECRETKEYBYTES mceliece_kem_8192128f_SECRETKEYBYTES
#define mceliece8192128f_PUBLICKEYBYTES mceliece_kem_8192128f_PUBLICKEYBYTES
#define mceliece8192128f_CIPHERTEXTBYTES mceliece_kem_8192128f_CIPHERTEXTBYTES
#define mceliece8192128f_BYTES mceliece_kem_8192128f_BYTES
#define mceliece8192128pc_keypair mceliece_kem_8192128pc_keypair
#define mceliece8192128pc_enc mceliece_kem_8192128pc_enc
#define mceliece8192128pc_dec mceliece_kem_8192128pc_dec
#define mceliece8192128pc_SECRETKEYBYTES mceliece_kem_8192128pc_SECRETKEYBYTES
#define mceliece8192128pc_PUBLICKEYBYTES mceliece_kem_8192128pc_PUBLICKEYBYTES
#define mceliece8192128pc_CIPHERTEXTBYTES mceliece_kem_8192128pc_CIPHERTEXTBYTES
#define mceliece8192128pc_BYTES mceliece_kem_8192128pc_BYTES
#define mceliece8192128pcf_keypair mceliece_kem_8192128pcf_keypair
#define mceliece8192128pcf_enc mceliece_kem_8192128pcf_enc
#define mceliece8192128pcf_dec mceliece_kem_8192128pcf_dec
#define mceliece8192128pcf_SECRETKEYBYTES mceliece_kem_8192128pcf_SECRETKEYBYTES
#define mceliece8192128pcf_PUBLICKEYBYTES mceliece_kem_8192128pcf_PUBLICKEYBYTES
#define mceliece8192128pcf_CIPHERTEXTBYTES mceliece_kem_8192128pcf_CIPHERTEXTBYTES
#define mceliece8192128pcf_BYTES mceliece_kem_8192128pcf_BYTES

/* randomness source of keypair/enc (replaces librandombytes' randombytes(),
   librandombytes include/randombytes.h:10); NULL restores getrandom() */
typedef void (*tmce_randombytes_fn)(void *out, long long len);
void tmce_set_randombytes(tmce_randombytes_fn fn);

/* ---------------------------------------------------------------- layer 2: batched engine */
enum {
  TMCE_SET_348864 = 0,
  TMCE_SET_460896 = 1,
  TMCE_SET_6688128 = 2,
  TMCE_SET_6960119 = 3,
  TMCE_SET_8192128 = 4
};
/* Variant flags, or-ed into a set id wherever a function below takes `set` (key handles, keypairs,
   pipelines, servers): TMCE_VARIANT_F selects the f variants' semi-systematic key generation
   (LM:crypto_kem/<set>f/vec/pk_gen.c; encapsulation and decapsulation are the base set's), and
   TMCE_VARIANT_PC the pc variants' plaintext confirmation (LM:crypto_kem/<set>pc/vec/kem_enc.c,
   kem_dec.c): ciphertexts are CIPHERTEXTBYTES = SYND_BYTES + 32 bytes.  Example: 6960119pcf is
   TMCE_SET_6960119 | TMCE_VARIANT_PC | TMCE_VARIANT_F.  The plain per-stage form (tmce_set_form)
   covers the variants without TMCE_VARIANT_PC. */
enum { TMCE_VARIANT_F = 8, TMCE_VARIANT_PC = 16 };

enum {
  TMCE_OK = 0,
  TMCE_EINVAL = -22,   /* bad argument (set id, null pointer, key of the wrong kind) */
  TMCE_ENODEV = -19,   /* no usable gfx950 device */
  TMCE_ENOMEM = -12,   /* device allocation failed */
  TMCE_EDEVICE = -5    /* a HIP call or kernel launch failed */
};

typedef struct tmce_key tmce_key;

/* initialise the engine on a device (implicit on first use with device 0) */
int tmce_init(int device);
const char *tmce_last_error(void);

/* device-resident key material: from a public key (encapsulation) or a secret key
   (decapsulation); pk/sk byte layouts as libmceliece */
int tmce_key_from_pk(int set, const unsigned char *pk, tmce_key **out);
int tmce_key_from_sk(int set, const unsigned char *sk, tmce_key **out);
void tmce_key_free(tmce_key *key);
/* bytes of device memory held by a key handle (for admission control) */
size_t tmce_key_device_bytes(const tmce_key *key);

/* seeded keypair (host): the keypair uses randombytes = the seeded chain above on seed[0..32) */
int tmce_keypair_seeded(int set, unsigned char *pk, unsigned char *sk, const unsigned char *seed);

/* batched seeded keypairs on the GPU: keypair i = tmce_keypair_seeded(set, pk_i, sk_i, seeds[32*i]),
   byte for byte.  Replaces n calls of libmceliece's crypto_kem_keypair
   (LM:crypto_kem/<set>/vec/kem_keypair.c; the reference generates its keys with it on the CPU,
   mceliece-gpu-params/main.cu:17-21,159-161).  pk: n*PUBLICKEYBYTES, sk: n*SECRETKEYBYTES.
   set may be any variant (the f variants' semi-systematic form included).  The Goppa polynomial
   (sk_gen.c), the systematic form (pk_gen.c) and the Benes control bits (controlbits.c) run on the
   GPU; seed expansion and the support sort on host threads.  Synchronous; concurrent calls are
   serialised (they share a 64 MB pinned staging buffer that the library allocates on the first
   call and keeps). */
int tmce_keypair_batch(int set, unsigned char *pk, unsigned char *sk, const unsigned char *seeds, size_t n);

/* batched KEM over host buffers: ct[n*CIPHERTEXTBYTES], k[n*32], status[n] (may be NULL),
   seeds[n*32] (encapsulation seeds, see above); synchronous */
int tmce_enc_batch(tmce_key *pk, unsigned char *ct, unsigned char *k, int *status,
                   const unsigned char *seeds, size_t n);
int tmce_dec_batch(tmce_key *sk, unsigned char *k, int *status, const unsigned char *ct, size_t n);

/* the same over device buffers, asynchronous on `stream` (a hipStream_t; NULL = the engine's
   stream); inputs resident in HBM.  The engine keeps one scratch set for encapsulations (and
   syndromes) and one for decapsulations: a call on a stream other than the set's previous user
   first waits for that use on the device, so calls of one direction never overlap, while an
   encapsulation and a decapsulation on two streams may run concurrently. */
int tmce_enc_batch_device(tmce_key *pk, unsigned char *d_ct, unsigned char *d_k, int *d_status,
                          const unsigned char *d_seeds, size_t n, void *stream);
int tmce_dec_batch_device(tmce_key *sk, unsigned char *d_k, int *d_status, const unsigned char *d_ct,
                          size_t n, void *stream);

/* batched syndrome (PKE encryption of given error vectors): ct_i = H e_i = e_i[0:mt) + T e_i[mt:n)
   over GF(2), bits >= mt of the last byte zero.  Replaces libmceliece's internal syndrome()
   (LM:crypto_kem/<set>/vec/encrypt.c:94-132) and the reference's batched syndrome GEMM + finalize
   (mceliece-gpu-bmma/tc_kernels.cuh:97-147, int8_tensor_enc.h:60-102).  e: n vectors of n_bits/8
   bytes in libmceliece's layout, any weight; ct: n * CIPHERTEXTBYTES.  Needs a public-key handle. */
enum {
  TMCE_SYND_SPARSE = 0, /* XOR of the columns of [I|T] at the set bits of e (engine default) */
  TMCE_SYND_MFMA = 1,   /* dense GF(2) GEMM on INT8 MFMA (bit in bit 0 of each byte), int32 accumulate, parity */
  TMCE_SYND_VALU = 2    /* dense bit-packed GEMM: 32-bit AND + XOR, popcount parity */
};
int tmce_syndrome_batch(tmce_key *pk, unsigned char *ct, const unsigned char *e, size_t n, int method);
int tmce_syndrome_batch_device(tmce_key *pk, unsigned char *d_ct, const unsigned char *d_e, size_t n,
                               int method, void *stream);

/* tuning knobs per parameter set (set by the autotuner, tensormceliece_amd.autotune):
   TMCE_TUNE_SYND_MFMA_CFG  tile shape of the MFMA syndrome GEMM (0..7)
   TMCE_TUNE_GEN_BLOCKS     persistent grid of the sampler (0 = built-in default)
   TMCE_TUNE_DECODE_BLOCKS  grid of the decoder (0 = built-in default)
   TMCE_TUNE_FRONT_MFMA     decapsulation front end S = F c: 1 = INT8 MFMA GEMM (default), 0 = VALU
   TMCE_TUNE_FRONT_CFG      tile shape of the MFMA front end (0..7)
   TMCE_TUNE_GEN_FORM       sampler: 1 = pipelined lane groups (default), 0 = one op per lane
   TMCE_TUNE_FFT_FORM       decoder's bottom-up additive FFT: 1 = bitsliced GF(2^m) arithmetic
                            (default), 0 = log/exp table lookups in LDS
   tmce_tune_get returns the value (>= 0) or a negative error code. */
enum {
  TMCE_TUNE_SYND_MFMA_CFG = 0,
  TMCE_TUNE_GEN_BLOCKS = 1,
  TMCE_TUNE_DECODE_BLOCKS = 2,
  TMCE_TUNE_FRONT_MFMA = 3,
  TMCE_TUNE_FRONT_CFG = 4,
  TMCE_TUNE_GEN_FORM = 5,
  TMCE_TUNE_FFT_FORM = 6
};
int tmce_tune_set(int set, int param, int value);
int tmce_tune_get(int set, int param);

/* ---------------------------------------------------------------- multi-batch pipelines
   Sub-batches of mixed sizes, under up to 16 distinct keys of one parameter set, run through one
   set of launches (no counterpart in the reference, whose harness launches one fixed-size batch at
   a time, mceliece-gpu-params/main.cu:213-262,326-370).  A pipeline has a fixed set, direction
   (0 = encapsulation, 1 = decapsulation) and capacity in ops (each distinct key's ops are padded to
   a multiple of 256); with use_graph the launch sequence is captured into a HIP graph at the second
   run and replayed afterwards.  All buffers are device buffers; runs are asynchronous on the
   pipeline's stream (its own non-blocking stream when `stream` is NULL); a run may start while the
   previous one is still executing on that stream (they serialise on it).  Alignment: k 16 bytes,
   status 4 bytes, in and ct 2 bytes. */
typedef struct tmce_sub_batch {
  tmce_key *key;            /* public key (encapsulation) or secret key (decapsulation) handle */
  const unsigned char *in;  /* seeds [n*32] (enc) or ciphertexts [n*CIPHERTEXTBYTES] (dec) */
  unsigned char *ct;        /* enc: ciphertexts out [n*CIPHERTEXTBYTES]; dec: unused */
  unsigned char *k;         /* session keys out [n*32] */
  int *status;              /* per-op status out [n], or NULL */
  size_t n;
} tmce_sub_batch;
typedef struct tmce_pipeline tmce_pipeline;
int tmce_pipeline_create(int set, int direction, size_t capacity, size_t max_sub_batches, int use_graph,
                         void *stream, tmce_pipeline **out);
int tmce_pipeline_run(tmce_pipeline *p, const tmce_sub_batch *subs, size_t nsub);
int tmce_pipeline_sync(tmce_pipeline *p);
void tmce_pipeline_free(tmce_pipeline *p);
size_t tmce_pipeline_device_bytes(const tmce_pipeline *p); /* device memory it holds */
size_t tmce_pipeline_capacity(const tmce_pipeline *p);     /* ops (rounded up to 256) */
void *tmce_pipeline_stream(tmce_pipeline *p);              /* its hipStream_t */

/* ---------------------------------------------------------------- serving loop
   Continuous batching of encapsulation / decapsulation requests of one parameter set (no
   counterpart in the reference).  A worker thread keeps `slots_per_direction` pipelines per
   direction busy, each on its own stream (so batches overlap on the device); pending requests join
   the next free slot FIFO while the batch fits the slot (ops after each key's padding to 256) and
   its device footprint fits `hbm_budget` bytes beside the batches in flight (0 = no limit).
   Requests use host buffers and key handles, which the caller keeps valid (and the keys
   unfreed) until the request is collected.
   tmce_server_wait blocks until a request completes and collects it; tmce_server_poll collects up
   to `max` completed requests.  A request being waited for is never returned by a concurrent
   poll, and a second concurrent wait for it fails.  Latency = completion - submission (host
   steady clock).  The worker learns of completions from a host function enqueued behind each
   batch on its slot's stream (no polling of the device).
   Zero-copy requests: when every buffer of a request (in, k, status if given, ct for
   encapsulation) lies inside memory from tmce_host_alloc, aligned as the pipelines require (k 16
   bytes, status 4, in and ct 2), the GPU reads the inputs and writes the outputs there in place,
   and the request skips the slot's staging (host copy, upload, download); other requests are
   staged.  Results are the same either way. */
typedef struct tmce_server tmce_server;
typedef struct tmce_request {
  int direction;            /* 0 = encapsulation, 1 = decapsulation */
  tmce_key *key;            /* public (enc) or secret (dec) key handle of the server's set */
  size_t n;                 /* ops: 1 .. slot_capacity - 255 */
  const unsigned char *in;  /* seeds [n*32] (enc) or ciphertexts [n*CIPHERTEXTBYTES] (dec) */
  unsigned char *ct;        /* enc: ciphertexts out; dec: unused */
  unsigned char *k;         /* session keys out [n*32] */
  int *status;              /* per-op status out [n], or NULL */
} tmce_request;
typedef struct tmce_server_stats {
  unsigned long long batches, requests, ops, pending;
  size_t in_flight_bytes, peak_in_flight_bytes, enc_bytes_per_op, dec_bytes_per_op;
  unsigned long long zero_copy_requests;  /* requests served in place from tmce_host_alloc slabs */
  unsigned long long max_batch_keys;      /* most distinct keys in one batch (at most 16) */
} tmce_server_stats;
int tmce_server_create(int set, size_t slot_capacity, int slots_per_direction, size_t hbm_budget,
                       tmce_server **out);
int tmce_server_submit(tmce_server *sv, const tmce_request *rq, unsigned long long *id);
int tmce_server_wait(tmce_server *sv, unsigned long long id, double *latency_s);
int tmce_server_poll(tmce_server *sv, unsigned long long *ids, double *latency_s, int max);
int tmce_server_get_stats(tmce_server *sv, tmce_server_stats *st);
void tmce_server_free(tmce_server *sv);
/* Open-loop load (no counterpart in the reference): submits rqs[i] at arrival_s[i] seconds after
   the call starts (arrivals sorted; a native thread sleeps, then spins to each arrival), collects
   every request as it completes and returns when all have: latency_s[i] = completion - submission
   of request i (may be NULL), *elapsed_s = last completion - start.  These requests are neither
   polled nor waited for by the other calls.  On an error the requests not yet collected go back
   to tmce_server_poll. */
int tmce_server_run_load(tmce_server *sv, const tmce_request *rqs, const double *arrival_s, size_t n,
                         double *latency_s, double *elapsed_s);

/* Decapsulation that allocates nothing (no counterpart in the reference: it is for callers that
   cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc).  Same arguments, semantics and results as tmce_dec_batch_device (engine
   form); every buffer beyond the inputs and outputs lives in a static device array of
   TMCE_NOALLOC_SCRATCH_MB MiB that each GPU holds until the process ends, so n is limited to
   TMCE_NOALLOC_MAX_OPS operations for every set and variant (larger n: TMCE_EINVAL).  All calls
   on one GPU share that array: calls that overlap on one GPU (several streams or host threads)
   race and must be ordered by the caller.  Calls no hipMalloc, hipFree or other allocating
   routine; the key handle and the library's set-up (tmce_init) allocate beforehand. */
#define TMCE_NOALLOC_SCRATCH_MB 40
#define TMCE_NOALLOC_MAX_OPS 16384
int tmce_dec_batch_device_noalloc(tmce_key *sk, unsigned char *d_k, int *d_status, const unsigned char *d_ct,
                                  size_t n, void *stream);

/* select the pipeline used by subsequent batched calls: 0 = engine form (default: fused,
   persistent sampler, row-parallel decoder), 1 = plain one-launch-per-stage form (the baseline
   mirroring the reference pipelines, mceliece-gpu-params/main.cu:213-262,326-370).  Both give
   byte-identical results. */
int tmce_set_form(int form);

/* device-memory helpers for callers without a HIP runtime binding */
void *tmce_dev_alloc(size_t bytes);
void tmce_dev_free(void *p);
/* pinned host memory mapped into the GPU's address space (a slab for zero-copy serving requests,
   see the serving loop); NULL on failure.  Free a slab only when no request using it is in flight. */
void *tmce_host_alloc(size_t bytes);
int tmce_host_free(void *p);
int tmce_memcpy_h2d(void *d, const void *h, size_t bytes);
int tmce_memcpy_d2h(void *h, const void *d, size_t bytes);
int tmce_memcpy_d2d(void *dst, const void *src, size_t bytes);
int tmce_sync(void);

/* streams and events for the asynchronous device entry points (callers without a HIP runtime
   binding): streams are non-blocking at the default priority; tmce_stream_wait_event makes later
   work on `stream` wait for the work recorded in `event` */
int tmce_stream_create(void **stream);
int tmce_stream_destroy(void *stream);
int tmce_stream_sync(void *stream);
int tmce_event_create(void **event);
int tmce_event_destroy(void *event);
int tmce_event_record(void *event, void *stream);
int tmce_stream_wait_event(void *stream, void *event);

/* kernel timing: while enabled, every kernel launch of the engine is bracketed by HIP events on
   its stream; tmce_timing_read() returns per-kernel launch counts and total milliseconds */
int tmce_timing_enable(int on);
int tmce_timing_read(int max, char *names /* max*48 bytes */, long *counts, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif /* TMCE_H */
