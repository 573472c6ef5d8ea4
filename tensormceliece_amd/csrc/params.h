// params.h -- compile-time description of the five Classic McEliece parameter sets and of the
// device-side data layouts derived from them.
//
// Base parameters follow libmceliece-20250507 (LM:crypto_kem/<set>/vec/params.h, gf_params.h;
// sizes as in LM:api); everything below "layouts" is this library's own choice.
#pragma once
#include <cstdint>

namespace tmce {

enum SetId : int { S348864 = 0, S460896 = 1, S6688128 = 2, S6960119 = 3, S8192128 = 4, NUM_SETS = 5 };

constexpr int ceil_div(int a, int b) { return (a + b - 1) / b; }
constexpr int round_up(int a, int b) { return ceil_div(a, b) * b; }
constexpr int next_pow2(int a) { int p = 1; while (p < a) p <<= 1; return p; }
constexpr int ilog2(int a) { int l = 0; while ((1 << l) < a) ++l; return l; }

template <int SET> struct SetBase;
template <> struct SetBase<S348864>  { static constexpr int M = 12, N = 3488, T = 64;  };
template <> struct SetBase<S460896>  { static constexpr int M = 13, N = 4608, T = 96;  };
template <> struct SetBase<S6688128> { static constexpr int M = 13, N = 6688, T = 128; };
template <> struct SetBase<S6960119> { static constexpr int M = 13, N = 6960, T = 119; };
template <> struct SetBase<S8192128> { static constexpr int M = 13, N = 8192, T = 128; };

template <int SET> struct Par : SetBase<SET> {
  using B = SetBase<SET>;
  static constexpr int id = SET;
  static constexpr int M = B::M, N = B::N, T = B::T;
  static constexpr int Q = (1 << M) - 1;            // multiplicative group order
  static constexpr int MT = M * T;                  // PK_NROWS
  static constexpr int K = N - MT;                  // PK_NCOLS
  static constexpr int PK_ROW_BYTES = (K + 7) / 8;
  static constexpr int SYND_BYTES = (MT + 7) / 8;   // ciphertext bytes
  static constexpr int IRR_BYTES = 2 * T;
  static constexpr int COND_BYTES = (1 << (M - 4)) * (2 * M - 1);
  static constexpr long PK_BYTES = (long)MT * PK_ROW_BYTES;
  static constexpr long SK_BYTES = 40L + IRR_BYTES + COND_BYTES + N / 8;
  static constexpr int E_BYTES = N / 8;

  // ---- layouts (this library) ----
  // gen_e: draws per attempt (the 8192128 code draws t values, the others 2t) and sort width
  static constexpr int NDRAW = (SET == S8192128) ? T : 2 * T;
  static constexpr int DRAW_BYTES = 2 * NDRAW;
  static constexpr int DRAW_BLOCKS = ceil_div(DRAW_BYTES, 136);
  static constexpr int TP = next_pow2(T);            // sort width of the plain per-thread sampler
  static constexpr int IND_STRIDE = round_up(T, 2);  // u16 entries per op of the support list
  static constexpr int MAX_ATTEMPTS = 256;           // P(256 rejections) < 2^-128 for every set
  // syndrome columns: column p (0 <= p < N) is the unit vector e_p for p < MT and column p-MT of
  // the public key T otherwise; CWP 32-bit words per column (multiple of 4 for 16-byte loads)
  static constexpr int CW = ceil_div(MT, 32);
  static constexpr int CWP = round_up(CW, 4);
  // message buffer of the session-key hash: [prefix][N/8 bytes][ciphertext][SHAKE padding]; the
  // ciphertext is SYND_BYTES long, or SYND_BYTES + 32 with plaintext confirmation (the pc
  // variants, LM:crypto_kem/<set>pc/vec/kem_enc.c), and rows are sized for the longer message
  static constexpr int MSG_LEN = 1 + N / 8 + SYND_BYTES;
  static constexpr int MSG_BLOCKS = MSG_LEN / 136 + 1;
  static constexpr int CT_BYTES_PC = SYND_BYTES + 32;
  static constexpr int MSG_LEN_PC = MSG_LEN + 32;
  static constexpr int MSG_BLOCKS_PC = MSG_LEN_PC / 136 + 1;
  static constexpr int MSG_WORDS = MSG_BLOCKS_PC * 34;
  // the pc variants' confirmation hash SHAKE256(2 || e) covers CONF_LEN bytes
  static constexpr int CONF_LEN = 1 + N / 8;
  // decapsulation front end: x = ciphertext bits (KBITS, incl. padding bits of the last byte);
  // output rows r = j*M + b hold bit b of syndrome j (syndrome-major; j < TT = 2t rounded up to 32,
  // rows of j >= 2t are zero)
  static constexpr int KBITS = 8 * SYND_BYTES;
  static constexpr int KW = ceil_div(SYND_BYTES, 4);
  static constexpr int KWP = round_up(KW, 4);
  static constexpr int TT = round_up(2 * T, 32);
  static constexpr int R = M * TT;
  static constexpr int RW = R / 32;
  static constexpr int FCW = round_up(RW, 4);  // words per column of the re-encryption's syndrome map
  // additive FFT root search: D coefficients (the x^t term of a monic degree-64/128 locator is
  // added from a power table), KD top-down levels
  static constexpr int D = (T == 64) ? 64 : 128;
  static constexpr int KD = ilog2(D);
  static constexpr bool TOP_TERM = (T == D);
  static constexpr int BM_SLOTS = ceil_div(T + 1, 64);
};

struct SetInfo {
  const char *name;
  int m, n, t, mt, k, pk_row_bytes, synd_bytes, cond_bytes;
  long pk_bytes, sk_bytes;
  int msg_words;  // 32-bit words per session-key message row
};

template <int SET> constexpr SetInfo make_info(const char *name) {
  using P = Par<SET>;
  return SetInfo{name, P::M, P::N, P::T, P::MT, P::K, P::PK_ROW_BYTES, P::SYND_BYTES, P::COND_BYTES, P::PK_BYTES,
                 P::SK_BYTES, P::MSG_WORDS};
}

inline const SetInfo &set_info(int set) {
  static const SetInfo tab[NUM_SETS] = {
      make_info<S348864>("348864"), make_info<S460896>("460896"), make_info<S6688128>("6688128"),
      make_info<S6960119>("6960119"), make_info<S8192128>("8192128")};
  return tab[set];
}

// Variant ids (TMCE_VARIANT_F / TMCE_VARIANT_PC of the C ABI): a base set | VAR_F (semi-systematic
// key generation, LM:crypto_kem/<set>f) | VAR_PC (plaintext confirmation, LM:crypto_kem/<set>pc).
// Encapsulation and decapsulation of an f variant are those of its base set.
constexpr int VAR_F = 8, VAR_PC = 16;
inline bool variant_valid(int v) { return v >= 0 && (v & 7) < NUM_SETS && (v & ~(7 | VAR_F | VAR_PC)) == 0; }
inline int variant_base(int v) { return v & 7; }
inline bool variant_f(int v) { return (v & VAR_F) != 0; }
inline bool variant_pc(int v) { return (v & VAR_PC) != 0; }
inline int variant_ct_bytes(int v) { return set_info(variant_base(v)).synd_bytes + (variant_pc(v) ? 32 : 0); }

}  // namespace tmce
