// kem_enc.hip -- batched Classic McEliece encapsulation on gfx950.
//
// Engine form (two launches):
//   k_gen_e    one op per lane; seeded fixed-weight sampling (SHAKE256 chain, range filter,
//              collision rejection) with a device-wide work queue so lanes whose attempt is
//              rejected retry while finished lanes take new ops.  Collision checks run
//              wave-cooperatively on an LDS bitmap; e and the support (in draw order -- only its
//              set matters downstream) leave through coalesced stores.
//   k_enc_finish one op per lane; c = XOR of t columns of [I | T] (sparse column gather from the
//              device-resident, column-major public key; L2 resident), then K = SHAKE256(1||e||c);
//              the pc variants append C = SHAKE256(2||e) to c first.
//
// Semantics: LM:crypto_kem/<set>/vec/encrypt.c (gen_e :41-92, syndrome :94-132) and kem_enc.c,
// with randombytes() replaced by the per-op seeded chain of the reference harness
// (mceliece-gpu-params/ref/348864.cuh:3434-3486): state = 0x40||seed, draws = SHAKE256(state),
// next state = first 32 bytes of the draws.
#include "kem_kernels.h"

namespace tmce {

constexpr int GEN_WG = 256;
constexpr int GEN_WAVES = GEN_WG / 64;

// Sampler.  Phase A (per lane): one SHAKE256 attempt of the lane's op; the first t draws below N
// go to a lane-private LDS list.  Phase B (per wave): the candidate ops of the wave are checked
// one after another by all 64 lanes -- each lane sets the bits of two of the op's indices in a
// per-wave LDS bitmap with ds_or and a returned bit already set means a repeated index (the
// predicate libmceliece evaluates by sorting and comparing neighbours).  An accepted op's bitmap
// is e: the wave writes it (shifted behind the 0x01 prefix byte of 1||e||c) and the support list
// to HBM with coalesced stores.  Rejected ops keep the lane and retry with the chained state.
template <int SET>
__global__ __launch_bounds__(GEN_WG) void k_gen_e(const uint8_t *__restrict__ seeds, uint16_t *__restrict__ ind_out,
                                                  uint32_t *__restrict__ msg, uint32_t *__restrict__ attempts,
                                                  int nops, unsigned *__restrict__ queue,
                                                  const PlanDev *__restrict__ plan) {
  using P = Par<SET>;
  constexpr uint32_t GFMASK = (1u << P::M) - 1u;
  constexpr int NEW = ceil_div(P::N, 32);  // 32-bit words of e
  __shared__ uint16_t lst[P::T * GEN_WG];  // [index][thread]
  __shared__ uint32_t bmp[GEN_WAVES][4][NEW + 1];
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid >> 6;
  const int grp = lane >> 4, g16 = lane & 15;  // phase B: four 16-lane groups
  uint32_t *bm = bmp[wave][grp];
  for (int i = lane; i < 4 * (NEW + 1); i += 64) bmp[wave][i / (NEW + 1)][i % (NEW + 1)] = 0;
  const KeccakRC rc;
  if (plan) nops = plan->total;

  int op = -1;
  bool done = false;
  uint32_t st[8];  // the 32-byte chained rng state
  uint32_t tries = 0;

  for (;;) {
    // ---- refill idle lanes from the device-wide queue
    const bool need = (op < 0) && !done;
    const uint64_t needmask = __ballot(need);
    if (needmask) {
      const int leader = __ffsll((long long)needmask) - 1;
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(queue, (unsigned)__popcll(needmask));
      base = (unsigned)__shfl((int)base, leader);
      if (need) {
        const unsigned r = base + mbcnt64(needmask);
        if (r < (unsigned)nops) {
          op = (int)r;
          tries = 0;
          const uint32_t *sp = reinterpret_cast<const uint32_t *>(seeds + 32 * (size_t)op);
#pragma unroll
          for (int i = 0; i < 8; ++i) st[i] = sp[i];
        } else {
          done = true;
        }
      }
    }
    if (!__any(op >= 0)) break;

    // ---- phase A: one attempt, draws = SHAKE256(0x40 || st, DRAW_BYTES)
    int count = 0;
    uint32_t nst[8];
    {
      tries += (op >= 0) ? 1u : 0u;
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
#pragma unroll
      for (int blk = 0; blk < P::DRAW_BLOCKS; ++blk) {
        keccak_f1600(A, rc);  // idle lanes compute too: the round constants need the full wave
        if (blk == 0) {
#pragma unroll
          for (int i = 0; i < 4; ++i) { nst[2 * i] = (uint32_t)A[i]; nst[2 * i + 1] = (uint32_t)(A[i] >> 32); }
        }
#pragma unroll
        for (int w = 0; w < 17; ++w)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int idx = blk * 68 + w * 4 + q;
            if (idx < P::NDRAW) {
              const uint32_t v = (uint32_t)(A[w] >> (16 * q)) & GFMASK;
              if (op >= 0 && v < (uint32_t)P::N && count < P::T) {
                lst[count * GEN_WG + tid] = (uint16_t)v;
                ++count;
              }
            }
          }
      }
    }
    wave_lds_sync();

    // ---- phase B: collision check of the wave's candidates, four at a time (one per 16-lane group)
    uint64_t cand = __ballot(op >= 0 && count == P::T);
    uint64_t okmask = 0;
    while (cand) {
      // candidate of this group: the grp-th lowest set bit (none if fewer remain)
      uint64_t c = cand;
      for (int k = 0; k < grp && c; ++k) c &= c - 1;
      const bool have = c != 0;
      const int j = have ? __ffsll((long long)c) - 1 : 0;
      for (int k = 0; k < 4 && cand; ++k) cand &= cand - 1;
      const int opj = __shfl(op, j);
      const uint16_t *lj = lst + (wave * 64 + j);
      bool coll = false;
      if (have) {
#pragma unroll
        for (int k = 0; k < ceil_div(P::T, 16); ++k) {
          const int idx = k * 16 + g16;
          if (idx < P::T) {
            const uint32_t v = lj[idx * GEN_WG];
            const uint32_t bit = 1u << (v & 31u);
            coll |= (atomicOr(bm + (v >> 5), bit) & bit) != 0;
          }
        }
      }
      const uint64_t cm = __ballot(coll);
      const bool ok = have && ((cm >> (16 * grp)) & 0xFFFFull) == 0;
      wave_lds_sync();
      if (ok) {
        // message row: the bitmap behind the prefix byte 1 (word k = bitmap word k << 8 | the top
        // byte of word k - 1), every read issued before the first store
        uint32_t *mrow = msg + (size_t)opj * P::MSG_WORDS;
        constexpr int MWL = ceil_div(NEW + 1, 16);
        uint32_t cw[MWL], pw[MWL];
#pragma unroll
        for (int i = 0; i < MWL; ++i) {
          const int k = 16 * i + g16;
          cw[i] = k < NEW ? bm[k] : 0u;
          pw[i] = k == 0 ? 0x01000000u : (k <= NEW ? bm[k - 1] : 0u);
        }
#pragma unroll
        for (int i = 0; i < MWL; ++i)
          if (16 * i + g16 <= NEW) mrow[16 * i + g16] = (cw[i] << 8) | (pw[i] >> 24);
        uint16_t *io = ind_out + (size_t)opj * P::IND_STRIDE;
        for (int idx = g16; idx < P::T; idx += 16) io[idx] = lj[idx * GEN_WG];
      }
      {
        // lanes g16 == 0 of accepting groups publish their candidate index
        const uint64_t acc = __ballot(ok && g16 == 0);
        for (int g = 0; g < 4; ++g)
          if ((acc >> (16 * g)) & 1ull) okmask |= 1ull << __shfl(j, 16 * g);
      }
      wave_lds_sync();
      if (have) {
#pragma unroll
        for (int k = 0; k < ceil_div(P::T, 16); ++k) {
          const int idx = k * 16 + g16;
          if (idx < P::T) bm[lj[idx * GEN_WG] >> 5] = 0;
        }
      }
      wave_lds_sync();
    }

    // ---- per-lane bookkeeping
    if (op >= 0) {
      if ((okmask >> lane) & 1ull) {
        if (attempts) attempts[op] = tries;
        op = -1;
      } else if (tries >= (uint32_t)P::MAX_ATTEMPTS) {
        ind_out[(size_t)op * P::IND_STRIDE] = 0xFFFFu;  // abandoned (never expected): status -2
        if (attempts) attempts[op] = tries;
        op = -1;
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) st[i] = nst[i];
      }
    }
  }
}

// Pipelined sampler.  An op's rejection chain is sequential only through the first SHAKE block of
// each attempt (attempt k+1's seed is the first 32 bytes of attempt k's first block), so a group of
// G >= B lanes works on one op and starts one attempt per step: at the group's step s, lane s mod G
// computes the first block of attempt s from its seed and hands the block's first 32 bytes -- the
// seed of attempt s + 1 -- to the next lane of the group; a lane that started attempt k squeezes
// its next B - 1 blocks in the following steps (one Keccak-f per lane and step) and collects the
// attempt's first t draws below n in its own list.  Attempts thus finish in order, one per step,
// B - 1 steps after their first block; a finished attempt with t draws goes through the collision
// check (as in k_gen_e), the first accepted one is the op's e (libmceliece's semantics), later
// speculative attempts are dropped and the group takes the next op.  The longest chain of a batch
// shrinks from (attempts x B) permutations to (attempts + B - 1).
#ifndef GEN_BITMAPS
#define GEN_BITMAPS 2
#endif
#ifndef GEN_G3
#define GEN_G3 1  // three-block attempts on three-lane groups (21 per wave, lane 63 idle)
#endif
// lanes per op of the pipelined sampler: at least the SHAKE blocks per attempt (B); two or four
// lanes keep the group inside a DPP quad, three leave every lane of the group busy every step for
// B = 3 (four-lane groups would leave one lane idle per step) at the cost of lane shuffles
constexpr int gen_group(int B) { return B == 2 ? 2 : (B == 3 && GEN_G3 ? 3 : 4); }
// DPP quad_perm: lane i of each group of four reads lane (CTRL >> 2 i) & 3 of the group
template <int CTRL>
__device__ __forceinline__ uint32_t quad_perm(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int SET>
__global__ __launch_bounds__(GEN_WG) void k_gen_e_pipe(const uint8_t *__restrict__ seeds,
                                                       uint16_t *__restrict__ ind_out, uint32_t *__restrict__ msg,
                                                       uint32_t *__restrict__ attempts, int nops,
                                                       unsigned *__restrict__ queue, const PlanDev *__restrict__ plan) {
  using P = Par<SET>;
  constexpr int B = P::DRAW_BLOCKS;
  constexpr int G = gen_group(B);  // lanes per op
  static_assert(B <= G, "a lane finishes its attempt before the rotation returns to it");
  constexpr uint32_t GFMASK = (1u << P::M) - 1u;
  constexpr int NEW = ceil_div(P::N, 32);
  constexpr int NBM = GEN_BITMAPS;                   // bitmaps (candidates per round) per 16-lane group
  constexpr int DPP_PREV = G == 2 ? 0xB1 : 0x93;    // quad_perm: lane r reads lane r - 1 of its group
  constexpr int DPP_GROUP0 = G == 2 ? 0xA0 : 0x00;  // quad_perm: each lane reads its group's lane 0
  (void)DPP_PREV;
  (void)DPP_GROUP0;
  constexpr int KV = ceil_div(P::T, 16);            // draws per lane of a candidate
  __shared__ uint16_t lst[(P::T + 1) * GEN_WG];     // [index][thread]; row T takes draws past the first T
  __shared__ uint32_t bmp[GEN_WAVES][4 * NBM][NEW + 1];
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid >> 6;
  const int grp = lane >> 4, g16 = lane & 15;  // collision check: four 16-lane groups
  const int r = lane % G, gl = lane - r;        // place in the op group
  const bool in_grp = gl + G <= 64;             // G = 3: lane 63 has no group and takes no op
  const uint64_t gmask = ((1ull << G) - 1ull) << gl;
  // lane r of a group reads lane r - 1 (mod G) of its group, or the group's lane 0
  const int src_prev = in_grp ? gl + (r == 0 ? G - 1 : r - 1) : lane, src_first = in_grp ? gl : lane;
  auto from_prev = [&](uint32_t v) -> uint32_t {
    if constexpr (G == 3) return (uint32_t)__shfl((int)v, src_prev);
    else return quad_perm<DPP_PREV>(v);
  };
  auto from_first = [&](uint32_t v) -> uint32_t {
    if constexpr (G == 3) return (uint32_t)__shfl((int)v, src_first);
    else return quad_perm<DPP_GROUP0>(v);
  };
  auto mod_g = [](int x) { return G == 3 ? (x + 3) % 3 : (x & (G - 1)); };  // x >= -(G - 1)
  (void)src_prev;
  (void)src_first;
  for (int i = lane; i < 4 * NBM * (NEW + 1); i += 64) bmp[wave][i / (NEW + 1)][i % (NEW + 1)] = 0;
  const KeccakRC rc;
  if (plan) nops = plan->total;
  // ops come from the queue as groups need them (workgroups that become resident late, behind a
  // grid larger than the chip holds, find it drained and leave)
  int op = -1;         // the group's op (every lane of the group holds it)
  bool done = false;   // the queue is exhausted (wave-uniform)
  int s = 0;           // the group's step in its op: attempt s starts now
  uint32_t st[8];      // seed of the attempt this lane starts next
  int count = 0;  // draws below N in this lane's list
  uint64_t A[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = 0;

  // draws of block `b` (0-based) of the lane's attempt, from the state A, appended to the lane's
  // list while it holds fewer than T (lanes with `on` false keep their count).  Branch-free: every
  // draw is written at the current count (row T is the spare) and the count advances only for a
  // draw below N that the list still takes.
  uint16_t *const col = lst + tid;
  auto extract = [&](int b, bool on) {
    const int cap = on ? P::T : 0;
    const int lim = P::NDRAW - 68 * b;  // draws in block b (the last block is partial)
#pragma unroll
    for (int w = 0; w < 17; ++w)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = 4 * w + q;
        const bool in_every_block = k < P::NDRAW - 68 * (B - 1);
        const uint32_t v = (uint32_t)(A[w] >> (16 * q)) & GFMASK;
        col[count * GEN_WG] = (uint16_t)v;
        count += (v < (uint32_t)P::N && count < cap && (in_every_block || k < lim)) ? 1 : 0;
      }
  };

  for (;;) {
    // ---- refill: group leaders without an op take the next ones from the queue (one atomic per
    //      wave and step); a group with a new op starts at step 0 with the op's seed in lane 0
    if (!done) {
      const bool need = r == 0 && op < 0 && in_grp;
      const uint64_t needmask = __ballot(need);
      if (needmask) {
        const int leader = __ffsll((long long)needmask) - 1;
        const unsigned cnt = (unsigned)__popcll(needmask);
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(queue, cnt);
        base = (unsigned)__builtin_amdgcn_readlane((int)base, leader);
        done = base + cnt >= (unsigned)nops;
        int nop = op;
        if (need) {
          const unsigned rr = base + mbcnt64(needmask);
          if (rr < (unsigned)nops) {
            nop = (int)rr;
            const uint32_t *sp = reinterpret_cast<const uint32_t *>(seeds + 32 * (size_t)rr);
#pragma unroll
            for (int i = 0; i < 8; ++i) st[i] = sp[i];
          }
        }
        const int gop = (int)from_first((uint32_t)nop);
        if (gop != op) s = 0;
        op = gop;
      }
    }
    if (!__any(op >= 0)) break;

    // ---- one Keccak-f per lane: this lane's attempt k = s - b, block b (b = 0: it starts attempt s)
    const int b = mod_g(s - r);
    const int k = s - b;
    const bool live = op >= 0 && k >= 0 && k < P::MAX_ATTEMPTS;
    const bool active = live && b < B;
    if (live && b == 0) {  // absorb 0x40 || seed (33 bytes) and the SHAKE256 padding
      const uint64_t s0 = st[0] | ((uint64_t)st[1] << 32), s1 = st[2] | ((uint64_t)st[3] << 32);
      const uint64_t s2 = st[4] | ((uint64_t)st[5] << 32), s3 = st[6] | ((uint64_t)st[7] << 32);
#pragma unroll
      for (int i = 0; i < 25; ++i) A[i] = 0;
      A[0] = 0x40ull | (s0 << 8);
      A[1] = (s0 >> 56) | (s1 << 8);
      A[2] = (s1 >> 56) | (s2 << 8);
      A[3] = (s2 >> 56) | (s3 << 8);
      A[4] = (s3 >> 56) | (0x1Full << 8);
      A[16] = 0x80ull << 56;
      count = 0;
    }
    keccak_f1600(A, rc);
    extract(b, active);
    const bool fin = active && b == B - 1;
    // the seed of attempt s + 1 (the first 32 bytes of attempt s's first block) to the next lane
    {
      uint32_t x[8];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        x[2 * i] = from_prev((uint32_t)A[i]);
        x[2 * i + 1] = from_prev((uint32_t)(A[i] >> 32));
      }
      if (mod_g(s + 1 - r) == 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) st[i] = x[i];
      }
    }
    wave_lds_sync();

    // ---- collision check of the attempts that finished this step, 4 NBM at a time (group g of 16
    //      lanes checks candidates g, g + 4, ... against its own bitmaps).  The candidates' lanes,
    //      ops and verdicts are kept in scalar registers; the LDS accesses of one wave execute in
    //      program order, so the bitmap reads of a message row see the group's atomics and the
    //      clears that follow them are seen by the next round's atomics without a wait.
    uint64_t cand = __ballot(fin && count == P::T);
    uint64_t okmask = 0;
    while (cand) {
      int jq[4 * NBM], opq[4 * NBM], nq = 0;
#pragma unroll
      for (int q = 0; q < 4 * NBM; ++q) {
        jq[q] = cand ? __ffsll((long long)cand) - 1 : 0;
        opq[q] = __builtin_amdgcn_readlane(op, jq[q]);
        nq += cand != 0;
        cand &= cand - 1;
      }
      uint32_t vv[NBM][KV];
      bool coll[NBM];
#pragma unroll
      for (int c = 0; c < NBM; ++c) {
        const int q0 = 4 * c;
        const int j = grp == 0 ? jq[q0] : grp == 1 ? jq[q0 + 1] : grp == 2 ? jq[q0 + 2] : jq[q0 + 3];
        const uint16_t *lj = lst + (wave * 64 + j);
#pragma unroll
        for (int kk = 0; kk < KV; ++kk) {
          const int idx = kk * 16 + g16;
          vv[c][kk] = idx < P::T ? (uint32_t)lj[idx * GEN_WG] : 0u;
        }
      }
#pragma unroll
      for (int c = 0; c < NBM; ++c) {
        coll[c] = false;
        uint32_t *bm = bmp[wave][4 * c + grp];
        if (4 * c + grp < nq) {
#pragma unroll
          for (int kk = 0; kk < KV; ++kk)
            if (kk * 16 + g16 < P::T) {
              const uint32_t bit = 1u << (vv[c][kk] & 31u);
              coll[c] |= (atomicOr(bm + (vv[c][kk] >> 5), bit) & bit) != 0;
            }
        }
      }
#pragma unroll
      for (int c = 0; c < NBM; ++c) {
        const uint64_t cm = __ballot(coll[c]);
#pragma unroll
        for (int g = 0; g < 4; ++g)
          if (4 * c + g < nq && ((cm >> (16 * g)) & 0xFFFFull) == 0) okmask |= 1ull << jq[4 * c + g];
        uint32_t *bm = bmp[wave][4 * c + grp];
        const int q = 4 * c + grp;
        if (q < nq && ((cm >> (16 * grp)) & 0xFFFFull) == 0) {
          const int opj = grp == 0 ? opq[4 * c] : grp == 1 ? opq[4 * c + 1] : grp == 2 ? opq[4 * c + 2] : opq[4 * c + 3];
          // message row: the bitmap behind the prefix byte 1 (word k = bitmap word k << 8 | the
          // top byte of word k - 1), every read issued before the first store
          uint32_t *mrow = msg + (size_t)opj * P::MSG_WORDS;
          constexpr int MWL = ceil_div(NEW + 1, 16);
          uint32_t cw[MWL], pw[MWL];
#pragma unroll
          for (int i = 0; i < MWL; ++i) {
            const int kw = 16 * i + g16;
            cw[i] = kw < NEW ? bm[kw] : 0u;
            pw[i] = kw == 0 ? 0x01000000u : (kw <= NEW ? bm[kw - 1] : 0u);
          }
#pragma unroll
          for (int i = 0; i < MWL; ++i)
            if (16 * i + g16 <= NEW) mrow[16 * i + g16] = (cw[i] << 8) | (pw[i] >> 24);
          uint16_t *io = ind_out + (size_t)opj * P::IND_STRIDE;
#pragma unroll
          for (int kk = 0; kk < KV; ++kk)
            if (kk * 16 + g16 < P::T) io[kk * 16 + g16] = (uint16_t)vv[c][kk];
        }
        if (q < nq) {
#pragma unroll
          for (int kk = 0; kk < KV; ++kk)
            if (kk * 16 + g16 < P::T) bm[vv[c][kk] >> 5] = 0;
        }
      }
    }
    wave_lds_sync();

    // ---- accepted or abandoned ops end; the group takes a new op next step
    const bool acc_me = (okmask >> lane) & 1ull;
    if (acc_me && attempts) attempts[op] = (uint32_t)(k + 1);
    const bool aband_me = fin && !acc_me && k == P::MAX_ATTEMPTS - 1;
    if (aband_me) {
      ind_out[(size_t)op * P::IND_STRIDE] = 0xFFFFu;  // never expected: status -2
      if (attempts) attempts[op] = P::MAX_ATTEMPTS;
    }
    const uint64_t endmask = okmask | __ballot(aband_me);
    if ((endmask & gmask) != 0) op = -1;
    else if (op >= 0) ++s;
  }
}

#ifndef FIN_COOP_MIN
#define FIN_COOP_MIN 960  // 16-byte column words per op above which the wave gathers cooperatively
#endif
#ifndef FIN_COOP_LOADS
// column loads in flight per round of the cooperative gather: 32 covers every set's op in one round
// (t columns, four per load instruction), so each op costs one L2 round trip instead of four
// (6960119 finish 0.267 -> 0.202 ms at N = 65536, 6688128 and 8192128 -7 to -8 %; 8 in flight before)
#define FIN_COOP_LOADS 32
#endif
template <int SET, bool PC>
__global__ __launch_bounds__(256) void k_enc_finish(const uint16_t *__restrict__ ind, const uint4 *__restrict__ cols,
                                                    uint32_t *__restrict__ msg, uint8_t *__restrict__ ct,
                                                    uint8_t *__restrict__ key, int *__restrict__ status, int nops,
                                                    const PlanDev *__restrict__ plan) {
  using P = Par<SET>;
  constexpr int C4 = P::CWP / 4;
  // Every lane stays active through the hash (the round constants are read with v_readlane):
  // lanes past the batch redo the last op and store nothing.
  const KeccakRC rc;
  if (plan) nops = plan->total;
  if ((int)(blockIdx.x * blockDim.x) >= nops) return;  // whole workgroup past the batch (uniform)
  const int gop = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = gop < nops;
  const int op = live ? gop : nops - 1;
  bool pad_bad = false;  // LM:crypto_kem/6960119/vec/kem_enc.c: ct and k cleared, return -1
  if (plan) {  // key groups are PLAN_ALIGN-aligned: one key per 256-op workgroup
    const KeyDev &kd = plan->keys[plan_group(plan, blockIdx.x * blockDim.x)];
    cols = uniform_ptr(reinterpret_cast<const uint4 *>(kd.cols));
    pad_bad = __builtin_amdgcn_readfirstlane(kd.pad_bad) != 0;
  }

  // ---- c = sum of columns over the support of e.  Long gathers (t columns of C4 16-byte words,
  //      t C4 > 960: the m = 13 sets of t >= 119) are done by the whole wave op after op: each span
  //      of SPN lanes loads one column (lane li of the span its word li), CPI columns per load
  //      instruction, coalesced; the spans' sums are combined across lanes and parked in LDS, from
  //      where each lane takes its own op's c for the hash below.  Short ones stay one op per lane
  //      (fewer instructions, and the key's columns stay in L2).
  constexpr bool COOP = C4 * P::T > FIN_COOP_MIN;
  constexpr int SPN = C4 <= 8 ? 8 : (C4 <= 16 ? 16 : 32);
  constexpr int CPI = 64 / SPN;
  constexpr int NJ = ceil_div(P::T, CPI);  // load instructions per op
  constexpr int FIN_LOADS = FIN_COOP_LOADS;  // load instructions in flight
  static_assert(C4 <= SPN, "a column fits its span");
  __shared__ uint4 cst[COOP ? 4 : 1][COOP ? 64 : 1][COOP ? C4 : 1];  // [wave][op of the wave][word]
  __shared__ uint16_t pst[COOP ? 4 : 1][round_up(P::T, 2) + 2];     // the support of the op gathered
  const uint32_t *ip = reinterpret_cast<const uint32_t *>(ind + (size_t)op * P::IND_STRIDE);
  const bool abandoned = (ip[0] & 0xFFFFu) == 0xFFFFu;  // sampler gave up on this op (never expected)
  uint4 c[C4];
  if constexpr (COOP) {
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int g = lane / SPN, li = lane % SPN, lq = li < C4 ? li : C4 - 1;
    const int wop0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63);
    const int jn = min(64, nops - wop0);  // the wave's ops in the batch (lanes past it store nothing)
#pragma unroll 1
    for (int jj = 0; jj < jn; ++jj) {
      const int opj = wop0 + jj;
      const uint32_t *ipj = reinterpret_cast<const uint32_t *>(ind + (size_t)opj * P::IND_STRIDE);
      for (int i = lane; i < P::IND_STRIDE / 2; i += 64) reinterpret_cast<uint32_t *>(pst[wv])[i] = ipj[i];
      wave_lds_sync();
      const bool ab = pst[wv][0] == 0xFFFFu;  // uniform
      uint4 acc = make_uint4(0, 0, 0, 0);
      if (!ab) {
#pragma unroll 1
        for (int i0 = 0; i0 < NJ; i0 += FIN_LOADS) {
          uint32_t pk[FIN_LOADS];
#pragma unroll
          for (int ii = 0; ii < FIN_LOADS; ++ii) {
            const int k = CPI * (i0 + ii) + g;
            pk[ii] = pst[wv][k < P::T ? k : P::T - 1];
          }
#pragma unroll
          for (int ii = 0; ii < FIN_LOADS; ++ii) {
            if (NJ % FIN_LOADS != 0 && i0 + ii >= NJ) break;
            const uint4 x = global_load16(cols, (size_t)pk[ii] * C4 + lq);
            // a span past the last position (t not a multiple of CPI) adds nothing
            const uint32_t keep = (P::T % CPI == 0 || CPI * (i0 + ii) + g < P::T) ? ~0u : 0u;
            acc.x ^= x.x & keep;
            acc.y ^= x.y & keep;
            acc.z ^= x.z & keep;
            acc.w ^= x.w & keep;
          }
        }
#pragma unroll
        for (int sh = SPN; sh < 64; sh *= 2) {
          acc.x ^= (uint32_t)__shfl_xor((int)acc.x, sh);
          acc.y ^= (uint32_t)__shfl_xor((int)acc.y, sh);
          acc.z ^= (uint32_t)__shfl_xor((int)acc.z, sh);
          acc.w ^= (uint32_t)__shfl_xor((int)acc.w, sh);
        }
      }
      if (lane < C4) cst[wv][jj][lane] = acc;
      wave_lds_sync();
    }
#pragma unroll
    for (int q = 0; q < C4; ++q) c[q] = cst[wv][lane][q];
  } else {
#pragma unroll
    for (int q = 0; q < C4; ++q) c[q] = make_uint4(0, 0, 0, 0);
    if (!abandoned) {
#pragma unroll 4
      for (int i = 0; i < P::T; i += 2) {
        const uint32_t pr = ip[i / 2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (i + h < P::T) {
            const uint32_t p = h ? (pr >> 16) : (pr & 0xFFFFu);
            const uint4 *col = cols + (size_t)p * C4;
#pragma unroll
            for (int q = 0; q < C4; ++q) {
              const uint4 x = global_load16(col, q);
              c[q].x ^= x.x; c[q].y ^= x.y; c[q].z ^= x.z; c[q].w ^= x.w;
            }
          }
        }
      }
    }
  }
  // ---- the ciphertext as words: c (SYND_BYTES; the bits past MT are zero), and for the pc
  // variants the confirmation C = SHAKE256(2 || e, 32) after it (LM:crypto_kem/<set>pc/vec/kem_enc.c:32)
  constexpr int CTB = PC ? P::CT_BYTES_PC : P::SYND_BYTES;  // even for every set
  constexpr int XW = PC ? ceil_div(CTB, 4) : P::CW;
  uint32_t *mrow = msg + (size_t)op * P::MSG_WORDS;  // [1][e] written by the sampler
  uint32_t cx[XW];
#pragma unroll
  for (int q = 0; q < C4; ++q) {
    const uint32_t w4[4] = {c[q].x, c[q].y, c[q].z, c[q].w};
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (4 * q + i < XW && 4 * q + i < P::CW) cx[4 * q + i] = w4[i];
  }
  if constexpr (PC) {
    uint64_t cf[4];
    shake256_prefix_32<P::CONF_LEN>(reinterpret_cast<const uint64_t *>(mrow), 2u, cf, rc);
    constexpr int SB4 = P::SYND_BYTES / 4, SO = P::SYND_BYTES % 4;  // SO is 0 or 2
    static_assert(SO == 0 || SO == 2, "ciphertext length");
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint32_t t = (uint32_t)(cf[i / 2] >> (32 * (i & 1)));
      if constexpr (SO == 0) {
        cx[SB4 + i] = t;
      } else {
        cx[SB4 + i] = (i == 0 ? (cx[SB4] & 0xFFFFu) : cx[SB4 + i]) | (t << 16);
        cx[SB4 + i + 1] = t >> 16;
      }
    }
  }

  // ---- ciphertext bytes
  uint16_t *ctp = reinterpret_cast<uint16_t *>(ct + (size_t)op * CTB);
  if (live) {
#pragma unroll
    for (int h = 0; h < CTB / 2; ++h) ctp[h] = (uint16_t)(cx[h / 2] >> (16 * (h & 1)));
  }

  // ---- finish the message: the ciphertext at byte 1+N/8, then SHAKE padding
  constexpr int MLEN = PC ? P::MSG_LEN_PC : P::MSG_LEN;
  constexpr int MBLK = PC ? P::MSG_BLOCKS_PC : P::MSG_BLOCKS;
  constexpr int OFF = 1 + P::N / 8;
  constexpr int KB = OFF / 4, RS = 8 * (OFF % 4);
  constexpr int PADW = MLEN / 4, PADS = 8 * (MLEN % 4);
  const uint32_t first = mrow[KB];
#pragma unroll
  for (int k = KB; k < 34 * MBLK; ++k) {
    const int j = k - KB;
    uint32_t v = 0;
    if (j < XW) v |= cx[j] << RS;
    if (RS != 0 && j >= 1 && j - 1 < XW) v |= cx[j - 1] >> ((32 - RS) & 31);
    if (k == KB) v |= first;
    if (k == PADW) v |= 0x1Fu << PADS;
    if (k == 34 * MBLK - 1) v |= 0x80u << 24;
    if (live) mrow[k] = v;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // this lane's message stores complete
  uint64_t out[4];
  // The next-block prefetch of the hash costs 34 VGPRs (128 instead of 90-95: four waves per SIMD
  // instead of five).  It pays where the column gather is short (348864: t = 64 columns of 24
  // words; 89 against 97 us at N = 65536) or done by the whole wave (the m = 13 sets with t >= 119:
  // 8192128 214 against 222 us), and loses where each lane's own long gather needs the occupancy
  // (460896: 96 columns of 40 words).
  constexpr bool PREFETCH = COOP || P::T * P::CWP <= 64 * 24;
  shake256_blocks_32<PREFETCH>(reinterpret_cast<const uint64_t *>(mrow), MBLK, out, rc);
  uint64_t *kp = reinterpret_cast<uint64_t *>(key + 32 * (size_t)op);
  if (live) {
    const bool clear = abandoned || pad_bad;
#pragma unroll
    for (int i = 0; i < 4; ++i) kp[i] = clear ? 0ull : out[i];
    if (clear) for (int h = 0; h < CTB / 2; ++h) ctp[h] = 0;
    if (status) status[op] = pad_bad ? -1 : abandoned ? -2 : 0;
  }
}

__global__ void k_fill_u32(unsigned *p, unsigned v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(unsigned *p, unsigned v, int n, hipStream_t s) {
  k_fill_u32<<<ceil_div(n, 256), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}

__global__ void k_fill_u8(uint8_t *p, uint8_t v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int *p, int v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
hipError_t launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s) {
  if (n) k_fill_u8<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_i32(int *p, int v, size_t n, hipStream_t s) {
  if (n) k_fill_i32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}

// ---- multi-batch gather / scatter: one thread per 16-bit unit of a row, adjacent threads on
// adjacent units, so each wave reads (writes) contiguous bytes of a sub-batch's rows -- in HBM, or
// across the host link when the sub-batch lies in a pinned host slab
__global__ void k_plan_gather(const PlanDev *__restrict__ plan, uint8_t *__restrict__ flat, int row_bytes) {
  const int H = row_bytes / 2;  // row_bytes even
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)plan->total * H) return;
  const int op = (int)(i / H), h = (int)(i - (size_t)op * H);
  const int sb = plan_sub_wave(plan, op);
  uint16_t v = 0;  // padding rows are zero
  if (sb >= 0) {
    const SubDev &d = plan->subs[sb];
    v = reinterpret_cast<const uint16_t *>(d.in + (size_t)(op - d.start) * row_bytes)[h];
  }
  reinterpret_cast<uint16_t *>(flat)[i] = v;
}
// units per op: the ciphertext's 16-bit halves (enc), at least two; units 0 and 1 also move the
// session key's two 16-byte halves, unit 0 the status
__global__ void k_plan_scatter(const PlanDev *__restrict__ plan, const uint8_t *__restrict__ ct, int ct_bytes,
                               const uint8_t *__restrict__ key, const int *__restrict__ status) {
  const int H = ct_bytes / 2, U = H > 2 ? H : 2;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)plan->total * U) return;
  const int op = (int)(i / U), u = (int)(i - (size_t)op * U);
  const int sb = plan_sub_wave(plan, op);
  if (sb < 0) return;
  const SubDev &d = plan->subs[sb];
  const int j = op - d.start;
  if (u < H)
    reinterpret_cast<uint16_t *>(d.ct + (size_t)j * ct_bytes)[u] =
        reinterpret_cast<const uint16_t *>(ct + (size_t)op * ct_bytes)[u];
  if (u < 2)
    reinterpret_cast<uint4 *>(d.key + 32 * (size_t)j)[u] = reinterpret_cast<const uint4 *>(key + 32 * (size_t)op)[u];
  if (u == 0 && d.status) d.status[j] = status[op];
}
hipError_t launch_plan_gather(const PlanDev *plan, uint8_t *flat, int row_bytes, int cap, hipStream_t s) {
  const size_t units = (size_t)cap * (row_bytes / 2);
  if (units > 0) k_plan_gather<<<(unsigned)((units + 255) / 256), 256, 0, s>>>(plan, flat, row_bytes);
  return hipGetLastError();
}
hipError_t launch_plan_scatter(const PlanDev *plan, const uint8_t *ct, int ct_bytes, const uint8_t *key,
                               const int *status, int cap, hipStream_t s) {
  const size_t units = (size_t)cap * (ct_bytes / 2 > 2 ? ct_bytes / 2 : 2);
  if (units > 0) k_plan_scatter<<<(unsigned)((units + 255) / 256), 256, 0, s>>>(plan, ct, ct_bytes, key, status);
  return hipGetLastError();
}

template <int SET>
hipError_t launch_enc(const EncArgs &a, hipStream_t s) {
  using P = Par<SET>;
  if (a.nops <= 0) return hipSuccess;
  k_fill_u32<<<1, 64, 0, s>>>(a.queue, 0u, 1);  // a kernel, not a memset: the pipeline may be graph-captured
  // persistent grid: enough resident waves to fill the chip, lanes (or lane groups) pull ops from the
  // queue
  if (a.gen_form == 1) {
    constexpr int G = gen_group(P::DRAW_BLOCKS);
    int grid = a.gen_blocks > 0 ? a.gen_blocks : 2048;
    const int maxgrid = ceil_div(a.nops, GEN_WAVES * (64 / G));  // 64 / G op groups per wave
    if (grid > maxgrid) grid = maxgrid;
    TMCE_TIMED("gen_e", s,
               k_gen_e_pipe<SET><<<grid, GEN_WG, 0, s>>>(a.seeds, a.ind, a.msg, a.attempts, a.nops, a.queue, a.plan));
  } else {
    int grid = a.gen_blocks > 0 ? a.gen_blocks : 1024;
    const int maxgrid = ceil_div(a.nops, GEN_WG);
    if (grid > maxgrid) grid = maxgrid;
    TMCE_TIMED("gen_e", s,
               k_gen_e<SET><<<grid, GEN_WG, 0, s>>>(a.seeds, a.ind, a.msg, a.attempts, a.nops, a.queue, a.plan));
  }
  if (a.pc)
    TMCE_TIMED("enc_finish", s,
               k_enc_finish<SET, true><<<ceil_div(a.nops, 256), 256, 0, s>>>(
                   a.ind, reinterpret_cast<const uint4 *>(a.cols), a.msg, a.ct, a.key, a.status, a.nops, a.plan));
  else
    TMCE_TIMED("enc_finish", s,
               k_enc_finish<SET, false><<<ceil_div(a.nops, 256), 256, 0, s>>>(
                   a.ind, reinterpret_cast<const uint4 *>(a.cols), a.msg, a.ct, a.key, a.status, a.nops, a.plan));
  return hipGetLastError();
}

template <int SET>
hipError_t enc_finish_only(const EncArgs &a, hipStream_t s) {
  if (a.nops <= 0) return hipSuccess;
  if (a.pc)
    k_enc_finish<SET, true><<<ceil_div(a.nops, 256), 256, 0, s>>>(a.ind, reinterpret_cast<const uint4 *>(a.cols),
                                                                  a.msg, a.ct, a.key, a.status, a.nops, a.plan);
  else
    k_enc_finish<SET, false><<<ceil_div(a.nops, 256), 256, 0, s>>>(a.ind, reinterpret_cast<const uint4 *>(a.cols),
                                                                   a.msg, a.ct, a.key, a.status, a.nops, a.plan);
  return hipGetLastError();
}

hipError_t launch_enc_finish_set(int set, const EncArgs &a, hipStream_t s) {
  switch (set) {
    case S348864: return enc_finish_only<S348864>(a, s);
    case S460896: return enc_finish_only<S460896>(a, s);
    case S6688128: return enc_finish_only<S6688128>(a, s);
    case S6960119: return enc_finish_only<S6960119>(a, s);
    case S8192128: return enc_finish_only<S8192128>(a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_enc_set(int set, const EncArgs &a, hipStream_t s) {
  switch (set) {
    case S348864: return launch_enc<S348864>(a, s);
    case S460896: return launch_enc<S460896>(a, s);
    case S6688128: return launch_enc<S6688128>(a, s);
    case S6960119: return launch_enc<S6960119>(a, s);
    case S8192128: return launch_enc<S8192128>(a, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace tmce
