// keygen.cpp -- tmce_keypair_batch: batched seeded key generation.
//
// Host threads run each key's attempts up to the Goppa polynomial (keygen_step: seed expansion, f,
// support by sorting) and write its secret key after the attempt succeeds.  The GPU computes the
// Goppa polynomial of every pending attempt (k_genpoly), the systematic form (k_pk_gen) and, for the
// keys done, the Benes control bits (k_controlbits), one workgroup per attempt or key.  A key whose candidate is
// singular continues its seed chain in the next round, as libmceliece's keypair loop does, so key
// i is byte-identical to tmce_keypair_seeded on seeds[32 i].  For the f variants the kernel also
// runs the semi-systematic column step and reports its pivot mask, which the host applies to the
// candidate's support permutation before the control bits.
#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "engine.h"

using namespace tmce;

namespace {

// fn(i) for i < n on the CPUs this process may run on
template <class Fn>
void parallel_for(size_t n, Fn fn) {
  cpu_set_t cs;
  size_t nt = sched_getaffinity(0, sizeof(cs), &cs) == 0 ? (size_t)CPU_COUNT(&cs) : 1;
  nt = std::max<size_t>(1, std::min(nt, n));
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
  };
  std::vector<std::thread> th;
  for (size_t t = 1; t < nt; ++t) th.emplace_back(work);
  work();
  for (auto &x : th) x.join();
}

// device buffers of one call (freed on every exit path)
struct KgBuffers {
  uint16_t *g = nullptr, *alpha = nullptr, *fv = nullptr;
  int *valid = nullptr;
  uint64_t *H = nullptr;
  uint8_t *pk = nullptr;
  int *ok = nullptr;
  uint64_t *piv = nullptr;
  uint16_t *pi = nullptr;  // support permutations of the keys whose control bits are computed
  uint8_t *cb = nullptr;   // their control bits
  hipStream_t s = nullptr;
  ~KgBuffers() {
    if (s) (void)hipStreamSynchronize(s);
    (void)hipFree(g);
    (void)hipFree(fv);
    (void)hipFree(valid);
    (void)hipFree(alpha);
    (void)hipFree(H);
    (void)hipFree(pk);
    (void)hipFree(ok);
    (void)hipFree(piv);
    (void)hipFree(pi);
    (void)hipFree(cb);
    if (s) (void)hipStreamDestroy(s);
  }
};

// candidates per launch (H of a candidate is mt * ceil(n/64) words: 0.34 MB for 348864, 1.7 MB for
// 8192128), and attempts of a key's chain evaluated per round: only ~29 % of candidates have a
// non-singular systematic block, so each round tries the next `ahead` attempts of every pending key
// (enough to fill a launch, 2..8) and keeps the first one in chain order that succeeds --
// libmceliece's outcome, in fewer rounds.  A launch takes about as long for one candidate per CU
// as for one candidate in all, so filling it costs little.
constexpr size_t KG_LAUNCH = 512;
constexpr size_t KG_CB_KEYS = 1024;  // control-bit workgroups per launch
constexpr size_t KG_AHEAD_MIN = 2, KG_AHEAD_MAX = 8;
// pinned staging for the winners' public keys: a download into pageable memory runs at a few GB/s
// (tens of milliseconds for a round's keys), into pinned memory at the link's rate
constexpr size_t KG_STAGE_BYTES = size_t(64) << 20;

}  // namespace

int tmce_keypair_batch(int variant, unsigned char *pk, unsigned char *sk, const unsigned char *seeds, size_t n) {
  if (!variant_valid(variant) || (n && (!pk || !sk || !seeds)))
    return fail(TMCE_EINVAL, "tmce_keypair_batch: bad argument");
  if (n == 0) return 0;
  const int set = variant_base(variant);
  const bool f = variant_f(variant);
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  const SetInfo &si = set_info(set);
  const size_t T1 = (size_t)si.t + 1, N = (size_t)si.n, PKB = (size_t)si.pk_bytes, SKB = (size_t)si.sk_bytes;
  const size_t hw = pk_gen_h_bytes(set) / 8;
  const int w = si.m - 12;

  // each key's chain: the keypair's seed is randombytes(32) of the seeded chain on seeds[32 i]
  // (state = 0x40 || seed; out = SHAKE256(state, 32)), preceded by the domain byte 64
  std::vector<std::array<uint8_t, 33>> chain(n);
  for (size_t i = 0; i < n; ++i) {
    uint8_t st[33];
    st[0] = 0x40;
    std::memcpy(st + 1, seeds + 32 * i, 32);
    chain[i][0] = 64;
    shake256(chain[i].data() + 1, 32, st, 33);
  }

  std::lock_guard<std::mutex> lock(e->kg_mu);
  if (!e->kg_stage && hipHostMalloc((void **)&e->kg_stage, KG_STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
    e->kg_stage = nullptr;
    return fail(TMCE_ENOMEM, "hipHostMalloc key generation staging");
  }
  const size_t slots = std::max<size_t>(1, KG_STAGE_BYTES / PKB);

  KgBuffers b;
  const size_t cap = std::min(n * KG_AHEAD_MAX, KG_LAUNCH);
  HIPCHK(hipStreamCreateWithFlags(&b.s, hipStreamNonBlocking));
  if (hipMalloc((void **)&b.g, cap * T1 * 2) != hipSuccess || hipMalloc((void **)&b.alpha, cap * N * 2) != hipSuccess ||
      hipMalloc((void **)&b.H, cap * hw * 8) != hipSuccess || hipMalloc((void **)&b.pk, cap * PKB) != hipSuccess ||
      hipMalloc((void **)&b.ok, cap * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&b.piv, cap * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc((void **)&b.fv, cap * (T1 - 1) * 2) != hipSuccess || hipMalloc((void **)&b.valid, cap * sizeof(int)) != hipSuccess)
    return fail(TMCE_ENOMEM, "hipMalloc key generation buffers");
  const size_t NQ = (size_t)1 << si.m, CBB = (size_t)si.cond_bytes, cbcap = std::min(n, KG_CB_KEYS);
  if (hipMalloc((void **)&b.pi, cbcap * NQ * 2) != hipSuccess || hipMalloc((void **)&b.cb, cbcap * CBB) != hipSuccess)
    return fail(TMCE_ENOMEM, "hipMalloc key generation buffers");

  // cand[i][k]: attempt k of key i in this round
  std::vector<std::vector<KeygenCandidate>> cand(n);
  std::vector<int> win(n, -1);
  std::vector<uint64_t> win_piv(n, 0);
  std::vector<size_t> pending(n);
  for (size_t i = 0; i < n; ++i) pending[i] = i;
  // host sides of the uploads and downloads, in the pinned staging buffer (each phase from its start)
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  uint16_t *ha = reinterpret_cast<uint16_t *>(e->kg_stage);
  uint16_t *hf = reinterpret_cast<uint16_t *>(e->kg_stage + align(cap * N * 2));
  int *hvalid = reinterpret_cast<int *>(e->kg_stage + align(cap * N * 2) + align(cap * (T1 - 1) * 2));
  uint16_t *hpi = reinterpret_cast<uint16_t *>(e->kg_stage);
  uint8_t *hcb = e->kg_stage + align(cbcap * NQ * 2);
  static_assert(KG_LAUNCH * (8192 + 128 + 2) * 2 + KG_CB_KEYS * (8192 * 2 + 25 * 4096 / 8) + 1024 <= KG_STAGE_BYTES,
                "staging holds a launch's inputs and a control-bit launch");
  std::vector<uint16_t> hg(cap * T1);
  std::vector<std::pair<size_t, size_t>> wins;  // (key, candidate slot of the launch)
  std::vector<std::vector<uint8_t>> sok(n);  // per key and attempt: the support's sort keys are distinct
  std::vector<int> hok(cap);
  std::vector<uint64_t> hpiv(cap);
  while (!pending.empty()) {
    const size_t ahead = std::max(KG_AHEAD_MIN, std::min(KG_AHEAD_MAX, KG_LAUNCH / pending.size()));
    parallel_for(pending.size(), [&](size_t k) {  // a key's attempts follow its chain, in order
      const size_t i = pending[k];
      cand[i].resize(ahead);
      sok[i].assign(ahead, 0);
      for (size_t a = 0; a < ahead; ++a) sok[i][a] = keygen_step(set, chain[i].data(), cand[i][a]) ? 1 : 0;
    });
    const size_t total = pending.size() * ahead;  // candidate c = (key pending[c / ahead], attempt c % ahead)
    std::vector<int> okv(total, 0);
    for (size_t off = 0; off < total; off += cap) {
      const size_t nc = std::min(cap, total - off);
      for (size_t k = 0; k < nc; ++k) {
        const size_t c = off + k;
        const KeygenCandidate &kc = cand[pending[c / ahead]][c % ahead];
        std::memcpy(&hf[k * (T1 - 1)], kc.f.data(), (T1 - 1) * 2);
        std::memcpy(&ha[k * N], kc.alpha.data(), N * 2);
        hvalid[k] = sok[pending[c / ahead]][c % ahead];
      }
      HIPCHK(hipMemcpyAsync(b.fv, hf, nc * (T1 - 1) * 2, hipMemcpyHostToDevice, b.s));
      HIPCHK(hipMemcpyAsync(b.alpha, ha, nc * N * 2, hipMemcpyHostToDevice, b.s));
      HIPCHK(hipMemcpyAsync(b.valid, hvalid, nc * sizeof(int), hipMemcpyHostToDevice, b.s));
      // Goppa polynomials on the GPU, then the systematic form of the attempts still valid
      HIPCHK(launch_genpoly_set(set, b.fv, e->lg[w], e->ex[w], b.g, b.valid, (int)nc, b.s));
      KeygenArgs a{b.g, b.alpha, e->lg[w], e->ex[w], b.H, b.pk, b.ok, f ? b.piv : nullptr, b.valid, (int)nc};
      HIPCHK(launch_pk_gen_set(set, a, b.s));
      HIPCHK(hipMemcpyAsync(hok.data(), b.ok, nc * sizeof(int), hipMemcpyDeviceToHost, b.s));
      HIPCHK(hipMemcpyAsync(hg.data(), b.g, nc * T1 * 2, hipMemcpyDeviceToHost, b.s));
      if (f) HIPCHK(hipMemcpyAsync(hpiv.data(), b.piv, nc * sizeof(uint64_t), hipMemcpyDeviceToHost, b.s));
      HIPCHK(hipStreamSynchronize(b.s));
      // the first successful attempt of each key in this launch (earlier launches hold earlier attempts)
      wins.clear();
      for (size_t k = 0; k < nc; ++k) {
        const size_t c = off + k, i = pending[c / ahead];
        okv[c] = hok[k];
        bool first = hok[k] != 0;
        for (size_t c2 = c - c % ahead; c2 < c && first; ++c2) first = !okv[c2];
        if (first) {
          win[i] = (int)(c % ahead);
          win_piv[i] = hpiv[k];
          std::memcpy(cand[i][c % ahead].g.data(), &hg[k * T1], T1 * 2);
          wins.emplace_back(i, k);
        }
      }
      // their public keys, through the pinned staging buffer
      for (size_t w0 = 0; w0 < wins.size(); w0 += slots) {
        const size_t nw = std::min(slots, wins.size() - w0);
        for (size_t q = 0; q < nw; ++q)
          HIPCHK(hipMemcpyAsync(e->kg_stage + q * PKB, b.pk + wins[w0 + q].second * PKB, PKB, hipMemcpyDeviceToHost, b.s));
        HIPCHK(hipStreamSynchronize(b.s));
        parallel_for(nw, [&](size_t q) { std::memcpy(pk + wins[w0 + q].first * PKB, e->kg_stage + q * PKB, PKB); });
      }
    }
    std::vector<size_t> done, failed;
    for (size_t i : pending) (win[i] >= 0 ? done : failed).push_back(i);
    // secret keys of the keys done: support permutations (the f variants' swaps applied) on the
    // host, Benes control bits on the GPU (one workgroup per key), the rest of sk on the host
    for (size_t off = 0; off < done.size(); off += cbcap) {
      const size_t nd = std::min(cbcap, done.size() - off);
      parallel_for(nd, [&](size_t k) {
        const size_t i = done[off + k];
        if (f) keygen_apply_pivots(set, cand[i][win[i]], win_piv[i]);
        std::memcpy(&hpi[k * NQ], cand[i][win[i]].pi.data(), NQ * 2);
      });
      HIPCHK(hipMemcpyAsync(b.pi, hpi, nd * NQ * 2, hipMemcpyHostToDevice, b.s));
      HIPCHK(launch_controlbits(si.m, b.pi, b.cb, (int)nd, b.s));
      HIPCHK(hipMemcpyAsync(hcb, b.cb, nd * CBB, hipMemcpyDeviceToHost, b.s));
      HIPCHK(hipStreamSynchronize(b.s));
      parallel_for(nd, [&](size_t k) {
        const size_t i = done[off + k];
        keygen_finish_with_cb(set, cand[i][win[i]], &hcb[k * CBB], sk + i * SKB);
        cand[i].clear();
      });
    }
    pending.swap(failed);
  }
  return 0;
}
