// pipeline.cpp -- multi-batch pipelines: sub-batches of mixed sizes (and keys) of one parameter set
// run through one set of engine launches, optionally replayed from a captured HIP graph.
//
// A pipeline has a fixed parameter set, direction and capacity (combined ops).  Each run builds a
// plan on the host -- key groups (one per distinct key, in order of first appearance, each padded to
// PLAN_ALIGN ops), the sub-batches laid out inside their group -- copies it to the device with one
// asynchronous copy from pinned memory, then issues
//     gather (sub-batch inputs -> flat buffer) -> stage kernels -> scatter (flat outputs -> callers)
// on the pipeline's stream.  Every kernel reads its live op count and per-op key material from the
// device plan, so the launch sequence does not depend on the run's contents: it is captured into a
// HIP graph once (at the second run -- the first runs eagerly -- or again after the tuning knobs
// change) and replayed after.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>

#include "engine.h"

struct tmce_pipeline {
  int set = -1, dir = 0;  // dir: 0 = encapsulation, 1 = decapsulation
  bool pc = false;        // plaintext-confirmation variant: ciphertexts are SYND_BYTES + 32 bytes
  int ctb = 0;            // ciphertext bytes
  int cap = 0, max_sub = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false, use_graph = false;
  uint8_t *plan_host = nullptr;  // pinned image of the plan
  tmce::PlanDev *plan_dev = nullptr;
  size_t plan_bytes = 0;
  hipEvent_t uploaded = nullptr;  // the pinned image may be rewritten once this has completed
  tmce::DevBuf in, ct, key, status, ind, msg, S, queue, xscratch, ct_synd;
  size_t dev_bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool warm = false;                // one eager run precedes the capture (first-use set-up, e.g.
                                    // kernel attributes, stays out of the captured stream)
  int tuned[6] = {-1, -1, -1, -1, -1, -1};  // knobs the captured graph was built with
};

namespace tmce {

namespace {

size_t plan_image_bytes(int max_sub) { return round_up((int)sizeof(PlanDev), 16) + (size_t)max_sub * sizeof(SubDev); }

// the stage launches of one run (everything after the plan upload)
int issue(tmce_pipeline *p, const int knobs[6]) {
  Engine *e = g_engine;
  const SetInfo &si = set_info(p->set);
  hipStream_t s = p->stream;
  if (p->dir == 0) {
    HIPCHK(launch_plan_gather(p->plan_dev, p->in.as<uint8_t>(), 32, p->cap, s));
    EncArgs a{};
    a.seeds = p->in.as<uint8_t>();
    a.ct = p->ct.as<uint8_t>();
    a.key = p->key.as<uint8_t>();
    a.status = p->status.as<int>();
    a.ind = p->ind.as<uint16_t>();
    a.msg = p->msg.as<uint32_t>();
    a.queue = p->queue.as<unsigned>();
    a.nops = p->cap;
    a.gen_blocks = knobs[0];
    a.gen_form = knobs[4];
    a.pc = p->pc ? 1 : 0;
    a.plan = p->plan_dev;
    HIPCHK(launch_enc_set(p->set, a, s));
    HIPCHK(launch_plan_scatter(p->plan_dev, p->ct.as<uint8_t>(), p->ctb, p->key.as<uint8_t>(),
                               p->status.as<int>(), p->cap, s));
  } else {
    HIPCHK(launch_plan_gather(p->plan_dev, p->in.as<uint8_t>(), p->ctb, p->cap, s));
    DecArgs a{};
    a.ct = p->in.as<uint8_t>();
    a.pc = p->pc ? 1 : 0;
    a.ct_synd = p->pc ? p->ct_synd.as<uint8_t>() : nullptr;
    const int w = si.m - 12;
    const SetTables &tb = e->st[p->set];
    a.ft = FieldTables{e->lg[w], e->ex[w], tb.twl, tb.plo, tb.phi, tb.ptlo, tb.pthi};
    a.key = p->key.as<uint8_t>();
    a.status = p->status.as<int>();
    a.S = p->S.as<uint32_t>();
    a.msg = p->msg.as<uint32_t>();
    a.queue = p->queue.as<unsigned>();
    a.nops = p->cap;
    a.decode_blocks = knobs[1];
    a.cus = e->cus;
    a.front_mfma = knobs[2];
    a.front_cfg = knobs[3];
    a.fft_form = knobs[5];
    a.xscratch = p->xscratch.as<uint32_t>();
    a.plan = p->plan_dev;
    HIPCHK(launch_dec_set(p->set, a, s));
    HIPCHK(launch_plan_scatter(p->plan_dev, nullptr, 0, p->key.as<uint8_t>(), p->status.as<int>(), p->cap, s));
  }
  return 0;
}

void destroy(tmce_pipeline *p) {
  if (!p) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->uploaded) (void)hipEventDestroy(p->uploaded);
  if (p->plan_host) (void)hipHostFree(p->plan_host);
  if (p->plan_dev) (void)hipFree(p->plan_dev);
  for (DevBuf *b : {&p->in, &p->ct, &p->key, &p->status, &p->ind, &p->msg, &p->S, &p->queue, &p->xscratch, &p->ct_synd})
    if (b->p) (void)hipFree(b->p);
  if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

}  // namespace

}  // namespace tmce

using namespace tmce;

extern "C" {

int tmce_pipeline_create(int variant, int direction, size_t capacity, size_t max_sub_batches, int use_graph,
                         void *stream, tmce_pipeline **out) {
  if (!variant_valid(variant) || (direction != 0 && direction != 1) || !out || capacity == 0 || max_sub_batches == 0)
    return fail(TMCE_EINVAL, "tmce_pipeline_create: bad argument");
  const int set = variant_base(variant);
  if (capacity > (size_t)(1 << 28) || max_sub_batches > (size_t)(1 << 24))
    return fail(TMCE_EINVAL, "tmce_pipeline_create: capacity too large");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  auto *p = new tmce_pipeline();
  p->set = set;
  p->pc = variant_pc(variant);
  p->ctb = variant_ct_bytes(variant);
  p->dir = direction;
  p->cap = round_up((int)capacity, PLAN_ALIGN);
  p->max_sub = (int)max_sub_batches;
  p->use_graph = use_graph != 0;
  auto bail = [&](int r) {
    destroy(p);
    return r;
  };
  if (stream) {
    p->stream = static_cast<hipStream_t>(stream);
  } else {
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
      return bail(fail(TMCE_EDEVICE, "hipStreamCreate"));
    p->own_stream = true;
  }
  p->plan_bytes = plan_image_bytes(p->max_sub);
  if (hipHostMalloc((void **)&p->plan_host, p->plan_bytes, hipHostMallocDefault) != hipSuccess)
    return bail(fail(TMCE_ENOMEM, "hipHostMalloc plan"));
  if (hipMalloc((void **)&p->plan_dev, p->plan_bytes) != hipSuccess) return bail(fail(TMCE_ENOMEM, "hipMalloc plan"));
  if (hipEventCreateWithFlags(&p->uploaded, hipEventDisableTiming) != hipSuccess)
    return bail(fail(TMCE_EDEVICE, "hipEventCreate"));
  const SetInfo &si = set_info(set);
  const size_t n = (size_t)p->cap;
  const int ind_stride = (si.t + 1) & ~1;
  const int rw = si.m * (((2 * si.t + 31) / 32) * 32) / 32;
  struct Need { DevBuf *b; size_t bytes; };
  std::vector<Need> need = {{&p->key, n * 32}, {&p->status, n * 4}, {&p->msg, n * si.msg_words * 4}};
  if (direction == 0) {
    need.push_back({&p->in, n * 32});
    need.push_back({&p->ct, n * p->ctb});
    need.push_back({&p->ind, n * ind_stride * 2});
    need.push_back({&p->queue, 64});
  } else {
    need.push_back({&p->in, n * p->ctb});
    if (p->pc) need.push_back({&p->ct_synd, n * si.synd_bytes});
    need.push_back({&p->S, n * rw * 4});
    need.push_back({&p->queue, 1024});
    const size_t xb = gf2_scratch_bytes(set, GF2_FRONT, n);
    if (xb) need.push_back({&p->xscratch, xb});
  }
  for (auto &q : need) {
    if (int r = q.b->ensure(q.bytes)) return bail(r);
    p->dev_bytes += q.bytes;
  }
  p->dev_bytes += p->plan_bytes;
  *out = p;
  return 0;
}

void tmce_pipeline_free(tmce_pipeline *p) { destroy(p); }

size_t tmce_pipeline_device_bytes(const tmce_pipeline *p) { return p ? p->dev_bytes : 0; }

size_t tmce_pipeline_capacity(const tmce_pipeline *p) { return p ? (size_t)p->cap : 0; }

void *tmce_pipeline_stream(tmce_pipeline *p) { return p ? p->stream : nullptr; }

int tmce_pipeline_run(tmce_pipeline *p, const tmce_sub_batch *subs, size_t nsub) {
  if (!p || (nsub && !subs)) return fail(TMCE_EINVAL, "tmce_pipeline_run: bad argument");
  if (nsub > (size_t)p->max_sub) return fail(TMCE_EINVAL, "tmce_pipeline_run: more sub-batches than the pipeline holds");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  const SetInfo &si = set_info(p->set);
  const int want_kind = p->dir == 0 ? 1 : 2;

  // ---- key groups and layout
  std::vector<const tmce_key *> keys;
  std::unordered_map<const tmce_key *, int> gidx;
  std::vector<long> gcount;
  for (size_t i = 0; i < nsub; ++i) {
    const tmce_sub_batch &b = subs[i];
    if (!b.key || b.key->set != p->set || b.key->pc != p->pc || b.key->kind != want_kind)
      return fail(TMCE_EINVAL, "tmce_pipeline_run: sub-batch key of the wrong set, variant or kind");
    if (b.n && (!b.in || !b.k || (p->dir == 0 && !b.ct)))
      return fail(TMCE_EINVAL, "tmce_pipeline_run: sub-batch without buffers");
    if (((uintptr_t)b.k & 15) || ((uintptr_t)b.status & 3) || ((uintptr_t)b.in & 1) || ((uintptr_t)b.ct & 1))
      return fail(TMCE_EINVAL, "tmce_pipeline_run: misaligned buffer (k: 16 bytes, status: 4, in/ct: 2)");
    auto it = gidx.find(b.key);
    if (it == gidx.end()) {
      if ((int)keys.size() == PLAN_MAX_GROUPS)
        return fail(TMCE_EINVAL, "tmce_pipeline_run: more distinct keys than one run holds (16)");
      it = gidx.emplace(b.key, (int)keys.size()).first;
      keys.push_back(b.key);
      gcount.push_back(0);
    }
    gcount[it->second] += (long)b.n;
  }
  std::vector<long> gstart(keys.size() + 1, 0);
  for (size_t g = 0; g < keys.size(); ++g)
    gstart[g + 1] = gstart[g] + (long)round_up((int)std::min<long>(gcount[g], 1L << 28), PLAN_ALIGN);
  const long total = gstart[keys.size()];
  if (total > p->cap) return fail(TMCE_EINVAL, "tmce_pipeline_run: sub-batches exceed the pipeline capacity");

  // ---- plan image (the previous upload must have completed before the pinned buffer is rewritten)
  HIPCHK(hipEventSynchronize(p->uploaded));
  PlanDev *ph = reinterpret_cast<PlanDev *>(p->plan_host);
  std::memset(ph, 0, sizeof(PlanDev));
  ph->total = (int)total;
  ph->ngroups = (int)keys.size();
  for (size_t g = 0; g <= keys.size(); ++g) ph->gstart[g] = (int)gstart[g];
  for (size_t g = 0; g < keys.size(); ++g) {
    const tmce_key *k = keys[g];
    KeyDev &kd = ph->keys[g];
    kd.cols = k->cols;
    kd.F = k->F;
    kd.Fsw = k->Fsw;
    kd.Fcol = k->Fcol;
    kd.alpha = k->alpha;
    kd.la = k->la;
    kd.lu = k->lu;
    kd.ainv = k->ainv;
    kd.supp = k->supp;
    kd.rej = k->rej;
    kd.pad_bad = k->pad_bad ? 1 : 0;
  }
  const size_t sub_off = round_up((int)sizeof(PlanDev), 16);
  SubDev *sh = reinterpret_cast<SubDev *>(p->plan_host + sub_off);
  ph->subs = reinterpret_cast<const SubDev *>(reinterpret_cast<uint8_t *>(p->plan_dev) + sub_off);
  std::vector<long> cursor(gstart.begin(), gstart.end() - 1);
  // sub-batches in start order: group by group, in call order within a group
  int ns = 0;
  for (size_t g = 0; g < keys.size(); ++g)
    for (size_t i = 0; i < nsub; ++i) {
      const tmce_sub_batch &b = subs[i];
      if (gidx[b.key] != (int)g || b.n == 0) continue;
      SubDev &d = sh[ns++];
      d.in = b.in;
      d.ct = b.ct;
      d.key = b.k;
      d.status = b.status;
      d.start = (int)cursor[g];
      d.n = (int)b.n;
      cursor[g] += (long)b.n;
    }
  ph->nsub = ns;
  const size_t bytes = sub_off + (size_t)ns * sizeof(SubDev);
  HIPCHK(hipMemcpyAsync(p->plan_dev, p->plan_host, bytes, hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipEventRecord(p->uploaded, p->stream));
  if (total == 0) return 0;

  // ---- launches: replay the captured graph, or capture / issue them
  int knobs[6];
  {
    std::lock_guard<std::mutex> lk(e->mu);
    knobs[0] = e->gen_blocks[p->set];
    knobs[1] = e->decode_blocks[p->set];
    knobs[2] = e->front_mfma[p->set];
    knobs[3] = e->front_cfg[p->set];
    knobs[4] = e->gen_form[p->set];
    knobs[5] = e->fft_form[p->set];
  }
  if (!p->use_graph || !p->warm) {
    p->warm = true;
    return issue(p, knobs);
  }
  if (p->exec && std::memcmp(knobs, p->tuned, sizeof knobs) != 0) {
    (void)hipGraphExecDestroy(p->exec);
    (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
  }
  if (!p->exec) {
    HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    g_ktimer_off = true;  // no per-kernel event records inside a captured graph (this thread only)
    const int r = issue(p, knobs);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(p->stream, &g);
    g_ktimer_off = false;
    if (r) {
      if (g) (void)hipGraphDestroy(g);
      return r;
    }
    if (ce != hipSuccess) return fail(TMCE_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {  // keep no half-built graph: the next run captures afresh
      (void)hipGraphDestroy(g);
      p->exec = nullptr;
      return fail(TMCE_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    p->graph = g;
    std::memcpy(p->tuned, knobs, sizeof knobs);
  }
  HIPCHK(hipGraphLaunch(p->exec, p->stream));
  (void)si;
  return 0;
}

int tmce_pipeline_sync(tmce_pipeline *p) {
  if (!p) return fail(TMCE_EINVAL, "tmce_pipeline_sync: null pipeline");
  HIPCHK(hipStreamSynchronize(p->stream));
  return 0;
}

}  // extern "C"
