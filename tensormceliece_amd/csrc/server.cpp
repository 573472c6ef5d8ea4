// server.cpp -- serving loop: continuous batching of encapsulation / decapsulation requests against
// device-resident keys of one parameter set.
//
// Requests (host buffers, any size up to a slot's capacity, any registered key handle) are queued
// per direction.  A worker thread keeps every slot busy: a slot is a multi-batch pipeline on its own
// stream plus pinned staging; when one is free and requests are pending, the worker forms a batch
// FIFO -- a request joins if the batch still fits the slot (ops after each key's padding to 256) and
// its device footprint fits the HBM budget left by the batches in flight -- stages the inputs,
// uploads them, runs the pipeline and queues the download, all asynchronously on the slot's stream.
// Slots run concurrently on the device (enc and dec slots, and several of each); requests that
// arrive while batches are in flight are admitted into the next free slot.  When a slot's event
// completes the worker copies each request's outputs back and stamps its completion time.
//
// Requests whose buffers all lie in pinned slabs from tmce_host_alloc skip the staging: their
// sub-batches point at the slabs' device addresses, so the pipeline's gather kernel reads the
// inputs and its scatter kernel writes the outputs across the host link directly.
//
// Completion: a host function enqueued behind each batch on its slot's stream wakes the worker, which
// then copies the outputs back (no polling).
//
// The host copies of a batch (request inputs into the slot's pinned staging, outputs back to the
// requests) are split over a few copier threads: one thread's memcpy bandwidth into and out of
// pinned memory, not the GPU, bounds a saturated server otherwise (a 32K-op batch moves ~7-10 MB).
//
// Latency = completion - submission (host steady clock), including queueing, transfers and compute.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "engine.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Req {
  unsigned long long id = 0;
  int dir = 0;
  tmce_key *key = nullptr;
  size_t n = 0;
  const uint8_t *in = nullptr;
  uint8_t *ct = nullptr, *k = nullptr;
  int *status = nullptr;
  size_t footprint = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  bool done = false;
  bool claimed = false;  // a tmce_server_wait call owns it: tmce_server_poll leaves it alone
  struct LoadRun *run = nullptr;  // owned by a tmce_server_run_load call (never polled or waited)
  size_t run_index = 0;
  // zero-copy: device addresses of the request's buffers inside tmce_host_alloc slabs
  bool zero_copy = false;
  const uint8_t *d_in = nullptr;
  uint8_t *d_ct = nullptr, *d_k = nullptr;
  int *d_status = nullptr;
};

// One copy of a batch's staging or unstaging: n bytes from src to dst.
struct CopyJob {
  void *dst;
  const void *src;
  size_t n;
};

// Copier threads: run() hands a list of copies to the pool, takes part itself, and returns when
// every copy is done and no pool thread still looks at the list.  One run() at a time (the worker
// thread is the only caller).
class CopyPool {
 public:
  explicit CopyPool(int threads) {
    for (int i = 0; i < threads; ++i) th_.emplace_back([this] { loop(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void run(const std::vector<CopyJob> &jobs) {
    if (jobs.empty()) return;
    if (th_.empty() || jobs.size() == 1) {
      for (const CopyJob &j : jobs) std::memcpy(j.dst, j.src, j.n);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      jobs_ = &jobs;
      next_.store(0, std::memory_order_relaxed);
      left_ = jobs.size();
      ++gen_;
    }
    cv_.notify_all();
    work(&jobs);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return left_ == 0 && active_ == 0; });
    jobs_ = nullptr;
  }

 private:
  // takes copies until none are left
  void work(const std::vector<CopyJob> *jl) {
    size_t mine = 0;
    for (;;) {
      const size_t i = next_.fetch_add(1, std::memory_order_relaxed);
      if (i >= jl->size()) break;
      const CopyJob &j = (*jl)[i];
      std::memcpy(j.dst, j.src, j.n);
      ++mine;
    }
    if (mine) {
      std::lock_guard<std::mutex> lk(mu_);
      left_ -= mine;
      if (left_ == 0 && active_ == 0) done_cv_.notify_all();
    }
  }
  void loop() {
    unsigned long long seen = 0;
    for (;;) {
      const std::vector<CopyJob> *jl;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || (gen_ != seen && jobs_ != nullptr); });
        if (stop_) return;
        seen = gen_;
        jl = jobs_;
        ++active_;
      }
      work(jl);
      std::lock_guard<std::mutex> lk(mu_);
      --active_;
      if (left_ == 0 && active_ == 0) done_cv_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::vector<CopyJob> *jobs_ = nullptr;
  std::atomic<size_t> next_{0};
  size_t left_ = 0;
  int active_ = 0;  // pool threads inside work() for the current list
  unsigned long long gen_ = 0;
  bool stop_ = false;
};

// A copy of n bytes appended to `jobs` in pieces of at most COPY_PIECE bytes.
constexpr size_t COPY_PIECE = 512 << 10;
void add_copy(std::vector<CopyJob> &jobs, void *dst, const void *src, size_t n) {
  for (size_t o = 0; o < n; o += COPY_PIECE)
    jobs.push_back({static_cast<uint8_t *>(dst) + o, static_cast<const uint8_t *>(src) + o, std::min(COPY_PIECE, n - o)});
}

// One tmce_server_run_load call: the indices of its requests that have completed, appended by the
// worker (under the server's mutex) and consumed by the call.
struct LoadRun {
  std::vector<size_t> done;
  std::vector<Req *> req;  // by request index
};

struct Slot {
  int dir = 0;
  tmce_pipeline *pipe = nullptr;
  hipStream_t stream = nullptr;
  size_t cap = 0, in_row = 0, out_row = 0;  // ops; bytes per op of input / output staging
  uint8_t *h_in = nullptr, *h_out = nullptr;  // pinned staging
  uint8_t *d_io = nullptr;                    // device: inputs [cap][in_row], outputs [cap][out_row]
  hipEvent_t done = nullptr;
  std::atomic<bool> signalled{false};  // set by the stream's host callback when the batch is done
  tmce_server *owner = nullptr;
  bool busy = false;
  size_t footprint = 0;
  std::vector<Req *> reqs;
};

}  // namespace

struct tmce_server {
  int set = -1;      // base parameter set
  bool pc = false;   // plaintext-confirmation variant
  size_t ctb = 0;    // ciphertext bytes
  size_t budget = 0, in_flight = 0, peak_in_flight = 0;
  std::deque<Slot> slots;  // stable addresses: the host callbacks hold Slot pointers
  std::deque<Req *> pending[2];
  std::unordered_map<unsigned long long, Req *> reqs;
  std::deque<unsigned long long> completed;
  unsigned long long next_id = 1;
  std::mutex mu;
  std::condition_variable cv_done;
  // Wake-ups of the worker: submissions, batch completions (host callbacks on the slot streams)
  // and stop each bump wake_seq under wake_mu.  The worker never calls HIP while holding wake_mu
  // and the callbacks take no other lock, so a callback never waits on the worker.
  std::mutex wake_mu;
  std::condition_variable cv_wake;
  unsigned long long wake_seq = 0;
  std::thread worker;
  std::unique_ptr<CopyPool> copiers;
  std::vector<CopyJob> copy_jobs;  // the worker's scratch list
  bool stop = false;
  int error = 0;  // first failure of the worker (guarded by mu)
  std::string error_msg;
  // statistics
  unsigned long long batches = 0, ops = 0, requests = 0, zero_copy = 0, max_batch_keys = 0;
  size_t per_op_bytes[2] = {0, 0};
};

namespace tmce {
namespace {

void wake(tmce_server *sv) {
  {
    std::lock_guard<std::mutex> lk(sv->wake_mu);
    ++sv->wake_seq;
  }
  sv->cv_wake.notify_one();
}

// host callback enqueued behind each batch on its slot's stream (runs on HIP's callback thread)
void slot_done_cb(void *arg) {
  Slot *sl = static_cast<Slot *>(arg);
  sl->signalled.store(true, std::memory_order_release);
  wake(sl->owner);
}

size_t round_up_sz(size_t a, size_t b) { return (a + b - 1) / b * b; }

// Fills `b` from the FIFO of direction `dir`; returns the requests taken (the caller holds mu).
std::vector<Req *> form_batch(tmce_server *sv, Slot &sl) {
  std::vector<Req *> out;
  std::unordered_map<tmce_key *, size_t> group;  // key -> ops in this batch
  size_t padded = 0, foot = 0;
  auto &q = sv->pending[sl.dir];
  while (!q.empty()) {
    Req *r = q.front();
    const size_t before = group.count(r->key) ? round_up_sz(group[r->key], PLAN_ALIGN) : 0;
    if (!group.count(r->key) && group.size() == (size_t)PLAN_MAX_GROUPS) break;
    const size_t after = round_up_sz((group.count(r->key) ? group[r->key] : 0) + r->n, PLAN_ALIGN);
    if (padded - before + after > sl.cap) break;
    if (sv->in_flight + foot + r->footprint > sv->budget && !(out.empty() && sv->in_flight == 0)) break;
    q.pop_front();
    group[r->key] += r->n;
    padded = padded - before + after;
    foot += r->footprint;
    out.push_back(r);
  }
  sl.footprint = foot;
  if (group.size() > sv->max_batch_keys) sv->max_batch_keys = group.size();
  return out;
}

int launch_slot(tmce_server *sv, Slot &sl, std::vector<Req *> &batch) {
  std::vector<tmce_sub_batch> subs;
  std::vector<CopyJob> &jobs = sv->copy_jobs;
  jobs.clear();
  size_t pos = 0, ops = 0;  // pos: staged ops
  uint8_t *d_in = sl.d_io, *d_out = sl.d_io + sl.cap * sl.in_row;
  const double t = now_s();
  for (Req *r : batch) {
    tmce_sub_batch b{};
    b.key = r->key;
    b.n = r->n;
    if (r->zero_copy) {  // the GPU reads and writes the request's slab buffers in place
      b.in = r->d_in;
      b.k = r->d_k;
      b.status = r->d_status;
      b.ct = r->d_ct;
    } else {
      add_copy(jobs, sl.h_in + pos * sl.in_row, r->in, r->n * sl.in_row);
      b.in = d_in + pos * sl.in_row;
      // outputs of the request: [k][status][ct (enc)] (k 16-byte aligned: out_row is a multiple of 16)
      uint8_t *o = d_out + pos * sl.out_row;
      b.k = o;
      b.status = reinterpret_cast<int *>(o + r->n * 32);
      b.ct = sl.dir == 0 ? o + r->n * 36 : nullptr;
      pos += r->n;
    }
    subs.push_back(b);
    r->t_start = t;
    ops += r->n;
  }
  sv->copiers->run(jobs);
  if (pos) HIPCHK(hipMemcpyAsync(d_in, sl.h_in, pos * sl.in_row, hipMemcpyHostToDevice, sl.stream));
  if (int rc = tmce_pipeline_run(sl.pipe, subs.data(), subs.size())) return rc;
  if (pos) HIPCHK(hipMemcpyAsync(sl.h_out, d_out, pos * sl.out_row, hipMemcpyDeviceToHost, sl.stream));
  HIPCHK(hipEventRecord(sl.done, sl.stream));
  sl.signalled.store(false, std::memory_order_relaxed);
  HIPCHK(hipLaunchHostFunc(sl.stream, slot_done_cb, &sl));
  sl.reqs = batch;
  sl.busy = true;
  sv->batches += 1;
  sv->requests += batch.size();
  sv->ops += ops;
  return 0;
}

void finish_slot(tmce_server *sv, Slot &sl) {
  const size_t out_ct = sl.dir == 0 ? sv->ctb : 0;
  size_t pos = 0;
  std::vector<CopyJob> &jobs = sv->copy_jobs;
  jobs.clear();
  for (Req *r : sl.reqs) {
    if (!r->zero_copy) {
      const uint8_t *o = sl.h_out + pos * sl.out_row;
      add_copy(jobs, r->k, o, r->n * 32);
      if (r->status) add_copy(jobs, r->status, o + r->n * 32, r->n * sizeof(int));
      if (sl.dir == 0) add_copy(jobs, r->ct, o + r->n * 36, r->n * out_ct);
      pos += r->n;
    }
  }
  sv->copiers->run(jobs);
  const double t = now_s();
  for (Req *r : sl.reqs) {
    r->t_done = t;
    r->done = true;
    if (r->run) r->run->done.push_back(r->run_index);
    else sv->completed.push_back(r->id);
  }
  sv->in_flight -= sl.footprint;
  sl.footprint = 0;
  sl.reqs.clear();
  sl.busy = false;
}

void worker_main(tmce_server *sv) {
  (void)hipSetDevice(g_engine->device);
  unsigned long long seen = 0;
  for (;;) {
    {
      std::lock_guard<std::mutex> wl(sv->wake_mu);
      seen = sv->wake_seq;  // events after this read wake the wait below
    }
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      if (sv->stop) break;
      bool progressed = false;
      for (Slot &sl : sv->slots) {
        // the callback ran after the batch's last operation; the event query confirms it
        if (sl.busy && sl.signalled.load(std::memory_order_acquire) && hipEventQuery(sl.done) == hipSuccess) {
          finish_slot(sv, sl);
          progressed = true;
        }
      }
      if (progressed) sv->cv_done.notify_all();
      for (Slot &sl : sv->slots) {
        if (!sl.busy && !sv->pending[sl.dir].empty()) {
          std::vector<Req *> batch = form_batch(sv, sl);
          if (batch.empty()) continue;
          sv->in_flight += sl.footprint;
          if (sv->in_flight > sv->peak_in_flight) sv->peak_in_flight = sv->in_flight;
          if (int rc = launch_slot(sv, sl, batch)) {
            sv->error = rc;
            sv->error_msg = tmce_last_error();
            sv->stop = true;
            break;
          }
        }
      }
      if (sv->stop) break;
    }
    std::unique_lock<std::mutex> wl(sv->wake_mu);
    sv->cv_wake.wait(wl, [&] { return sv->wake_seq != seen; });
  }
  sv->cv_done.notify_all();
}

void destroy_server(tmce_server *sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  wake(sv);
  if (sv->worker.joinable()) sv->worker.join();
  for (Slot &sl : sv->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.pipe) tmce_pipeline_free(sl.pipe);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.h_in) (void)hipHostFree(sl.h_in);
    if (sl.h_out) (void)hipHostFree(sl.h_out);
    if (sl.d_io) (void)hipFree(sl.d_io);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  for (auto &kv : sv->reqs) delete kv.second;
  delete sv;
}

// validates and queues one request; with `run` it belongs to a tmce_server_run_load call
int submit_request(tmce_server *sv, const tmce_request *rq, LoadRun *run, size_t run_index, unsigned long long *id) {
  if (!sv || !rq || !rq->key || (rq->direction != 0 && rq->direction != 1))
    return fail(TMCE_EINVAL, "tmce_server_submit: bad argument");
  if (rq->key->set != sv->set || rq->key->pc != sv->pc || rq->key->kind != (rq->direction == 0 ? 1 : 2))
    return fail(TMCE_EINVAL, "tmce_server_submit: key of the wrong set, variant or kind");
  if (rq->n == 0 || rq->n + PLAN_ALIGN - 1 > sv->slots[0].cap || !rq->in || !rq->k || (rq->direction == 0 && !rq->ct))
    return fail(TMCE_EINVAL, "tmce_server_submit: empty request, request larger than a slot, or missing buffers");
  auto *r = new Req();
  r->dir = rq->direction;
  r->key = rq->key;
  r->n = rq->n;
  r->in = rq->in;
  r->ct = rq->ct;
  r->k = rq->k;
  r->status = rq->status;
  r->footprint = rq->n * sv->per_op_bytes[r->dir];
  r->run = run;
  r->run_index = run_index;
  r->claimed = run != nullptr;
  {  // zero-copy when every buffer lies in a slab with the pipelines' alignment
    const size_t n = rq->n;
    r->d_in = slab_device_ptr(rq->in, n * (r->dir == 0 ? 32 : sv->ctb));
    r->d_k = slab_device_ptr(rq->k, n * 32);
    r->d_status = rq->status ? reinterpret_cast<int *>(slab_device_ptr(rq->status, n * sizeof(int))) : nullptr;
    r->d_ct = r->dir == 0 ? slab_device_ptr(rq->ct, n * sv->ctb) : nullptr;
    r->zero_copy = r->d_in && r->d_k && (!rq->status || r->d_status) && (r->dir == 1 || r->d_ct) &&
                   !((uintptr_t)r->d_k & 15) && !((uintptr_t)r->d_status & 3) && !((uintptr_t)r->d_in & 1) &&
                   !((uintptr_t)r->d_ct & 1);
  }
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) {
      delete r;
      return fail(sv->error ? sv->error : TMCE_EINVAL, sv->error ? sv->error_msg : "server stopped");
    }
    r->id = sv->next_id++;
    r->t_submit = now_s();
    if (r->zero_copy) sv->zero_copy += 1;
    sv->reqs[r->id] = r;
    if (run) run->req[run_index] = r;
    sv->pending[r->dir].push_back(r);
    if (id) *id = r->id;
  }
  wake(sv);
  return 0;
}

}  // namespace
}  // namespace tmce

using namespace tmce;

extern "C" {

int tmce_server_create(int variant, size_t slot_capacity, int slots_per_direction, size_t hbm_budget,
                       tmce_server **out) {
  if (!variant_valid(variant) || !out || slot_capacity == 0 || slots_per_direction < 1 || slots_per_direction > 8)
    return fail(TMCE_EINVAL, "tmce_server_create: bad argument");
  Engine *e = engine();
  if (!e) return TMCE_ENODEV;
  auto *sv = new tmce_server();
  sv->set = variant_base(variant);
  sv->pc = variant_pc(variant);
  sv->ctb = (size_t)variant_ct_bytes(variant);
  sv->budget = hbm_budget ? hbm_budget : (size_t)-1;
  auto bail = [&](int r) {
    destroy_server(sv);
    return r;
  };
  for (int dir = 0; dir < 2; ++dir)
    for (int i = 0; i < slots_per_direction; ++i) {
      sv->slots.emplace_back();
      Slot &sl = sv->slots.back();
      sl.dir = dir;
      sl.owner = sv;
      if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(TMCE_EDEVICE, "hipStreamCreate"));
      if (int r = tmce_pipeline_create(variant, dir, slot_capacity, PLAN_MAX_GROUPS * 64 + 1024, 1, sl.stream, &sl.pipe))
        return bail(r);
      sl.cap = tmce_pipeline_capacity(sl.pipe);
      sl.in_row = dir == 0 ? 32 : sv->ctb;
      sl.out_row = round_up_sz((dir == 0 ? sv->ctb : 0) + 32 + sizeof(int), 16);
      if (hipHostMalloc((void **)&sl.h_in, sl.cap * sl.in_row, hipHostMallocDefault) != hipSuccess ||
          hipHostMalloc((void **)&sl.h_out, sl.cap * sl.out_row, hipHostMallocDefault) != hipSuccess)
        return bail(fail(TMCE_ENOMEM, "hipHostMalloc staging"));
      if (hipMalloc((void **)&sl.d_io, sl.cap * (sl.in_row + sl.out_row)) != hipSuccess)
        return bail(fail(TMCE_ENOMEM, "hipMalloc slot buffers"));
      if (hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
        return bail(fail(TMCE_EDEVICE, "hipEventCreate"));
      // per-op device footprint of this direction: pipeline buffers + slot I/O, per op of capacity
      const size_t per_op = (tmce_pipeline_device_bytes(sl.pipe) + sl.cap * (sl.in_row + sl.out_row)) / sl.cap;
      sv->per_op_bytes[dir] = per_op;
    }
  // copier threads: a quarter of the host's cores, 2 to 8
  const unsigned hc = std::thread::hardware_concurrency();
  sv->copiers = std::make_unique<CopyPool>((int)std::min(8u, std::max(2u, hc / 4)) - 1);
  sv->worker = std::thread(worker_main, sv);
  *out = sv;
  return 0;
}

int tmce_server_submit(tmce_server *sv, const tmce_request *rq, unsigned long long *id) {
  if (!id) return fail(TMCE_EINVAL, "tmce_server_submit: bad argument");
  return submit_request(sv, rq, nullptr, 0, id);
}

int tmce_server_wait(tmce_server *sv, unsigned long long id, double *latency_s) {
  if (!sv) return fail(TMCE_EINVAL, "tmce_server_wait: null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(id);
  if (it == sv->reqs.end()) return fail(TMCE_EINVAL, "tmce_server_wait: unknown request (or already collected)");
  Req *r = it->second;
  if (r->claimed) return fail(TMCE_EINVAL, "tmce_server_wait: another thread is waiting for this request");
  r->claimed = true;  // from here on only this call collects (and frees) it
  sv->cv_done.wait(lk, [&] { return r->done || sv->stop; });
  if (!r->done) {
    r->claimed = false;
    return fail(sv->error ? sv->error : TMCE_EDEVICE, sv->error ? sv->error_msg : "server stopped");
  }
  if (latency_s) *latency_s = r->t_done - r->t_submit;
  sv->reqs.erase(id);
  delete r;
  return 0;
}

int tmce_server_poll(tmce_server *sv, unsigned long long *ids, double *latency_s, int max) {
  if (!sv || (max > 0 && !ids)) return fail(TMCE_EINVAL, "tmce_server_poll: bad argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sv->error) return fail(sv->error, sv->error_msg);
  int n = 0;
  while (n < max && !sv->completed.empty()) {
    const unsigned long long id = sv->completed.front();
    sv->completed.pop_front();
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) continue;  // already collected by tmce_server_wait
    if (it->second->claimed) continue;   // a tmce_server_wait call collects it
    ids[n] = id;
    if (latency_s) latency_s[n] = it->second->t_done - it->second->t_submit;
    delete it->second;
    sv->reqs.erase(it);
    ++n;
  }
  return n;
}

int tmce_server_run_load(tmce_server *sv, const tmce_request *rqs, const double *arrival_s, size_t n,
                         double *latency_s, double *elapsed_s) {
  if (!sv || (n && (!rqs || !arrival_s))) return fail(TMCE_EINVAL, "tmce_server_run_load: bad argument");
  for (size_t i = 1; i < n; ++i)
    if (arrival_s[i] < arrival_s[i - 1]) return fail(TMCE_EINVAL, "tmce_server_run_load: arrivals not sorted");
  LoadRun run;
  run.req.assign(n, nullptr);
  const auto t0 = std::chrono::steady_clock::now();
  const double t0s = now_s();
  int sub_rc = 0;
  size_t submitted = 0;
  std::thread submitter([&] {
    for (size_t i = 0; i < n; ++i) {
      const auto due = t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                std::chrono::duration<double>(arrival_s[i]));
      // sleep to about 100 us before the arrival, then spin (sleep wake-ups are coarser than the
      // gaps between arrivals at high rates)
      const auto early = due - std::chrono::microseconds(100);
      if (std::chrono::steady_clock::now() < early) std::this_thread::sleep_until(early);
      while (std::chrono::steady_clock::now() < due) {
      }
      const int rc = submit_request(sv, &rqs[i], &run, i, nullptr);
      std::lock_guard<std::mutex> lk(sv->mu);
      if (rc) {
        sub_rc = rc;
        sv->cv_done.notify_all();
        return;
      }
      ++submitted;
    }
  });
  int rc = 0;
  size_t collected = 0;
  double t_last = t0s;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
      sv->cv_done.wait(lk, [&] { return !run.done.empty() || sv->stop || sub_rc || collected == n; });
      for (size_t idx : run.done) {
        Req *r = run.req[idx];
        if (latency_s) latency_s[idx] = r->t_done - r->t_submit;
        if (r->t_done > t_last) t_last = r->t_done;
        sv->reqs.erase(r->id);
        delete r;
        run.req[idx] = nullptr;
        ++collected;
      }
      run.done.clear();
      if (collected == n) break;
      if (sv->stop) {
        rc = fail(sv->error ? sv->error : TMCE_EDEVICE, sv->error ? sv->error_msg : "server stopped");
        break;
      }
      if (sub_rc && collected == submitted) {
        rc = sub_rc;
        break;
      }
    }
  }
  submitter.join();
  if (rc) {  // requests still queued or in flight go back to the server (tmce_server_poll collects them)
    std::lock_guard<std::mutex> lk(sv->mu);
    for (size_t idx : run.done) {
      sv->reqs.erase(run.req[idx]->id);
      delete run.req[idx];
      run.req[idx] = nullptr;
    }
    run.done.clear();
    for (Req *r : run.req)
      if (r) {
        r->run = nullptr;
        r->claimed = false;
      }
    return rc;
  }
  if (elapsed_s) *elapsed_s = t_last - t0s;
  return 0;
}

int tmce_server_get_stats(tmce_server *sv, tmce_server_stats *st) {
  if (!sv || !st) return fail(TMCE_EINVAL, "tmce_server_get_stats: bad argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  st->batches = sv->batches;
  st->requests = sv->requests;
  st->ops = sv->ops;
  st->pending = sv->pending[0].size() + sv->pending[1].size();
  st->in_flight_bytes = sv->in_flight;
  st->peak_in_flight_bytes = sv->peak_in_flight;
  st->enc_bytes_per_op = sv->per_op_bytes[0];
  st->dec_bytes_per_op = sv->per_op_bytes[1];
  st->zero_copy_requests = sv->zero_copy;
  st->max_batch_keys = sv->max_batch_keys;
  return 0;
}

void tmce_server_free(tmce_server *sv) { destroy_server(sv); }

}  // extern "C"
