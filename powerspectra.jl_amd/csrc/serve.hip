// Persistent request loop: many mcm / coupledcov requests of mixed lmax and
// kind in flight at once on one device.
//
// Slots: `slots` worker threads, each owning a non-blocking HIP stream and
// two pinned staging buffers.  A worker takes the request at the head of the
// queue as soon as the device arena can hold it (outputs + record-table
// scratch), enqueues its kernels on its own stream, waits for that stream
// only, copies the result out if the caller gave host outputs, and frees the
// memory.  Several requests therefore run on the GPU concurrently, and a
// request that arrives while others run is admitted as soon as a slot and
// arena space free up.  The request to admit next is the oldest (FIFO) or the
// one with the earliest deadline (arrival + slack x estimated run time); it
// waits for arena space and no other request overtakes it meanwhile, so no
// request starves.  Requests big enough to fill the GPU alone run one at a
// time: while one runs, the next small request is admitted instead.
//
// Spectra live on the device under caller-chosen 64-bit keys (put once,
// referenced by many requests; reference counted while requests use them).
// The kernels and their launch geometry are the single-call ones, so every
// result is bit-identical alone and in a mix.
#include "psx_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

namespace psx {

namespace {

using Clock = std::chrono::steady_clock;

thread_local std::string t_server_error;

// First-fit allocator over one device allocation, in 256-byte granules.
class Arena {
 public:
  void init(char* base, size_t bytes) {
    base_ = base;
    size_ = bytes & ~size_t(255);
    free_.clear();
    used_.clear();
    if (size_) free_[0] = size_;
  }
  static size_t granule(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~size_t(255); }
  size_t size() const { return size_; }
  char* alloc(size_t bytes) {
    bytes = granule(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, sz = it->second;
      free_.erase(it);
      if (sz > bytes) free_[off + bytes] = sz - bytes;
      used_[off] = bytes;
      return base_ + off;
    }
    return nullptr;
  }
  void release(char* p) {
    if (!p) return;
    const size_t off = (size_t)(p - base_);
    auto u = used_.find(off);
    if (u == used_.end()) return;
    size_t sz = u->second;
    used_.erase(u);
    size_t start = off;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && off + sz == nx->first) {
      sz += nx->second;
      nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        start = pv->first;
        sz += pv->second;
        free_.erase(pv);
      }
    }
    free_[start] = sz;
  }

 private:
  char* base_ = nullptr;
  size_t size_ = 0;
  std::map<size_t, size_t> free_, used_;
};

struct Resident {
  double* d = nullptr;
  int64_t len = 0;
  int refs = 0;
};

struct Job {
  uint64_t ticket = 0;
  psx_request rq{};
  ContractRequest crq;
  int nout = 0;
  int n = 0;
  bool host_out = false;
  size_t out_bytes = 0, scratch_bytes = 0;
  char* out_mem = nullptr;
  char* scratch_mem = nullptr;
  std::vector<uint64_t> keys;
  Clock::time_point t_submit, t_start, t_done;
  double est = 0.0;       // estimated seconds alone
  double deadline = 0.0;  // seconds since the server started (PSX_SCHED_DEADLINE)
  int status = PSX_OK;
  std::string msg;
  bool done = false;
};

struct Slot {
  hipStream_t stream = nullptr;
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::thread th;
};

}  // namespace

struct Server {
  Ctx* ctx = nullptr;
  int lmax_max = 0;
  char* arena_mem = nullptr;
  Arena arena;
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Job>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs;
  std::unordered_map<uint64_t, Resident> resident;
  uint64_t next_ticket = 1;
  bool stop = false;
  int policy = PSX_SCHED_FIFO;
  double slack = 2.0;
  Clock::time_point t0 = Clock::now();
  // requests whose kernels fill the GPU alone run one at a time (two of them
  // sharing the device would each take twice as long); smaller ones run
  // beside them in the other slots
  int big_running = 0;
};

// A request is "big" when its kernels alone keep every CU busy for longer
// than this (estimated seconds).
constexpr double kBigSeconds = 5e-4;

namespace {

double secs(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

// Enqueue, wait, copy out.  Runs on the worker's thread, outside the server
// lock; the context lock is held only while the kernels are enqueued (tile
// lists and tables are shared context state).
void execute(Server* S, Slot* sl, Job& j) {
  const int n = j.n;
  const size_t blk = (size_t)n * n;
  double* outs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int k = 0; k < j.nout; ++k) outs[k] = (double*)j.out_mem + k * blk;
  ContractRequest rq = j.crq;
  for (int k = 0; k < j.nout; ++k) rq.dout[k] = outs[k];
  rq.ldo = n;
  int* err = nullptr;
  {
    std::lock_guard<std::recursive_mutex> lk(S->ctx->mu);
    err = run_contraction_on(S->ctx, rq, sl->stream, j.scratch_mem, j.scratch_bytes);
  }
  int h = 0;
  if (err) PSX_HIP(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, sl->stream));
  PSX_HIP(hipStreamSynchronize(sl->stream));
  PSX_REQUIRE(h == 0, PSX_ERR_DOMAIN, "sqrt of a negative spectrum product (Julia DomainError)");
  if (j.host_out) {
    D2HMatrix m[5];
    for (int k = 0; k < j.nout; ++k) m[k] = {j.rq.out[k], j.rq.ld, outs[k], n, n, n};
    d2h_matrices_on(sl->stream, sl->stage, sl->ev, m, j.nout);
  }
}

void worker(Server* S, Slot* sl) {
  (void)hipSetDevice(S->ctx->device);
  for (;;) {
    std::shared_ptr<Job> j;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        if (S->stop) return;
        if (!S->queue.empty()) {
          // next request by policy; while a big request runs, the next small
          // one (big ones wait their turn, ahead of later big ones)
          auto better = [&](const std::shared_ptr<Job>& x, const std::shared_ptr<Job>& y) {
            return S->policy == PSX_SCHED_DEADLINE ? x->deadline < y->deadline
                                                   : x->ticket < y->ticket;
          };
          auto pick = S->queue.end();
          for (auto it = S->queue.begin(); it != S->queue.end(); ++it) {
            if (S->big_running > 0 && (*it)->est > kBigSeconds) continue;
            if (pick == S->queue.end() || better(*it, *pick)) pick = it;
          }
          if (pick == S->queue.end()) {
            S->cv_work.wait(lk);
            continue;
          }
          auto& h = *pick;
          char* a = S->arena.alloc(h->scratch_bytes);
          if (a) {
            char* b = S->arena.alloc(h->out_bytes);
            if (b) {
              h->scratch_mem = a;
              h->out_mem = b;
              j = h;
              S->queue.erase(pick);
              if (j->est > kBigSeconds) S->big_running++;
              break;
            }
            S->arena.release(a);
          }
        }
        S->cv_work.wait(lk);
      }
      j->t_start = Clock::now();
    }
    int status = PSX_OK;
    std::string msg;
    try {
      execute(S, sl, *j);
    } catch (const Error& e) {
      status = e.code;
      msg = e.msg;
    } catch (const std::exception& e) {
      status = PSX_ERR_DEVICE;
      msg = e.what();
    }
    // nothing of this request may still run when its memory is handed back
    if (status != PSX_OK) (void)hipStreamSynchronize(sl->stream);
    {
      std::lock_guard<std::mutex> lk(S->mu);
      j->t_done = Clock::now();
      j->status = status;
      j->msg = msg;
      S->arena.release(j->scratch_mem);
      j->scratch_mem = nullptr;
      if (j->host_out || status != PSX_OK) {
        S->arena.release(j->out_mem);
        j->out_mem = nullptr;
      }
      for (uint64_t k : j->keys) {
        auto it = S->resident.find(k);
        if (it != S->resident.end()) it->second.refs--;
      }
      if (j->est > kBigSeconds) S->big_running--;
      j->done = true;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
}

const Resident& lookup(Server* S, uint64_t key, int64_t need, const char* what) {
  auto it = S->resident.find(key);
  PSX_REQUIRE(it != S->resident.end(), PSX_ERR_ARGUMENT,
              std::string(what) + " key " + std::to_string(key) + " is not resident");
  PSX_REQUIRE(it->second.len >= need, PSX_ERR_ARGUMENT,
              std::string(what) + " key " + std::to_string(key) + " holds fewer values than needed");
  return it->second;
}

// Validate a request and build its contraction (under the server lock).
std::shared_ptr<Job> build_job(Server* S, const psx_request& r) {
  PSX_REQUIRE(r.lmin >= 0 && r.lmax >= r.lmin, PSX_ERR_ARGUMENT, "need 0 <= lmin <= lmax");
  PSX_REQUIRE(r.lmax <= S->lmax_max, PSX_ERR_ARGUMENT, "lmax above the server's lmax_max");
  PSX_REQUIRE(r.LW >= 1, PSX_ERR_ARGUMENT, "W must have at least one entry");
  auto j = std::make_shared<Job>();
  j->rq = r;
  j->n = r.lmax - r.lmin + 1;
  ContractRequest& rq = j->crq;
  rq.lmin = r.lmin;
  rq.lmax = r.lmax;
  rq.LW = r.LW;
  auto W = [&](int i) {
    j->keys.push_back(r.W[i]);
    return lookup(S, r.W[i], r.LW, "W").d;
  };
  switch (r.type) {
    case PSX_REQ_MCM:
      rq.nfam[mcm_family(r.kind)] = 1;
      rq.dW[0] = W(0);
      rq.epi = EPI_MCM;
      j->nout = 1;
      break;
    case PSX_REQ_MCM_PM:
      rq.nfam[FAM_PE] = 1;
      rq.nfam[FAM_PO] = 1;
      rq.dW[0] = rq.dW[1] = W(0);
      rq.epi = EPI_MCM;
      j->nout = 2;
      break;
    case PSX_REQ_MCM_SET: {
      rq.nfam[FAM_00] = 1;
      rq.nfam[FAM_02] = 2;
      rq.nfam[FAM_PE] = 1;
      rq.nfam[FAM_PO] = 1;
      const double* w[4] = {W(0), W(1), W(2), W(3)};
      const double* order[5] = {w[0], w[1], w[2], w[3], w[3]};
      for (int k = 0; k < 5; ++k) rq.dW[k] = order[k];
      rq.epi = EPI_MCM;
      j->nout = 5;
      break;
    }
    case PSX_REQ_COV: {
      int nW = 0, nS = 0, nR = 0;
      PSX_REQUIRE(psx_cov_arity(r.kind, &nW, &nS, &nR) == PSX_OK, PSX_ERR_ARGUMENT,
                  "unknown covariance kind");
      const double* w[8] = {};
      for (int i = 0; i < nW; ++i) w[i] = W(i);
      cov_layout(r.kind, rq, w);
      rq.epi = EPI_COV;
      rq.cov_kind = r.kind;
      for (int i = 0; i < 4; ++i) {
        rq.dspec[i] = nullptr;
        rq.dratio[i] = nullptr;
      }
      for (int i = 0; i < nS; ++i) {
        j->keys.push_back(r.spectra[i]);
        rq.dspec[i] = lookup(S, r.spectra[i], r.lmax + 1, "spectrum").d;
      }
      for (int i = 0; i < nR; ++i) {
        j->keys.push_back(r.ratios[i]);
        rq.dratio[i] = lookup(S, r.ratios[i], r.lmax + 1, "noise ratio").d;
      }
      rq.precision = PSX_FP64;
      j->nout = 1;
      break;
    }
    default:
      throw Error{PSX_ERR_ARGUMENT, "unknown request type"};
  }
  int nnull = 0;
  for (int k = 0; k < j->nout; ++k) nnull += r.out[k] == nullptr;
  PSX_REQUIRE(nnull == 0 || nnull == j->nout, PSX_ERR_ARGUMENT,
              "give all host outputs of the request or none");
  j->host_out = nnull == 0;
  if (j->host_out)
    PSX_REQUIRE(r.ld >= j->n, PSX_ERR_AXES, "leading dimension smaller than lmax-lmin+1");
  j->out_bytes = (size_t)j->nout * j->n * (size_t)j->n * sizeof(double);
  j->scratch_bytes = contraction_scratch_bytes(rq);
  // run time alone: ~n^3/8 (l1 <= l2, c) steps at the FP64 rate the
  // contraction kernels sustain, ~1.6 VALU ops per step per table, plus the
  // launch and copy-back overheads
  {
    const double n3 = (double)j->n * j->n * j->n;
    const int nv = rq.nfam[0] + rq.nfam[1] + rq.nfam[2] + rq.nfam[3];
    const double ops = (2.0 + 2.5 * nv) * n3 / 8.0;
    j->est = ops / 1.6e13 + 3e-5 +
             (j->host_out ? (double)j->out_bytes / 3e10 : 0.0);
  }
  PSX_REQUIRE(Arena::granule(j->out_bytes) + Arena::granule(j->scratch_bytes) <= S->arena.size(),
              PSX_ERR_BUDGET, "request does not fit in the server arena");
  return j;
}

template <class F>
int sguard(Server* S, F&& f) {
  if (!S) return PSX_ERR_ARGUMENT;
  try {
    f();
    return PSX_OK;
  } catch (const Error& e) {
    t_server_error = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    t_server_error = e.what();
    return PSX_ERR_DEVICE;
  }
}

void free_server(Server* S) {
  for (auto& sl : S->slots) {
    if (sl->stream) (void)hipStreamSynchronize(sl->stream);
    if (sl->stream) (void)hipStreamDestroy(sl->stream);
    for (int i = 0; i < 2; ++i) {
      if (sl->ev[i]) (void)hipEventDestroy(sl->ev[i]);
      if (sl->stage[i]) (void)hipHostFree(sl->stage[i]);
    }
  }
  std::lock_guard<std::recursive_mutex> lk(S->ctx->mu);
  for (auto& kv : S->resident) {
    try {
      ctx_free(S->ctx, kv.second.d);
    } catch (...) {
    }
  }
  if (S->arena_mem) {
    try {
      ctx_free(S->ctx, S->arena_mem);
    } catch (...) {
    }
  }
}

}  // namespace
}  // namespace psx

using namespace psx;

extern "C" {

int psx_server_create(psx_ctx* hctx, int slots, size_t arena_bytes, int lmax_max,
                      psx_server** out) {
  Ctx* c = reinterpret_cast<Ctx*>(hctx);
  if (!c || !out) return PSX_ERR_ARGUMENT;
  *out = nullptr;
  auto* S = new Server();
  S->ctx = c;
  const int st = sguard(S, [&] {
    PSX_REQUIRE(slots >= 1 && slots <= 16, PSX_ERR_ARGUMENT, "slots must be 1..16");
    PSX_REQUIRE(lmax_max >= 0 && lmax_max <= 8191, PSX_ERR_ARGUMENT, "lmax_max must be 0..8191");
    PSX_REQUIRE(arena_bytes >= 4096, PSX_ERR_ARGUMENT, "arena too small");
    S->lmax_max = lmax_max;
    {
      std::lock_guard<std::recursive_mutex> lk(c->mu);
      PSX_HIP(hipSetDevice(c->device));
      // tables for the largest request, built now: they must not be
      // reallocated while requests run
      ensure_tables(c, (2 * lmax_max + 1) + 2 * lmax_max + 512);
      S->arena_mem = (char*)ctx_alloc(c, arena_bytes);
    }
    S->arena.init(S->arena_mem, arena_bytes);
    for (int i = 0; i < slots; ++i) {
      auto sl = std::make_unique<Slot>();
      PSX_HIP(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
      for (int b = 0; b < 2; ++b) {
        PSX_HIP(hipHostMalloc(&sl->stage[b], kStageBytes, hipHostMallocDefault));
        PSX_HIP(hipEventCreateWithFlags(&sl->ev[b], hipEventDisableTiming));
      }
      S->slots.push_back(std::move(sl));
    }
    for (auto& sl : S->slots) sl->th = std::thread(worker, S, sl.get());
  });
  if (st != PSX_OK) {
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& sl : S->slots)
      if (sl->th.joinable()) sl->th.join();
    free_server(S);
    delete S;
    return st;
  }
  {
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->live_servers++;
  }
  *out = reinterpret_cast<psx_server*>(S);
  return PSX_OK;
}

int psx_server_destroy(psx_server* hs) {
  Server* S = reinterpret_cast<Server*>(hs);
  if (!S) return PSX_ERR_ARGUMENT;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (auto& sl : S->slots)
    if (sl->th.joinable()) sl->th.join();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto& j : S->queue) {
      j->status = PSX_ERR_ARGUMENT;
      j->msg = "server destroyed before the request ran";
      j->done = true;
    }
    S->queue.clear();
  }
  S->cv_done.notify_all();
  free_server(S);
  {
    std::lock_guard<std::recursive_mutex> lk(S->ctx->mu);
    S->ctx->live_servers--;
  }
  delete S;
  return PSX_OK;
}

int psx_server_set_policy(psx_server* hs, int policy, double slack) {
  Server* S = reinterpret_cast<Server*>(hs);
  return sguard(S, [&] {
    PSX_REQUIRE(policy == PSX_SCHED_FIFO || policy == PSX_SCHED_DEADLINE, PSX_ERR_ARGUMENT,
                "unknown scheduling policy");
    PSX_REQUIRE(slack > 0.0, PSX_ERR_ARGUMENT, "slack must be positive");
    std::lock_guard<std::mutex> lk(S->mu);
    S->policy = policy;
    S->slack = slack;
  });
}

int psx_resident_put(psx_server* hs, uint64_t key, const double* values, int64_t len) {
  Server* S = reinterpret_cast<Server*>(hs);
  return sguard(S, [&] {
    PSX_REQUIRE(values && len >= 1, PSX_ERR_ARGUMENT, "need at least one value");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->resident.find(key);
    PSX_REQUIRE(it == S->resident.end() || it->second.refs == 0, PSX_ERR_ARGUMENT,
                "key is used by a pending request");
    std::lock_guard<std::recursive_mutex> lc(S->ctx->mu);
    PSX_HIP(hipSetDevice(S->ctx->device));
    double* d = (double*)ctx_alloc(S->ctx, sizeof(double) * (size_t)len);
    hipError_t e = hipMemcpy(d, values, sizeof(double) * (size_t)len, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      ctx_free(S->ctx, d);
      throw Error{PSX_ERR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e)};
    }
    if (it != S->resident.end()) ctx_free(S->ctx, it->second.d);
    S->resident[key] = Resident{d, len, 0};
  });
}

int psx_resident_evict(psx_server* hs, uint64_t key) {
  Server* S = reinterpret_cast<Server*>(hs);
  return sguard(S, [&] {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->resident.find(key);
    PSX_REQUIRE(it != S->resident.end(), PSX_ERR_ARGUMENT, "key is not resident");
    PSX_REQUIRE(it->second.refs == 0, PSX_ERR_ARGUMENT, "key is used by a pending request");
    std::lock_guard<std::recursive_mutex> lc(S->ctx->mu);
    ctx_free(S->ctx, it->second.d);
    S->resident.erase(it);
  });
}

int psx_submit(psx_server* hs, const psx_request* rq, uint64_t* ticket) {
  Server* S = reinterpret_cast<Server*>(hs);
  return sguard(S, [&] {
    PSX_REQUIRE(rq && ticket, PSX_ERR_ARGUMENT, "null pointer");
    {
      std::lock_guard<std::mutex> lk(S->mu);
      PSX_REQUIRE(!S->stop, PSX_ERR_ARGUMENT, "server is shutting down");
      auto j = build_job(S, *rq);
      j->ticket = S->next_ticket++;
      j->t_submit = Clock::now();
      j->deadline = secs(S->t0, j->t_submit) + S->slack * j->est;
      for (uint64_t k : j->keys) S->resident[k].refs++;
      S->jobs[j->ticket] = j;
      S->queue.push_back(j);
      *ticket = j->ticket;
    }
    S->cv_work.notify_one();
  });
}

int psx_poll(psx_server* hs, uint64_t ticket, int* done) {
  Server* S = reinterpret_cast<Server*>(hs);
  return sguard(S, [&] {
    PSX_REQUIRE(done, PSX_ERR_ARGUMENT, "null pointer");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(ticket);
    PSX_REQUIRE(it != S->jobs.end(), PSX_ERR_ARGUMENT, "unknown ticket");
    *done = it->second->done ? 1 : 0;
  });
}

int psx_wait(psx_server* hs, uint64_t ticket, int* status, double* times) {
  Server* S = reinterpret_cast<Server*>(hs);
  return sguard(S, [&] {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(ticket);
    PSX_REQUIRE(it != S->jobs.end(), PSX_ERR_ARGUMENT, "unknown ticket");
    std::shared_ptr<Job> j = it->second;
    S->cv_done.wait(lk, [&] { return j->done; });
    if (status) *status = j->status;
    if (j->status != PSX_OK) t_server_error = j->msg;
    if (times) {
      const bool ran = j->t_start.time_since_epoch().count() != 0;
      times[0] = ran ? secs(j->t_submit, j->t_start) : 0.0;
      times[1] = ran ? secs(j->t_start, j->t_done) : 0.0;
      times[2] = ran ? secs(j->t_submit, j->t_done) : 0.0;
    }
  });
}

int psx_fetch(psx_server* hs, uint64_t ticket, double* const* out, int64_t ld) {
  Server* S = reinterpret_cast<Server*>(hs);
  return sguard(S, [&] {
    PSX_REQUIRE(out, PSX_ERR_ARGUMENT, "null pointer");
    std::shared_ptr<Job> j;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      auto it = S->jobs.find(ticket);
      PSX_REQUIRE(it != S->jobs.end(), PSX_ERR_ARGUMENT, "unknown ticket");
      j = it->second;
      PSX_REQUIRE(j->done, PSX_ERR_NOT_READY, "request has not finished");
      PSX_REQUIRE(j->status == PSX_OK, j->status, j->msg);
      PSX_REQUIRE(j->out_mem, PSX_ERR_ARGUMENT, "result was copied to host outputs at submit");
    }
    PSX_REQUIRE(ld >= j->n, PSX_ERR_AXES, "leading dimension smaller than lmax-lmin+1");
    for (int k = 0; k < j->nout; ++k) PSX_REQUIRE(out[k], PSX_ERR_ARGUMENT, "null output");
    std::lock_guard<std::recursive_mutex> lc(S->ctx->mu);
    PSX_HIP(hipSetDevice(S->ctx->device));
    const size_t blk = (size_t)j->n * j->n;
    D2HMatrix m[5];
    for (int k = 0; k < j->nout; ++k)
      m[k] = {out[k], ld, (const double*)j->out_mem + k * blk, j->n, j->n, j->n};
    d2h_matrices(S->ctx, m, j->nout);
  });
}

int psx_release(psx_server* hs, uint64_t ticket) {
  Server* S = reinterpret_cast<Server*>(hs);
  const int st = sguard(S, [&] {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(ticket);
    PSX_REQUIRE(it != S->jobs.end(), PSX_ERR_ARGUMENT, "unknown ticket");
    PSX_REQUIRE(it->second->done, PSX_ERR_NOT_READY, "request has not finished");
    S->arena.release(it->second->out_mem);
    it->second->out_mem = nullptr;
    S->jobs.erase(it);
  });
  if (st == PSX_OK) S->cv_work.notify_all();
  return st;
}

const char* psx_server_last_error(psx_server* hs) {
  (void)hs;
  return t_server_error.c_str();
}

}  // extern "C"
