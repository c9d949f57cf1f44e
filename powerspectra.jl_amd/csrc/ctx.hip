// Context, device memory, kernel timing and the Wigner tables.
#include "psx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <thread>
#include <vector>

namespace psx {

void* ctx_alloc(Ctx* c, size_t bytes) {
  if (bytes == 0) bytes = 8;
  // the idle cached buffers are caches: give them back before failing
  for (CachedSlot* sl : {&c->xfer, &c->work}) {
    if (c->in_use + bytes <= c->budget) break;
    if (sl->p && !sl->busy) {
      PSX_HIP(hipStreamSynchronize(c->stream));
      void* p = sl->p;
      sl->p = nullptr;
      sl->bytes = 0;
      ctx_free(c, p);
    }
  }
  PSX_REQUIRE(c->in_use + bytes <= c->budget, PSX_ERR_BUDGET,
              "allocation of " + std::to_string(bytes) +
                  " bytes exceeds the context HBM budget");
  void* p = nullptr;
  PSX_HIP(hipMalloc(&p, bytes));
  c->allocations[p] = bytes;
  c->in_use += bytes;
  return p;
}

void ctx_free(Ctx* c, void* p) {
  if (!p) return;
  auto it = c->allocations.find(p);
  PSX_REQUIRE(it != c->allocations.end(), PSX_ERR_ARGUMENT,
              "pointer not owned by this context");
  c->in_use -= it->second;
  c->allocations.erase(it);
  PSX_HIP(hipFree(p));
}

void* ctx_scratch(Ctx* c, size_t bytes) {
  if (bytes <= c->scratch_bytes) return c->scratch;
  if (c->scratch) {
    PSX_HIP(hipStreamSynchronize(c->stream));
    ctx_free(c, c->scratch);
    c->scratch = nullptr;
    c->scratch_bytes = 0;
  }
  size_t want = bytes + bytes / 4;
  c->scratch = ctx_alloc(c, want);
  c->scratch_bytes = want;
  return c->scratch;
}

HeldBuf::HeldBuf(Ctx* cc, CachedSlot& s, size_t bytes) : c(cc), slot(&s) {
  PSX_REQUIRE(!s.busy, PSX_ERR_ARGUMENT, "cached device buffer already in use");
  if (bytes > s.bytes) {
    if (s.p) {
      PSX_HIP(hipStreamSynchronize(c->stream));
      void* p = s.p;
      s.p = nullptr;
      s.bytes = 0;
      ctx_free(c, p);
    }
    s.p = ctx_alloc(c, bytes);
    s.bytes = bytes;
  }
  s.busy = true;
  cap = bytes;
}

HeldBuf::~HeldBuf() { slot->busy = false; }

double* HeldBuf::take(size_t bytes) {
  const size_t at = (used + 255) & ~size_t(255);
  PSX_REQUIRE(at + bytes <= cap, PSX_ERR_ARGUMENT, "cached buffer carve overflow");
  used = at + bytes;
  return reinterpret_cast<double*>(static_cast<char*>(slot->p) + at);
}

TimedLaunch::TimedLaunch(Ctx* ctx, const char* kname) : c(ctx), name(kname) {
  if (c->timing) {
    PSX_HIP(hipEventCreate(&e0));
    PSX_HIP(hipEventCreate(&e1));
    PSX_HIP(hipEventRecord(e0, c->stream));
  }
}

TimedLaunch::~TimedLaunch() noexcept(false) {
  hipError_t le = hipGetLastError();
  if (e0) {
    (void)hipEventRecord(e1, c->stream);
    c->timers[name].pending.emplace_back(e0, e1);
  }
  // while another error unwinds the stack (a failed call inside the timed
  // scope), that error is the one reported: a second throw would terminate
  if (le != hipSuccess && std::uncaught_exceptions() == 0)
    throw Error{PSX_ERR_DEVICE, std::string("launch of ") + name + ": " +
                                    hipGetErrorString(le)};
}

// Wait for the recorded launches and fold their durations into the totals.
void resolve_timers(Ctx* c) {
  for (auto& kv : c->timers) {
    for (auto& ev : kv.second.pending) {
      PSX_HIP(hipEventSynchronize(ev.second));
      float ms = 0;
      PSX_HIP(hipEventElapsedTime(&ms, ev.first, ev.second));
      kv.second.total_ms += ms;
      kv.second.launches += 1;
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    kv.second.pending.clear();
  }
}

// gamma_m = prod_{k<=m} (2k-1)/(2k) accumulated in long double (64-bit
// mantissa); the rounding to double of each entry is then the dominant error.
int table_entries(int nmax_index) { return (nmax_index + 2 * kTabOff + 1023) & ~1023; }

void build_tables_host(int n, double* G, double* Z, double* D, double* IG) {
  for (int i = 0; i < n; ++i) G[i] = Z[i] = D[i] = IG[i] = 0.0;
  long double g = 1.0L;
  for (int m = 0; m + kTabOff < n; ++m) {
    if (m > 0) g = g * (long double)(2 * m - 1) / (long double)(2 * m);
    int i = m + kTabOff;
    G[i] = (double)g;
    Z[i] = (double)((long double)(2 * m + 1) * g);
    D[i] = (double)(1.0L / ((long double)(2 * m + 1) * g));
    IG[i] = (double)(1.0L / g);
  }
}

void ensure_tables(Ctx* c, int nmax_index) {
  Tables& t = c->tables;
  if (t.n >= nmax_index + 2 * kTabOff) return;
  const int n = table_entries(nmax_index);
  std::vector<double> G(n), Z(n), D(n), IG(n);
  build_tables_host(n, G.data(), Z.data(), D.data(), IG.data());
  // other streams (a server's slots) may read the old tables
  PSX_HIP(hipDeviceSynchronize());
  if (t.dG) {
    ctx_free(c, t.dG);
    ctx_free(c, t.dZ);
    ctx_free(c, t.dD);
    ctx_free(c, t.dIG);
  }
  size_t b = sizeof(double) * n;
  t.dG = (double*)ctx_alloc(c, b);
  t.dZ = (double*)ctx_alloc(c, b);
  t.dD = (double*)ctx_alloc(c, b);
  t.dIG = (double*)ctx_alloc(c, b);
  PSX_HIP(hipMemcpy(t.dG, G.data(), b, hipMemcpyHostToDevice));
  PSX_HIP(hipMemcpy(t.dZ, Z.data(), b, hipMemcpyHostToDevice));
  PSX_HIP(hipMemcpy(t.dD, D.data(), b, hipMemcpyHostToDevice));
  PSX_HIP(hipMemcpy(t.dIG, IG.data(), b, hipMemcpyHostToDevice));
  t.n = n;
}

}  // namespace psx

using namespace psx;

extern "C" {

int psx_ctx_create(int device, size_t hbm_budget, psx_ctx** out) {
  if (!out) return PSX_ERR_ARGUMENT;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PSX_ERR_DEVICE;
  if (device < 0 || device >= ndev) return PSX_ERR_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return PSX_ERR_DEVICE;
  Ctx* c = new Ctx();
  c->device = device;
  size_t freeb = 0, total = 0;
  if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
    delete c;
    return PSX_ERR_DEVICE;
  }
  c->budget = hbm_budget ? hbm_budget : (size_t)(0.9 * (double)freeb);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return PSX_ERR_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    c->cu_count = prop.multiProcessorCount;
  try {
    noalloc_init_device(device);
  } catch (const Error&) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return PSX_ERR_DEVICE;
  }
  *out = reinterpret_cast<psx_ctx*>(c);
  return PSX_OK;
}

int psx_ctx_destroy(psx_ctx* h) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PSX_ERR_ARGUMENT;
  {
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (c->live_servers > 0) {
      c->last_error = "context still has " + std::to_string(c->live_servers) +
                      " live server(s): destroy them first (psx_server_destroy)";
      return PSX_ERR_ARGUMENT;
    }
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    try {
      resolve_timers(c);
    } catch (...) {
    }
    // library handles first, while their stream and the memory they may
    // still reference exist; then the plans, the memory and the streams
    release_blas(c);
    release_tile_cache(c);
    release_mix_tiles(c);
    release_sht_plans(c);
    for (auto& kv : c->allocations) (void)hipFree(kv.first);
    c->allocations.clear();
    release_stage(c);
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
  return PSX_OK;
}

int psx_ctx_synchronize(psx_ctx* h) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_HIP(hipSetDevice(c->device));
    PSX_HIP(hipStreamSynchronize(c->stream));
  });
}

const char* psx_ctx_last_error(psx_ctx* h) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return "null context";
  return c->last_error.c_str();
}

size_t psx_ctx_bytes_in_use(psx_ctx* h) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return c ? c->in_use : 0;
}

int psx_ctx_set_timing(psx_ctx* h, int enable) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] { c->timing = enable != 0; });
}

int psx_ctx_kernel_time(psx_ctx* h, const char* name, int reset,
                        double* total_ms, int64_t* launches) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_REQUIRE(name != nullptr, PSX_ERR_ARGUMENT, "null kernel name");
    resolve_timers(c);
    auto it = c->timers.find(name);
    double ms = 0;
    int64_t n = 0;
    if (it != c->timers.end()) {
      ms = it->second.total_ms;
      n = it->second.launches;
      if (reset) c->timers.erase(it);
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = n;
  });
}

int psx_dev_alloc(psx_ctx* h, size_t bytes, void** dptr) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_REQUIRE(dptr != nullptr, PSX_ERR_ARGUMENT, "null output pointer");
    PSX_HIP(hipSetDevice(c->device));
    *dptr = ctx_alloc(c, bytes);
  });
}

int psx_dev_free(psx_ctx* h, void* dptr) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_HIP(hipSetDevice(c->device));
    PSX_HIP(hipStreamSynchronize(c->stream));
    ctx_free(c, dptr);
  });
}

int psx_memcpy_h2d(psx_ctx* h, void* dst, const void* src, size_t bytes) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_HIP(hipSetDevice(c->device));
    PSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    PSX_HIP(hipStreamSynchronize(c->stream));
  });
}

int psx_memcpy_d2h(psx_ctx* h, void* dst, const void* src, size_t bytes) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded(c, [&] {
    PSX_HIP(hipSetDevice(c->device));
    PSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    PSX_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"

// ---- staged host copies ---------------------------------------------------
// Copies between device memory and pageable host memory run far below PCIe
// speed; large ones go through two pinned buffers so that the DMA into (or
// out of) one overlaps host threads copying the other.
namespace psx {

namespace {
constexpr size_t kDirectBytes = size_t(1) << 20;
// A contiguous range longer than a stage buffer (one long column, or columns
// packed on both sides: a map, a set of a_lm) is recut into columns of this
// many doubles (512 KB), so that it goes through the staged pipeline and its
// host copies spread over threads.
constexpr int64_t kLineDoubles = int64_t(1) << 16;

void ensure_stage(Ctx* c) {
  if (c->stage[0]) return;
  for (int i = 0; i < 2; ++i) {
    PSX_HIP(hipHostMalloc(&c->stage[i], kStageBytes, hipHostMallocDefault));
    PSX_HIP(hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming));
  }
  PSX_HIP(hipStreamCreateWithFlags(&c->xstream, hipStreamNonBlocking));
}

// dst(:, j) = src(:, j) for j < cols, column-major, split over host threads
void copy_cols(double* dst, int64_t ldd, const double* src, int64_t lds, int64_t rows,
               int64_t cols) {
  const size_t bytes = sizeof(double) * (size_t)rows * (size_t)cols;
  const int nt = (int)std::min<int64_t>(
      {8, (int64_t)std::max(1u, std::thread::hardware_concurrency()), cols,
       (int64_t)(bytes >> 22) + 1});
  auto work = [&](int t) {
    const int64_t j0 = cols * t / nt, j1 = cols * (t + 1) / nt;
    if (ldd == rows && lds == rows) {
      std::memcpy(dst + j0 * rows, src + j0 * rows, sizeof(double) * rows * (j1 - j0));
    } else {
      for (int64_t j = j0; j < j1; ++j)
        std::memcpy(dst + j * ldd, src + j * lds, sizeof(double) * rows);
    }
  };
  if (nt <= 1) {
    work(0);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
}
}  // namespace

void release_stage(Ctx* c) {
  if (c->xstream) (void)hipStreamSynchronize(c->xstream);
  for (int i = 0; i < 2; ++i) {
    if (c->stage_ev[i]) (void)hipEventDestroy(c->stage_ev[i]);
    if (c->stage[i]) (void)hipHostFree(c->stage[i]);
    c->stage_ev[i] = nullptr;
    c->stage[i] = nullptr;
  }
  if (c->xstream) (void)hipStreamDestroy(c->xstream);
  c->xstream = nullptr;
  for (hipEvent_t& e : c->band_ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
}

void ensure_band_events(Ctx* c) {
  for (hipEvent_t& e : c->band_ev)
    if (!e) PSX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

namespace {
// one column block of a staged copy: nc columns of `rows` doubles
struct Chunk {
  double* host;
  int64_t ldh;
  const double* dev;
  int64_t ldd;
  int64_t rows, nc;
  hipEvent_t after;
};

void chunk_to_stage(hipStream_t s, void* stage, const Chunk& k) {
  const size_t colb = sizeof(double) * (size_t)k.rows;
  if (k.ldd == k.rows)
    PSX_HIP(hipMemcpyAsync(stage, k.dev, colb * (size_t)k.nc, hipMemcpyDeviceToHost, s));
  else
    PSX_HIP(hipMemcpy2DAsync(stage, colb, k.dev, k.ldd * sizeof(double), colb, (size_t)k.nc,
                             hipMemcpyDeviceToHost, s));
}
}  // namespace

void d2h_matrices_on(hipStream_t s, void* const stage[2], hipEvent_t const ev[2],
                     const D2HMatrix* m, int nm) {
  // Every matrix is cut into stage-sized column blocks and all blocks go
  // through one two-buffer pipeline, so the DMA stays busy across matrix
  // boundaries.  Small matrices, and columns longer than a stage buffer, are
  // copied directly, ahead of the pipeline.
  std::vector<Chunk> ch;
  for (int i = 0; i < nm; ++i) {
    const D2HMatrix& x = m[i];
    if (x.rows <= 0 || x.cols <= 0) continue;
    const size_t colb = sizeof(double) * (size_t)x.rows;
    if (colb > kStageBytes && (x.cols == 1 || (x.ldh == x.rows && x.ldd == x.rows))) {
      const int64_t total = x.rows * x.cols, full = total / kLineDoubles;
      const int64_t rem = total % kLineDoubles;
      const int64_t cpc = (int64_t)(kStageBytes / (sizeof(double) * kLineDoubles));
      for (int64_t j0 = 0; j0 < full; j0 += cpc)
        ch.push_back({x.host + j0 * kLineDoubles, kLineDoubles, x.dev + j0 * kLineDoubles,
                      kLineDoubles, kLineDoubles, std::min(cpc, full - j0), x.after});
      if (rem)
        ch.push_back({x.host + full * kLineDoubles, rem, x.dev + full * kLineDoubles, rem, rem,
                      1, x.after});
      continue;
    }
    // (a piece that waits for an event is always staged: a direct copy to
    // pageable memory would hold this thread until the event fires)
    if ((!x.after && colb * (size_t)x.cols <= kDirectBytes) || colb > kStageBytes) {
      if (x.after) PSX_HIP(hipStreamWaitEvent(s, x.after, 0));
      PSX_HIP(hipMemcpy2DAsync(x.host, x.ldh * sizeof(double), x.dev, x.ldd * sizeof(double),
                               colb, (size_t)x.cols, hipMemcpyDeviceToHost, s));
      continue;
    }
    const int64_t cpc = (int64_t)(kStageBytes / colb);  // columns per block
    for (int64_t j0 = 0; j0 < x.cols; j0 += cpc)
      ch.push_back({x.host + j0 * x.ldh, x.ldh, x.dev + j0 * x.ldd, x.ldd, x.rows,
                    std::min(cpc, x.cols - j0), x.after});
  }
  const size_t nch = ch.size();
  auto issue = [&](size_t k) {
    if (ch[k].after) PSX_HIP(hipStreamWaitEvent(s, ch[k].after, 0));
    chunk_to_stage(s, stage[k & 1], ch[k]);
    PSX_HIP(hipEventRecord(ev[k & 1], s));
  };
  if (nch > 0) issue(0);
  if (nch > 1) issue(1);
  for (size_t k = 0; k < nch; ++k) {
    const Chunk& x = ch[k];
    PSX_HIP(hipEventSynchronize(ev[k & 1]));
    copy_cols(x.host, x.ldh, (const double*)stage[k & 1], x.rows, x.rows, x.nc);
    if (k + 2 < nch) issue(k + 2);
  }
  PSX_HIP(hipStreamSynchronize(s));
}

void d2h_matrices(Ctx* c, const D2HMatrix* m, int nm) {
  for (int i = 0; i < nm; ++i)
    if ((size_t)m[i].rows * (size_t)m[i].cols * sizeof(double) > kDirectBytes || m[i].after)
      ensure_stage(c);
  d2h_matrices_on(c->stream, c->stage, c->stage_ev, m, nm);
}

void d2h_matrix(Ctx* c, double* host, int64_t ldh, const double* dev, int64_t ldd,
                int64_t rows, int64_t cols) {
  const D2HMatrix m{host, ldh, dev, ldd, rows, cols};
  d2h_matrices(c, &m, 1);
}

void d2h_matrices_overlapped(Ctx* c, const D2HMatrix* m, int nm) {
  ensure_stage(c);
  d2h_matrices_on(c->xstream, c->stage, c->stage_ev, m, nm);
}

namespace {
void h2d_matrix_on(Ctx* c, hipStream_t s, double* dev, int64_t ldd, const double* host,
                   int64_t ldh, int64_t rows, int64_t cols) {
  const size_t colb = sizeof(double) * (size_t)rows;
  if (rows <= 0 || cols <= 0) return;
  if (colb > kStageBytes && (cols == 1 || (ldd == rows && ldh == rows))) {
    const int64_t total = rows * cols, full = total / kLineDoubles, rem = total % kLineDoubles;
    if (full) h2d_matrix_on(c, s, dev, kLineDoubles, host, kLineDoubles, kLineDoubles, full);
    if (rem)
      h2d_matrix_on(c, s, dev + full * kLineDoubles, rem, host + full * kLineDoubles, rem, rem,
                    1);
    return;
  }
  if (colb * (size_t)cols <= kDirectBytes || colb > kStageBytes) {
    PSX_HIP(hipMemcpy2DAsync(dev, ldd * sizeof(double), host, ldh * sizeof(double), colb,
                             (size_t)cols, hipMemcpyHostToDevice, s));
    PSX_HIP(hipStreamSynchronize(s));
    return;
  }
  const int64_t cpc = (int64_t)(kStageBytes / colb);
  const int64_t nch = (cols + cpc - 1) / cpc;
  for (int64_t k = 0; k < nch; ++k) {
    const int b = (int)(k & 1);
    const int64_t j0 = k * cpc, nc = std::min(cpc, cols - j0);
    PSX_HIP(hipEventSynchronize(c->stage_ev[b]));  // buffer b free again
    copy_cols((double*)c->stage[b], rows, host + j0 * ldh, ldh, rows, nc);
    if (ldd == rows)
      PSX_HIP(hipMemcpyAsync(dev + j0 * ldd, c->stage[b], colb * (size_t)nc,
                             hipMemcpyHostToDevice, s));
    else
      PSX_HIP(hipMemcpy2DAsync(dev + j0 * ldd, ldd * sizeof(double), c->stage[b], colb, colb,
                               (size_t)nc, hipMemcpyHostToDevice, s));
    PSX_HIP(hipEventRecord(c->stage_ev[b], s));
  }
  PSX_HIP(hipStreamSynchronize(s));
}
}  // namespace

void h2d_matrix(Ctx* c, double* dev, int64_t ldd, const double* host, int64_t ldh,
                int64_t rows, int64_t cols) {
  if ((size_t)rows * (size_t)cols * sizeof(double) > kDirectBytes) ensure_stage(c);
  h2d_matrix_on(c, c->stream, dev, ldd, host, ldh, rows, cols);
}

void h2d_matrix_x(Ctx* c, double* dev, int64_t ldd, const double* host, int64_t ldh,
                  int64_t rows, int64_t cols) {
  ensure_stage(c);
  h2d_matrix_on(c, c->xstream, dev, ldd, host, ldh, rows, cols);
}

}  // namespace psx
