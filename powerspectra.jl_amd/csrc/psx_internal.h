// Internal declarations of the gfx950 engine (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <initializer_list>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/powerspectra_hip.h"

namespace psx {

// ---- error plumbing ------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};

#define PSX_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      throw ::psx::Error{PSX_ERR_DEVICE, std::string(#call) + ": " +        \
                                             hipGetErrorString(e_)};         \
  } while (0)

#define PSX_REQUIRE(cond, code, text)                 \
  do {                                                \
    if (!(cond)) throw ::psx::Error{(code), (text)}; \
  } while (0)

// ---- Wigner tables -------------------------------------------------------
// Host-built (long double) tables of the central-binomial ratio
//   gamma_m = C(2m, m) / 4^m = prod_{k=1..m} (2k-1)/(2k)
// and the derived tables the closed-form 3j products need:
//   G[m]  = gamma_m                     (0 for m < 0)
//   Z[m]  = (2m+1) gamma_m              (0 for m < 0)
//   D[g]  = 1 / ((2g+1) gamma_g)
//   IG[g] = 1 / gamma_g
// Index m is stored at offset m + kTabOff.
constexpr int kTabOff = 256;

struct Tables {
  int n = 0;            // entries per table (indices -kTabOff .. n-kTabOff-1)
  double* dG = nullptr;
  double* dZ = nullptr;
  double* dD = nullptr;
  double* dIG = nullptr;
};

// ---- context -------------------------------------------------------------
struct KernelTimer {
  double total_ms = 0;
  int64_t launches = 0;
  // event pairs recorded on the context stream, resolved on query
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

// A device buffer the context keeps between calls; busy while held.
struct CachedSlot {
  void* p = nullptr;
  size_t bytes = 0;
  bool busy = false;
};

struct Ctx {
  int device = 0;
  size_t budget = 0;
  size_t in_use = 0;
  hipStream_t stream = nullptr;
  std::recursive_mutex mu;
  std::string last_error;
  bool timing = false;
  std::map<std::string, KernelTimer> timers;
  std::map<void*, size_t> allocations;
  Tables tables;
  int cu_count = 0;
  // reusable device scratch (grown on demand)
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // rocSOLVER/rocBLAS handle (created lazily)
  void* blas = nullptr;
  // two device buffers kept between calls (grown on demand, given back when
  // an allocation would otherwise exceed the budget): the device copies of a
  // host-pointer call's inputs and outputs, and the work area of a transform
  CachedSlot xfer, work;
  // two pinned host staging buffers for large host <-> device copies, and the
  // stream that copies device results out while later kernels still run
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  hipStream_t xstream = nullptr;
  // one event per l1 band of a banded contraction
  hipEvent_t band_ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  // extra streams, rocBLAS handles and events for independent factorizations
  // run side by side (lu_solve_concurrent), created on first use
  hipStream_t lstream[3] = {nullptr, nullptr, nullptr};
  void* lblas[3] = {nullptr, nullptr, nullptr};
  hipEvent_t lane_ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // servers created on this context and not yet destroyed: their worker
  // threads, arenas and resident spectra live on the context, so it is not
  // destroyed while any is alive
  int live_servers = 0;
};
constexpr size_t kMaxBands = 5;
void ensure_band_events(Ctx* c);

void resolve_timers(Ctx* c);
void* ctx_alloc(Ctx* c, size_t bytes);
void ctx_free(Ctx* c, void* p);
void* ctx_scratch(Ctx* c, size_t bytes);
// One of the context's cached buffers held for the duration of a call;
// take() carves 256-byte aligned pieces, and the size passed to the
// constructor must cover them (xfer_bytes(...) sums pieces with their
// padding).  Holding a slot that is already held raises.
struct HeldBuf {
  Ctx* c;
  CachedSlot* slot;
  size_t cap = 0, used = 0;
  HeldBuf(Ctx* cc, CachedSlot& s, size_t bytes);
  ~HeldBuf();
  HeldBuf(const HeldBuf&) = delete;
  HeldBuf& operator=(const HeldBuf&) = delete;
  double* take(size_t bytes);
  void* base() const { return slot->p; }
};
// the transfer buffer of a host-pointer call
struct XferBuf : HeldBuf {
  XferBuf(Ctx* cc, size_t bytes) : HeldBuf(cc, cc->xfer, bytes) {}
};
// the work area of a transform (spherical-harmonic passes, cross-spectra)
struct WorkBuf : HeldBuf {
  WorkBuf(Ctx* cc, size_t bytes) : HeldBuf(cc, cc->work, bytes) {}
};
inline size_t xfer_bytes(std::initializer_list<size_t> pieces) {
  size_t t = 0;
  for (size_t b : pieces) t += (b + 255) & ~size_t(255);
  return t;
}
void ensure_tables(Ctx* c, int nmax_index);
// entries per table covering indices up to nmax_index, and the host-side
// build of the four tables (n entries each, index m at m + kTabOff)
int table_entries(int nmax_index);
void build_tables_host(int n, double* G, double* Z, double* D, double* IG);
// The allocation-free mode-coupling path (psx_mcm_d_noalloc): the device
// address of this device's static scratch array, looked up once per device
// (called by psx_ctx_create with the device current).
void noalloc_init_device(int device);
void release_blas(Ctx* c);
void release_stage(Ctx* c);
void release_tile_cache(Ctx* c);
void release_sht_plans(Ctx* c);
// spherical-harmonic transforms on device buffers (context stream)
// `before`, if given, is called with map index q just before map q is first
// read (its ring transforms are enqueued): a caller can still be copying the
// later maps in while the first ones are transformed.
void map2alm_dev(Ctx* c, int nside, int lmax, int niter, int nmaps, const double* dmaps,
                 double* dalms, const std::function<void(int)>* before = nullptr);
void alm2map_dev(Ctx* c, int nside, int lmax, int nmaps, const double* dalms, double* dmaps);
// (spin0 = false: the spin-2 part only, the I maps' alms left as they are;
// `between`, if given, is called once the spin-0 work is enqueued and before
// Q and U are first read: a caller can still be copying them in)
void map2alm_pol_dev(Ctx* c, int nside, int lmax, int niter, int nmaps, const double* diqu,
                     double* dteb, bool spin0 = true,
                     const std::function<void()>* between = nullptr);
void alm2map_pol_dev(Ctx* c, int nside, int lmax, int nmaps, const double* dteb, double* diqu);

// Column-major matrix copies between host memory (any pageable buffer,
// leading dimension ldh) and device memory (leading dimension ldd), in
// stream order on the context stream; they return when the host side is
// complete.  Large copies are staged through the context's pinned buffers.
void d2h_matrix(Ctx* c, double* host, int64_t ldh, const double* dev, int64_t ldd,
                int64_t rows, int64_t cols);
// size of each of the two pinned staging buffers (the context's, and each
// server slot's)
constexpr size_t kStageBytes = size_t(32) << 20;
// several matrices through one pipeline (the DMA does not drain between them)
struct D2HMatrix {
  double* host;
  int64_t ldh;
  const double* dev;
  int64_t ldd;
  int64_t rows, cols;
  hipEvent_t after = nullptr;  // the copy waits for this event when set
};
void d2h_matrices(Ctx* c, const D2HMatrix* m, int nm);
// the same on the context's copy stream, so that the pieces waiting for
// events of the context stream copy out while later work there still runs
void d2h_matrices_overlapped(Ctx* c, const D2HMatrix* m, int nm);
// the same on stream s through two caller-owned pinned buffers of
// kStageBytes each and their events
void d2h_matrices_on(hipStream_t s, void* const stage[2], hipEvent_t const ev[2],
                     const D2HMatrix* m, int nm);
void h2d_matrix(Ctx* c, double* dev, int64_t ldd, const double* host, int64_t ldh,
                int64_t rows, int64_t cols);
// The same copy through the staging pipeline on the context's copy stream:
// returns when the data is on the device, without waiting for (or ordering
// against) work on the context stream.  The caller keeps both streams off the
// destination meanwhile.
void h2d_matrix_x(Ctx* c, double* dev, int64_t ldd, const double* host, int64_t ldh,
                  int64_t rows, int64_t cols);

// Launch helper: records HIP events around the launch when timing is on.
struct TimedLaunch {
  Ctx* c;
  const char* name;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TimedLaunch(Ctx* ctx, const char* kname);
  ~TimedLaunch() noexcept(false);
};

// ---- contraction engine --------------------------------------------------
// Family products a V table is contracted against (closed forms, see
// contract.hip):
enum Family : int {
  FAM_00 = 0,  // (l3 l1 l2; 0 0 0)^2                          even l1+l2+l3
  FAM_02 = 1,  // (l3 l1 l2; 0 0 0)(l3 l1 l2; 0 -2 2)          even
  FAM_PE = 2,  // (l3 l1 l2; 0 -2 2)^2                         even
  FAM_PO = 3   // (l3 l1 l2; 0 -2 2)^2                         odd
};

enum EpiMode : int {
  EPI_MCM = 0,   // out_k[l1,l2] = (2l2+1) kappa Ξ_k, out_k[l2,l1] = (2l1+1) kappa Ξ_k
  EPI_COV = 1,   // out[l1,l2] = out[l2,l1] = sum_k coef_k(l1,l2) Ξ_k
  EPI_RMAX = 2   // one V table taken raw (no (2 l3 + 1) / 4 pi): out[doff[S - 2 lmin] + l1 -
                 // max(lmin, S - lmax)] = max over the family of |term x V(l3)|
};

struct ContractRequest {
  int lmin = 0, lmax = 0;
  int64_t LW = 0;
  // V tables in family order: all FAM_00, then FAM_02, FAM_PE, FAM_PO.
  int nfam[4] = {0, 0, 0, 0};
  const double* dW[16];      // source W (device) for each V table, in order
  EpiMode epi = EPI_MCM;
  // EPI_MCM: one output per V table, ld = ldo
  double* dout[16];
  int64_t ldo = 0;
  // EPI_COV
  int cov_kind = 0;
  const double* dspec[4];
  const double* dratio[4];
  int precision = PSX_FP64;
  // EPI_RMAX: the anti-diagonal row offsets of the output table
  const int64_t* doff = nullptr;
};

void run_contraction(Ctx* c, const ContractRequest& rq);

// A matrix-core pass (mixed.hip): ncol W spectra against one family, into
// ncol mode-coupling blocks (EPI_MCM) or nblk covariance blocks of nW
// spectra each (EPI_COV), in precision PSX_MIXED or PSX_FP64_MFMA.
struct MixRequest {
  int lmin = 0, lmax = 0;
  int64_t LW = 0;
  int fam = FAM_00;
  int ncol = 0;
  const double* dW[64] = {};
  EpiMode epi = EPI_MCM;
  double* dout[64] = {};
  int64_t ldo = 0;
  int cov_kind = 0, nblk = 0, nW = 0;
  const double* dspec[16][4] = {};
  const double* dratio[16][4] = {};
  int mode = PSX_MIXED;
};
void run_mixed(Ctx* c, const MixRequest& rq);
// the one family a covariance kind contracts (ArgumentError for TETE)
int mix_family_of_cov(int kind);
void release_mix_tiles(Ctx* c);
// The same through the device's static scratch array only: no device
// allocation (psx_mcm_d_noalloc).
void run_contraction_noalloc(Ctx* c, const ContractRequest& rq);
// Enqueue a contraction on stream s with caller-provided scratch (at least
// contraction_scratch_bytes(rq)).  Returns the device address of the
// covariance domain-error flag inside the scratch (nonzero after the kernel
// when a spectrum product under a square root was negative), or nullptr for
// mode-coupling requests.  On the context stream the flag is checked here.
int* run_contraction_on(Ctx* c, const ContractRequest& rq, hipStream_t s,
                        void* scratch, size_t scratch_bytes);
size_t contraction_scratch_bytes(const ContractRequest& rq);
// A contraction into packed device outputs (rq.ldo = n) followed by their
// copy to the host blocks host[k] (leading dimension ldh).  Large outputs run
// in l1 bands: the columns each band completes are copied out while the
// later bands compute.  Returns with the host blocks complete; raises the
// covariance domain error after the copy.
void run_contraction_to_host(Ctx* c, const ContractRequest& rq, double* const* host,
                             int64_t ldh);
// request building shared by the entry points and the server
int mcm_family(int kind);
void cov_layout(int kind, ContractRequest& rq, const double* const* W);

// ---- linear algebra ------------------------------------------------------
// C (m x n) = A (m x k) B (k x n), column-major FP64, rocBLAS on the context
// stream
void dgemm_dev(Ctx* c, int m, int n, int k, const double* dA, int64_t lda, const double* dB,
               int64_t ldb, double* dC, int64_t ldc);
void lu_solve_dev(Ctx* c, int n, double* dA, int64_t lda, double* dB,
                  int64_t ldb, int nrhs);  // dA overwritten by LU
void decouple_dev(Ctx* c, int n, const double* dY, double* dB1, double* dB2,
                  double* dOut);
// One of several independent systems A X = B on the device (A overwritten by
// its LU factors, B by X), with 1 to 4 right-hand sides and ipiv a device
// array of n + 64 ints (LAPACK interchanges, info slot at n + 32).
struct LuSystem {
  int n;
  double* dA;
  int64_t lda;
  double* dB;
  int64_t ldb;
  int nrhs;
  int* ipiv;
};
constexpr int kLuExtraLanes = 3;
// Factor and solve up to 1 + kLuExtraLanes systems at the same time, in
// stream order on the context stream; returns when all are solved and raises
// SingularException for a zero pivot.
void lu_solve_concurrent(Ctx* c, const LuSystem* sys, int count);

// status wrapper used by every extern "C" entry point
template <class F>
int guarded(Ctx* c, F&& f) {
  if (!c) return PSX_ERR_ARGUMENT;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  try {
    f();
    return PSX_OK;
  } catch (const Error& e) {
    c->last_error = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    c->last_error = e.what();
    return PSX_ERR_DEVICE;
  }
}

}  // namespace psx
