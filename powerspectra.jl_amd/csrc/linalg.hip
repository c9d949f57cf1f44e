// FP64 LU decoupling: `lu` and `A \ B` (reference
// src/blockspectralmatrix.jl:89-129) and decouple_covmat
// (src/covariance.jl:8-14).  Julia's `lu` is LAPACK getrf with partial
// pivoting and throws SingularException on an exact zero pivot.  The
// factorisation is this file's blocked getrf (HIP leaf kernels + rocBLAS
// updates) in LAPACK layout; the solves apply its interchanges (a gather for
// many columns, rocSOLVER's laswp for a few) and run this file's forward and
// back substitution (rocBLAS updates between diagonal blocks).
#include "psx_internal.h"

#include <utility>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace psx {

static rocblas_handle blas(Ctx* c) {
  if (!c->blas) {
    rocblas_handle h;
    PSX_REQUIRE(rocblas_create_handle(&h) == rocblas_status_success, PSX_ERR_DEVICE,
                "rocblas_create_handle failed");
    rocblas_set_stream(h, c->stream);
    c->blas = h;
  }
  return static_cast<rocblas_handle>(c->blas);
}

void release_lanes(Ctx* c);

void release_blas(Ctx* c) {
  release_lanes(c);
  if (c->blas) {
    rocblas_destroy_handle(static_cast<rocblas_handle>(c->blas));
    c->blas = nullptr;
  }
}

#define PSX_ROC(call)                                                         \
  do {                                                                        \
    rocblas_status s_ = (call);                                               \
    PSX_REQUIRE(s_ == rocblas_status_success, PSX_ERR_DEVICE,                 \
                std::string(#call) + " failed: " + rocblas_status_to_string(s_)); \
  } while (0)

void dgemm_dev(Ctx* c, int m, int n, int k, const double* dA, int64_t lda, const double* dB,
               int64_t ldb, double* dC, int64_t ldc) {
  const double one = 1.0, zero = 0.0;
  PSX_ROC(rocblas_dgemm(blas(c), rocblas_operation_none, rocblas_operation_none, m, n, k, &one,
                        dA, (rocblas_int)lda, dB, (rocblas_int)ldb, &zero, dC,
                        (rocblas_int)ldc));
}

// ---- LU with partial pivoting ---------------------------------------------
// Right-looking blocked getrf.  Outer blocks of kLuPanel columns end with a
// trailing update by rocBLAS (dtrsm + dgemm of depth kLuPanel).  Inside a
// block, leaves of kLuLeaf columns are factored one column per launch: the
// kernel picks the pivot from the candidates the previous launch published
// (each with the candidate row's remaining leaf entries, so the pivot row
// needs no dependent read), eliminates it from every other unflagged row of
// the leaf and publishes the next column's candidates.  The rest of the
// block is then updated by a dtrsm + dgemm of depth kLuLeaf.  Within a leaf
// rows are not moved: the pivot row is flagged and left in place, so no
// launch reads a row another workgroup writes.  After the leaf the pivot
// sequence is turned into LAPACK's row interchanges (ipiv, 1-based) and
// applied to every column, which leaves the matrix exactly as getrf's
// physical swaps would.  Ties in |pivot| go to the smallest row index.  A
// zero pivot column is skipped and reported (first one, 1-based), as getrf
// does.  (A single cooperative launch per leaf with a grid barrier per
// column measured slower: the cross-XCD barrier costs more than a launch.)
constexpr int kLuLeaf = 32;
constexpr int kLuPanel = 256;
constexpr int kLuRows = 128;     // rows per workgroup in the leaf kernels (one per thread)

struct LuWork {
  // per-workgroup pivot candidates, two buffers of cstride entries: the
  // launch for column j reads buffer j & 1 and writes the other, since the
  // workgroups of one launch read every candidate while others write theirs.
  // Each candidate is |value|, its row, and that row's entries from the
  // candidate's column to the leaf's end (kLuLeaf slots).
  double* candv;
  int* candi;
  double* candrow;
  int cstride;
  unsigned char* done;// rows already chosen as pivots in the current leaf
  int* piv;           // the leaf's pivot rows (original positions)
  int* ipiv;          // LAPACK interchanges, 1-based, n entries
  int* info;          // first zero pivot column (1-based) or 0
  // the leaf's net row movement: position mkey[k] receives the row at
  // position mval[k], k < *nmove (at most 2 kLuLeaf)
  int* mkey;
  int* mval;
  int* nmove;
};

// |x| for the pivot search; NaN ranks above everything, so a NaN column
// still yields a row (and propagates) and the search never comes back empty.
__device__ inline double lu_mag(double x) { return x != x ? INFINITY : fabs(x); }

__device__ inline void lu_better(double& v, int& r, double ov, int orr) {
  if (ov > v || (ov == v && orr < r)) {
    v = ov;
    r = orr;
  }
}

// (value, row) maximum with ties to the smaller row over the workgroup; every
// thread returns with the result.
__device__ inline void lu_argmax_block(double& v, int& r) {
  __shared__ double sv[kLuRows / 64];
  __shared__ int sr[kLuRows / 64];
  for (int o = 32; o >= 1; o >>= 1) lu_better(v, r, __shfl_xor(v, o), __shfl_xor(r, o));
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sv[w] = v;
    sr[w] = r;
  }
  __syncthreads();
  v = sv[0];
  r = sr[0];
  for (int k = 1; k < kLuRows / 64; ++k) lu_better(v, r, sv[k], sr[k]);
}

// Publish the workgroup's candidate into buffer `buf`: the winning thread
// copies its row entries vals[first .. first+count).
__device__ inline void lu_publish(const LuWork& w, int buf, double v, int r, int i,
                                  const double* vals, int first, int count) {
  const int slot = buf * w.cstride + (int)blockIdx.x;
  if (threadIdx.x == 0) {
    w.candv[slot] = v;
    w.candi[slot] = r;
  }
  if (i == r) {
    double* dst = w.candrow + (int64_t)slot * kLuLeaf;
#pragma unroll
    for (int k = 0; k < kLuLeaf; ++k)
      if (k < count) dst[k] = vals[first + k];
  }
}

// Rows [c, n): clear the pivot flags and publish the candidates of column c.
__global__ void __launch_bounds__(kLuRows) lu_leaf_init(double* A, int64_t lda, int n, int c,
                                                        int cend, LuWork w) {
  const int i = c + (int)blockIdx.x * kLuRows + (int)threadIdx.x;
  const int nc = cend - c;
  double vals[kLuLeaf];
  double v = -1.0;
  int r = 0x7fffffff;
  if (i < n) {
    w.done[i] = 0;
#pragma unroll
    for (int k = 0; k < kLuLeaf; ++k)
      if (k < nc) vals[k] = A[i + (int64_t)(c + k) * lda];
    v = lu_mag(vals[0]);
    r = i;
  }
  lu_argmax_block(v, r);
  lu_publish(w, 0, v, r, i, vals, 0, nc);
}

// Column cj = c + j of the leaf [c, cend): choose the pivot among the
// candidates, eliminate it from every other unflagged row of the leaf, and
// publish the candidates of column cj + 1.
__global__ void __launch_bounds__(kLuRows) lu_leaf_col(double* A, int64_t lda, int n, int c,
                                                       int j, int cend, int G, LuWork w) {
  __shared__ double u[kLuLeaf];
  __shared__ int ps, gs;
  const int cj = c + j;
  const int ncol = cend - cj;                      // columns cj .. cend-1
  const int i = c + (int)blockIdx.x * kLuRows + (int)threadIdx.x;
  // independent loads first: this thread's row and flag, the candidates
  double vals[kLuLeaf];
  bool mine = false;
  if (i < n) {
    mine = !w.done[i];
#pragma unroll
    for (int k = 0; k < kLuLeaf; ++k)
      if (k < ncol) vals[k] = A[i + (int64_t)(cj + k) * lda];
  }
  const int rb = (j & 1) * w.cstride;
  if (threadIdx.x < 64) {
    double v = -1.0;
    int r = 0x7fffffff, g = 0;
    for (int t = (int)threadIdx.x; t < G; t += 64) {
      const double tv = w.candv[rb + t];
      const int tr = w.candi[rb + t];
      if (tv > v || (tv == v && tr < r)) {
        v = tv;
        r = tr;
        g = t;
      }
    }
    for (int o = 32; o >= 1; o >>= 1) {
      const double ov = __shfl_xor(v, o);
      const int orr = __shfl_xor(r, o), og = __shfl_xor(g, o);
      if (ov > v || (ov == v && orr < r)) {
        v = ov;
        r = orr;
        g = og;
      }
    }
    if (threadIdx.x == 0) {
      ps = r;
      gs = g;
    }
  }
  __syncthreads();
  // rows [cj, n) hold n - cj >= 1 unflagged rows, so the search always
  // returns one; the clamp only keeps the accesses below in bounds regardless
  const int p = ps < n ? ps : cj;
  if ((int)threadIdx.x < ncol)
    u[threadIdx.x] = w.candrow[(int64_t)(rb + gs) * kLuLeaf + threadIdx.x];
  __syncthreads();
  const double piv = u[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    w.piv[j] = p;
    if (piv == 0.0 && *w.info == 0) *w.info = cj + 1;
  }
  double v = -1.0;
  int r = 0x7fffffff;
  if (i < n && mine) {
    if (i == p) {
      w.done[i] = 1;
    } else {
      if (piv != 0.0) {
        const double l = vals[0] / piv;
        vals[0] = l;
        A[i + (int64_t)cj * lda] = l;
#pragma unroll
        for (int k = 1; k < kLuLeaf; ++k)
          if (k < ncol) {
            vals[k] = fma(-l, u[k], vals[k]);
            A[i + (int64_t)(cj + k) * lda] = vals[k];
          }
      }
      if (ncol > 1) {
        v = lu_mag(vals[1]);
        r = i;
      }
    }
  }
  if (ncol > 1) {
    lu_argmax_block(v, r);
    lu_publish(w, (j + 1) & 1, v, r, i, vals, 1, ncol - 1);
  }
}

// The leaf factored by one workgroup when its rows fit in registers: 1024
// threads, thread t holding rows c + t + 1024 k (k < NR) of the leaf's W
// columns (kLuLeaf with one row per thread, kLuWgLeaf with more).  Per column: a workgroup argmax (ties to the lower row,
// NaN first, as lu_better), the pivot row through LDS, and the elimination
// of every other unpivoted row -- the arithmetic of lu_leaf_col, with two
// barriers in place of a launch.  Rows stay in place; the pivots go to
// w.piv for lu_leaf_ipiv and lu_leaf_move, as on the multi-workgroup path.
constexpr int kLuWgLeaf = 16;
constexpr int kLuWgThreads = 1024;
constexpr int kLuWgRowsMax = 3 * kLuWgThreads;   // leaf rows one workgroup takes

// LAPACK interchanges and net row movement of a leaf from its pivot rows, by
// one wavefront (defined below).
__device__ __forceinline__ void lu_leaf_ipiv_wave(int c, int nc, const int* piv, const LuWork& w);

// One column step of lu_leaf_wg (J static, so every register index is).
template <int NR, int W, int J>
__device__ __forceinline__ void lu_wg_step(double (&vals)[NR][W], bool (&live)[NR],
                                           int t, int c, int n, int nc, double* sv, int* sr,
                                           double* u, int* spiv, const LuWork& w) {
  if (J >= nc) return;  // uniform
  const int lane = t & 63, wv = t >> 6;
  double v = -1.0;
  int r = 0x7fffffff;
#pragma unroll
  for (int k = 0; k < NR; ++k)
    if (live[k]) lu_better(v, r, lu_mag(vals[k][J]), c + t + kLuWgThreads * k);
  for (int o = 32; o >= 1; o >>= 1) lu_better(v, r, __shfl_xor(v, o), __shfl_xor(r, o));
  if (lane == 0) {
    sv[wv] = v;
    sr[wv] = r;
  }
  __syncthreads();
  // every wave combines the 16 wave winners: one LDS read per lane, then a
  // 16-lane butterfly, and lane 0's result to all lanes
  {
    constexpr int kW = kLuWgThreads / 64;
    v = lane < kW ? sv[lane] : -1.0;
    r = lane < kW ? sr[lane] : 0x7fffffff;
    for (int o = kW / 2; o >= 1; o >>= 1) lu_better(v, r, __shfl_xor(v, o), __shfl_xor(r, o));
    r = __shfl(r, 0);
  }
  // a candidate always exists; the clamp keeps accesses in bounds regardless
  const int p = r < n ? r : c + J;
  // the pivot row's owner publishes its entries and drops out
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    if (live[k] && c + t + kLuWgThreads * k == p) {
#pragma unroll
      for (int jj = J; jj < W; ++jj)
        if (jj < nc) u[jj] = vals[k][jj];
      live[k] = false;
    }
  }
  __syncthreads();
  const double piv = u[J];
  if (t == 0) {
    w.piv[J] = p;
    spiv[J] = p;
    if (piv == 0.0 && *w.info == 0) *w.info = c + J + 1;
  }
  if (piv != 0.0) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      if (live[k]) {  // the pivot rows are no longer live
        const double l = vals[k][J] / piv;
        vals[k][J] = l;
#pragma unroll
        for (int jj = J + 1; jj < W; ++jj)
          if (jj < nc) vals[k][jj] = fma(-l, u[jj], vals[k][jj]);
      }
    }
  }
}

template <int NR, int W, int... J>
__device__ __forceinline__ void lu_wg_steps(double (&vals)[NR][W], bool (&live)[NR],
                                            int t, int c, int n, int nc, double* sv, int* sr,
                                            double* u, int* spiv, const LuWork& w,
                                            std::integer_sequence<int, J...>) {
  (lu_wg_step<NR, W, J>(vals, live, t, c, n, nc, sv, sr, u, spiv, w), ...);
}

template <int NR, int W>
__global__ void __launch_bounds__(kLuWgThreads) lu_leaf_wg(double* A, int64_t lda, int n, int c,
                                                           int cend, LuWork w) {
  __shared__ double sv[kLuWgThreads / 64];
  __shared__ int sr[kLuWgThreads / 64];
  __shared__ double u[W];
  __shared__ int spiv[W];
  const int t = threadIdx.x;
  const int nc = cend - c;
  double vals[NR][W];
  bool live[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int r = c + t + kLuWgThreads * k;
    live[k] = r < n;
#pragma unroll
    for (int jj = 0; jj < W; ++jj)
      vals[k][jj] = (live[k] && jj < nc) ? A[r + (int64_t)(c + jj) * lda] : 0.0;
  }
  lu_wg_steps<NR, W>(vals, live, t, c, n, nc, sv, sr, u, spiv, w,
                     std::make_integer_sequence<int, W>{});
  // the leaf's interchanges and row movement (what lu_leaf_ipiv does on the
  // other path), by wave 0 from the pivots in LDS
  __syncthreads();
  if (t < 64) lu_leaf_ipiv_wave(c, nc, spiv, w);
  // pivot rows keep the values they had when chosen
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int r = c + t + kLuWgThreads * k;
    if (r < n) {
#pragma unroll
      for (int jj = 0; jj < W; ++jj)
        if (jj < nc) A[r + (int64_t)(c + jj) * lda] = vals[k][jj];
    }
  }
}

// LAPACK interchanges of the leaf: at step j the row chosen (original
// position piv[j]) is swapped with whatever row is at position c + j.  Only
// positions touched by this leaf's swaps differ from the identity; one
// wavefront keeps them as (position, row) pairs, one pair per lane, and finds
// entries by ballot.
__device__ __forceinline__ void lu_leaf_ipiv_wave(int c, int nc, const int* piv, const LuWork& w) {
  const int lane = (int)(threadIdx.x & 63);
  int key = -1, val = -1, m = 0;
  for (int j = 0; j < nc; ++j) {
    const int want = piv[j];
    const int pj = c + j;
    const unsigned long long hv = __ballot(key >= 0 && val == want);
    const int q = hv ? __shfl(key, __ffsll((long long)hv) - 1) : want;
    const unsigned long long hp = __ballot(key == pj);
    const int at_pj = hp ? __shfl(val, __ffsll((long long)hp) - 1) : pj;
    if (lane == 0) w.ipiv[pj] = q + 1;
    // position pj now holds `want`, position q the row that was at pj
    if (hp) {
      if (key == pj) val = want;
    } else {
      if (lane == m) {
        key = pj;
        val = want;
      }
      ++m;
    }
    if (q != pj) {
      const unsigned long long hq = __ballot(key == q);
      if (hq) {
        if (key == q) val = at_pj;
      } else {
        if (lane == m) {
          key = q;
          val = at_pj;
        }
        ++m;
      }
    }
  }
  // publish the positions whose row changed, compacted
  const bool moved = lane < m && key != val;
  const unsigned long long hm = __ballot(moved);
  const int slot = __popcll(hm & ((1ull << lane) - 1ull));
  if (moved) {
    w.mkey[slot] = key;
    w.mval[slot] = val;
  }
  if (lane == 0) *w.nmove = __popcll(hm);
}

__global__ void __launch_bounds__(64) lu_leaf_ipiv(int c, int nc, LuWork w) {
  lu_leaf_ipiv_wave(c, nc, w.piv, w);
}

// Apply the leaf's row movement to every column, one thread per column: the
// moved rows' values are all loaded, then all stored (the same result as the
// leaf's interchanges applied one after another).
__global__ void __launch_bounds__(256) lu_leaf_move(double* A, int64_t lda, int ncols, LuWork w) {
  __shared__ int sk[2 * kLuLeaf], sv[2 * kLuLeaf];
  __shared__ int sn;
  const int t = threadIdx.x;
  if (t == 0) sn = *w.nmove;
  __syncthreads();
  const int nm = sn;
  if (t < nm) {
    sk[t] = w.mkey[t];
    sv[t] = w.mval[t];
  }
  __syncthreads();
  const int col = (int)blockIdx.x * 256 + t;
  if (col >= ncols || nm == 0) return;
  double* a = A + (int64_t)col * lda;
  double v[2 * kLuLeaf];
#pragma unroll
  for (int k = 0; k < 2 * kLuLeaf; ++k)
    if (k < nm) v[k] = a[sv[k]];
#pragma unroll
  for (int k = 0; k < 2 * kLuLeaf; ++k)
    if (k < nm) a[sk[k]] = v[k];
}

// Where an LU runs: a stream, the rocBLAS handle bound to it and a device work
// area of lu_work_bytes(n) bytes (pivot candidates, flags, the leaf's row
// movement).  Factorizations on different lanes may run at the same time.
struct LuLane {
  hipStream_t s;
  rocblas_handle h;
  void* work;
};

size_t lu_work_bytes(int n) {
  const int Gmax = (n + kLuRows - 1) / kLuRows;
  const size_t rows_bytes = 2 * (size_t)Gmax * kLuLeaf * sizeof(double);
  const size_t move_off = (rows_bytes + 24 * (size_t)Gmax + 4 * kLuLeaf + (size_t)n + 255) &
                          ~size_t(255);
  return move_off + sizeof(int) * (4 * kLuLeaf + 1) + 256;
}

void lu_factor_on(const LuLane& L, int n, double* dA, int64_t lda, int* dipiv, int* dinfo) {
  rocblas_handle h = L.h;
  hipStream_t s = L.s;
  const int Gmax = (n + kLuRows - 1) / kLuRows;
  const size_t rows_bytes = 2 * (size_t)Gmax * kLuLeaf * sizeof(double);
  const size_t move_off = (rows_bytes + 24 * (size_t)Gmax + 4 * kLuLeaf + (size_t)n + 255) &
                          ~size_t(255);
  char* base = (char*)L.work;
  LuWork w;
  w.cstride = Gmax;
  w.candrow = (double*)base;                                  // 2 Gmax kLuLeaf doubles
  w.candv = (double*)(base + rows_bytes);                     // 2 Gmax doubles
  w.candi = (int*)(base + rows_bytes + 16 * (size_t)Gmax);    // 2 Gmax ints
  w.piv = (int*)(base + rows_bytes + 24 * (size_t)Gmax);      // kLuLeaf ints
  w.done = (unsigned char*)(base + rows_bytes + 24 * (size_t)Gmax + 4 * kLuLeaf);  // n flags
  w.mkey = (int*)(base + move_off);                           // 2 kLuLeaf ints
  w.mval = w.mkey + 2 * kLuLeaf;                              // 2 kLuLeaf ints
  w.nmove = w.mval + 2 * kLuLeaf;                             // 1 int
  w.ipiv = dipiv;
  w.info = dinfo;
  PSX_HIP(hipMemsetAsync(dinfo, 0, sizeof(int), s));
  const double one = 1.0, mone = -1.0;
  for (int k = 0; k < n; k += kLuPanel) {
    const int kend = std::min(n, k + kLuPanel);
    for (int cl = k; cl < kend;) {
      // the trailing leaves, once their rows fit one workgroup, in one launch each
      const int rows = n - cl;
      const bool one_wg = rows <= kLuWgRowsMax;
      const int nr = (rows + kLuWgThreads - 1) / kLuWgThreads;
      // one row per thread leaves room for kLuLeaf columns, more rows for kLuWgLeaf
      const int width = !one_wg || nr == 1 ? kLuLeaf : kLuWgLeaf;
      const int cend = std::min(kend, cl + width);
      const int nc = cend - cl;
      if (one_wg) {
        if (nr == 1)
          hipLaunchKernelGGL((lu_leaf_wg<1, kLuLeaf>), dim3(1), dim3(kLuWgThreads), 0, s, dA, lda,
                             n, cl, cend, w);
        else if (nr == 2)
          hipLaunchKernelGGL((lu_leaf_wg<2, kLuWgLeaf>), dim3(1), dim3(kLuWgThreads), 0, s, dA,
                             lda, n, cl, cend, w);
        else
          hipLaunchKernelGGL((lu_leaf_wg<3, kLuWgLeaf>), dim3(1), dim3(kLuWgThreads), 0, s, dA,
                             lda, n, cl, cend, w);
        // (the kernel converts its pivots itself)
      } else {
        const int G = (n - cl + kLuRows - 1) / kLuRows;
        hipLaunchKernelGGL(lu_leaf_init, dim3(G), dim3(kLuRows), 0, s, dA, lda, n, cl, cend, w);
        for (int j = 0; j < nc; ++j)
          hipLaunchKernelGGL(lu_leaf_col, dim3(G), dim3(kLuRows), 0, s, dA, lda, n, cl, j, cend,
                             G, w);
      }
      if (!one_wg) hipLaunchKernelGGL(lu_leaf_ipiv, dim3(1), dim3(64), 0, s, cl, nc, w);
      hipLaunchKernelGGL(lu_leaf_move, dim3((n + 255) / 256), dim3(256), 0, s, dA, lda, n, w);
      PSX_HIP(hipGetLastError());
      if (cend < kend) {   // rest of the block: U = L11^-1 A, then the update
        double* L11 = dA + cl + (int64_t)cl * lda;
        double* A12 = dA + cl + (int64_t)cend * lda;
        PSX_ROC(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                              rocblas_diagonal_unit, nc, kend - cend, &one, L11, (rocblas_int)lda,
                              A12, (rocblas_int)lda));
        if (cend < n)
          PSX_ROC(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n - cend,
                                kend - cend, nc, &mone, dA + cend + (int64_t)cl * lda,
                                (rocblas_int)lda, A12, (rocblas_int)lda, &one,
                                dA + cend + (int64_t)cend * lda, (rocblas_int)lda));
      }
      cl = cend;
    }
    if (kend < n) {        // trailing matrix
      double* L11 = dA + k + (int64_t)k * lda;
      double* A12 = dA + k + (int64_t)kend * lda;
      PSX_ROC(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                            rocblas_diagonal_unit, kend - k, n - kend, &one, L11, (rocblas_int)lda,
                            A12, (rocblas_int)lda));
      PSX_ROC(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n - kend, n - kend,
                            kend - k, &mone, dA + kend + (int64_t)k * lda, (rocblas_int)lda, A12,
                            (rocblas_int)lda, &one, dA + kend + (int64_t)kend * lda,
                            (rocblas_int)lda));
    }
  }
}

void lu_factor_dev(Ctx* c, int n, double* dA, int64_t lda, int* dipiv, int* dinfo) {
  lu_factor_on(LuLane{c->stream, blas(c), ctx_scratch(c, lu_work_bytes(n))}, n, dA, lda, dipiv,
               dinfo);
}

// ---- triangular solves with the LU factors, one right-hand side ------------
// A blocked substitution in kTrsRows-row blocks, one launch per block step.
// Forward (unit lower L): launch k updates every row block from k on with the
// solution of block k-1, then workgroup 0, which holds block k, solves block
// k's diagonal triangle.  Back (U, diagonal included): launch k, from the last
// block down, updates row blocks k, k-1, .., 0 with block k+1's solution, and
// workgroup 0 solves block k.  Each launch reads its columns once, so a
// triangle costs n^2/2 reads.
//   The update is a column-major GEMV with two threads per row, each taking
// 64 columns with all its loads in flight.  Workgroup 0 stages the whole
// diagonal block in LDS (132 KB) while it runs its own rows' update, then
// solves the block from LDS in two 64-row sub-blocks, one wave each, the
// solved value broadcast by v_readlane.
constexpr int kTrsRows = 128;
constexpr int kTrsLd = kTrsRows + 1;   // LDS column stride of the diagonal block
constexpr int kTrsMaxRhs = 4;          // more right-hand sides take the blocked trsm below

// lane j's value in every lane (j uniform: two v_readlane_b32)
__device__ inline double trs_bcast(double v, int j) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, j);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), j);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <bool UPPER>
__global__ void __launch_bounds__(256) trs_step(const double* __restrict__ A, int64_t lda, int n,
                                                int k, double* __restrict__ x) {
  __shared__ double D[kTrsRows * kTrsLd];  // diagonal block, column-major (workgroup 0)
  __shared__ double xs[kTrsRows];          // source block's solution
  __shared__ double part[2][kTrsRows];     // the two half-row partial sums
  const int nb = (n + kTrsRows - 1) / kTrsRows;
  const int t = threadIdx.x;
  const int rl = t & (kTrsRows - 1), h = t >> 7;  // row in the block, column half
  const int rb = UPPER ? k - (int)blockIdx.x : k + (int)blockIdx.x;
  const int R0 = rb * kTrsRows;
  const int r = R0 + rl;
  const bool diag = blockIdx.x == 0;
  const int cn = min(kTrsRows, n - R0);  // rows in this block (= columns of the diagonal)
  if (diag) {
    // thread t fills row t & 127 of columns 2m + (t >> 7), m = 0..63, 32 loads
    // in flight at a time
    const bool rin = rl < cn;
    const double* a = A + (R0 + (rin ? rl : 0)) + (int64_t)R0 * lda;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      double v[32];
#pragma unroll
      for (int q = 0; q < 32; ++q) {
        const int ic = 2 * (half * 32 + q) + h;
        v[q] = (rin && ic < cn) ? a[(int64_t)ic * lda] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 32; ++q) D[(2 * (half * 32 + q) + h) * kTrsLd + rl] = v[q];
    }
  }
  double y = (h == 0 && r < n) ? x[r] : 0.0;
  const int src = UPPER ? k + 1 : k - 1;
  if (src >= 0 && src < nb) {
    const int c0 = src * kTrsRows, sn = min(kTrsRows, n - c0);
    if (t < sn) xs[t] = x[c0 + t];
    __syncthreads();
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
    if (r < n) {
      const int j0 = h * 64;
      const double* a = A + r + (int64_t)(c0 + j0) * lda;
      if (sn == kTrsRows) {  // a full source block: all 64 loads in flight
        double v[64];
#pragma unroll
        for (int q = 0; q < 64; ++q) v[q] = a[(int64_t)q * lda];
#pragma unroll
        for (int q = 0; q < 64; q += 4) {
          p0 = fma(v[q], xs[j0 + q], p0);
          p1 = fma(v[q + 1], xs[j0 + q + 1], p1);
          p2 = fma(v[q + 2], xs[j0 + q + 2], p2);
          p3 = fma(v[q + 3], xs[j0 + q + 3], p3);
        }
      } else {
        const int m = min(sn, j0 + 64) - j0;
        for (int q = 0; q < m; ++q) p0 = fma(a[(int64_t)q * lda], xs[j0 + q], p0);
      }
    }
    part[h][rl] = (p0 + p1) + (p2 + p3);
    __syncthreads();
    if (h == 0) y -= part[0][rl] + part[1][rl];
  }
  if (!diag) {
    if (h == 0 && r < n) x[r] = y;
    return;
  }
  __syncthreads();  // D complete
  // the diagonal block: rows 0..63 are wave 0, rows 64..127 wave 1
  const int w = t >> 6, lane = t & 63;
  const int nsub = (cn + 63) >> 6;
  for (int q = 0; q < nsub; ++q) {
    const int s = UPPER ? nsub - 1 - q : q;
    const int s0 = s * 64, sn = min(64, cn - s0);
    if (w == s) {
      const double* d = D + s0 * kTrsLd + s0 + lane;  // column j at d[j * kTrsLd]
      double T[64];
#pragma unroll
      for (int j = 0; j < 64; ++j) T[j] = d[j * kTrsLd];
      if (!UPPER) {
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          if (j < sn) {
            const double xj = trs_bcast(y, j);
            if (lane > j) y = fma(-T[j], xj, y);
          }
        }
      } else {
#pragma unroll
        for (int j = 63; j >= 0; --j) {
          if (j < sn) {
            if (lane == j) y = y / T[j];
            const double xj = trs_bcast(y, j);
            if (lane < j) y = fma(-T[j], xj, y);
          }
        }
      }
      if (lane < sn) xs[s0 + lane] = y;
    }
    __syncthreads();
    const bool later = t < cn && (UPPER ? (w < s) : (w > s));
    if (later) {
      const double* d = D + s0 * kTrsLd + t;
      double p0 = 0.0, p1 = 0.0;
      int j = 0;
      for (; j + 2 <= sn; j += 2) {
        p0 = fma(d[j * kTrsLd], xs[s0 + j], p0);
        p1 = fma(d[(j + 1) * kTrsLd], xs[s0 + j + 1], p1);
      }
      if (j < sn) p0 = fma(d[j * kTrsLd], xs[s0 + j], p0);
      y -= p0 + p1;
    }
  }
  if (t < cn) x[r] = y;
}

// x = L^{-1} x, then x = U^{-1} x, with the packed LU factors of dA, on stream s.
void lu_trs_on(hipStream_t s, int n, const double* dA, int64_t lda, double* dx) {
  const int nb = (n + kTrsRows - 1) / kTrsRows;
  // (the first step of each sweep has no source block: only the diagonal)
  for (int k = 0; k < nb; ++k)
    hipLaunchKernelGGL(trs_step<false>, dim3(k == 0 ? 1 : nb - k), dim3(256), 0, s, dA, lda, n,
                       k, dx);
  for (int k = nb - 1; k >= 0; --k)
    hipLaunchKernelGGL(trs_step<true>, dim3(k == nb - 1 ? 1 : k + 1), dim3(256), 0, s, dA, lda,
                       n, k, dx);
  PSX_HIP(hipGetLastError());
}

void lu_trs_dev(Ctx* c, int n, const double* dA, int64_t lda, double* dx) {
  lu_trs_on(c->stream, n, dA, lda, dx);
}

// ---- triangular solves with many right-hand sides ---------------------------
// Blocked left substitution X = T^{-1} B for the LU factors (T = unit lower L
// or upper U) and m right-hand sides, in kTrsRows-row blocks: the diagonal
// block is solved by substitution (trsm_diag, one thread per right-hand side,
// the block staged in LDS), and the rows still to come are updated by a
// rocBLAS GEMM.  Substitution keeps LAPACK's backward error on ill-conditioned
// U, which a trsm built on inverted diagonal blocks does not.
constexpr int kTsmRows = 64;               // diagonal block of the many-column solve
constexpr int kTsmLd = kTsmRows + 1;        // its LDS column stride
constexpr int kTsmCols = 192;               // right-hand sides per workgroup (one per thread)

template <bool UPPER>
__global__ void __launch_bounds__(kTsmCols) trsm_diag(const double* __restrict__ A, int64_t lda,
                                                      int n, int k, double* __restrict__ B,
                                                      int64_t ldb, int m) {
  __shared__ double D[kTsmRows * kTsmLd];         // the diagonal block, column-major
  __shared__ double Bs[kTsmRows][kTsmCols];       // this workgroup's rows of B
  const int t = threadIdx.x;
  const int R0 = k * kTsmRows, cn = min(kTsmRows, n - R0);
  const int c0 = (int)blockIdx.x * kTsmCols, cm = min(kTsmCols, m - c0);
  // stage D and B: 64 consecutive rows per wave, one column per wave and pass
  const int lane = t & 63, wv = t >> 6;
  {
    const bool rin = lane < cn;
    const double* a = A + (R0 + (rin ? lane : 0)) + (int64_t)R0 * lda;
    constexpr int kW = kTsmCols / 64;                       // waves
    constexpr int kPer = (kTsmRows + kW - 1) / kW;          // columns per wave, rounded up
    double v[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int ic = q * kW + wv;
      v[q] = (rin && ic < cn) ? a[(int64_t)ic * lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int ic = q * kW + wv;
      if (ic < kTsmRows) D[ic * kTsmLd + lane] = v[q];
    }
  }
  {
    const bool rin = lane < cn;
    const double* b = B + (R0 + (rin ? lane : 0)) + (int64_t)c0 * ldb;
    constexpr int kPer = kTsmCols / (kTsmCols / 64);  // 64 columns per wave
#pragma unroll 16
    for (int q = 0; q < kPer; ++q) {
      const int jc = q * (kTsmCols / 64) + wv;
      Bs[lane][jc] = (rin && jc < cm) ? b[(int64_t)jc * ldb] : 0.0;
    }
  }
  __syncthreads();
  if (t < cm) {
    // two 32-row sub-blocks; the one being solved sits in registers
    constexpr int S = 32;
    const int nsub = (cn + S - 1) / S;
#pragma unroll 1
    for (int q = 0; q < nsub; ++q) {
      const int sb = UPPER ? nsub - 1 - q : q;
      const int s0 = sb * S, sn = min(S, cn - s0);
      double x[S];
#pragma unroll
      for (int i = 0; i < S; ++i) x[i] = Bs[(s0 + i) & (kTsmRows - 1)][t];
      if (!UPPER) {
#pragma unroll
        for (int j = 0; j < S; ++j) {
          if (j < sn) {
#pragma unroll
            for (int i = j + 1; i < S; ++i)
              x[i] = fma(-D[(s0 + j) * kTsmLd + s0 + i], x[j], x[i]);
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      } else {
#pragma unroll
        for (int j = S - 1; j >= 0; --j) {
          if (j < sn) {
            x[j] = x[j] / D[(s0 + j) * kTsmLd + s0 + j];
#pragma unroll
            for (int i = 0; i < j; ++i) x[i] = fma(-D[(s0 + j) * kTsmLd + s0 + i], x[j], x[i]);
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
#pragma unroll
      for (int i = 0; i < S; ++i)
        if (i < sn) Bs[s0 + i][t] = x[i];
      // the other sub-block's rows take this one's contribution
      const int r0 = UPPER ? 0 : s0 + sn, r1 = UPPER ? s0 : cn;
#pragma unroll 4
      for (int i = r0; i < r1; ++i) {
        double p0 = 0.0, p1 = 0.0;
#pragma unroll
        for (int j = 0; j < S; j += 2) {
          p0 = fma(D[(s0 + j) * kTsmLd + i], x[j], p0);
          p1 = fma(D[(s0 + j + 1) * kTsmLd + i], x[j + 1], p1);
        }
        Bs[i][t] -= p0 + p1;
      }
    }
  }
  __syncthreads();
  {
    const bool rin = lane < cn;
    double* b = B + (R0 + (rin ? lane : 0)) + (int64_t)c0 * ldb;
    constexpr int kPer = kTsmCols / (kTsmCols / 64);
#pragma unroll 16
    for (int q = 0; q < kPer; ++q) {
      const int jc = q * (kTsmCols / 64) + wv;
      if (rin && jc < cm) b[(int64_t)jc * ldb] = Bs[lane][jc];
    }
  }
}

// B (n x m, ldb) = T^{-1} B with T the unit lower (UPPER false) or upper
// factor packed in dA.
static void trsm_left_dev(Ctx* c, bool upper, int n, const double* dA, int64_t lda, double* dB,
                          int64_t ldb, int m) {
  rocblas_handle h = blas(c);
  hipStream_t s = c->stream;
  const int nb = (n + kTsmRows - 1) / kTsmRows;
  const unsigned g = (unsigned)((m + kTsmCols - 1) / kTsmCols);
  const double one = 1.0, mone = -1.0;
  for (int q = 0; q < nb; ++q) {
    const int k = upper ? nb - 1 - q : q;
    const int R0 = k * kTsmRows, cn = std::min(kTsmRows, n - R0);
    if (upper)
      hipLaunchKernelGGL(trsm_diag<true>, dim3(g), dim3(kTsmCols), 0, s, dA, lda, n, k, dB, ldb,
                         m);
    else
      hipLaunchKernelGGL(trsm_diag<false>, dim3(g), dim3(kTsmCols), 0, s, dA, lda, n, k, dB, ldb,
                         m);
    PSX_HIP(hipGetLastError());
    if (!upper && R0 + cn < n)  // rows below: B2 -= L21 X1
      PSX_ROC(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n - R0 - cn, m, cn,
                            &mone, dA + (R0 + cn) + (int64_t)R0 * lda, (rocblas_int)lda,
                            dB + R0, (rocblas_int)ldb, &one, dB + R0 + cn, (rocblas_int)ldb));
    if (upper && R0 > 0)        // rows above: B0 -= U01 X1
      PSX_ROC(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, R0, m, cn, &mone,
                            dA + (int64_t)R0 * lda, (rocblas_int)lda, dB + R0, (rocblas_int)ldb,
                            &one, dB, (rocblas_int)ldb));
  }
}

// dA (n x n, lda) is overwritten by its LU factors; dB (n x nrhs) by X.
// Factor dA in place; ipiv (device, n + 64 ints) receives the LAPACK
// interchanges.  Raises SingularException on a zero pivot.
static void lu_factor_checked(Ctx* c, int n, double* dA, int64_t lda, int* ipiv) {
  int* info = ipiv + n + 32;
  lu_factor_dev(c, n, dA, lda, ipiv, info);
  int hinfo = 0;
  PSX_HIP(hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, c->stream));
  PSX_HIP(hipStreamSynchronize(c->stream));
  PSX_REQUIRE(hinfo == 0, PSX_ERR_SINGULAR,
              "LU: zero pivot at position " + std::to_string(hinfo) +
                  " (Julia SingularException)");
}

// A device int array of n + 64 entries (LAPACK interchanges and their info
// slot, or a row permutation), freed on scope exit after the stream has
// drained.
struct Pivots {
  Ctx* c;
  int* p;
  Pivots(Ctx* cc, int n) : c(cc), p((int*)ctx_alloc(cc, sizeof(int) * ((size_t)n + 64))) {}
  ~Pivots() {
    (void)hipStreamSynchronize(c->stream);
    try {
      ctx_free(c, p);
    } catch (...) {
    }
  }
};

// The row order LAPACK's interchanges produce, perm[i] = the original row
// that ends at position i, built by one thread in LDS (the interchanges are
// sequential by definition) and written out by the workgroup.
constexpr int kPermMax = 16384;
__global__ void __launch_bounds__(256) lu_perm(const int* __restrict__ ipiv, int n,
                                               int* __restrict__ perm) {
  __shared__ int pr[kPermMax], iv[kPermMax];
  for (int i = threadIdx.x; i < n; i += 256) {
    pr[i] = i;
    iv[i] = ipiv[i] - 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < n; ++i) {
      const int q = iv[i];
      if (q != i && q > i && q < n) {
        const int t = pr[i];
        pr[i] = pr[q];
        pr[q] = t;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 256) perm[i] = pr[i];
}

// Column blockIdx.x of B (n rows, ldb) reordered in place: b[i] = b[perm[i]],
// through LDS.
__global__ void __launch_bounds__(256) lu_permute_cols(const int* __restrict__ perm, int n,
                                                       double* __restrict__ B, int64_t ldb) {
  __shared__ double col[kPermMax];
  double* b = B + (int64_t)blockIdx.x * ldb;
  for (int i = threadIdx.x; i < n; i += 256) col[i] = b[i];
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 256) b[i] = col[perm[i]];
}

// B = A^{-1} B from the factors of A in dA and its interchanges ipiv.
static void lu_getrs_dev(Ctx* c, int n, const double* dA, int64_t lda, const int* ipiv,
                         double* dB, int64_t ldb, int nrhs) {
  if (nrhs <= 0) return;
  // P b by the LAPACK interchanges, then the two substitutions.  Many
  // columns share one permutation, built once and applied as a gather;
  // a few columns take rocSOLVER's laswp (the sequential chain is the same).
  if (nrhs > kTrsMaxRhs && n <= kPermMax) {
    Pivots perm(c, n);
    hipLaunchKernelGGL(lu_perm, dim3(1), dim3(256), 0, c->stream, ipiv, n, perm.p);
    hipLaunchKernelGGL(lu_permute_cols, dim3((unsigned)nrhs), dim3(256), 0, c->stream,
                       (const int*)perm.p, n, dB, ldb);
    PSX_HIP(hipGetLastError());
  } else {
    PSX_ROC(rocsolver_dlaswp(blas(c), nrhs, dB, (rocblas_int)ldb, 1, n,
                             (const rocblas_int*)ipiv, 1));
  }
  if (nrhs <= kTrsMaxRhs) {
    for (int j = 0; j < nrhs; ++j) lu_trs_dev(c, n, dA, lda, dB + (int64_t)j * ldb);
  } else {
    trsm_left_dev(c, false, n, dA, lda, dB, ldb, nrhs);
    trsm_left_dev(c, true, n, dA, lda, dB, ldb, nrhs);
  }
}


void lu_solve_dev(Ctx* c, int n, double* dA, int64_t lda, double* dB, int64_t ldb,
                  int nrhs) {
  Pivots pv(c, n);
  lu_factor_checked(c, n, dA, lda, pv.p);
  lu_getrs_dev(c, n, dA, lda, pv.p, dB, ldb, nrhs);
  PSX_HIP(hipStreamSynchronize(c->stream));
}

// Independent systems side by side.  System i runs on lane i: lane 0 is the
// context stream, lanes 1.. the context's extra streams, each with its own
// rocBLAS handle and its own part of the scratch.  The lanes start after the
// work already queued on the context stream, and the context stream waits
// for all of them, so the call keeps stream order for its caller.  The
// factorizations are latency-bound (a leaf column runs on one or a few
// workgroups), so running them at once overlaps their latencies.
void ensure_lanes(Ctx* c, int nextra) {
  for (int i = 0; i < nextra; ++i) {
    if (c->lstream[i]) continue;
    PSX_HIP(hipStreamCreateWithFlags(&c->lstream[i], hipStreamNonBlocking));
    rocblas_handle h;
    PSX_REQUIRE(rocblas_create_handle(&h) == rocblas_status_success, PSX_ERR_DEVICE,
                "rocblas_create_handle failed");
    rocblas_set_stream(h, c->lstream[i]);
    c->lblas[i] = h;
  }
  for (hipEvent_t& e : c->lane_ev)
    if (!e) PSX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

void lu_solve_concurrent(Ctx* c, const LuSystem* sys, int count) {
  PSX_REQUIRE(count >= 1 && count <= 1 + kLuExtraLanes, PSX_ERR_ARGUMENT,
              "too many concurrent systems");
  ensure_lanes(c, count - 1);
  std::vector<size_t> off(count + 1, 0);
  for (int i = 0; i < count; ++i) {
    PSX_REQUIRE(sys[i].nrhs >= 1 && sys[i].nrhs <= kTrsMaxRhs, PSX_ERR_ARGUMENT,
                "concurrent systems take 1 to 4 right-hand sides");
    off[i + 1] = off[i] + ((lu_work_bytes(sys[i].n) + 255) & ~size_t(255));
  }
  char* work = (char*)ctx_scratch(c, off[count]);
  blas(c);
  PSX_HIP(hipEventRecord(c->lane_ev[0], c->stream));
  try {
    for (int i = 0; i < count; ++i) {
      const LuSystem& q = sys[i];
      const LuLane L = i == 0 ? LuLane{c->stream, blas(c), work}
                              : LuLane{c->lstream[i - 1], (rocblas_handle)c->lblas[i - 1],
                                       work + off[i]};
      if (i > 0) PSX_HIP(hipStreamWaitEvent(L.s, c->lane_ev[0], 0));
      lu_factor_on(L, q.n, q.dA, q.lda, q.ipiv, q.ipiv + q.n + 32);
      PSX_ROC(rocsolver_dlaswp(L.h, q.nrhs, q.dB, (rocblas_int)q.ldb, 1, q.n,
                               (const rocblas_int*)q.ipiv, 1));
      for (int j = 0; j < q.nrhs; ++j)
        lu_trs_on(L.s, q.n, q.dA, q.lda, q.dB + (int64_t)j * q.ldb);
      PSX_HIP(hipGetLastError());
      if (i > 0) {
        PSX_HIP(hipEventRecord(c->lane_ev[i], L.s));
        PSX_HIP(hipStreamWaitEvent(c->stream, c->lane_ev[i], 0));
      }
    }
  } catch (...) {
    // nothing may still run on the caller's buffers once the error is out
    for (int i = 0; i + 1 < count; ++i) (void)hipStreamSynchronize(c->lstream[i]);
    (void)hipStreamSynchronize(c->stream);
    throw;
  }
  // SingularException for the first system with a zero pivot
  std::vector<int> info(count, 0);
  for (int i = 0; i < count; ++i)
    PSX_HIP(hipMemcpyAsync(&info[i], sys[i].ipiv + sys[i].n + 32, sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
  PSX_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < count; ++i)
    PSX_REQUIRE(info[i] == 0, PSX_ERR_SINGULAR,
                "LU: zero pivot at position " + std::to_string(info[i]) + " of system " +
                    std::to_string(i) + " (Julia SingularException)");
}

void release_lanes(Ctx* c) {
  for (int i = 0; i < kLuExtraLanes; ++i) {
    if (c->lstream[i]) (void)hipStreamSynchronize(c->lstream[i]);
    if (c->lblas[i]) rocblas_destroy_handle(static_cast<rocblas_handle>(c->lblas[i]));
    if (c->lstream[i]) (void)hipStreamDestroy(c->lstream[i]);
    c->lblas[i] = nullptr;
    c->lstream[i] = nullptr;
  }
  for (hipEvent_t& e : c->lane_ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
}

// B[r0 + i, j] = -B[r0 + i, j] for i < nr, j < gridDim.y
__global__ void __launch_bounds__(256) negate_rows(double* __restrict__ B, int64_t ldb, int r0,
                                                   int nr) {
  const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (i < nr) {
    double* p = B + (int64_t)blockIdx.y * ldb + r0 + i;
    *p = -*p;
  }
}

static void transpose(Ctx* c, int n, const double* dIn, double* dOut) {
  const double one = 1.0, zero = 0.0;
  PSX_ROC(rocblas_dgeam(blas(c), rocblas_operation_transpose, rocblas_operation_none, n, n,
                        &one, dIn, n, &zero, dOut, n, dOut, n));
}

// dOut = B1^{-1} Y B2^{-T}; dB1, dB2 are overwritten by their LU factors.
// dB2 == dB1 (one matrix on both sides) is factored once.
void decouple_dev(Ctx* c, int n, const double* dY, double* dB1, double* dB2,
                  double* dOut) {
  const size_t nn = (size_t)n * n;
  WorkBuf wk(c, sizeof(double) * nn);
  double* tmp = (double*)wk.base();
  PSX_HIP(hipMemcpyAsync(tmp, dY, sizeof(double) * nn, hipMemcpyDeviceToDevice, c->stream));
  Pivots pv(c, n);
  lu_factor_checked(c, n, dB1, n, pv.p);
  lu_getrs_dev(c, n, dB1, n, pv.p, tmp, n, n);   // tmp = B1^{-1} Y
  transpose(c, n, tmp, dOut);                     // dOut = tmp^T
  if (dB2 != dB1) lu_factor_checked(c, n, dB2, n, pv.p);
  lu_getrs_dev(c, n, dB2, n, pv.p, dOut, n, n);  // dOut = B2^{-1} tmp^T
  transpose(c, n, dOut, tmp);                     // tmp = (B2^{-1} tmp^T)^T
  PSX_HIP(hipMemcpyAsync(dOut, tmp, sizeof(double) * nn, hipMemcpyDeviceToDevice, c->stream));
  PSX_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace psx

using namespace psx;

extern "C" {

int psx_lu_solve(psx_ctx* hctx, int n, const double* A, int64_t lda, double* B,
                 int64_t ldb, int nrhs) {
  Ctx* c = reinterpret_cast<Ctx*>(hctx);
  return guarded(c, [&] {
    PSX_REQUIRE(n >= 1 && nrhs >= 0 && lda >= n && ldb >= n, PSX_ERR_ARGUMENT,
                "bad sizes for lu_solve");
    PSX_REQUIRE(A && (B || nrhs == 0), PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    const size_t ab = sizeof(double) * n * (size_t)n,
                 bb = sizeof(double) * n * (size_t)std::max(nrhs, 1);
    XferBuf x(c, xfer_bytes({ab, bb}));
    double* dA = x.take(ab);
    double* dB = x.take(bb);
    h2d_matrix(c, dA, n, A, lda, n, n);
    if (nrhs > 0) h2d_matrix(c, dB, n, B, ldb, n, nrhs);
    lu_solve_dev(c, n, dA, n, dB, n, nrhs);
    if (nrhs > 0) d2h_matrix(c, B, ldb, dB, n, n, nrhs);
    PSX_HIP(hipStreamSynchronize(c->stream));
  });
}

int psx_lu_solve_pm(psx_ctx* hctx, int n, int sign, const double* Mpp, const double* Mmm,
                    int64_t ldM, double* B, int64_t ldb, int nrhs) {
  Ctx* c = reinterpret_cast<Ctx*>(hctx);
  return guarded(c, [&] {
    PSX_REQUIRE(n >= 1 && nrhs >= 1 && ldM >= n && ldb >= 2 * (int64_t)n, PSX_ERR_ARGUMENT,
                "bad sizes for lu_solve_pm");
    PSX_REQUIRE(sign == 1 || sign == -1, PSX_ERR_ARGUMENT, "sign must be +1 or -1");
    PSX_REQUIRE(Mpp && Mmm && B, PSX_ERR_ARGUMENT, "null pointer");
    PSX_REQUIRE(2 * (int64_t)n <= INT32_MAX / 2, PSX_ERR_ARGUMENT, "n too large");
    PSX_HIP(hipSetDevice(c->device));
    const int N = 2 * n;
    const size_t ab = sizeof(double) * N * (size_t)N, bb = sizeof(double) * N * (size_t)nrhs;
    XferBuf x(c, xfer_bytes({ab, bb}));
    double* dA = x.take(ab);
    double* dB = x.take(bb);
    // [M++ M--; M-- M++]: each block crosses the bus once, its twin is a
    // device copy
    const int64_t off = (int64_t)n * N + n;   // the bottom-right block
    h2d_matrix(c, dA, N, Mpp, ldM, n, n);
    h2d_matrix(c, dA + n, N, Mmm, ldM, n, n);
    PSX_HIP(hipMemcpy2DAsync(dA + off, sizeof(double) * N, dA, sizeof(double) * N,
                             sizeof(double) * n, n, hipMemcpyDeviceToDevice, c->stream));
    PSX_HIP(hipMemcpy2DAsync(dA + (int64_t)n * N, sizeof(double) * N, dA + n,
                             sizeof(double) * N, sizeof(double) * n, n,
                             hipMemcpyDeviceToDevice, c->stream));
    h2d_matrix(c, dB, N, B, ldb, N, nrhs);
    // [M++ -M--; -M-- M++] = D [M++ M--; M-- M++] D, D = diag(I, -I): solve
    // the first against D B and return D X
    const dim3 ng((unsigned)((n + 255) / 256), (unsigned)nrhs);
    if (sign < 0) {
      hipLaunchKernelGGL(negate_rows, ng, dim3(256), 0, c->stream, dB, (int64_t)N, n, n);
      PSX_HIP(hipGetLastError());
    }
    lu_solve_dev(c, N, dA, N, dB, N, nrhs);
    if (sign < 0) {
      hipLaunchKernelGGL(negate_rows, ng, dim3(256), 0, c->stream, dB, (int64_t)N, n, n);
      PSX_HIP(hipGetLastError());
    }
    d2h_matrix(c, B, ldb, dB, N, N, nrhs);
    PSX_HIP(hipStreamSynchronize(c->stream));
  });
}

int psx_lu_solve_d(psx_ctx* hctx, int n, double* dA, int64_t lda, double* dB, int64_t ldb,
                   int nrhs) {
  Ctx* c = reinterpret_cast<Ctx*>(hctx);
  return guarded(c, [&] {
    PSX_REQUIRE(n >= 1 && nrhs >= 0 && lda >= n && ldb >= n, PSX_ERR_ARGUMENT,
                "bad sizes for lu_solve");
    PSX_REQUIRE(dA && (dB || nrhs == 0), PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    lu_solve_dev(c, n, dA, lda, dB, ldb, nrhs);
  });
}

int psx_lu_factor(psx_ctx* hctx, int n, const double* A, int64_t lda, double* LU, int64_t ldlu,
                  int* ipiv) {
  Ctx* c = reinterpret_cast<Ctx*>(hctx);
  return guarded(c, [&] {
    PSX_REQUIRE(n >= 1 && lda >= n && ldlu >= n, PSX_ERR_ARGUMENT, "bad sizes for lu_factor");
    PSX_REQUIRE(A && LU && ipiv, PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    const size_t ab = sizeof(double) * n * (size_t)n, pb = sizeof(int) * ((size_t)n + 64);
    XferBuf x(c, xfer_bytes({ab, pb}));
    double* dA = x.take(ab);
    int* dp = reinterpret_cast<int*>(x.take(pb));
    h2d_matrix(c, dA, n, A, lda, n, n);
    lu_factor_dev(c, n, dA, n, dp, dp + n + 32);
    int hinfo = 0;
    PSX_HIP(hipMemcpyAsync(ipiv, dp, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    PSX_HIP(hipMemcpyAsync(&hinfo, dp + n + 32, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    d2h_matrix(c, LU, ldlu, dA, n, n, n);
    PSX_HIP(hipStreamSynchronize(c->stream));
    PSX_REQUIRE(hinfo == 0, PSX_ERR_SINGULAR,
                "LU: zero pivot at position " + std::to_string(hinfo) +
                    " (Julia SingularException)");
  });
}

int psx_decouple_covmat(psx_ctx* hctx, int n, const double* Y, const double* B1,
                        const double* B2, double* out) {
  Ctx* c = reinterpret_cast<Ctx*>(hctx);
  return guarded(c, [&] {
    PSX_REQUIRE(n >= 1, PSX_ERR_ARGUMENT, "n must be positive");
    PSX_REQUIRE(Y && B1 && B2 && out, PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    // one matrix passed for both B1 and B2 is copied and factored once
    const bool same = B1 == B2;
    const size_t nn = (size_t)n * n, b = sizeof(double) * nn;
    XferBuf x(c, xfer_bytes({4 * b}));
    double* d = x.take(4 * b);
    double* dB2 = same ? d + nn : d + 2 * nn;
    h2d_matrix(c, d, n, Y, n, n, n);
    h2d_matrix(c, d + nn, n, B1, n, n, n);
    if (!same) h2d_matrix(c, dB2, n, B2, n, n, n);
    decouple_dev(c, n, d, d + nn, dB2, d + 3 * nn);
    d2h_matrix(c, out, n, d + 3 * nn, n, n, n);
    PSX_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"
