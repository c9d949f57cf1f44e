// FP64 LU decoupling on rocSOLVER: `A \ B` (reference
// src/blockspectralmatrix.jl:89-129) and decouple_covmat
// (src/covariance.jl:8-14).  Julia's `lu` is LAPACK getrf with partial
// pivoting and throws SingularException on an exact zero pivot; rocSOLVER's
// getrf reports the same condition through `info`.
#include "psx_internal.h"

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

void release_blas(Ctx* c) {
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

// dA (n x n, lda) is overwritten by its LU factors; dB (n x nrhs) by X.
void lu_solve_dev(Ctx* c, int n, double* dA, int64_t lda, double* dB, int64_t ldb,
                  int nrhs) {
  rocblas_handle h = blas(c);
  char* s = (char*)ctx_scratch(c, sizeof(rocblas_int) * (n + 64));
  rocblas_int* ipiv = (rocblas_int*)s;
  rocblas_int* info = ipiv + n + 32;
  PSX_ROC(rocsolver_dgetrf(h, n, n, dA, (rocblas_int)lda, ipiv, info));
  rocblas_int hinfo = 0;
  PSX_HIP(hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, c->stream));
  PSX_HIP(hipStreamSynchronize(c->stream));
  PSX_REQUIRE(hinfo == 0, PSX_ERR_SINGULAR,
              "LU: zero pivot at position " + std::to_string(hinfo) +
                  " (Julia SingularException)");
  if (nrhs > 0)
    PSX_ROC(rocsolver_dgetrs(h, rocblas_operation_none, n, nrhs, dA, (rocblas_int)lda,
                             ipiv, dB, (rocblas_int)ldb));
}

static void transpose(Ctx* c, int n, const double* dIn, double* dOut) {
  const double one = 1.0, zero = 0.0;
  PSX_ROC(rocblas_dgeam(blas(c), rocblas_operation_transpose, rocblas_operation_none, n, n,
                        &one, dIn, n, &zero, dOut, n, dOut, n));
}

// dOut = B1^{-1} Y B2^{-T}; dB1, dB2 are overwritten by their LU factors.
void decouple_dev(Ctx* c, int n, const double* dY, double* dB1, double* dB2,
                  double* dOut) {
  const size_t nn = (size_t)n * n;
  double* tmp = (double*)ctx_alloc(c, sizeof(double) * nn);
  try {
    PSX_HIP(hipMemcpyAsync(tmp, dY, sizeof(double) * nn, hipMemcpyDeviceToDevice,
                           c->stream));
    lu_solve_dev(c, n, dB1, n, tmp, n, n);  // tmp = B1^{-1} Y
    transpose(c, n, tmp, dOut);             // dOut = tmp^T
    lu_solve_dev(c, n, dB2, n, dOut, n, n); // dOut = B2^{-1} tmp^T
    transpose(c, n, dOut, tmp);             // tmp = (B2^{-1} tmp^T)^T
    PSX_HIP(hipMemcpyAsync(dOut, tmp, sizeof(double) * nn, hipMemcpyDeviceToDevice,
                           c->stream));
    PSX_HIP(hipStreamSynchronize(c->stream));
  } catch (...) {
    (void)hipStreamSynchronize(c->stream);
    ctx_free(c, tmp);
    throw;
  }
  ctx_free(c, tmp);
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
    const size_t nn = (size_t)n * n, nb = (size_t)n * std::max(nrhs, 1);
    double* dA = (double*)ctx_alloc(c, sizeof(double) * nn);
    double* dB = (double*)ctx_alloc(c, sizeof(double) * nb);
    try {
      h2d_matrix(c, dA, n, A, lda, n, n);
      if (nrhs > 0) h2d_matrix(c, dB, n, B, ldb, n, nrhs);
      lu_solve_dev(c, n, dA, n, dB, n, nrhs);
      if (nrhs > 0) d2h_matrix(c, B, ldb, dB, n, n, nrhs);
      PSX_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
      (void)hipStreamSynchronize(c->stream);
      ctx_free(c, dA);
      ctx_free(c, dB);
      throw;
    }
    ctx_free(c, dA);
    ctx_free(c, dB);
  });
}

int psx_decouple_covmat(psx_ctx* hctx, int n, const double* Y, const double* B1,
                        const double* B2, double* out) {
  Ctx* c = reinterpret_cast<Ctx*>(hctx);
  return guarded(c, [&] {
    PSX_REQUIRE(n >= 1, PSX_ERR_ARGUMENT, "n must be positive");
    PSX_REQUIRE(Y && B1 && B2 && out, PSX_ERR_ARGUMENT, "null pointer");
    PSX_HIP(hipSetDevice(c->device));
    const size_t nn = (size_t)n * n, b = sizeof(double) * nn;
    double* d = (double*)ctx_alloc(c, 4 * b);
    try {
      h2d_matrix(c, d, n, Y, n, n, n);
      h2d_matrix(c, d + nn, n, B1, n, n, n);
      h2d_matrix(c, d + 2 * nn, n, B2, n, n, n);
      decouple_dev(c, n, d, d + nn, d + 2 * nn, d + 3 * nn);
      d2h_matrix(c, out, n, d + 3 * nn, n, n, n);
      PSX_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
      (void)hipStreamSynchronize(c->stream);
      ctx_free(c, d);
      throw;
    }
    ctx_free(c, d);
  });
}

}  // extern "C"
