// apply.hip -- applying the factors (src/id.jl:94-331 for IDPackedV / ID, src/pqr.jl:86-175,
// src/psvd.jl:61-124, src/permute.jl): the interpolation matrix V = [I T] P' (k x n) and its
// adjoint are a row gather / scatter plus one FP64 MFMA GEMM with T; the dense factors of
// PartialQR / PartialSVD / PartialHermEigen are plain GEMMs (lra_dgemm).
#include "lra_internal.h"

#include <algorithm>

namespace {

// D (k x nrhs) = X[idx-1, :] (row gather, idx 1-based)
__global__ void gather_rows_kernel(i64 k, i64 nrhs, const double* __restrict__ X, i64 ldx,
                                   const i64* __restrict__ idx, double* __restrict__ D, i64 ldd) {
  const i64 total = k * nrhs;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;
    D[i + j * ldd] = X[(idx[i] - 1) + j * ldx];
  }
}

// Y[idx-1, :] = S (k x nrhs) (row scatter)
__global__ void scatter_rows_kernel(i64 k, i64 nrhs, const double* __restrict__ S, i64 lds,
                                    const i64* __restrict__ idx, double* __restrict__ Y, i64 ldy) {
  const i64 total = k * nrhs;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;
    Y[(idx[i] - 1) + j * ldy] = S[i + j * lds];
  }
}

// X (m x n) = diag(d) X
__global__ void scale_rows_kernel(i64 m, i64 n, const double* __restrict__ d, double* X, i64 ldx) {
  const i64 total = m * n;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % m, j = e / m;
    X[i + j * ldx] *= d[i];
  }
}

int grid_of(i64 work) { return (int)std::max<i64>(1, std::min<i64>((work + 255) / 256, 8192)); }

}  // namespace

// C = alpha op(A) op(B) + beta C, FP64 (the dgemm behind the reference's mul! on dense
// factors).  trans: 0 = N, 1 = T.  BLAS argument order.
extern "C" int lra_dgemm(lra_ctx* c, int transa, int transb, int64_t m, int64_t n, int64_t k,
                         double alpha, const double* A, int64_t lda, const double* B,
                         int64_t ldb, double beta, double* C, int64_t ldc) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(transa == 0 || transa == 1, 2);
  LRA_ARG(transb == 0 || transb == 1, 3);
  LRA_ARG(m >= 0, 4);
  LRA_ARG(n >= 0, 5);
  LRA_ARG(k >= 0, 6);
  LRA_ARG(lda >= std::max<int64_t>(1, transa ? k : m), 9);
  LRA_ARG(ldb >= std::max<int64_t>(1, transb ? n : k), 11);
  LRA_ARG(ldc >= std::max<int64_t>(1, m), 14);
  LRA_HIP(hipSetDevice(c->device));
  if (m == 0 || n == 0) return LRA_OK;
  GemmArgs g;
  g.M = m; g.N = n; g.K = k; g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.transA = transa; g.B = B; g.ldb = ldb; g.transB = transb;
  g.C = C; g.ldc = ldc; g.bgather = nullptr; g.prec = LRA_PREC_F64;
  return lra_gemm(c, g);
}

// Y = V X (trans 0; X n x nrhs, Y k x nrhs) or Y = V' X (trans 1; X k x nrhs, Y n x nrhs)
// for the interpolation matrix V = [I T] P' of an ID (sk, rd 1-based, T k x (n-k)):
// V X = X[sk, :] + T X[rd, :];  V' X = P [X; T' X].
extern "C" int lra_id_apply(lra_ctx* c, int trans, int64_t k, int64_t n, const int64_t* sk,
                            const int64_t* rd, const double* T, int64_t ldt, int64_t nrhs,
                            const double* X, int64_t ldx, double* Y, int64_t ldy) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == 0 || trans == 1, 2);
  LRA_ARG(k >= 0, 3);
  LRA_ARG(n >= k, 4);
  LRA_ARG(ldt >= std::max<int64_t>(1, k), 8);
  LRA_ARG(nrhs >= 0, 9);
  LRA_ARG(ldx >= std::max<int64_t>(1, trans ? k : n), 11);
  LRA_ARG(ldy >= std::max<int64_t>(1, trans ? n : k), 13);
  LRA_HIP(hipSetDevice(c->device));
  const i64 nr = n - k;
  if (nrhs == 0) return LRA_OK;
  GemmArgs g;
  g.bgather = nullptr; g.prec = LRA_PREC_F64;
  if (trans == 0) {
    if (k == 0) return LRA_OK;
    // Y = X[sk, :], then Y += T X[rd, :]
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_of(k * nrhs)), dim3(256), 0, c->stream, k,
                       nrhs, X, ldx, sk, Y, ldy);
    LRA_HIP(hipGetLastError());
    if (nr > 0) {
      double* Xr = nullptr;
      LRA_CHECK(lra_wsT(c, WS_TMP0, (size_t)nr * nrhs, &Xr));
      hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_of(nr * nrhs)), dim3(256), 0, c->stream,
                         nr, nrhs, X, ldx, rd, Xr, nr);
      LRA_HIP(hipGetLastError());
      g.M = k; g.N = nrhs; g.K = nr; g.alpha = 1.0; g.beta = 1.0;
      g.A = T; g.lda = ldt; g.transA = 0; g.B = Xr; g.ldb = nr; g.transB = 0;
      g.C = Y; g.ldc = ldy;
      LRA_CHECK(lra_gemm(c, g));
    }
  } else {
    // Y[sk, :] = X;  Y[rd, :] = T' X
    if (k > 0) {
      hipLaunchKernelGGL(scatter_rows_kernel, dim3(grid_of(k * nrhs)), dim3(256), 0, c->stream,
                         k, nrhs, X, ldx, sk, Y, ldy);
      LRA_HIP(hipGetLastError());
    }
    if (nr > 0) {
      double* Tx = nullptr;
      LRA_CHECK(lra_wsT(c, WS_TMP0, (size_t)nr * nrhs, &Tx));
      if (k > 0) {
        g.M = nr; g.N = nrhs; g.K = k; g.alpha = 1.0; g.beta = 0.0;
        g.A = T; g.lda = ldt; g.transA = 1; g.B = X; g.ldb = ldx; g.transB = 0;
        g.C = Tx; g.ldc = nr;
        LRA_CHECK(lra_gemm(c, g));
      } else {
        LRA_HIP(hipMemsetAsync(Tx, 0, sizeof(double) * nr * nrhs, c->stream));
      }
      hipLaunchKernelGGL(scatter_rows_kernel, dim3(grid_of(nr * nrhs)), dim3(256), 0, c->stream,
                         nr, nrhs, Tx, nr, rd, Y, ldy);
      LRA_HIP(hipGetLastError());
    }
  }
  return LRA_OK;
}

// X (m x n) = diag(d) X (the diagonal factor of PartialSVD / PartialHermEigen products)
extern "C" int lra_scale_rows(lra_ctx* c, int64_t m, int64_t n, const double* d, double* X,
                              int64_t ldx) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(ldx >= std::max<int64_t>(1, m), 6);
  LRA_HIP(hipSetDevice(c->device));
  if (m == 0 || n == 0) return LRA_OK;
  hipLaunchKernelGGL(scale_rows_kernel, dim3(grid_of(m * n)), dim3(256), 0, c->stream, m, n, d,
                     X, ldx);
  LRA_HIP(hipGetLastError());
  return LRA_OK;
}
