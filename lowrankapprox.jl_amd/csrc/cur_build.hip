// cur_build.hip -- CUR(A, CURPackedU) (src/cur.jl:85-93): the factors of A ~ C U R from a
// row and column skeleton: C = A[:, cols], R = A[rows, :], and U the pseudo-inverse of the
// intersection W = A[rows, cols] as PartialSVD(V, 1 ./ sigma, U') from W = U diag(sigma) V'
// (svd!, here the one-sided Jacobi SVD of the k x k intersection).  The Hermitian and
// symmetric forms (HermCUR / SymCUR, :102-116) invert the same intersection.
#include "lra_internal.h"

#include <algorithm>

int lra_gesvj_impl(lra_ctx* c, i64 m, i64 n, double* X, i64 ldx, double* U, i64 ldu,
                   double* s, double* V, i64 ldv, int* sweeps);

namespace {

// out (k x n) = A[rows, :] (rows 1-based)
__global__ void gather_rows_kernel(i64 k, i64 n, const double* __restrict__ A, i64 lda,
                                   const i64* __restrict__ rows, double* __restrict__ out,
                                   i64 ldo) {
  const i64 total = k * n;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;
    out[i + j * ldo] = A[(rows[i] - 1) + j * lda];
  }
}

// out (k x k) = A[rows, cols]
__global__ void gather_2d_kernel(i64 k, const double* __restrict__ A, i64 lda,
                                 const i64* __restrict__ rows, const i64* __restrict__ cols,
                                 double* __restrict__ out) {
  const i64 total = k * k;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;
    out[e] = A[(rows[i] - 1) + (cols[j] - 1) * lda];
  }
}

// sinv = 1 ./ s (1/0 = Inf, as 1 ./ sigma in the reference); V (k x k, ldv) = Vw;
// Ut (k x k, ldut) = Uw'
__global__ void cur_u_kernel(i64 k, const double* __restrict__ s, const double* __restrict__ Uw,
                             const double* __restrict__ Vw, double* __restrict__ sinv,
                             double* __restrict__ V, i64 ldv, double* __restrict__ Ut,
                             i64 ldut) {
  const i64 total = k * k;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;
    V[i + j * ldv] = Vw[e];
    Ut[i + j * ldut] = Uw[j + i * k];
    if (j == 0) sinv[i] = 1.0 / s[i];
  }
}

int grid_of(i64 work) { return (int)std::max<i64>(1, std::min<i64>((work + 255) / 256, 8192)); }

}  // namespace

extern "C" int lra_cur_build(lra_ctx* c, int64_t m, int64_t n, const double* A, int64_t lda,
                             const int64_t* rows, const int64_t* cols, int64_t k, double* C,
                             int64_t ldc, double* R, int64_t ldr, double* V, int64_t ldv,
                             double* sinv, double* Ut, int64_t ldut) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(A != nullptr || m * n == 0, 4);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 5);
  LRA_ARG(k >= 0 && k <= std::min(m, n), 8);
  LRA_ARG(rows != nullptr || k == 0, 6);
  LRA_ARG(cols != nullptr || k == 0, 7);
  LRA_ARG(C == nullptr || ldc >= std::max<int64_t>(1, m), 10);
  LRA_ARG(R == nullptr || ldr >= std::max<int64_t>(1, k), 12);
  LRA_ARG(V != nullptr || k == 0, 13);
  LRA_ARG(ldv >= std::max<int64_t>(1, k), 14);
  LRA_ARG(sinv != nullptr || k == 0, 15);
  LRA_ARG(Ut != nullptr || k == 0, 16);
  LRA_ARG(ldut >= std::max<int64_t>(1, k), 17);
  LRA_HIP(hipSetDevice(c->device));
  if (k == 0) return LRA_OK;
  if (C) LRA_CHECK(lra_launch_gather_cols(c->stream, m, k, A, lda, cols, C, ldc));
  if (R) {
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_of(k * n)), dim3(256), 0, c->stream, k, n,
                       A, lda, rows, R, ldr);
    LRA_HIP(hipGetLastError());
  }
  double* W = nullptr;
  LRA_CHECK(lra_wsT(c, WS_CU0, (size_t)3 * k * k + k, &W));
  double* Uw = W + k * k;
  double* Vw = Uw + k * k;
  double* s = Vw + k * k;
  hipLaunchKernelGGL(gather_2d_kernel, dim3(grid_of(k * k)), dim3(256), 0, c->stream, k, A, lda,
                     rows, cols, W);
  LRA_HIP(hipGetLastError());
  int sweeps = 0;
  LRA_CHECK(lra_gesvj_impl(c, k, k, W, k, Uw, k, s, Vw, k, &sweeps));
  hipLaunchKernelGGL(cur_u_kernel, dim3(grid_of(k * k)), dim3(256), 0, c->stream, k, s, Uw, Vw,
                     sinv, V, ldv, Ut, ldut);
  LRA_HIP(hipGetLastError());
  return LRA_OK;
}
