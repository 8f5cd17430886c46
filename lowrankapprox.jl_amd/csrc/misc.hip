// misc.hip -- data-movement kernels of the pipeline (gathers, transposes, index splits).
#include "lra_internal.h"

namespace {

inline int grid_for(i64 work, int block = 256) {
  i64 g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

// out(:, j) = A(:, cols[j]-1)   (getcols(:n, A, sk), src/util.jl:25-29)
__global__ void gather_cols_kernel(i64 m, i64 k, const double* A, i64 lda, const i64* cols,
                                   double* out, i64 ldo) {
  const i64 j = blockIdx.y;
  if (j >= k) return;
  const double* src = A + (cols[j] - 1) * lda;
  double* dst = out + j * ldo;
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < m; i += (i64)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// out(:, j) = A(rows[j]-1, :)'   (getcols(:c, A, sk) for real A)
__global__ void gather_rows_t_kernel(i64 n, i64 k, const double* A, i64 lda, const i64* rows,
                                     double* out, i64 ldo) {
  const i64 j = blockIdx.y;
  if (j >= k) return;
  const i64 r = rows[j] - 1;
  double* dst = out + j * ldo;
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
    dst[i] = A[r + i * lda];
}

// out (n x m) = A (m x n)'  via 32x32 LDS tiles
__global__ void transpose_kernel(i64 m, i64 n, const double* A, i64 lda, double* out, i64 ldo) {
  __shared__ double tile[32][33];
  const i64 i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const i64 i = i0 + tx, j = j0 + r;
    tile[r][tx] = (i < m && j < n) ? A[i + j * lda] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const i64 j = j0 + tx, i = i0 + r;
    if (i < m && j < n) out[j + i * ldo] = tile[tx][r];
  }
}

__global__ void copy2d_kernel(i64 m, i64 n, const double* A, i64 lda, double* out, i64 ldo) {
  const i64 j = blockIdx.y;
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < m; i += (i64)gridDim.x * blockDim.x)
    out[i + j * ldo] = A[i + j * lda];
}

// S (rows x k) = [e_idx[0], ..., e_idx[k-1]]: the one-hot block a matrix-free operator is
// applied to for its columns (getindex(L, :, cols), src/linop.jl:54-60)
__global__ void onehot_kernel(i64 rows, i64 k, const i64* idx, double* S, i64 lds) {
  const i64 j = blockIdx.y;
  if (j >= k) return;
  const i64 hot = idx[j] - 1;
  double* dst = S + j * lds;
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < rows; i += (i64)gridDim.x * blockDim.x)
    dst[i] = i == hot ? 1.0 : 0.0;
}

// out (k x n) = A(rows - 1, :)
__global__ void gather_rows_kernel(i64 k, i64 n, const double* A, i64 lda, const i64* rows,
                                   double* out, i64 ldo) {
  const i64 j = blockIdx.y;
  if (j >= n) return;
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < k; i += (i64)gridDim.x * blockDim.x)
    out[i + j * ldo] = A[(rows[i] - 1) + j * lda];
}

__global__ void split_ids_kernel(i64 n, i64 k, const i64* jpvt, i64* sk, i64* rd) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
    if (i < k) { if (sk) sk[i] = jpvt[i]; }
    else if (rd) rd[i - k] = jpvt[i];
  }
}

}  // namespace

#define LRA_LAUNCH_CHECK()                                                                 \
  do {                                                                                     \
    hipError_t _e = hipGetLastError();                                                     \
    if (_e != hipSuccess) {                                                                \
      lra_set_error("%s:%d launch: %s", __FILE__, __LINE__, hipGetErrorString(_e));       \
      return LRA_E_HIP;                                                                    \
    }                                                                                      \
  } while (0)

int lra_launch_gather_cols(hipStream_t st, i64 m, i64 k, const double* A, i64 lda,
                           const i64* cols, double* out, i64 ldo) {
  if (m <= 0 || k <= 0) return LRA_OK;
  if (k > 65535) { lra_set_error("gather_cols: k too large"); return LRA_E_UNSUPPORTED; }
  dim3 grid((unsigned)std::min<i64>((m + 255) / 256, 64), (unsigned)k);
  hipLaunchKernelGGL(gather_cols_kernel, grid, dim3(256), 0, st, m, k, A, lda, cols, out, ldo);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_gather_rows_t(hipStream_t st, i64 n, i64 k, const double* A, i64 lda,
                             const i64* rows, double* out, i64 ldo) {
  if (n <= 0 || k <= 0) return LRA_OK;
  if (k > 65535) { lra_set_error("gather_rows_t: k too large"); return LRA_E_UNSUPPORTED; }
  dim3 grid((unsigned)std::min<i64>((n + 255) / 256, 64), (unsigned)k);
  hipLaunchKernelGGL(gather_rows_t_kernel, grid, dim3(256), 0, st, n, k, A, lda, rows, out, ldo);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_transpose(hipStream_t st, i64 m, i64 n, const double* A, i64 lda, double* out,
                         i64 ldo) {
  if (m <= 0 || n <= 0) return LRA_OK;
  if ((m + 31) / 32 > 2147483647 || (n + 31) / 32 > 65535) {
    lra_set_error("transpose: shape too large");
    return LRA_E_UNSUPPORTED;
  }
  dim3 grid((unsigned)((m + 31) / 32), (unsigned)((n + 31) / 32));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, st, m, n, A, lda, out, ldo);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_copy2d(hipStream_t st, i64 m, i64 n, const double* A, i64 lda, double* out,
                      i64 ldo) {
  if (m <= 0 || n <= 0) return LRA_OK;
  if (n > 65535) {
    for (i64 j0 = 0; j0 < n; j0 += 65535) {
      const i64 nn = std::min<i64>(65535, n - j0);
      LRA_CHECK(lra_launch_copy2d(st, m, nn, A + j0 * lda, lda, out + j0 * ldo, ldo));
    }
    return LRA_OK;
  }
  dim3 grid((unsigned)std::min<i64>((m + 255) / 256, 64), (unsigned)n);
  hipLaunchKernelGGL(copy2d_kernel, grid, dim3(256), 0, st, m, n, A, lda, out, ldo);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_onehot(hipStream_t st, i64 rows, i64 k, const i64* idx, double* S, i64 lds) {
  if (rows <= 0 || k <= 0) return LRA_OK;
  if (k > 65535) { lra_set_error("onehot: k too large"); return LRA_E_UNSUPPORTED; }
  dim3 grid((unsigned)std::min<i64>((rows + 255) / 256, 64), (unsigned)k);
  hipLaunchKernelGGL(onehot_kernel, grid, dim3(256), 0, st, rows, k, idx, S, lds);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_gather_rows(hipStream_t st, i64 k, i64 n, const double* A, i64 lda,
                           const i64* rows, double* out, i64 ldo) {
  if (k <= 0 || n <= 0) return LRA_OK;
  if (n > 65535) { lra_set_error("gather_rows: n too large"); return LRA_E_UNSUPPORTED; }
  dim3 grid((unsigned)std::min<i64>((k + 255) / 256, 64), (unsigned)n);
  hipLaunchKernelGGL(gather_rows_kernel, grid, dim3(256), 0, st, k, n, A, lda, rows, out, ldo);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_split_ids(hipStream_t st, i64 n, i64 k, const i64* jpvt, i64* sk, i64* rd) {
  if (n <= 0) return LRA_OK;
  hipLaunchKernelGGL(split_ids_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, k, jpvt, sk, rd);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

// flag = 0 when the n x n matrix A is not exactly symmetric (ishermitian for real A); the
// caller sets flag = 1 first.  Each thread compares one (i, j), i > j.
__global__ void not_symmetric_kernel(i64 n, const double* __restrict__ A, i64 lda, int* flag) {
  const i64 j = blockIdx.y;
  for (i64 i = j + 1 + blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n;
       i += (i64)gridDim.x * blockDim.x)
    if (A[i + j * lda] != A[j + i * lda])
      __hip_atomic_store(flag, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__global__ void set_flag_kernel(int* flag, int v) { *flag = v; }

int lra_launch_is_symmetric(hipStream_t st, i64 n, const double* A, i64 lda, int* flag) {
  hipLaunchKernelGGL(set_flag_kernel, dim3(1), dim3(1), 0, st, flag, 1);
  LRA_LAUNCH_CHECK();
  if (n <= 1) return LRA_OK;
  if (n > 65535) {
    lra_set_error("symmetry test: n=%lld exceeds the grid", (long long)n);
    return LRA_E_UNSUPPORTED;
  }
  const unsigned bx = (unsigned)std::min<i64>(64, (n + 255) / 256);
  hipLaunchKernelGGL(not_symmetric_kernel, dim3(bx, (unsigned)n), dim3(256), 0, st, n, A, lda,
                     flag);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}
