// trsm.hip -- T = R11 \ R12 (maxdet_t, src/pqr.jl:438-442; the dtrsm behind
// ldiv!(UpperTriangular(R[:,1:k]), R[:,k+1:n])).
//
// One workgroup per slab of `nc` right-hand sides; the k x nc slab stays in LDS for the whole
// back substitution, the diagonal block column R(0:i, i) is read from L2 once per row.
#include "lra_internal.h"

#include <algorithm>

namespace {

__global__ __launch_bounds__(256) void id_T_kernel(i64 k, i64 nr, const double* __restrict__ R,
                                                   i64 ldr, double* __restrict__ T, i64 ldt,
                                                   int nc) {
  extern __shared__ double slab[];  // k x nc, column-major with leading dimension k
  const i64 c0 = (i64)blockIdx.x * nc;
  const int ncols = (int)std::min<i64>(nc, nr - c0);
  const int tid = threadIdx.x;
  // thread -> (column c, row group g); nc divides 256
  const int c = tid % nc, g = tid / nc, ng = 256 / nc;
  const bool live = c < ncols;
  double* x = slab + (i64)c * k;
  if (live)
    for (i64 r = g; r < k; r += ng) x[r] = R[r + (k + c0 + c) * ldr];
  __syncthreads();
  for (i64 i = k - 1; i >= 0; --i) {
    // x(i) = b(i) / r(i,i); b(0:i-1) -= R(0:i-1, i) * x(i)
    const double xi = live ? x[i] / R[i + i * ldr] : 0.0;
    if (live)
      for (i64 r = g; r < i; r += ng) x[r] -= R[r + i * ldr] * xi;
    __syncthreads();
    if (live && g == 0) x[i] = xi;
    __syncthreads();
  }
  if (live)
    for (i64 r = g; r < k; r += ng) T[r + (c0 + c) * ldt] = x[r];
}

// Blocked back substitution on MFMA: one workgroup per slab of NC right-hand sides, the
// (k rounded up to 16) x NC slab in LDS (leading dimension kp+1 keeps the 16 lanes of an
// MFMA operand on distinct banks).  From the bottom, per block of 16 rows: wave 0 solves
// the 16 x 16 diagonal triangle lane-per-column (dtrsm's division and column-oriented
// update order inside the block), then all waves subtract R(0:r0, block) * X(block) from the
// rows above with v_mfma_f64_16x16x4f64.  Rows beyond k are padding with R = I and zero
// right-hand sides, so they solve to zero and never touch the real rows.
typedef double d4 __attribute__((ext_vector_type(4)));

template <int NC>
__global__ __launch_bounds__(256) void id_T_mfma_kernel(i64 k, i64 nr, const double* __restrict__ R,
                                                        i64 ldr, const double* Bsrc, i64 ldb,
                                                        double* T, i64 ldt) {
  extern __shared__ double slab[];  // (kp+1) x NC, column-major
  __shared__ double s_rbb[16][17];
  const int kp = (int)((k + 15) / 16) * 16;
  const int ld = kp + 1;
  const i64 c0 = (i64)blockIdx.x * NC;
  const int ncols = (int)std::min<i64>(NC, nr - c0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // load the slab (padding rows and missing columns are zero); the right-hand sides are
  // Bsrc(0:k, c0:c0+NC), which may be T itself (each workgroup reads all of its columns before
  // it writes any)
  for (int e = tid; e < kp * NC; e += 256) {
    const int r = e % kp, c = e / kp;
    slab[r + c * ld] = (r < k && c < ncols) ? Bsrc[r + (c0 + c) * ldb] : 0.0;
  }
  for (int r0 = kp - 16; r0 >= 0; r0 -= 16) {
    {  // stage the diagonal block R(r0:r0+16, r0:r0+16) (identity in the padding)
      const int i = tid & 15, j = tid >> 4;
      const i64 gi = r0 + i, gj = r0 + j;
      s_rbb[i][j] = (gi < k && gj < k) ? R[gi + gj * ldr] : (gi == gj ? 1.0 : 0.0);
    }
    __syncthreads();
    if (wave == 0 && lane < NC) {
      // the column's 16 rows in registers (compile-time indices)
      double* x = slab + r0 + lane * ld;
      double xv[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) xv[i] = x[i];
#pragma unroll
      for (int i = 15; i >= 0; --i) {
        const double xi = xv[i] / s_rbb[i][i];
        xv[i] = xi;
#pragma unroll
        for (int r = 0; r < i; ++r) xv[r] -= s_rbb[r][i] * xi;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = xv[i];
    }
    __syncthreads();
    // S(0:r0, :) -= R(0:r0, r0:r0+16) * X(r0:r0+16, :).  Wave w takes row tiles w, w+4, ...:
    // it loads all its R operands first (one memory latency per block instead of one per
    // tile), then runs the MFMAs for every column tile.
    const int ntiles = r0 / 16;
    constexpr int MAXT = 8;
    for (int tb = wave; tb < ntiles; tb += 4 * MAXT) {
      double av[MAXT][4];
#pragma unroll
      for (int j = 0; j < MAXT; ++j) {
        const int t = tb + 4 * j;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const i64 gi = (i64)t * 16 + (lane & 15), gj = r0 + 4 * q + (lane >> 4);
          // gi < r0 <= gj: never diagonal.  Unconditional load from a clamped address (a
          // load under a branch would wait for the one before it), masked afterwards.
          const bool ok = t < ntiles && gj < k;
          const double v = R[(ok ? gi : 0) + (ok ? gj : 0) * ldr];
          av[j][q] = ok ? v : 0.0;
        }
      }
#pragma unroll
      for (int j = 0; j < MAXT; ++j) {
        const int t = tb + 4 * j;
        if (t >= ntiles) break;
#pragma unroll
        for (int ct = 0; ct < NC / 16; ++ct) {
          d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double b = slab[(r0 + 4 * q + (lane >> 4)) + (ct * 16 + (lane & 15)) * ld];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[j][q], b, acc, 0, 0, 0);
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            double* d = slab + (t * 16 + (lane >> 4) + 4 * r) + (ct * 16 + (lane & 15)) * ld;
            *d -= acc[r];
          }
        }
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < (int)k * NC; e += 256) {
    const int r = e % (int)k, c = e / (int)k;
    if (c < ncols) T[r + (c0 + c) * ldt] = slab[r + c * ld];
  }
}

}  // namespace

namespace {
constexpr size_t kTMfmaLds = 150 * 1024;  // the MFMA solve's slab budget (dynamic LDS)

// X (k x nr, ldt) = R(0:k, 0:k) \ Bsrc (k x nr, ldb) on MFMA, slabs of 32 (or 16) right-hand
// sides in LDS; false when even 16 of them do not fit (k above 1199)
bool launch_T_mfma(hipStream_t st, i64 k, i64 nr, const double* R, i64 ldr, const double* Bsrc,
                   i64 ldb, double* T, i64 ldt, hipError_t* err) {
  const i64 kp = (k + 15) / 16 * 16;
  for (int nc : {32, 16}) {
    const size_t bytes = (size_t)(kp + 1) * nc * sizeof(double);
    if (bytes > kTMfmaLds) continue;
    const i64 blocks = (nr + nc - 1) / nc;
    static std::atomic<uint64_t> lds32{0}, lds16{0};  // dynamic LDS beyond 64 KiB
    lra_lds_limit_once(lds32, (const void*)id_T_mfma_kernel<32>, (int)kTMfmaLds);
    lra_lds_limit_once(lds16, (const void*)id_T_mfma_kernel<16>, (int)kTMfmaLds);
    if (nc == 32)
      hipLaunchKernelGGL(id_T_mfma_kernel<32>, dim3((unsigned)blocks), dim3(256), bytes, st, k,
                         nr, R, ldr, Bsrc, ldb, T, ldt);
    else
      hipLaunchKernelGGL(id_T_mfma_kernel<16>, dim3((unsigned)blocks), dim3(256), bytes, st, k,
                         nr, R, ldr, Bsrc, ldb, T, ldt);
    *err = hipGetLastError();
    return true;
  }
  return false;
}
}  // namespace

int lra_launch_id_T(hipStream_t st, i64 k, i64 n, const double* R, i64 ldr, double* T, i64 ldt) {
  const i64 nr = n - k;
  if (k <= 0 || nr <= 0) return LRA_OK;
  {
    hipError_t e = hipSuccess;
    if (launch_T_mfma(st, k, nr, R, ldr, R + k * ldr, ldr, T, ldt, &e)) {
      if (e != hipSuccess) {
        lra_set_error("id_T launch: %s", hipGetErrorString(e));
        return LRA_E_HIP;
      }
      return LRA_OK;
    }
  }
  const size_t lds_cap = 64 * 1024;  // dynamic LDS budget per block
  int nc = 32;  // power of two <= 32 (divides the 256 threads)
  while (nc > 1 && (size_t)k * nc * sizeof(double) > lds_cap) nc >>= 1;
  if ((size_t)k * nc * sizeof(double) > lds_cap) {
    lra_set_error("id_T: k=%lld too large for the LDS slab", (long long)k);
    return LRA_E_UNSUPPORTED;
  }
  const i64 blocks = (nr + nc - 1) / nc;
  hipLaunchKernelGGL(id_T_kernel, dim3((unsigned)blocks), dim3(256),
                     (size_t)k * nc * sizeof(double), st, k, nr, R, ldr, T, ldt, nc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    lra_set_error("id_T launch: %s", hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
}

// T = R11 \ R12 for a k beyond one MFMA slab (k > 1199): block back substitution over row
// blocks of 1184 from the bottom.  Block j's right-hand side R12(j) - R(j, below) T(below) is
// formed in T by a copy and one GEMM, then solved in place on MFMA.  The slab kernel alone
// would hold 2 columns of k doubles per workgroup and run k barrier pairs per column.
// X = R \ X in place for a k x k upper-triangular R (ldr): T holds the nr right-hand sides
// on entry and the solution on exit.  Row blocks of KB from the bottom, as lra_id_T: each
// block's right-hand side loses R(block, below) * X(below) in one GEMM, then the MFMA kernel
// solves it in place.  (At most 64 blocks: LRA_E_UNSUPPORTED beyond.)
int lra_trsm_upper_inplace(lra_ctx* c, i64 k, i64 nr, const double* R, i64 ldr, double* T,
                           i64 ldt) {
  if (k <= 0 || nr <= 0) return LRA_OK;
  constexpr i64 KB = 1184;
  const i64 nblk = (k + KB - 1) / KB;
  if (nblk > 64) {
    lra_set_error("trsm: k=%lld beyond %lld row blocks", (long long)k, (long long)(64 * KB));
    return LRA_E_UNSUPPORTED;
  }
  for (i64 j = nblk - 1; j >= 0; --j) {
    const i64 r0 = j * KB, r1 = std::min(k, r0 + KB), kb = r1 - r0;
    if (r1 < k) {
      GemmArgs g{};
      g.M = kb; g.N = nr; g.K = k - r1;
      g.alpha = -1.0; g.beta = 1.0;
      g.A = R + r0 + r1 * ldr; g.lda = ldr; g.transA = 0;
      g.B = T + r1; g.ldb = ldt; g.transB = 0;
      g.C = T + r0; g.ldc = ldt;
      g.bgather = nullptr;
      g.prec = LRA_PREC_F64;
      LRA_CHECK(lra_gemm(c, g));
    }
    hipError_t e = hipSuccess;
    const bool ok = launch_T_mfma(c->stream, kb, nr, R + r0 + r0 * ldr, ldr, T + r0, ldt, T + r0,
                                  ldt, &e);
    if (!ok || e != hipSuccess) {
      lra_set_error("trsm block launch: %s", ok ? hipGetErrorString(e) : "slab too large");
      return LRA_E_HIP;
    }
  }
  return LRA_OK;
}

int lra_id_T(lra_ctx* c, i64 k, i64 n, const double* R, i64 ldr, double* T, i64 ldt) {
  const i64 nr = n - k;
  if (k <= 0 || nr <= 0) return LRA_OK;
  constexpr i64 KB = 1184;  // (1184 + 1) x 16 doubles fit the slab budget
  if (k <= KB || (k + KB - 1) / KB > 64) return lra_launch_id_T(c->stream, k, n, R, ldr, T, ldt);
  for (i64 j = (k + KB - 1) / KB - 1; j >= 0; --j) {
    const i64 r0 = j * KB, r1 = std::min(k, r0 + KB), kb = r1 - r0;
    const double* Rjj = R + r0 + r0 * ldr;
    hipError_t e = hipSuccess;
    bool ok;
    if (r1 == k) {
      ok = launch_T_mfma(c->stream, kb, nr, Rjj, ldr, R + r0 + k * ldr, ldr, T + r0, ldt, &e);
    } else {
      LRA_CHECK(lra_launch_copy2d(c->stream, kb, nr, R + r0 + k * ldr, ldr, T + r0, ldt));
      GemmArgs g{};
      g.M = kb; g.N = nr; g.K = k - r1;
      g.alpha = -1.0; g.beta = 1.0;
      g.A = R + r0 + r1 * ldr; g.lda = ldr; g.transA = 0;
      g.B = T + r1; g.ldb = ldt; g.transB = 0;
      g.C = T + r0; g.ldc = ldt;
      g.bgather = nullptr;
      g.prec = LRA_PREC_F64;
      LRA_CHECK(lra_gemm(c, g));
      ok = launch_T_mfma(c->stream, kb, nr, Rjj, ldr, T + r0, ldt, T + r0, ldt, &e);
    }
    if (!ok || e != hipSuccess) {
      lra_set_error("id_T block launch: %s", ok ? hipGetErrorString(e) : "slab too large");
      return LRA_E_HIP;
    }
  }
  return LRA_OK;
}

// ---------------------------------------------------------------------------------------
// In-place solve op(U) X = B for the k x k upper triangle U of R (ldr), B k x nrhs (ldb);
// op = U (trans 0, back substitution) or U' (trans 1, forward substitution).  Same slab
// scheme as id_T.
namespace {
__global__ __launch_bounds__(256) void trsm_upper_kernel(i64 k, i64 nrhs,
                                                         const double* __restrict__ R, i64 ldr,
                                                         double* __restrict__ B, i64 ldb,
                                                         int trans, int nc) {
  extern __shared__ double slab[];
  const i64 c0 = (i64)blockIdx.x * nc;
  const int ncols = (int)std::min<i64>(nc, nrhs - c0);
  const int tid = threadIdx.x;
  const int c = tid % nc, g = tid / nc, ng = 256 / nc;
  const bool live = c < ncols;
  double* x = slab + (i64)c * k;
  if (live)
    for (i64 r = g; r < k; r += ng) x[r] = B[r + (c0 + c) * ldb];
  __syncthreads();
  if (!trans) {
    for (i64 i = k - 1; i >= 0; --i) {
      const double xi = live ? x[i] / R[i + i * ldr] : 0.0;
      if (live)
        for (i64 r = g; r < i; r += ng) x[r] -= R[r + i * ldr] * xi;
      __syncthreads();
      if (live && g == 0) x[i] = xi;
      __syncthreads();
    }
  } else {
    // U' is lower triangular: (U')(r, i) = U(i, r)
    for (i64 i = 0; i < k; ++i) {
      const double xi = live ? x[i] / R[i + i * ldr] : 0.0;
      if (live)
        for (i64 r = i + 1 + g; r < k; r += ng) x[r] -= R[i + r * ldr] * xi;
      __syncthreads();
      if (live && g == 0) x[i] = xi;
      __syncthreads();
    }
  }
  if (live)
    for (i64 r = g; r < k; r += ng) B[r + (c0 + c) * ldb] = x[r];
}

__global__ void axpy2d_kernel(i64 m, i64 n, double alpha, const double* X, i64 ldx, double* Y,
                              i64 ldy) {
  const i64 total = m * n;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % m, j = e / m;
    Y[i + j * ldy] += alpha * X[i + j * ldx];
  }
}
}  // namespace

int lra_launch_trsm_upper(hipStream_t st, i64 k, i64 nrhs, const double* R, i64 ldr, double* B,
                          i64 ldb, int trans) {
  if (k <= 0 || nrhs <= 0) return LRA_OK;
  const size_t lds_cap = 64 * 1024;
  int nc = 32;
  while (nc > 1 && (size_t)k * nc * sizeof(double) > lds_cap) nc >>= 1;
  if ((size_t)k * nc * sizeof(double) > lds_cap) {
    lra_set_error("trsm: k=%lld too large for the LDS slab", (long long)k);
    return LRA_E_UNSUPPORTED;
  }
  const i64 blocks = (nrhs + nc - 1) / nc;
  hipLaunchKernelGGL(trsm_upper_kernel, dim3((unsigned)blocks), dim3(256),
                     (size_t)k * nc * sizeof(double), st, k, nrhs, R, ldr, B, ldb, trans, nc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    lra_set_error("trsm launch: %s", hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
}

int lra_launch_axpy2d(hipStream_t st, i64 m, i64 n, double alpha, const double* X, i64 ldx,
                      double* Y, i64 ldy) {
  if (m <= 0 || n <= 0) return LRA_OK;
  const i64 total = m * n;
  const int blocks = (int)std::min<i64>((total + 255) / 256, 16384);
  hipLaunchKernelGGL(axpy2d_kernel, dim3(blocks), dim3(256), 0, st, m, n, alpha, X, ldx, Y, ldy);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    lra_set_error("axpy launch: %s", hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
}
