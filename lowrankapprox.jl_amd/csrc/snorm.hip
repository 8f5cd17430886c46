// snorm.hip -- spectral norm by power iteration (snorm / snormdiff, src/snorm.jl:14-53).
//
// The reference iterates xn <- A'(A xn) (A itself when Hermitian) with s = sqrt(|xn|) and
// stops when |s - t| <= max(atol, t*rtol) or after snorm_niter steps (LRAOptions default 32).
// Each step is two HBM-bound matrix-vector products; they run as GEMV kernels that read A
// once per product.  snormdiff(A, F) applies A - F matrix-free (src/linop.jl:155-187); for an
// ID F = A[:,sk] [I T] P', (A - F) x = A x - A[:,sk] (x[sk] + T x[rd]) and
// (A - F)' y = A' y - P [I T]' A[:,sk]' y.
#include "lra_internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace {

constexpr int GV_ROWS = 256;

// partial[c][i] = sum_{j in chunk c} A[i, cols? cols[j]-1 : j] * x[j]
__global__ __launch_bounds__(256) void gemv_n_kernel(i64 m, i64 n, const double* __restrict__ A,
                                                     i64 lda, const i64* __restrict__ cols,
                                                     const double* __restrict__ x, i64 chunk,
                                                     double* __restrict__ partial) {
  __shared__ double sx[512];
  const i64 i = (i64)blockIdx.x * GV_ROWS + threadIdx.x;
  const i64 j0 = (i64)blockIdx.y * chunk, j1 = std::min<i64>(n, j0 + chunk);
  double acc = 0.0;
  for (i64 jb = j0; jb < j1; jb += 512) {
    const i64 je = std::min<i64>(j1, jb + 512);
    __syncthreads();
    for (i64 j = jb + threadIdx.x; j < je; j += 256) sx[j - jb] = x[j];
    __syncthreads();
    if (i < m) {
      for (i64 j = jb; j < je; ++j) {
        const i64 cj = cols ? cols[j] - 1 : j;
        acc += A[i + cj * lda] * sx[j - jb];
      }
    }
  }
  if (i < m) partial[(i64)blockIdx.y * m + i] = acc;
}

__global__ void sum_partials_kernel(i64 m, int nchunks, const double* partial, double alpha,
                                    double* y, double beta) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < m; i += (i64)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += partial[(i64)c * m + i];
    y[i] = beta == 0.0 ? alpha * s : alpha * s + beta * y[i];
  }
}

// y[j] = alpha * dot(A[:, cols? cols[j]-1 : j], x) + beta * y[j]   (one wave per column)
__global__ __launch_bounds__(256) void gemv_t_kernel(i64 m, i64 n, const double* __restrict__ A,
                                                     i64 lda, const i64* __restrict__ cols,
                                                     const double* __restrict__ x, double alpha,
                                                     double* __restrict__ y, double beta) {
  const int lane = threadIdx.x & 63;
  const i64 j = (i64)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  const double* a = A + (cols ? cols[j] - 1 : j) * lda;
  double s = 0.0;
  for (i64 i = lane; i < m; i += 64) s += a[i] * x[i];
  s = lra_wave_sum(s);
  if (lane == 0) y[j] = beta == 0.0 ? alpha * s : alpha * s + beta * y[j];
}

// z[r] = x[sk[r]-1] + sum_c T[r,c] x[rd[c]-1]   (k outputs, one wave each)
__global__ void id_apply_n_kernel(i64 k, i64 nr, const double* __restrict__ T, i64 ldt,
                                  const i64* __restrict__ sk, const i64* __restrict__ rd,
                                  const double* __restrict__ x, double* __restrict__ z) {
  const int lane = threadIdx.x & 63;
  const i64 r = (i64)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= k) return;
  double s = 0.0;
  for (i64 c = lane; c < nr; c += 64) s += T[r + c * ldt] * x[rd[c] - 1];
  s = lra_wave_sum(s);
  if (lane == 0) z[r] = x[sk[r] - 1] + s;
}

// y[sk[r]-1] -= w[r];  y[rd[c]-1] -= sum_r T[r,c] w[r]   (w = A[:,sk]' v)
__global__ void id_apply_t_sub_kernel(i64 k, i64 nr, const double* __restrict__ T, i64 ldt,
                                      const i64* __restrict__ sk, const i64* __restrict__ rd,
                                      const double* __restrict__ w, double* __restrict__ y) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < k + nr;
       e += (i64)gridDim.x * blockDim.x) {
    if (e < k) {
      y[sk[e] - 1] -= w[e];
    } else {
      const i64 c = e - k;
      double s = 0.0;
      for (i64 r = 0; r < k; ++r) s += T[r + c * ldt] * w[r];
      y[rd[c] - 1] -= s;
    }
  }
}

// out[0] = |x|^2 (one block)
__global__ void sumsq_kernel(i64 n, const double* x, double* out) {
  __shared__ double red[4];
  double s = 0.0;
  for (i64 i = threadIdx.x; i < n; i += 256) s += x[i] * x[i];
  s = lra_wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = red[0] + red[1] + red[2] + red[3];
}

__global__ void scale_kernel(i64 n, double* x, const double* nrm2) {
  const double inv = 1.0 / sqrt(nrm2[0]);
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
    x[i] *= inv;
}

__global__ void sub_kernel(i64 n, double* y, const double* z) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
    y[i] -= z[i];
}

inline int blocks_for(i64 n) { return (int)std::min<i64>(std::max<i64>((n + 255) / 256, 1), 4096); }

}  // namespace

// column chunks of gemv_n for an m x n product (the partial sums are nchunks x m)
static i64 gemv_chunks(const lra_ctx* c, i64 m, i64 n, i64* chunk) {
  const i64 rt = (m + GV_ROWS - 1) / GV_ROWS;
  i64 nchunks = std::max<i64>(1, std::min<i64>((n + 255) / 256, (4ll * c->num_cu + rt - 1) / rt));
  *chunk = (n + nchunks - 1) / std::max<i64>(nchunks, 1);
  return n > 0 ? (n + *chunk - 1) / *chunk : 1;
}

// y = alpha * A[:, cols] x + beta * y  (cols optional, 1-based).  The partial sums go to
// `part` (at least gemv_chunks(...) * m doubles) when given, else to a workspace slot.
int lra_gemv_n(lra_ctx* c, i64 m, i64 n, const double* A, i64 lda, const i64* cols,
               const double* x, double alpha, double* y, double beta, double* part = nullptr) {
  if (m <= 0) return LRA_OK;
  const i64 rt = (m + GV_ROWS - 1) / GV_ROWS;
  i64 chunk = 0;
  const i64 nchunks = gemv_chunks(c, m, n, &chunk);
  if (!part) LRA_CHECK(lra_wsT(c, WS_X0, (size_t)nchunks * m, &part));
  if (n > 0) {
    hipLaunchKernelGGL(gemv_n_kernel, dim3((unsigned)rt, (unsigned)nchunks), dim3(256), 0, c->stream,
                       m, n, A, lda, cols, x, chunk, part);
  }
  hipLaunchKernelGGL(sum_partials_kernel, dim3(blocks_for(m)), dim3(256), 0, c->stream, m,
                     n > 0 ? (int)nchunks : 0, part, alpha, y, beta);
  LRA_HIP(hipGetLastError());
  return LRA_OK;
}

// y = alpha * A[:, cols]' x + beta * y
int lra_gemv_t(lra_ctx* c, i64 m, i64 n, const double* A, i64 lda, const i64* cols,
               const double* x, double alpha, double* y, double beta) {
  if (n <= 0) return LRA_OK;
  hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, c->stream, m, n,
                     A, lda, cols, x, alpha, y, beta);
  LRA_HIP(hipGetLastError());
  return LRA_OK;
}

namespace {

struct Op {
  // (A - F) where F is an ID of A (k > 0) or nothing (k < 0); `part`: caller-provided
  // partial-sum buffer of the GEMVs (else a workspace slot)
  i64 m, n;
  const double* A;
  i64 lda;
  i64 k;
  const i64* sk;
  const i64* rd;
  const double* T;
  i64 ldt;
  const double* B;  // second dense operand (snormdiff of two matrices) or nullptr
  i64 ldb;
  lra_apply_fn apply = nullptr;  // matrix-free operator (replaces A when set)
  void* user = nullptr;
  double* part = nullptr;
};

int op_apply(lra_ctx* c, const Op& o, const double* x, double* y, double* wk) {
  if (o.apply) {
    const int st = o.apply(o.user, 0, 1, x, std::max<i64>(o.n, 1), y, std::max<i64>(o.m, 1),
                           (void*)c->stream);
    if (st != 0) { lra_set_error("operator apply returned %d", st); return LRA_E_STATE; }
    return LRA_OK;
  }
  LRA_CHECK(lra_gemv_n(c, o.m, o.n, o.A, o.lda, nullptr, x, 1.0, y, 0.0, o.part));
  if (o.B) LRA_CHECK(lra_gemv_n(c, o.m, o.n, o.B, o.ldb, nullptr, x, -1.0, y, 1.0, o.part));
  if (o.k > 0) {
    const i64 nr = o.n - o.k;
    hipLaunchKernelGGL(id_apply_n_kernel, dim3((unsigned)((o.k + 3) / 4)), dim3(256), 0, c->stream,
                       o.k, nr, o.T, o.ldt, o.sk, o.rd, x, wk);
    LRA_HIP(hipGetLastError());
    LRA_CHECK(lra_gemv_n(c, o.m, o.k, o.A, o.lda, o.sk, wk, -1.0, y, 1.0, o.part));
  }
  return LRA_OK;
}

int op_apply_t(lra_ctx* c, const Op& o, const double* x, double* y, double* wk) {
  if (o.apply) {
    const int st = o.apply(o.user, 1, 1, x, std::max<i64>(o.m, 1), y, std::max<i64>(o.n, 1),
                           (void*)c->stream);
    if (st != 0) { lra_set_error("operator apply returned %d", st); return LRA_E_STATE; }
    return LRA_OK;
  }
  LRA_CHECK(lra_gemv_t(c, o.m, o.n, o.A, o.lda, nullptr, x, 1.0, y, 0.0));
  if (o.B) LRA_CHECK(lra_gemv_t(c, o.m, o.n, o.B, o.ldb, nullptr, x, -1.0, y, 1.0));
  if (o.k > 0) {
    const i64 nr = o.n - o.k;
    LRA_CHECK(lra_gemv_t(c, o.m, o.k, o.A, o.lda, o.sk, x, 1.0, wk, 0.0));
    hipLaunchKernelGGL(id_apply_t_sub_kernel, dim3(blocks_for(o.k + nr)), dim3(256), 0,
                       c->stream, o.k, nr, o.T, o.ldt, o.sk, o.rd, wk, y);
    LRA_HIP(hipGetLastError());
  }
  return LRA_OK;
}

// Vectors of the power iteration: xn (n), xm (m), wk (max(k,1) + 64; its tail holds the
// squared norm).
struct PowerBufs {
  double *xn, *xm, *wk;
};

// power iteration loop of snorm (src/snorm.jl:14-45); buffers from the workspace unless given
int power_iteration(lra_ctx* c, const Op& o, int isherm, const lra_opts* opts, const double* x0,
                    double* result, const PowerBufs* given = nullptr) {
  const i64 m = o.m, n = o.n;
  double *xn = nullptr, *xm = nullptr, *wk = nullptr, *scal = nullptr;
  if (given) {
    xn = given->xn; xm = given->xm; wk = given->wk;
  } else {
    LRA_CHECK(lra_wsT(c, WS_QR0, (size_t)std::max<i64>(n, 1), &xn));
    LRA_CHECK(lra_wsT(c, WS_QR1, (size_t)std::max<i64>(m, 1), &xm));
    LRA_CHECK(lra_wsT(c, WS_QR2, (size_t)std::max<i64>(std::max<i64>(o.k, 1), 1) + 64, &wk));
  }
  scal = wk + std::max<i64>(o.k, 1);
  if (x0) LRA_HIP(hipMemcpyAsync(xn, x0, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  else LRA_CHECK(lra_launch_gaussian_fill(c->stream, opts->seed, 0x5A0A5A0Aull, n, 1, xn, n));
  hipLaunchKernelGGL(sumsq_kernel, dim3(1), dim3(256), 0, c->stream, n, xn, scal);
  LRA_HIP(hipGetLastError());
  double s = 1.0, t = 0.0;
  i64 niter = 0;
  double* hs = (double*)c->host_pinned;
  while (s > 0 && std::fabs(s - t) > std::max(opts->atol, t * opts->rtol)) {
    if (niter == opts->snorm_niter) break;
    ++niter;
    hipLaunchKernelGGL(scale_kernel, dim3(blocks_for(n)), dim3(256), 0, c->stream, n, xn, scal);
    LRA_HIP(hipGetLastError());
    if (isherm) {
      LRA_CHECK(op_apply(c, o, xn, xm, wk));
      LRA_HIP(hipMemcpyAsync(xn, xm, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    } else {
      LRA_CHECK(op_apply(c, o, xn, xm, wk));
      LRA_CHECK(op_apply_t(c, o, xm, xn, wk));
    }
    hipLaunchKernelGGL(sumsq_kernel, dim3(1), dim3(256), 0, c->stream, n, xn, scal);
    LRA_HIP(hipGetLastError());
    LRA_HIP(hipMemcpyAsync(hs, scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LRA_HIP(hipStreamSynchronize(c->stream));
    const double xnrm = std::sqrt(hs[0]);
    t = s;
    s = isherm ? xnrm : std::sqrt(xnrm);
  }
  *result = s;
  return LRA_OK;
}

}  // namespace

extern "C" int lra_snorm(lra_ctx* c, int64_t m, int64_t n, const double* A, int64_t lda,
                         int isherm, const lra_opts* opts, const double* x0, double* result) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 5);
  LRA_ARG(!isherm || m == n, 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(result != nullptr, 9);
  Op o{m, n, A, lda, -1, nullptr, nullptr, nullptr, 0, nullptr, 0};
  return power_iteration(c, o, isherm, opts, x0, result);
}

extern "C" int lra_snormdiff(lra_ctx* c, int64_t m, int64_t n, const double* A, int64_t lda,
                             const double* B, int64_t ldb, int isherm, const lra_opts* opts,
                             const double* x0, double* result) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 5);
  LRA_ARG(B != nullptr, 6);
  LRA_ARG(ldb >= std::max<int64_t>(1, m), 7);
  LRA_ARG(opts != nullptr, 9);
  LRA_ARG(result != nullptr, 11);
  Op o{m, n, A, lda, -1, nullptr, nullptr, nullptr, 0, B, ldb};
  return power_iteration(c, o, isherm, opts, x0, result);
}

extern "C" int lra_snormdiff_id(lra_ctx* c, int64_t m, int64_t n, const double* A, int64_t lda,
                                int64_t k, const int64_t* sk, const int64_t* rd, const double* T,
                                int64_t ldt, const lra_opts* opts, const double* x0,
                                double* result) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 5);
  LRA_ARG(k >= 0 && k <= n, 6);
  LRA_ARG(ldt >= std::max<int64_t>(1, k), 10);
  LRA_ARG(opts != nullptr, 11);
  LRA_ARG(result != nullptr, 13);
  Op o{m, n, A, lda, k, sk, rd, T, ldt, nullptr, 0};
  if (k == 0) o.k = -1;
  return power_iteration(c, o, 0, opts, x0, result);
}

// snorm of a matrix-free operator (snorm(L::LinearOperator), src/snorm.jl:14-45): the power
// iteration applies the caller's L and L' to one vector per step.
extern "C" int lra_snorm_op(lra_ctx* c, int64_t m, int64_t n, lra_apply_fn apply, void* user,
                            int isherm, const lra_opts* opts, const double* x0, double* result) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(apply != nullptr, 4);
  LRA_ARG(!isherm || m == n, 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(result != nullptr, 9);
  Op o{m, n, nullptr, 1, -1, nullptr, nullptr, nullptr, 0, nullptr, 0};
  o.apply = apply;
  o.user = user;
  return power_iteration(c, o, isherm, opts, x0, result);
}

// ---------------------------------------------------------------------------------------
// Allocation-free variant of lra_snormdiff_id.  All of its device memory beyond the caller's
// inputs is `scratch`, a file-scope device array of kScratchDoubles doubles (32 MiB) that
// every GPU the library runs on holds until the process ends.  On the host the array's name
// is only a stand-in: its device address is fetched per device with hipGetSymbolAddress on the
// first call on that device (with the device current) and kept in g_scratch; only that
// address is handed to kernels.  Calls on one GPU share its array (the last call's contents
// stay behind; every element a call reads it has written first, so nothing is cleared), and
// calls that overlap on one GPU race.

static constexpr size_t kScratchDoubles = (size_t)4 << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
double* g_scratch[64] = {};

int scratch_on(int device, double** out) {
  if (device < 0 || device >= 64) {
    lra_set_error("scratch: device %d out of range", device);
    return LRA_E_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    LRA_HIP(hipSetDevice(device));
    void* p = nullptr;
    LRA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return LRA_OK;
}
}  // namespace

extern "C" int lra_snormdiff_id_noalloc(lra_ctx* c, int64_t m, int64_t n, const double* A,
                                        int64_t lda, int64_t k, const int64_t* sk,
                                        const int64_t* rd, const double* T, int64_t ldt,
                                        const lra_opts* opts, const double* x0,
                                        double* result) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 5);
  LRA_ARG(k >= 0 && k <= n, 6);
  LRA_ARG(ldt >= std::max<int64_t>(1, k), 10);
  LRA_ARG(opts != nullptr, 11);
  LRA_ARG(result != nullptr, 13);
  // scratch layout: xn (n) | xm (m) | wk (max(k,1) + 64) | GEMV partial sums (chunks x m)
  i64 chunk = 0;
  const i64 nparts = std::max(gemv_chunks(c, m, n, &chunk), gemv_chunks(c, m, k, &chunk));
  const size_t need = (size_t)std::max<i64>(n, 1) + (size_t)std::max<i64>(m, 1) +
                      (size_t)std::max<i64>(k, 1) + 64 + (size_t)nparts * (size_t)m;
  if (need > kScratchDoubles) {
    lra_set_error("snormdiff_id_noalloc: %zu doubles of scratch needed for m=%lld n=%lld "
                  "k=%lld, the maximum is %zu (use lra_snormdiff_id)", need, (long long)m,
                  (long long)n, (long long)k, kScratchDoubles);
    return LRA_E_UNSUPPORTED;
  }
  double* base = nullptr;
  LRA_CHECK(scratch_on(c->device, &base));
  LRA_HIP(hipSetDevice(c->device));
  PowerBufs b;
  b.xn = base;
  b.xm = b.xn + std::max<i64>(n, 1);
  b.wk = b.xm + std::max<i64>(m, 1);
  Op o{m, n, A, lda, k, sk, rd, T, ldt, nullptr, 0};
  o.part = b.wk + std::max<i64>(k, 1) + 64;
  if (k == 0) o.k = -1;
  return power_iteration(c, o, 0, opts, x0, result, &b);
}
