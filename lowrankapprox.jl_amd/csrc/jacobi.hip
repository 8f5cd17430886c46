// jacobi.hip -- one-sided (Hestenes) Jacobi SVD of a small dense matrix: the svd!/eigen! of
// the k x k cores of psvdfact and pheigfact (src/psvd.jl:245,259, src/pheig.jl:284).
//
// X (m x n, m >= n) is orthogonalised column pair by column pair: for a pair (p, q) with
// a = |x_p|^2, b = |x_q|^2, g = x_p'x_q, the rotation that diagonalises [a g; g b] is applied
// to the two columns of X and of V.  A sweep visits every pair once in the round-robin
// ("circle") order: n/2 disjoint pairs per round, one workgroup per pair, n-1 rounds per
// sweep; sweeps repeat until a sweep rotates nothing (|g| <= tol sqrt(a b)).  At the end the
// column norms of X are the singular values, X's normalised columns the left vectors, V the
// right ones (X_in = U diag(s) V'), sorted by decreasing s.  One-sided Jacobi computes the
// small singular values to high relative accuracy.
#include "lra_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

// pair (p, q) of slot i in round r for N (even) players: player 0 fixed, the others rotate
LRA_DEV void rr_pair(int N, int r, int i, int* p, int* q) {
  *p = i == 0 ? 0 : 1 + (r + i - 1) % (N - 1);
  *q = 1 + (r + N - 2 - i) % (N - 1);
}

__global__ __launch_bounds__(256) void jacobi_round_kernel(int m, int n, int N, int r,
                                                           double* __restrict__ X, i64 ldx,
                                                           double* __restrict__ V, i64 ldv,
                                                           double tol, int* __restrict__ rotated) {
  int p, q;
  rr_pair(N, r, blockIdx.x, &p, &q);
  if (p >= n || q >= n) return;  // the dummy player of an odd n
  if (p > q) { const int t = p; p = q; q = t; }
  double* xp = X + (i64)p * ldx;
  double* xq = X + (i64)q * ldx;
  double a = 0.0, b = 0.0, g = 0.0;
  for (int i = threadIdx.x; i < m; i += 256) {
    const double u = xp[i], w = xq[i];
    a += u * u;
    b += w * w;
    g += u * w;
  }
  __shared__ double sa[4], sb[4], sg[4];
  a = lra_wave_sum(a);
  b = lra_wave_sum(b);
  g = lra_wave_sum(g);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { sa[wave] = a; sb[wave] = b; sg[wave] = g; }
  __syncthreads();
  a = (sa[0] + sa[1]) + (sa[2] + sa[3]);
  b = (sb[0] + sb[1]) + (sb[2] + sb[3]);
  g = (sg[0] + sg[1]) + (sg[2] + sg[3]);
  if (g == 0.0 || !(fabs(g) > tol * sqrt(a) * sqrt(b))) return;
  // t = tan(theta): the smaller root of t^2 + 2 zeta t - 1 = 0, zeta = (b - a) / (2 g)
  const double zeta = (b - a) / (2.0 * g);
  const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
  for (int i = threadIdx.x; i < m; i += 256) {
    const double u = xp[i], w = xq[i];
    xp[i] = cs * u - sn * w;
    xq[i] = sn * u + cs * w;
  }
  double* vp = V + (i64)p * ldv;
  double* vq = V + (i64)q * ldv;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double u = vp[i], w = vq[i];
    vp[i] = cs * u - sn * w;
    vq[i] = sn * u + cs * w;
  }
  if (threadIdx.x == 0) __hip_atomic_fetch_add(rotated, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void jacobi_eye_kernel(int n, double* V, i64 ldv, int* rotated) {
  const i64 total = (i64)n * n;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x)
    V[(e % n) + (e / n) * ldv] = (e % n) == (e / n) ? 1.0 : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *rotated = 0;
}

__global__ void zero_int_kernel(int* x) { *x = 0; }

// column norms of X into s
__global__ __launch_bounds__(256) void colnorm_kernel(int m, const double* __restrict__ X,
                                                      i64 ldx, double* __restrict__ s) {
  const double* x = X + (i64)blockIdx.x * ldx;
  double a = 0.0;
  for (int i = threadIdx.x; i < m; i += 256) a += x[i] * x[i];
  __shared__ double sa[4];
  a = lra_wave_sum(a);
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) s[blockIdx.x] = sqrt((sa[0] + sa[1]) + (sa[2] + sa[3]));
}

// U(:, j) = X(:, perm[j]) / s(perm[j]);  Vo(:, j) = V(:, perm[j]);  so(j) = s(perm[j])
__global__ void jacobi_finish_kernel(int m, int n, const int* __restrict__ perm,
                                     const double* __restrict__ X, i64 ldx,
                                     const double* __restrict__ V, i64 ldv,
                                     const double* __restrict__ s, double* __restrict__ U,
                                     i64 ldu, double* __restrict__ Vo, i64 ldvo,
                                     double* __restrict__ so) {
  const int j = blockIdx.x;
  const int pj = perm[j];
  const double sj = s[pj];
  const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
  for (int i = threadIdx.x; i < m; i += 256)
    U[i + (i64)j * ldu] = sj > 0.0 ? X[i + (i64)pj * ldx] * inv : (i == j ? 1.0 : 0.0);
  for (int i = threadIdx.x; i < n; i += 256) Vo[i + (i64)j * ldvo] = V[i + (i64)pj * ldv];
  if (threadIdx.x == 0) so[j] = sj;
}

}  // namespace

// X (m x n, m >= n) = U diag(s) V': X is overwritten, U (m x n), s (n, decreasing), Vt-less
// V (n x n).  Returns the sweep count in *sweeps.
int lra_gesvj_impl(lra_ctx* c, i64 m, i64 n, double* X, i64 ldx, double* U, i64 ldu,
                   double* s, double* V, i64 ldv, int* sweeps) {
  if (n <= 0) return LRA_OK;
  if (m < n || n > 65536) {
    lra_set_error("gesvj: shape %lld x %lld", (long long)m, (long long)n);
    return LRA_E_UNSUPPORTED;
  }
  double *Vw = nullptr, *sw = nullptr;
  int* rot = nullptr;
  LRA_CHECK(lra_wsT(c, WS_J0, (size_t)n * n + 2 * (size_t)n + 64, &Vw));
  sw = Vw + (size_t)n * n;
  rot = (int*)(sw + n);
  const int N = (int)(n + (n & 1));
  hipLaunchKernelGGL(jacobi_eye_kernel, dim3((unsigned)std::min<i64>((n * n + 255) / 256, 4096)),
                     dim3(256), 0, c->stream, (int)n, Vw, n, rot);
  LRA_HIP(hipGetLastError());
  const double tol = std::sqrt((double)m) * 1.1102230246251565e-16;
  int sw_done = 0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, c->stream, rot);
    for (int r = 0; r < N - 1; ++r)
      hipLaunchKernelGGL(jacobi_round_kernel, dim3(N / 2), dim3(256), 0, c->stream, (int)m,
                         (int)n, N, r, X, ldx, Vw, n, tol, rot);
    LRA_HIP(hipGetLastError());
    int h = 0;
    LRA_HIP(hipMemcpyAsync(&h, rot, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    LRA_HIP(hipStreamSynchronize(c->stream));
    sw_done = sweep + 1;
    if (h == 0) break;
  }
  if (sweeps) *sweeps = sw_done;
  hipLaunchKernelGGL(colnorm_kernel, dim3((unsigned)n), dim3(256), 0, c->stream, (int)m, X, ldx,
                     sw);
  LRA_HIP(hipGetLastError());
  std::vector<double> hs(n);
  LRA_HIP(hipMemcpyAsync(hs.data(), sw, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  LRA_HIP(hipStreamSynchronize(c->stream));
  std::vector<int> perm(n);
  for (i64 i = 0; i < n; ++i) perm[i] = (int)i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return hs[a] > hs[b]; });
  int* dperm = nullptr;
  LRA_CHECK(lra_wsT(c, WS_J1, (size_t)n, &dperm));
  LRA_HIP(hipMemcpyAsync(dperm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(jacobi_finish_kernel, dim3((unsigned)n), dim3(256), 0, c->stream, (int)m,
                     (int)n, dperm, X, ldx, Vw, n, sw, U, ldu, V, ldv, s);
  LRA_HIP(hipGetLastError());
  LRA_HIP(hipStreamSynchronize(c->stream));  // perm (host) must outlive the copy
  return LRA_OK;
}

extern "C" int lra_gesvj_f64(lra_ctx* c, int64_t m, int64_t n, double* X, int64_t ldx,
                             double* U, int64_t ldu, double* s, double* V, int64_t ldv) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0 && n <= m, 3);
  LRA_ARG(X != nullptr || n == 0, 4);
  LRA_ARG(ldx >= std::max<int64_t>(1, m), 5);
  LRA_ARG(U != nullptr || n == 0, 6);
  LRA_ARG(ldu >= std::max<int64_t>(1, m), 7);
  LRA_ARG(s != nullptr || n == 0, 8);
  LRA_ARG(V != nullptr || n == 0, 9);
  LRA_ARG(ldv >= std::max<int64_t>(1, n), 10);
  LRA_HIP(hipSetDevice(c->device));
  int sweeps = 0;
  return lra_gesvj_impl(c, m, n, X, ldx, U, ldu, s, V, ldv, &sweeps);
}
