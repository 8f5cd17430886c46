// factor.hip -- the partial factorizations built on the ID: pqrfact (src/pqr.jl:285-320),
// psvdfact/psvdvals (src/psvd.jl:238-308) and pheigfact/pheigvals (src/pheig.jl:276-364).
// Every step runs on the device: the ID (lra_idfact), the thin QR of the skeleton
// (lra_geqrf_impl / lra_orgqr_impl), the small cores on FP64 MFMA GEMMs and the one-sided
// Jacobi SVD (lra_gesvj_impl); only ranks and a few scalars come back to the host.
#include "lra_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

int lra_gesvj_impl(lra_ctx* c, i64 m, i64 n, double* X, i64 ldx, double* U, i64 ldu,
                   double* s, double* V, i64 ldv, int* sweeps);

namespace {

int grid_of(i64 work) { return (int)std::max<i64>(1, std::min<i64>((work + 255) / 256, 8192)); }

// D (k x n) = triu(S(0:k, 0:n))
__global__ void triu_copy_kernel(i64 k, i64 n, const double* __restrict__ S, i64 lds,
                                 double* __restrict__ D, i64 ldd) {
  const i64 total = k * n;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;
    D[i + j * ldd] = i <= j ? S[i + j * lds] : 0.0;
  }
}

// D (n x k) = S(0:k, 0:n)' with only S's upper triangle (i <= j) kept: D = triu(S)'
__global__ void triu_transpose_kernel(i64 k, i64 n, const double* __restrict__ S, i64 lds,
                                      double* __restrict__ D, i64 ldd) {
  const i64 total = k * n;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;  // S(i, j) -> D(j, i)
    D[j + i * ldd] = i <= j ? S[i + j * lds] : 0.0;
  }
}

// Y (N x k): row p[i]-1 of Y = column i of Z (k x N) -- the rows of (Z P')' for the
// permutation p (1-based), i.e. Y = P Z'.
__global__ void perm_rows_kernel(i64 k, i64 N, const double* __restrict__ Z, i64 ldz,
                                 const i64* __restrict__ p, double* __restrict__ Y, i64 ldy) {
  const i64 total = k * N;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 cidx = e % k, i = e / k;
    Y[(p[i] - 1) + cidx * ldy] = Z[cidx + i * ldz];
  }
}

// Z (k x N) = [I_k, T]  (the interpolation matrix in skeleton order)
__global__ void id_matrix_kernel(i64 k, i64 N, const double* __restrict__ T, i64 ldt,
                                 double* __restrict__ Z, i64 ldz) {
  const i64 total = k * N;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;
    Z[i + j * ldz] = j < k ? (i == j ? 1.0 : 0.0) : T[i + (j - k) * ldt];
  }
}

// D (k x k) = A[r, c] for 1-based index lists r, c
__global__ void gather2d_kernel(i64 k, const double* __restrict__ A, i64 lda,
                                const i64* __restrict__ r, const i64* __restrict__ cidx,
                                double* __restrict__ D, i64 ldd) {
  const i64 total = k * k;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;
    D[i + j * ldd] = A[(r[i] - 1) + (cidx[j] - 1) * lda];
  }
}

// S = Hermitian(B, :U) + mu I: S(i,j) = S(j,i) = (B(i,j) + B(j,i)) / 2 (hermitianize!,
// src/util.jl:36-49), shifted by mu on the diagonal
__global__ void herm_shift_kernel(i64 k, const double* __restrict__ B, i64 ldb, double mu,
                                  double* __restrict__ S, i64 lds) {
  const i64 total = k * k;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;
    const i64 a = std::min(i, j), b = std::max(i, j);
    double v = 0.5 * (B[a + b * ldb] + B[b + a * ldb]);
    if (i == j) v += mu;
    S[i + j * lds] = v;
  }
}

// sum of squares of a k x k matrix (one block)
__global__ __launch_bounds__(256) void frob2_kernel(i64 k, const double* __restrict__ B, i64 ldb,
                                                    double* out) {
  double a = 0.0;
  for (i64 e = threadIdx.x; e < k * k; e += 256) {
    const double v = B[(e % k) + (e / k) * ldb];
    a += v * v;
  }
  __shared__ double sa[4];
  a = lra_wave_sum(a);
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) *out = (sa[0] + sa[1]) + (sa[2] + sa[3]);
}

// D(:, j) = S(:, sel[j]) for j < nsel (column selection)
__global__ void select_cols_kernel(i64 m, i64 nsel, const int* __restrict__ sel,
                                   const double* __restrict__ S, i64 lds,
                                   double* __restrict__ D, i64 ldd) {
  const i64 total = m * nsel;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % m, j = e / m;
    D[i + j * ldd] = S[i + (i64)sel[j] * lds];
  }
}

// pheigorth! (src/pheig.jl:340-364) on one cluster [a, b] of columns of Vm (k rows):
// for i in a:b, j in i+1:b: v_j -= (v_i . v_j) v_i.  One workgroup, sequential over pairs.
__global__ __launch_bounds__(256) void cluster_orth_kernel(i64 k, int a, int b,
                                                           double* __restrict__ Vm, i64 ldv) {
  __shared__ double sred[4];
  for (int i = a; i <= b; ++i) {
    const double* vi = Vm + (i64)i * ldv;
    for (int j = i + 1; j <= b; ++j) {
      double* vj = Vm + (i64)j * ldv;
      double d = 0.0;
      for (i64 r = threadIdx.x; r < k; r += 256) d += vi[r] * vj[r];
      d = lra_wave_sum(d);
      if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = d;
      __syncthreads();
      d = (sred[0] + sred[1]) + (sred[2] + sred[3]);
      for (i64 r = threadIdx.x; r < k; r += 256) vj[r] -= d * vi[r];
      __syncthreads();
    }
  }
}

int gemm(lra_ctx* c, i64 M, i64 N, i64 K, double alpha, const double* A, i64 lda, int tA,
         const double* B, i64 ldb, int tB, double beta, double* C, i64 ldc) {
  GemmArgs g;
  g.M = M; g.N = N; g.K = K; g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.transA = tA; g.B = B; g.ldb = ldb; g.transB = tB;
  g.C = C; g.ldc = ldc; g.bgather = nullptr; g.prec = LRA_PREC_F64;
  return lra_gemm(c, g);
}

// The ID of op(A) with workspace-owned sk, rd (WS_F0, WS_F1) and T (library-owned).
struct IdOut {
  i64 k, N, P;
  i64 *sk, *rd, *p;  // p = [sk; rd] (WS_F2)
  double* T;
  i64 ldt;
  i64 rounds;
};

int run_id(lra_ctx* c, int trans, const LraSrc& src, const lra_opts* o, IdOut* r) {
  const i64 m = src.m, n = src.n;
  const i64 N = trans == LRA_TRANS_N ? n : m;
  r->N = N;
  r->P = trans == LRA_TRANS_N ? m : n;
  i64 *sk = nullptr, *rd = nullptr, *p = nullptr;
  LRA_CHECK(lra_wsT(c, WS_F0, (size_t)std::max<i64>(N, 1), &sk));
  LRA_CHECK(lra_wsT(c, WS_F1, (size_t)std::max<i64>(N, 1), &rd));
  LRA_CHECK(lra_wsT(c, WS_F2, (size_t)std::max<i64>(N, 1), &p));
  lra_id_out out{};
  out.sk = sk; out.rd = rd; out.T = nullptr; out.ldt = 0;
  LRA_CHECK(lra_idfact_src(c, trans, src, o, &out));
  r->k = out.k;
  r->sk = sk; r->rd = rd; r->p = p;
  r->T = out.T; r->ldt = out.ldt; r->rounds = out.rounds;
  if (r->k > 0)
    LRA_HIP(hipMemcpyAsync(p, sk, sizeof(i64) * r->k, hipMemcpyDeviceToDevice, c->stream));
  if (N > r->k)
    LRA_HIP(hipMemcpyAsync(p + r->k, rd, sizeof(i64) * (N - r->k), hipMemcpyDeviceToDevice,
                           c->stream));
  return LRA_OK;
}

// Thin QR of the skeleton columns: Cq (P x k) overwritten by the geqrf factors, tau (k),
// and (when Q != nullptr) the explicit Q (P x k).
int skeleton_qr(lra_ctx* c, int trans, const LraSrc& src, const IdOut& id, double* Cq,
                double* tau, double* Q, i64 ldq) {
  LRA_CHECK(lra_src_getcols(c, src, trans, id.sk, id.k, Cq, id.P));
  LRA_CHECK(lra_geqrf_impl(c, id.P, id.k, Cq, id.P, tau));
  if (Q) LRA_CHECK(lra_orgqr_impl(c, id.P, id.k, Cq, id.P, tau, Q, ldq));
  return LRA_OK;
}

// psvdrank (src/psvd.jl:302-308)
i64 psvdrank(const std::vector<double>& s, const lra_opts& o) {
  const i64 k = (i64)s.size();
  if (k == 0) return 0;
  const double ptol = std::max(o.atol, o.rtol * s[0]);
  for (i64 i = 1; i < k; ++i)
    if (s[i] <= ptol) return i;
  return k;
}

// pheigrank1 (src/pheig.jl:329-337) on w in the given order
i64 pheigrank1(const std::vector<double>& w, const lra_opts& o, double wmax) {
  i64 k = (i64)w.size();
  if (o.rank >= 0) k = std::min<i64>(o.rank, k);
  const double ptol = std::max(o.atol, o.rtol * wmax);
  for (i64 i = 1; i < k; ++i)
    if (std::fabs(w[i]) <= ptol) return i;
  return k;
}

template <class T>
int fetch(lra_ctx* c, const T* dev, std::vector<T>* host, i64 n) {
  host->resize(n);
  if (n > 0) LRA_HIP(hipMemcpyAsync(host->data(), dev, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
  LRA_HIP(hipStreamSynchronize(c->stream));
  return LRA_OK;
}

// the singular triplets of op(A) ~ Uc diag(s) Vc' through the ID (the m >= n branch of
// psvdfact applied to op(A)); s sorted decreasingly, all k of them (truncation by the caller)
struct SvdCore {
  i64 k, P, N;
  double *Q, *Q2, *Us, *W, *s;  // Q (P x k), Q2 (N x k), Us, W (k x k), s (k)
};

int svd_core(lra_ctx* c, int trans, const LraSrc& src, const lra_opts* o, bool vectors,
             SvdCore* out) {
  IdOut id;
  LRA_CHECK(run_id(c, trans, src, o, &id));
  const i64 k = id.k, P = id.P, N = id.N;
  out->k = k; out->P = P; out->N = N;
  if (k == 0) return LRA_OK;
  double *Cq = nullptr, *tau = nullptr, *Q = nullptr, *Z = nullptr, *Y = nullptr, *X = nullptr;
  LRA_CHECK(lra_wsT(c, WS_F3, (size_t)P * k, &Cq));
  LRA_CHECK(lra_wsT(c, WS_F4, (size_t)2 * k + 16, &tau));
  if (vectors) LRA_CHECK(lra_wsT(c, WS_F5, (size_t)P * k, &Q));
  LRA_CHECK(skeleton_qr(c, trans, src, id, Cq, tau, Q, P));
  // Z = R1 [I T] (k x N): R1 = triu(Cq(0:k, 0:k)), Z(:, k:N) = R1 T
  LRA_CHECK(lra_wsT(c, WS_F6, (size_t)k * N, &Z));
  hipLaunchKernelGGL(triu_copy_kernel, dim3(grid_of(k * k)), dim3(256), 0, c->stream, k, k, Cq,
                     P, Z, k);
  LRA_HIP(hipGetLastError());
  if (N > k) LRA_CHECK(gemm(c, k, N - k, k, 1.0, Z, k, 0, id.T, id.ldt, 0, 0.0, Z + k * k, k));
  // Y = P Z' (N x k): M' for M = R1 [I T] P'
  LRA_CHECK(lra_wsT(c, WS_F7, (size_t)N * k, &Y));
  hipLaunchKernelGGL(perm_rows_kernel, dim3(grid_of(k * N)), dim3(256), 0, c->stream, k, N, Z, k,
                     id.p, Y, N);
  LRA_HIP(hipGetLastError());
  // M' = Q2 R2; M = R2' Q2'
  double* tau2 = tau + k + 8;
  LRA_CHECK(lra_geqrf_impl(c, N, k, Y, N, tau2));
  double* Q2 = nullptr;
  if (vectors) {
    LRA_CHECK(lra_wsT(c, WS_F8, (size_t)N * k, &Q2));
    LRA_CHECK(lra_orgqr_impl(c, N, k, Y, N, tau2, Q2, N));
  }
  // X = R2' (k x k), then R2' = Us diag(s) W'
  double *Us = nullptr, *W = nullptr, *s = nullptr;
  LRA_CHECK(lra_wsT(c, WS_F9, (size_t)3 * k * k + k, &X));
  Us = X + k * k;
  W = Us + k * k;
  s = W + k * k;
  hipLaunchKernelGGL(triu_transpose_kernel, dim3(grid_of(k * k)), dim3(256), 0, c->stream, k, k,
                     Y, N, X, k);
  LRA_HIP(hipGetLastError());
  int sweeps = 0;
  LRA_CHECK(lra_gesvj_impl(c, k, k, X, k, Us, k, s, W, k, &sweeps));
  out->Q = Q; out->Q2 = Q2; out->Us = Us; out->W = W; out->s = s;
  return LRA_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// pqrfact (src/pqr.jl:285-320): the ID of op(A), then qr! of the skeleton columns:
// Q (P x k), R = [R1, R1 T] (k x N, pqrr :330-340), p = [sk; rd], T.  Which of Q, R, T are
// formed follows pqrfact_retval (bits LRA_RET_Q/R/T); sketch = none runs the rank-adaptive
// QRCP on op(A) itself (pqrfact_none -> pqrback_postproc, :323-327, :420-436).
static int pqrfact_src(lra_ctx* c, int trans, const LraSrc& src, const lra_opts* opts,
                       lra_pqr_out* out) {
  lra_opts o = *opts;
  if (src.op()) lra_opts_for_operator(&o);
  if (o.sketch == LRA_SKETCH_NONE && o.maxdet_tol >= 0) {
    lra_set_error("pqrfact: sketch = none with maxdet swaps is not supported");
    return LRA_E_UNSUPPORTED;
  }
  LRA_HIP(hipSetDevice(c->device));
  const i64 m = src.m, n = src.n;
  const double* A = src.A;
  const i64 lda = src.lda;
  const int ret = o.pqrfact_retval;
  const i64 N = trans == LRA_TRANS_N ? n : m, P = trans == LRA_TRANS_N ? m : n;
  out->k = 0;
  if (o.sketch == LRA_SKETCH_NONE) {
    // QRCP of op(A) itself; Q from its reflectors, R = triu(B(0:k, :)), T = R11 \ R12
    double* B = nullptr;
    i64 *jpvt = nullptr;
    double* tau = nullptr;
    LRA_CHECK(lra_wsT(c, WS_F3, (size_t)std::max<i64>(P * N, 1), &B));
    if (trans == LRA_TRANS_N) LRA_CHECK(lra_launch_copy2d(c->stream, m, n, A, lda, B, std::max<i64>(m, 1)));
    else LRA_CHECK(lra_launch_transpose(c->stream, m, n, A, lda, B, std::max<i64>(n, 1)));
    LRA_CHECK(lra_wsT(c, WS_F2, (size_t)std::max<i64>(N, 1), &jpvt));
    LRA_CHECK(lra_wsT(c, WS_F4, (size_t)std::max<i64>(std::min(P, N), 1), &tau));
    i64 k = 0;
    LRA_CHECK(lra_geqp3_adap_f64(c, P, N, B, std::max<i64>(P, 1), o.rank, o.atol, o.rtol, o.nb,
                                 jpvt, tau, &k));
    out->k = k;
    out->rounds = 1;
    if (!out->p) { i64* pp = nullptr; LRA_CHECK(lra_wsT(c, WS_F6, (size_t)std::max<i64>(N, 1), &pp)); out->p = pp; }
    if (N > 0) LRA_HIP(hipMemcpyAsync(out->p, jpvt, sizeof(i64) * N, hipMemcpyDeviceToDevice, c->stream));
    if ((ret & LRA_RET_Q) && k > 0) {
      if (!out->Q) { double* q = nullptr; LRA_CHECK(lra_wsT(c, WS_F5, (size_t)P * k, &q)); out->Q = q; out->ldq = std::max<i64>(P, 1); }
      LRA_CHECK(lra_orgqr_impl(c, P, k, B, std::max<i64>(P, 1), tau, out->Q, out->ldq));
    }
    if ((ret & LRA_RET_R) && k > 0) {
      if (!out->R) { double* r = nullptr; LRA_CHECK(lra_wsT(c, WS_F7, (size_t)k * N, &r)); out->R = r; out->ldr = k; }
      hipLaunchKernelGGL(triu_copy_kernel, dim3(grid_of(k * N)), dim3(256), 0, c->stream, k, N,
                         B, std::max<i64>(P, 1), out->R, out->ldr);
      LRA_HIP(hipGetLastError());
    }
    if ((ret & LRA_RET_T) && k > 0 && N > k) {
      if (!out->T) { double* t = nullptr; LRA_CHECK(lra_wsT(c, WS_F8, (size_t)k * (N - k), &t)); out->T = t; out->ldt = k; }
      LRA_CHECK(lra_id_T(c, k, N, B, std::max<i64>(P, 1), out->T, out->ldt));
    }
    return lra_check_device_error(c);
  }
  IdOut id;
  LRA_CHECK(run_id(c, trans, src, &o, &id));
  const i64 k = id.k;
  out->k = k;
  out->rounds = id.rounds;
  if (!out->p) { i64* pp = nullptr; LRA_CHECK(lra_wsT(c, WS_F6, (size_t)std::max<i64>(N, 1), &pp)); out->p = pp; }
  if (N > 0) LRA_HIP(hipMemcpyAsync(out->p, id.p, sizeof(i64) * N, hipMemcpyDeviceToDevice, c->stream));
  if (k == 0) return LRA_OK;
  if (ret & (LRA_RET_Q | LRA_RET_R)) {
    double *Cq = nullptr, *tau = nullptr;
    LRA_CHECK(lra_wsT(c, WS_F3, (size_t)P * k, &Cq));
    LRA_CHECK(lra_wsT(c, WS_F4, (size_t)k + 8, &tau));
    double* Q = nullptr;
    if (ret & LRA_RET_Q) {
      if (!out->Q) { double* q = nullptr; LRA_CHECK(lra_wsT(c, WS_F5, (size_t)P * k, &q)); out->Q = q; out->ldq = P; }
      Q = out->Q;
    }
    LRA_CHECK(skeleton_qr(c, trans, src, id, Cq, tau, Q, out->ldq));
    if (ret & LRA_RET_R) {
      if (!out->R) { double* r = nullptr; LRA_CHECK(lra_wsT(c, WS_F7, (size_t)k * N, &r)); out->R = r; out->ldr = k; }
      hipLaunchKernelGGL(triu_copy_kernel, dim3(grid_of(k * k)), dim3(256), 0, c->stream, k, k,
                         Cq, P, out->R, out->ldr);
      LRA_HIP(hipGetLastError());
      if (N > k)
        LRA_CHECK(gemm(c, k, N - k, k, 1.0, out->R, out->ldr, 0, id.T, id.ldt, 0, 0.0,
                       out->R + k * out->ldr, out->ldr));
    }
  }
  if ((ret & LRA_RET_T) && N > k) {
    if (!out->T) { out->T = id.T; out->ldt = id.ldt; }
    else LRA_CHECK(lra_launch_copy2d(c->stream, k, N - k, id.T, id.ldt, out->T, out->ldt));
  }
  return lra_check_device_error(c);
}

extern "C" int lra_pqrfact(lra_ctx* c, int trans, int64_t m, int64_t n, const double* A,
                           int64_t lda, const lra_opts* opts, lra_pqr_out* out) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(A != nullptr || m * n == 0, 5);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(out != nullptr, 8);
  return pqrfact_src(c, trans, lra_src_dense(m, n, A, lda), opts, out);
}

// pqrfact on an operator (src/pqr.jl:285-320 with A a LinearOperator): the skeleton columns
// come from one application to one-hot vectors; sketch none falls back to randn as chkopts!
// does for operators (src/LowRankApprox.jl:142-148).
extern "C" int lra_pqrfact_op(lra_ctx* c, int trans, int64_t m, int64_t n, lra_apply_fn apply,
                              void* user, int isherm, const lra_opts* opts, lra_pqr_out* out) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(apply != nullptr, 5);
  LRA_ARG(!isherm || m == n, 7);
  LRA_ARG(opts != nullptr, 8);
  LRA_ARG(out != nullptr, 9);
  return pqrfact_src(c, trans, lra_src_op(m, n, apply, user, isherm != 0), opts, out);
}

// ---------------------------------------------------------------------------------------
// psvdfact / psvdvals (src/psvd.jl:238-308)
static int psvd_impl(lra_ctx* c, const LraSrc& src, const lra_opts* opts, bool vectors,
                     lra_psvd_out* out) {
  LRA_HIP(hipSetDevice(c->device));
  const lra_opts o = *opts;
  const i64 m = src.m, n = src.n;
  const int trans = m >= n ? LRA_TRANS_N : LRA_TRANS_C;
  SvdCore sv{};
  LRA_CHECK(svd_core(c, trans, src, &o, vectors, &sv));
  out->k = 0;
  if (sv.k == 0) return LRA_OK;
  std::vector<double> hs;
  LRA_CHECK(fetch(c, sv.s, &hs, sv.k));
  const i64 kk = psvdrank(hs, o);
  out->k = kk;
  if (!out->S) { double* s = nullptr; LRA_CHECK(lra_wsT(c, WS_OP0, (size_t)sv.k, &s)); out->S = s; }
  LRA_HIP(hipMemcpyAsync(out->S, sv.s, sizeof(double) * kk, hipMemcpyDeviceToDevice, c->stream));
  if (!vectors) return lra_check_device_error(c);
  // op(A) ~ (Q Us) diag(s) (Q2 W)'
  const i64 k = sv.k, P = sv.P, N = sv.N;
  double *QU = nullptr, *QW = nullptr;
  LRA_CHECK(lra_wsT(c, WS_OP1, (size_t)P * kk, &QU));
  LRA_CHECK(lra_wsT(c, WS_CUR0, (size_t)N * kk, &QW));
  LRA_CHECK(gemm(c, P, kk, k, 1.0, sv.Q, P, 0, sv.Us, k, 0, 0.0, QU, P));
  LRA_CHECK(gemm(c, N, kk, k, 1.0, sv.Q2, N, 0, sv.W, k, 0, 0.0, QW, N));
  // trans n (m >= n): U = Q Us, Vt = (Q2 W)';  trans c: U = Q2 W, Vt = (Q Us)'
  const double* Uf = trans == LRA_TRANS_N ? QU : QW;
  const double* Vf = trans == LRA_TRANS_N ? QW : QU;
  const i64 ldUf = trans == LRA_TRANS_N ? P : N, ldVf = trans == LRA_TRANS_N ? N : P;
  if (!out->U) { double* u = nullptr; LRA_CHECK(lra_wsT(c, WS_MD0, (size_t)m * kk, &u)); out->U = u; out->ldu = std::max<i64>(m, 1); }
  if (!out->Vt) { double* v = nullptr; LRA_CHECK(lra_wsT(c, WS_MD1, (size_t)kk * n, &v)); out->Vt = v; out->ldvt = std::max<i64>(kk, 1); }
  LRA_CHECK(lra_launch_copy2d(c->stream, m, kk, Uf, ldUf, out->U, out->ldu));
  LRA_CHECK(lra_launch_transpose(c->stream, n, kk, Vf, ldVf, out->Vt, out->ldvt));
  return lra_check_device_error(c);
}

#define LRA_DENSE_ARGS(first)                          \
  LRA_ARG(c != nullptr, 1);                            \
  LRA_ARG(m >= 0, first);                              \
  LRA_ARG(n >= 0, first + 1);                          \
  LRA_ARG(A != nullptr || m * n == 0, first + 2);      \
  LRA_ARG(lda >= std::max<int64_t>(1, m), first + 3)

extern "C" int lra_psvdfact(lra_ctx* c, int64_t m, int64_t n, const double* A, int64_t lda,
                            const lra_opts* opts, lra_psvd_out* out) {
  LRA_DENSE_ARGS(2);
  LRA_ARG(opts != nullptr, 6);
  LRA_ARG(out != nullptr, 7);
  return psvd_impl(c, lra_src_dense(m, n, A, lda), opts, true, out);
}

extern "C" int lra_psvdvals(lra_ctx* c, int64_t m, int64_t n, const double* A, int64_t lda,
                            const lra_opts* opts, double* S, int64_t* k) {
  LRA_DENSE_ARGS(2);
  LRA_ARG(opts != nullptr, 6);
  LRA_ARG(k != nullptr, 8);
  lra_psvd_out o{};
  o.S = S;
  const int st = psvd_impl(c, lra_src_dense(m, n, A, lda), opts, false, &o);
  *k = o.k;
  return st;
}

// psvdfact / psvdvals on an operator (src/psvd.jl:238-290 with A a LinearOperator)
extern "C" int lra_psvdfact_op(lra_ctx* c, int64_t m, int64_t n, lra_apply_fn apply,
                               void* user, int isherm, const lra_opts* opts,
                               lra_psvd_out* out) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(apply != nullptr, 4);
  LRA_ARG(!isherm || m == n, 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(out != nullptr, 8);
  lra_opts o = *opts;
  lra_opts_for_operator(&o);
  return psvd_impl(c, lra_src_op(m, n, apply, user, isherm != 0), &o, true, out);
}

extern "C" int lra_psvdvals_op(lra_ctx* c, int64_t m, int64_t n, lra_apply_fn apply,
                               void* user, int isherm, const lra_opts* opts, double* S,
                               int64_t* k) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(apply != nullptr, 4);
  LRA_ARG(!isherm || m == n, 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(k != nullptr, 9);
  lra_opts o = *opts;
  lra_opts_for_operator(&o);
  lra_psvd_out r{};
  r.S = S;
  const int st = psvd_impl(c, lra_src_op(m, n, apply, user, isherm != 0), &o, false, &r);
  *k = r.k;
  return st;
}

// ---------------------------------------------------------------------------------------
// pheigfact / pheigvals (src/pheig.jl:276-364): Hermitian A (n x n).  V = idfact(A); QR of
// the interpolation matrix V' = P [I T]' (n x k); B = R A[sk,sk] R'; eigenpairs of
// Hermitian(B) by one-sided Jacobi on the positive definite Hermitian(B) + mu I (mu = |B|_F,
// so the singular pairs are the eigenpairs shifted by mu); pheigrank selection, the cluster
// re-orthogonalisation pheigorth!, and vectors = Q * vectors.
static int pheig_impl(lra_ctx* c, const LraSrc& src, const lra_opts* opts, bool vectors,
                      lra_pheig_out* out) {
  LRA_HIP(hipSetDevice(c->device));
  const lra_opts o = *opts;
  const i64 n = src.n;
  IdOut id;
  LRA_CHECK(run_id(c, LRA_TRANS_N, src, &o, &id));
  const i64 k = id.k, N = n;
  out->k = out->kn = out->kp = 0;
  if (k == 0) return LRA_OK;
  // Y = V' (n x k), its QR
  double *Z = nullptr, *Y = nullptr, *tau = nullptr, *Q = nullptr;
  LRA_CHECK(lra_wsT(c, WS_F6, (size_t)k * N, &Z));
  LRA_CHECK(lra_wsT(c, WS_F7, (size_t)N * k, &Y));
  LRA_CHECK(lra_wsT(c, WS_F4, (size_t)k + 8, &tau));
  hipLaunchKernelGGL(id_matrix_kernel, dim3(grid_of(k * N)), dim3(256), 0, c->stream, k, N, id.T,
                     id.ldt, Z, k);
  LRA_HIP(hipGetLastError());
  hipLaunchKernelGGL(perm_rows_kernel, dim3(grid_of(k * N)), dim3(256), 0, c->stream, k, N, Z, k,
                     id.p, Y, N);
  LRA_HIP(hipGetLastError());
  LRA_CHECK(lra_geqrf_impl(c, N, k, Y, N, tau));
  if (vectors) {
    LRA_CHECK(lra_wsT(c, WS_F5, (size_t)N * k, &Q));
    LRA_CHECK(lra_orgqr_impl(c, N, k, Y, N, tau, Q, N));
  }
  // B = R (A[sk,sk] R') with R = triu(Y(0:k, 0:k))
  double *R = nullptr, *Ask = nullptr, *W1 = nullptr, *B = nullptr, *S = nullptr;
  LRA_CHECK(lra_wsT(c, WS_F9, (size_t)8 * k * k + 2 * k + 64, &R));
  Ask = R + k * k; W1 = Ask + k * k; B = W1 + k * k; S = B + k * k;
  double* Uj = S + k * k;
  double* Vj = Uj + k * k;
  double* sj = Vj + k * k;
  double* scratch = sj + k;
  hipLaunchKernelGGL(triu_copy_kernel, dim3(grid_of(k * k)), dim3(256), 0, c->stream, k, k, Y, N,
                     R, k);
  LRA_HIP(hipGetLastError());
  LRA_CHECK(lra_src_principal(c, src, id.sk, k, Ask, k));
  LRA_CHECK(gemm(c, k, k, k, 1.0, Ask, k, 0, R, k, 1, 0.0, W1, k));
  LRA_CHECK(gemm(c, k, k, k, 1.0, R, k, 0, W1, k, 0, 0.0, B, k));
  hipLaunchKernelGGL(frob2_kernel, dim3(1), dim3(256), 0, c->stream, k, B, k, scratch);
  LRA_HIP(hipGetLastError());
  std::vector<double> hf;
  LRA_CHECK(fetch(c, scratch, &hf, 1));
  const double mu = std::sqrt(hf[0]) * (1.0 + 1e-3) + 1e-300;
  hipLaunchKernelGGL(herm_shift_kernel, dim3(grid_of(k * k)), dim3(256), 0, c->stream, k, B, k,
                     mu, S, k);
  LRA_HIP(hipGetLastError());
  int sweeps = 0;
  LRA_CHECK(lra_gesvj_impl(c, k, k, S, k, Uj, k, sj, Vj, k, &sweeps));
  std::vector<double> hs;
  LRA_CHECK(fetch(c, sj, &hs, k));
  // eigenvalues ascending: the decreasing singular values reversed, minus the shift
  std::vector<double> w(k);
  std::vector<int> order(k);
  for (i64 i = 0; i < k; ++i) {
    w[i] = hs[k - 1 - i] - mu;
    order[i] = (int)(k - 1 - i);
  }
  // pheigrank (src/pheig.jl:316-327)
  const double wmax = std::max(std::fabs(w[0]), std::fabs(w[k - 1]));
  i64 first = 0;  // first index with w >= 0 (searchsorted(w, 0))
  while (first < k && w[first] < 0.0) ++first;
  i64 last = first;  // one past the last index with w <= 0
  while (last < k && w[last] <= 0.0) ++last;
  std::vector<double> neg(w.begin(), w.begin() + first), pos;
  for (i64 i = k - 1; i >= last; --i) pos.push_back(w[i]);
  const i64 kn = pheigrank1(neg, o, wmax), kp = pheigrank1(pos, o, wmax);
  std::vector<int> sel;
  std::vector<double> vals;
  if (kn + kp < k) {
    for (i64 i = 0; i < kn; ++i) sel.push_back((int)i);
    for (i64 i = k - kp; i < k; ++i) sel.push_back((int)i);
  } else {
    for (i64 i = 0; i < k; ++i) sel.push_back((int)i);
  }
  for (int i : sel) vals.push_back(w[i]);
  const i64 ks = (i64)sel.size();
  out->k = ks; out->kn = kn; out->kp = kp;
  if (!out->values) { double* v = nullptr; LRA_CHECK(lra_wsT(c, WS_OP0, (size_t)ks + 1, &v)); out->values = v; }
  LRA_HIP(hipMemcpyAsync(out->values, vals.data(), sizeof(double) * ks, hipMemcpyHostToDevice, c->stream));
  if (vectors && ks > 0) {
    // selected eigenvectors (columns of the Jacobi V in ascending-eigenvalue order)
    std::vector<int> cols(ks);
    for (i64 j = 0; j < ks; ++j) cols[j] = order[sel[j]];
    int* dsel = nullptr;
    double* Ev = nullptr;
    LRA_CHECK(lra_wsT(c, WS_J1, (size_t)ks + 1, &dsel));
    LRA_CHECK(lra_wsT(c, WS_OP1, (size_t)k * ks, &Ev));
    LRA_HIP(hipMemcpyAsync(dsel, cols.data(), sizeof(int) * ks, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(select_cols_kernel, dim3(grid_of(k * ks)), dim3(256), 0, c->stream, k, ks,
                       dsel, Vj, k, Ev, k);
    LRA_HIP(hipGetLastError());
    // pheigorth!: clusters of consecutive values with symrelerr <= pheig_orthtol
    for (i64 a = 0; a < ks;) {
      i64 b = a + 1;
      // (a NaN relative error, 0 vs 0, continues the cluster as Julia's `>` test does)
      while (b < ks && !(2.0 * std::fabs((vals[a] - vals[b]) / (vals[a] + vals[b])) > o.pheig_orthtol)) ++b;
      --b;
      if (b > a)
        hipLaunchKernelGGL(cluster_orth_kernel, dim3(1), dim3(256), 0, c->stream, k, (int)a,
                           (int)b, Ev, k);
      a = b + 1;
    }
    LRA_HIP(hipGetLastError());
    if (!out->vectors) { double* v = nullptr; LRA_CHECK(lra_wsT(c, WS_MD0, (size_t)N * ks, &v)); out->vectors = v; out->ldv = N; }
    LRA_CHECK(gemm(c, N, ks, k, 1.0, Q, N, 0, Ev, k, 0, 0.0, out->vectors, out->ldv));
  }
  LRA_HIP(hipStreamSynchronize(c->stream));  // host value/selection arrays go out of scope
  return lra_check_device_error(c);
}

extern "C" int lra_pheigfact(lra_ctx* c, int64_t n, const double* A, int64_t lda,
                             const lra_opts* opts, lra_pheig_out* out) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(n >= 0, 2);
  LRA_ARG(A != nullptr || n == 0, 3);
  LRA_ARG(lda >= std::max<int64_t>(1, n), 4);
  LRA_ARG(opts != nullptr, 5);
  LRA_ARG(out != nullptr, 6);
  return pheig_impl(c, lra_src_dense(n, n, A, lda), opts, true, out);
}

extern "C" int lra_pheigvals(lra_ctx* c, int64_t n, const double* A, int64_t lda,
                             const lra_opts* opts, double* values, int64_t* k) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(n >= 0, 2);
  LRA_ARG(A != nullptr || n == 0, 3);
  LRA_ARG(lda >= std::max<int64_t>(1, n), 4);
  LRA_ARG(opts != nullptr, 5);
  LRA_ARG(k != nullptr, 7);
  lra_pheig_out o{};
  o.values = values;
  const int st = pheig_impl(c, lra_src_dense(n, n, A, lda), opts, false, &o);
  *k = o.k;
  return st;
}

// pheigfact / pheigvals on a Hermitian operator (src/pheig.jl:276-314 with A a
// HermitianLinearOperator): A[sk, sk] comes from one application to one-hot vectors
// (src/linop.jl:70-74).  The caller guarantees that the operator is Hermitian.
extern "C" int lra_pheigfact_op(lra_ctx* c, int64_t n, lra_apply_fn apply, void* user,
                                const lra_opts* opts, lra_pheig_out* out) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(n >= 0, 2);
  LRA_ARG(apply != nullptr, 3);
  LRA_ARG(opts != nullptr, 5);
  LRA_ARG(out != nullptr, 6);
  lra_opts o = *opts;
  lra_opts_for_operator(&o);
  return pheig_impl(c, lra_src_op(n, n, apply, user, true), &o, true, out);
}

extern "C" int lra_pheigvals_op(lra_ctx* c, int64_t n, lra_apply_fn apply, void* user,
                                const lra_opts* opts, double* values, int64_t* k) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(n >= 0, 2);
  LRA_ARG(apply != nullptr, 3);
  LRA_ARG(opts != nullptr, 5);
  LRA_ARG(k != nullptr, 7);
  lra_opts o = *opts;
  lra_opts_for_operator(&o);
  lra_pheig_out r{};
  r.values = values;
  const int st = pheig_impl(c, lra_src_op(n, n, apply, user, true), &o, false, &r);
  *k = r.k;
  return st;
}

// D (k x k) = A[idx, idx]: a 2-D gather for a dense source, the operator applied to k one-hot
// vectors and the rows idx of the result kept (getindex(L, rows, cols), src/linop.jl:70-74)
int lra_src_principal(lra_ctx* c, const LraSrc& s, const i64* idx, i64 k, double* D, i64 ldd) {
  if (k <= 0) return LRA_OK;
  if (!s.op()) {
    hipLaunchKernelGGL(gather2d_kernel, dim3(grid_of(k * k)), dim3(256), 0, c->stream, k, s.A,
                       s.lda, idx, idx, D, ldd);
    LRA_HIP(hipGetLastError());
    return LRA_OK;
  }
  double* C = nullptr;
  LRA_CHECK(lra_wsT(c, WS_OP1, (size_t)std::max<i64>(s.m * k, 1), &C));
  LRA_CHECK(lra_src_getcols(c, s, LRA_TRANS_N, idx, k, C, std::max<i64>(s.m, 1)));
  return lra_launch_gather_rows(c->stream, k, k, C, std::max<i64>(s.m, 1), idx, D, ldd);
}
