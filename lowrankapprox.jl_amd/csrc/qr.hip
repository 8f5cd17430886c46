// qr.hip -- Householder QR of a tall matrix and the explicit Q (the qr! / LAPACK.orgqr! calls
// of pqrfact, psvdfact and pheigfact: src/pqr.jl:298, src/psvd.jl:244,258, src/pheig.jl:282).
//
// Blocked right-looking Householder QR with LAPACK's reflector convention (dlarfg), so R's
// diagonal signs are dgeqrf's:
//   panel (nb <= 32 columns, all remaining rows): a cooperative grid keeps the panel in
//     registers, one row per thread (rows strided over the block for coalescing); every
//     column needs one grid-wide all-reduction (the column norm, the pivot entry, and for
//     each column to its right the dot product with this column below the pivot and its
//     pivot-row entry, from which v'A follows), done with tagged 16-byte chunks in uncached
//     memory and summed in block order, so every block holds bit-identical reflectors;
//   trailing update: compact WY (I - V T V'), T from V'V (dlarft), with FP64 MFMA GEMMs:
//     W = V' X2, W = T' W, X2 -= V W.
// orgqr applies the blocks in reverse to [I; 0].
#include "lra_internal.h"

#include <algorithm>

namespace {

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
constexpr int QR_NB = 32;
constexpr int QR_NV = 2 * QR_NB;  // values of one column's exchange (2 + 2 (nb - 1) <= 64)

struct PanelDev {
  double* A;
  i64 lda;
  int m, j0, nb;  // panel rows j0..m-1, columns j0..j0+nb-1
  int G;          // blocks; block g holds rows j0 + g*256*RPR ...
  double* tau;    // tau[j0 + c]
  u64x2* xchg;    // [2][G][QR_NV + 2] chunks
  uint32_t gen;
  int uncached;
  int* err;
  int* abort_flag;
};

LRA_DEV double lapy2_qr(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

LRA_DEV void qr_raise(const PanelDev& p, int code) {
  __hip_atomic_store(p.err, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// read chunk until its tag matches (bounded)
LRA_DEV double qr_read(const PanelDev& p, const u64x2* src, uint64_t tag, bool* ok) {
  const volatile u64x2* v = (const volatile u64x2*)src;
  u64x2 w = *v;
  unsigned spins = 0;
  while (w.y != tag) {
    if ((++spins & 63) == 0 &&
        (spins > LRA_SPIN_LIMIT ||
         __hip_atomic_load(p.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      *ok = false;
      return 0.0;
    }
    __builtin_amdgcn_s_sleep(1);
    if (!p.uncached) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    w = *v;
  }
  return __longlong_as_double((long long)w.x);
}

LRA_DEV void qr_put(const PanelDev& p, u64x2* dst, double val, uint64_t tag) {
  u64x2 w;
  w.x = (uint64_t)__double_as_longlong(val);
  w.y = tag;
  if (p.uncached) *dst = w;
  else *(volatile u64x2*)dst = w;
}

template <int RPR>
__global__ __launch_bounds__(256) void geqr2_coop_kernel(PanelDev p) {
  constexpr int NB = QR_NB;
  constexpr int NV = QR_NV;
  const int tid = threadIdx.x;
  const int g = blockIdx.x;
  const int nb = p.nb;
  __shared__ double s_w[NV][257];  // per-value thread partials, transposed
  __shared__ double s_red[NV];
  __shared__ int s_ok;

  // rows held: r(i) = j0 + g*256*RPR + i*256 + tid
  double x[RPR][NB];
  int row[RPR];
#pragma unroll
  for (int i = 0; i < RPR; ++i) {
    row[i] = p.j0 + g * 256 * RPR + i * 256 + tid;
    const bool in = row[i] < p.m;
#pragma unroll
    for (int j = 0; j < NB; ++j)
      x[i][j] = (in && j < nb) ? p.A[row[i] + (i64)(p.j0 + j) * p.lda] : 0.0;
  }

  int seq = 0;
  // Grid all-reduction of the nv (<= NV) values each thread left in s_w[0..nv)[tid]: the
  // block sums each value over its 256 threads (4 threads per value, 64 partials each, then a
  // fixed butterfly), publishes the block partials as tagged chunks, and every block adds the
  // G partials of each value in a fixed order (4 interleaved strands per value, combined by
  // the same butterfly).  The result, bit-identical in every block, is left in s_red.
  auto allreduce = [&](int nv) {
    __syncthreads();
    const int j = tid >> 2, q = tid & 3;
    const uint64_t tag = ((uint64_t)p.gen << 32) | (uint64_t)(seq + 1);
    u64x2* region = p.xchg + (size_t)(seq & 1) * p.G * (NV + 2);
    double bp = 0.0;
    if (j < nv) {
#pragma unroll 8
      for (int r = 0; r < 64; ++r) bp += s_w[j][q * 64 + r];
    }
    bp += __shfl_xor(bp, 1, 64);
    bp += __shfl_xor(bp, 2, 64);
    if (j < nv && q == 0) qr_put(p, region + (size_t)g * (NV + 2) + j, bp, tag);
    if (!p.uncached) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    double sum = 0.0;
    bool lok = true;
    if (j < nv) {
      // strand q: blocks q, q+4, q+8, ... in ascending order, loads issued in batches
      for (int b0 = q; b0 < p.G; b0 += 32) {
        u64x2 t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int b = b0 + 4 * u;
          const u64x2* src = region + (size_t)b * (NV + 2) + j;
          if (b < p.G) t[u] = p.uncached ? *src : *(const volatile u64x2*)src;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int b = b0 + 4 * u;
          if (b >= p.G) continue;
          double val = __longlong_as_double((long long)t[u].x);
          if (t[u].y != tag) val = qr_read(p, region + (size_t)b * (NV + 2) + j, tag, &lok);
          sum += val;
        }
      }
    }
    sum += __shfl_xor(sum, 1, 64);
    sum += __shfl_xor(sum, 2, 64);
    if (!lok) {
      qr_raise(p, LRA_E_TIMEOUT);
      s_ok = 0;
    }
    if (j < nv && q == 0) s_red[j] = sum;
    __syncthreads();
    ++seq;
  };
  if (tid == 0) s_ok = 1;
  __syncthreads();

  for (int c = 0; c < nb; ++c) {
    const int rc = p.j0 + c;
    const int nq = nb - c - 1;  // panel columns right of c
    // ---- one exchange per column: the norm below the pivot row, the pivot entry, and for
    // every column j right of c the dot product with column c below the pivot (d_j) and the
    // pivot-row entry (p_j); v'A(:, j) = p_j + d_j / (alpha - beta) follows from them.
    {
      // column c below the pivot (zero elsewhere) and the pivot-row indicator, per row held
      double xb[RPR], pr[RPR];
      double ss = 0.0, al = 0.0;
#pragma unroll
      for (int i = 0; i < RPR; ++i) {
        double xc = 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j) xc = (j == c) ? x[i][j] : xc;
        const bool below = row[i] > rc && row[i] < p.m;
        xb[i] = below ? xc : 0.0;
        pr[i] = row[i] == rc ? 1.0 : 0.0;
        ss += xb[i] * xb[i];
        if (row[i] == rc) al = xc;
      }
      s_w[0][tid] = ss;
      s_w[1][tid] = al;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j <= c || j >= nb) continue;
        double d = 0.0, pj = 0.0;
#pragma unroll
        for (int i = 0; i < RPR; ++i) {
          d += xb[i] * x[i][j];
          pj += pr[i] * x[i][j];
        }
        s_w[2 + 2 * (j - c - 1)][tid] = d;
        s_w[3 + 2 * (j - c - 1)][tid] = pj;
      }
    }
    allreduce(2 + 2 * nq);
    if (!s_ok) return;
    const double ss = s_red[0], alpha0 = s_red[1];
    // ---- dlarfg (identical in every block)
    double alpha = alpha0, beta = alpha0, tau = 0.0, scal = 1.0, fac = 1.0;
    const int len = p.m - rc;
    double xnorm = sqrt(ss);
    if (len > 1 && xnorm != 0.0) {
      beta = -copysign(lapy2_qr(alpha, xnorm), alpha);
      const double safmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;
      int knt = 0;
      if (fabs(beta) < safmin) {
        const double rsafmn = 1.0 / safmin;  // power of two: exact rescaling
        do {
          ++knt;
          fac *= rsafmn;
          beta *= rsafmn;
          alpha *= rsafmn;
          xnorm *= rsafmn;
        } while (fabs(beta) < safmin && knt < 20);
        beta = -copysign(lapy2_qr(alpha, xnorm), alpha);
      }
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
      for (int j = 0; j < knt; ++j) beta *= safmin;
    }
    if (g == 0 && tid == 0) p.tau[rc] = tau;
    // v below the pivot, beta on the pivot row
    double v[RPR];
#pragma unroll
    for (int i = 0; i < RPR; ++i) {
      double xc = 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) xc = (j == c) ? x[i][j] : xc;
      double vi = 0.0;
      if (row[i] > rc && row[i] < p.m) vi = (xc * fac) * scal;
      else if (row[i] == rc) vi = 1.0;
      v[i] = vi;
      const double nx = (row[i] == rc) ? beta : ((row[i] > rc) ? vi : xc);
#pragma unroll
      for (int j = 0; j < NB; ++j) x[i][j] = (j == c) ? nx : x[i][j];
    }
    if (nq == 0) break;
    // ---- A(:, j) -= tau (v'A(:, j)) v for the panel columns right of c
    const double fs = fac * scal;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j > c && j < nb) {
        const double w = s_red[3 + 2 * (j - c - 1)] + fs * s_red[2 + 2 * (j - c - 1)];
        const double tw = tau * w;
#pragma unroll
        for (int i = 0; i < RPR; ++i) x[i][j] -= tw * v[i];
      }
    }
    __syncthreads();  // s_red is rewritten by the next exchange
  }
#pragma unroll
  for (int i = 0; i < RPR; ++i) {
    if (row[i] >= p.m) continue;
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < nb) p.A[row[i] + (i64)(p.j0 + j) * p.lda] = x[i][j];
  }
}

// V (mp x nb, ld mp) = unit lower trapezoid of the factored panel A(j0:m, j0:j0+nb)
__global__ void extract_v_kernel(int mp, int nb, const double* __restrict__ A, i64 lda,
                                 double* __restrict__ V) {
  const i64 total = (i64)mp * nb;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const int r = (int)(e % mp), c = (int)(e / mp);
    V[e] = r > c ? A[r + (i64)c * lda] : (r == c ? 1.0 : 0.0);
  }
}

// dlarft (forward, columnwise): T (nb x nb, ld nb) upper triangular with
// I - V T V' = H_1 ... H_nb, from S = V'V and tau.  dlarft's recurrence T(0:c, c) =
// -tau_c T(0:c, 0:c) S(0:c, c) runs column after column; the same T is the inverse of
// U = diag(1/tau) + triu(S, 1), so T(c, c) = tau_c and, for i < c,
// T(i, c) = -tau_i sum_{l > i} S(i, l) T(l, c) (also with tau_i = 0: H_i = I, row i zero).
// Lane c back-substitutes column c on its own: one wave, S in LDS (broadcast reads), no
// barrier between columns.
__global__ __launch_bounds__(64) void larft_kernel(int nb, const double* __restrict__ S,
                                                   const double* __restrict__ tau,
                                                   double* __restrict__ T) {
  __shared__ double sS[QR_NB][QR_NB];  // S(i, l), zero outside nb x nb
  __shared__ double st[QR_NB];
  const int c = threadIdx.x;
  double sv[QR_NB * QR_NB / 64];
#pragma unroll
  for (int u = 0; u < QR_NB * QR_NB / 64; ++u) {
    const int e = c + 64 * u, i = e % QR_NB, l = e / QR_NB;
    const bool ok = i < nb && l < nb;
    const double x = S[ok ? i + l * nb : 0];
    sv[u] = ok ? x : 0.0;
  }
  const double tv = c < nb ? tau[c] : 0.0;
#pragma unroll
  for (int u = 0; u < QR_NB * QR_NB / 64; ++u) {
    const int e = c + 64 * u;
    sS[e % QR_NB][e / QR_NB] = sv[u];
  }
  if (c < QR_NB) st[c] = tv;
  __syncthreads();
  double x[QR_NB];
#pragma unroll
  for (int i = QR_NB - 1; i >= 0; --i) {
    double s = 0.0;
#pragma unroll
    for (int l = i + 1; l < QR_NB; ++l) s += sS[i][l] * x[l];
    x[i] = i < c ? -st[i] * s : (i == c ? st[i] : 0.0);
  }
  if (c < nb) {
#pragma unroll
    for (int i = 0; i < QR_NB; ++i)
      if (i < nb) T[i + c * nb] = x[i];
  }
}

// C (m x n) = [I; 0] (the first n columns of the m x m identity)
__global__ void eye_kernel(i64 m, i64 n, double* C, i64 ldc) {
  const i64 total = m * n;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 r = e % m, c = e / m;
    C[r + c * ldc] = r == c ? 1.0 : 0.0;
  }
}

int grid_of(i64 work) { return (int)std::max<i64>(1, std::min<i64>((work + 255) / 256, 8192)); }

template <int RPR>
hipError_t launch_panel(const PanelDev& d, hipStream_t st) {
  void* args[] = {(void*)&d};
  return hipLaunchCooperativeKernel((const void*)geqr2_coop_kernel<RPR>, dim3(d.G), dim3(256),
                                    args, 0, st);
}

// one panel: rows j0..m-1, columns j0..j0+nb-1 of A
// `first`: the first panel of a factorization clears the abort word (a raised abort ends the
// factorization with an error, so later panels find it clear)
int factor_panel(lra_ctx* c, i64 m, i64 j0, int nb, double* A, i64 lda, double* tau,
                 bool first) {
  const i64 mp = m - j0;
  int rpr = 1;
  while (rpr < 2 && (mp + 256 * rpr - 1) / (256 * rpr) > 256) rpr *= 2;
  const i64 G = (mp + 256 * rpr - 1) / (256 * rpr);
  if (G > 256) {
    lra_set_error("geqrf: %lld rows exceed the panel kernel", (long long)m);
    return LRA_E_UNSUPPORTED;
  }
  const size_t bytes = 256 + 2 * (size_t)G * (QR_NV + 2) * 16;
  void* area = nullptr;
  bool uc = false;
  LRA_CHECK(lra_coop_area(c, bytes, &area, &uc));
  PanelDev d;
  d.A = A; d.lda = lda; d.m = (int)m; d.j0 = (int)j0; d.nb = nb; d.G = (int)G; d.tau = tau;
  d.abort_flag = (int*)area;
  d.xchg = (u64x2*)((char*)area + 256);
  d.gen = lra_next_qrcp_gen();  // chunk tags of this launch (shared counter with the QRCP)
  d.uncached = uc ? 1 : 0;
  d.err = c->d_err;
  if (first) LRA_CHECK(lra_launch_zero_u32(c->stream, (unsigned*)area, 256 / 4));
  hipError_t e;
  {
    SpinSection gate(c);
    e = rpr == 1 ? launch_panel<1>(d, c->stream) : launch_panel<2>(d, c->stream);
  }
  if (e != hipSuccess) {
    lra_set_error("geqrf panel launch (m=%lld G=%lld): %s", (long long)m, (long long)G,
                  hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
}

// V and T of the panel starting at column j0 (nb columns) into Vw (mp x nb) and Tw (nb x nb)
int panel_wy(lra_ctx* c, i64 m, i64 j0, int nb, const double* A, i64 lda, const double* tau,
             double* Vw, double* Sw, double* Tw) {
  const i64 mp = m - j0;
  hipLaunchKernelGGL(extract_v_kernel, dim3(grid_of(mp * nb)), dim3(256), 0, c->stream, (int)mp,
                     nb, A + j0 + j0 * lda, lda, Vw);
  LRA_HIP(hipGetLastError());
  GemmArgs g;
  g.M = nb; g.N = nb; g.K = mp; g.alpha = 1.0; g.beta = 0.0;
  g.A = Vw; g.lda = mp; g.transA = 1; g.B = Vw; g.ldb = mp; g.transB = 0;
  g.C = Sw; g.ldc = nb; g.bgather = nullptr; g.prec = LRA_PREC_F64;
  LRA_CHECK(lra_gemm(c, g));
  hipLaunchKernelGGL(larft_kernel, dim3(1), dim3(64), 0, c->stream, nb, Sw, tau + j0, Tw);
  LRA_HIP(hipGetLastError());
  return LRA_OK;
}

// X (mp x n2, ldx) = (I - V T' V')' ... op: trans = 1 applies Q' = (I - V T V')' = I - V T' V'
// (geqrf's trailing update), trans = 0 applies Q = I - V T V' (orgqr).
int apply_block(lra_ctx* c, i64 mp, i64 n2, int nb, const double* Vw, const double* Tw,
                double* X, i64 ldx, int trans, double* W1, double* W2) {
  GemmArgs g;
  g.bgather = nullptr; g.prec = LRA_PREC_F64;
  // W1 = V' X  (nb x n2)
  g.M = nb; g.N = n2; g.K = mp; g.alpha = 1.0; g.beta = 0.0;
  g.A = Vw; g.lda = mp; g.transA = 1; g.B = X; g.ldb = ldx; g.transB = 0;
  g.C = W1; g.ldc = nb;
  LRA_CHECK(lra_gemm(c, g));
  // W2 = op(T) W1
  g.M = nb; g.N = n2; g.K = nb; g.alpha = 1.0; g.beta = 0.0;
  g.A = Tw; g.lda = nb; g.transA = trans; g.B = W1; g.ldb = nb; g.transB = 0;
  g.C = W2; g.ldc = nb;
  LRA_CHECK(lra_gemm(c, g));
  // X -= V W2
  g.M = mp; g.N = n2; g.K = nb; g.alpha = -1.0; g.beta = 1.0;
  g.A = Vw; g.lda = mp; g.transA = 0; g.B = W2; g.ldb = nb; g.transB = 0;
  g.C = X; g.ldc = ldx;
  LRA_CHECK(lra_gemm(c, g));
  return LRA_OK;
}

}  // namespace

// Householder QR of A (m x n, m >= n), in place in LAPACK's geqrf layout (R on and above
// the diagonal, the reflectors below it, tau[0..n)).
int lra_geqrf_impl(lra_ctx* c, i64 m, i64 n, double* A, i64 lda, double* tau) {
  if (n <= 0 || m <= 0) return LRA_OK;
  if (m < n) {
    lra_set_error("geqrf: m=%lld < n=%lld", (long long)m, (long long)n);
    return LRA_E_UNSUPPORTED;
  }
  double *Vw = nullptr, *Sw = nullptr, *Tw = nullptr, *W1 = nullptr, *W2 = nullptr;
  LRA_CHECK(lra_wsT(c, WS_QR0, (size_t)m * QR_NB, &Vw));
  LRA_CHECK(lra_wsT(c, WS_QR1, (size_t)2 * QR_NB * QR_NB, &Sw));
  Tw = Sw + QR_NB * QR_NB;
  LRA_CHECK(lra_wsT(c, WS_QR2, (size_t)2 * QR_NB * std::max<i64>(n, 1), &W1));
  W2 = W1 + (size_t)QR_NB * std::max<i64>(n, 1);
  for (i64 j0 = 0; j0 < n; j0 += QR_NB) {
    const int nb = (int)std::min<i64>(QR_NB, n - j0);
    LRA_CHECK(factor_panel(c, m, j0, nb, A, lda, tau, j0 == 0));
    const i64 n2 = n - j0 - nb;
    if (n2 > 0) {
      LRA_CHECK(panel_wy(c, m, j0, nb, A, lda, tau, Vw, Sw, Tw));
      LRA_CHECK(apply_block(c, m - j0, n2, nb, Vw, Tw, A + j0 + (j0 + nb) * lda, lda, 1, W1, W2));
    }
  }
  return lra_check_device_error(c);
}

// Q (m x k, ldq) = the first k columns of H_1 ... H_k from a geqrf factorization in A.
int lra_orgqr_impl(lra_ctx* c, i64 m, i64 k, const double* A, i64 lda, const double* tau,
                   double* Q, i64 ldq) {
  if (m <= 0 || k <= 0) return LRA_OK;
  hipLaunchKernelGGL(eye_kernel, dim3(grid_of(m * k)), dim3(256), 0, c->stream, m, k, Q, ldq);
  LRA_HIP(hipGetLastError());
  double *Vw = nullptr, *Sw = nullptr, *Tw = nullptr, *W1 = nullptr, *W2 = nullptr;
  LRA_CHECK(lra_wsT(c, WS_QR0, (size_t)m * QR_NB, &Vw));
  LRA_CHECK(lra_wsT(c, WS_QR1, (size_t)2 * QR_NB * QR_NB, &Sw));
  Tw = Sw + QR_NB * QR_NB;
  LRA_CHECK(lra_wsT(c, WS_QR2, (size_t)2 * QR_NB * k, &W1));
  W2 = W1 + (size_t)QR_NB * k;
  const i64 nblk = (k + QR_NB - 1) / QR_NB;
  for (i64 b = nblk - 1; b >= 0; --b) {
    const i64 j0 = b * QR_NB;
    const int nb = (int)std::min<i64>(QR_NB, k - j0);
    LRA_CHECK(panel_wy(c, m, j0, nb, A, lda, tau, Vw, Sw, Tw));
    // Q(j0:m, j0:k) = (I - V T V') Q(j0:m, j0:k)
    LRA_CHECK(apply_block(c, m - j0, k - j0, nb, Vw, Tw, Q + j0 + j0 * ldq, ldq, 0, W1, W2));
  }
  return LRA_OK;
}

extern "C" int lra_geqrf_f64(lra_ctx* c, int64_t m, int64_t n, double* A, int64_t lda,
                             double* tau) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0 && n <= m, 3);
  LRA_ARG(A != nullptr || m * n == 0, 4);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 5);
  LRA_ARG(tau != nullptr || n == 0, 6);
  LRA_HIP(hipSetDevice(c->device));
  return lra_geqrf_impl(c, m, n, A, lda, tau);
}

extern "C" int lra_orgqr_f64(lra_ctx* c, int64_t m, int64_t k, const double* A, int64_t lda,
                             const double* tau, double* Q, int64_t ldq) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(k >= 0 && k <= m, 3);
  LRA_ARG(A != nullptr || m * k == 0, 4);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 5);
  LRA_ARG(tau != nullptr || k == 0, 6);
  LRA_ARG(Q != nullptr || m * k == 0, 7);
  LRA_ARG(ldq >= std::max<int64_t>(1, m), 8);
  LRA_HIP(hipSetDevice(c->device));
  return lra_orgqr_impl(c, m, k, A, lda, tau, Q, ldq);
}
