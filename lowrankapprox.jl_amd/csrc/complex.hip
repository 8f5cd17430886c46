// complex.hip -- the ComplexF64 path of idfact (src/id.jl:434-447 for T <: Complex): complex
// left sketches (Gaussian crandn, SRFT with a complex FFT, random subset, sparse Gaussian;
// src/util.jl:4-11, src/sketch.jl:87-110, 338-376, 452-472, 244-281, 566-590), the adaptive
// loop (src/sketch.jl:223-240 and siblings), the rank-adaptive pivoted QR with zlaqps'
// arithmetic (qrcp_global.hip) and T = R11 \ R12 (src/pqr.jl:438-442).
//
// A complex sketch Phi * op(A) runs as ONE real FP64 MFMA GEMM: with op(A) viewed as the real
// 2P x N matrix X of its interleaved (re, im) rows, the 2l x 2P real matrix
//   W[2p, 2i] = Re Phi[p,i],  W[2p, 2i+1] = -Im Phi[p,i],
//   W[2p+1, 2i] = Im Phi[p,i], W[2p+1, 2i+1] = Re Phi[p,i]
// gives W X = the l x N complex sketch in the same interleaved layout.  Every sketch kind is
// formed as its dense Phi (l x P) first; the SRFT's Phi[p, t] = d_t exp(-2 pi i f_p t / P) is
// the sampled DFT the reference's six-step FFT evaluates (f_p = idx_p - 1).
#include "lra_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

int lra_zqrcp_global(lra_ctx* c, double* Bz, i64 ldb, int m, int n, int kmax, int nb,
                     double atol, double rtol, int stop_after, i64* jpvt, double* tau,
                     i64* k_out);

namespace {

typedef double2 zc;

inline int grid_for(i64 work) {
  i64 g = (work + 255) / 256;
  return (int)std::max<i64>(1, std::min<i64>(g, 65536));
}

#define LRA_LAUNCH_CHECK()                                                                 \
  do {                                                                                     \
    hipError_t _e = hipGetLastError();                                                     \
    if (_e != hipSuccess) {                                                                \
      lra_set_error("%s:%d launch: %s", __FILE__, __LINE__, hipGetErrorString(_e));       \
      return LRA_E_HIP;                                                                    \
    }                                                                                      \
  } while (0)

// crandn(ComplexF64, l, P) (src/util.jl:4-11): the reinterpretation of a real 2l x P
// Gaussian, so Phi[p, i] = G[2p, i] + i G[2p+1, i] with G[r, c] normal #(c 2l + r): the two
// normals of Box-Muller pair c l + p.
__global__ void zphi_randn_kernel(uint64_t seed, uint64_t stream, i64 l, i64 P, zc* Phi) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < l * P;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 p = e % l, i = e / l;
    double z0, z1;
    lra_normal2(seed, stream, (uint64_t)(i * l + p), &z0, &z1);
    Phi[e] = make_double2(z0, z1);
  }
}

// complex SRFT signs (srft_rand for Complex, src/sketch.jl:345-351): crandn entries scaled to
// unit modulus
__global__ void zsrft_d_kernel(uint64_t seed, uint64_t stream, i64 P, zc* d) {
  for (i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x; t < P;
       t += (i64)gridDim.x * blockDim.x) {
    double z0, z1;
    lra_normal2(seed, stream, (uint64_t)t, &z0, &z1);
    const double a = hypot(z0, z1);
    d[t] = make_double2(z0 / a, z1 / a);
  }
}

// Phi[p, t] = d_t exp(-2 pi i ((f_p t) mod P) / P), f_p = idx[p] - 1
__global__ void zphi_srft_kernel(i64 l, i64 P, const zc* d, const i64* idx, zc* Phi) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < l * P;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 p = e % l, t = e / l;
    const i64 f = idx[p] - 1;
    // f, t < P < 2^31: the product fits 64 bits exactly
    const i64 r = (i64)(((uint64_t)f * (uint64_t)t) % (uint64_t)P);
    double sn, cs;
    sincospi(-2.0 * (double)r / (double)P, &sn, &cs);
    const zc dd = d[t];
    Phi[e] = make_double2(dd.x * cs - dd.y * sn, dd.x * sn + dd.y * cs);
  }
}

__global__ void zfill_zero_kernel(i64 count, zc* X) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < count;
       e += (i64)gridDim.x * blockDim.x)
    X[e] = make_double2(0.0, 0.0);
}

// random subset (src/sketch.jl:248-257): Phi[p, r_p - 1] = 1
__global__ void zphi_sub_kernel(i64 l, const i64* r, zc* Phi) {
  const i64 p = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (p < l) Phi[p + (r[p] - 1) * l] = make_double2(1.0, 0.0);
}

// sparse Gaussian (src/sketch.jl:571-589): row p has q_p = fld(P - p - 1, l) + 1 entries
// s[o + j] at columns r[o + j] - 1, o the sum of the earlier rows' counts (closed form)
__global__ void zphi_sprn_kernel(i64 l, i64 P, const i64* r, const zc* s, zc* Phi) {
  const i64 p = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (p >= l) return;
  const i64 q = P - p - 1 >= 0 ? (P - p - 1) / l + 1 : 0;
  // o = sum_{i<p} (fld(P - i - 1, l) + 1): each of the first P mod l rows gets one extra
  const i64 base = P / l, extra = P % l;
  const i64 o = p * base + std::min<i64>(p, extra);
  for (i64 j = 0; j < q; ++j) Phi[p + (r[o + j] - 1) * l] = s[o + j];
}

__global__ void zsprn_s_kernel(uint64_t seed, uint64_t stream, i64 P, zc* s) {
  for (i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x; t < P;
       t += (i64)gridDim.x * blockDim.x) {
    double z0, z1;
    lra_normal2(seed, stream ^ 0xA5A5A5A5ull, (uint64_t)t, &z0, &z1);
    s[t] = make_double2(z0, z1);
  }
}

// Wt (2P x 2l, real, ld 2P) = W' for W the real form of left multiplication by Phi (l x P,
// complex: each entry x + iy is the 2 x 2 block [x -y; y x] of W), stored transposed so the
// sketch product reads its k index contiguously (the k-contiguous GEMM kernels); consecutive
// threads write consecutive rows of Wt.
__global__ void zexpand_t_kernel(i64 l, i64 P, const zc* Phi, double* Wt) {
  const i64 kw = 2 * P;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < l * P;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % P, p = e / P;
    const zc z = Phi[p + i * l];
    double* c0 = Wt + (2 * p) * kw + 2 * i;      // W(2p, 2i .. 2i+1)
    double* c1 = Wt + (2 * p + 1) * kw + 2 * i;  // W(2p+1, 2i .. 2i+1)
    c0[0] = z.x;  c0[1] = -z.y;
    c1[0] = z.y;  c1[1] = z.x;
  }
}

// out (n x m) = A^H for the m x n complex A (the operand of trans = c)
__global__ void zadjoint_kernel(i64 m, i64 n, const zc* A, i64 lda, zc* out, i64 ldo) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < m * n;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 j = e % n, i = e / n;  // out(j, i) = conj(A(i, j)); consecutive threads write
    const zc a = A[i + j * lda];       // consecutive rows of out
    out[j + i * ldo] = make_double2(a.x, -a.y);
  }
}

__global__ void zcopy2d_kernel(i64 m, i64 n, const zc* A, i64 lda, zc* out, i64 ldo) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < m * n;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % m, j = e / m;
    out[i + j * ldo] = A[i + j * lda];
  }
}

// T (k x nr) = R11 \ R12 for the complex upper-triangular R11 = R(0:k, 0:k) and R12 =
// R(0:k, k:k+nr): one thread per right-hand side, the substitution order of ztrsm
__global__ void ztrsm_T_kernel(i64 k, i64 nr, const zc* R, i64 ldr, zc* T, i64 ldt) {
  const i64 j = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (j >= nr) return;
  zc* t = T + j * ldt;
  for (i64 i = 0; i < k; ++i) t[i] = R[i + (k + j) * ldr];
  for (i64 i = k - 1; i >= 0; --i) {
    const zc d = R[i + i * ldr];
    const zc b = t[i];
    // b / d (Smith)
    zc x;
    if (fabs(d.x) >= fabs(d.y)) {
      const double rr = d.y / d.x, den = d.x + d.y * rr;
      x = make_double2((b.x + b.y * rr) / den, (b.y - b.x * rr) / den);
    } else {
      const double rr = d.x / d.y, den = d.y + d.x * rr;
      x = make_double2((b.x * rr + b.y) / den, (b.y * rr - b.x) / den);
    }
    t[i] = x;
    for (i64 r = 0; r < i; ++r) {
      const zc a = R[r + i * ldr];
      t[r] = make_double2(t[r].x - (a.x * x.x - a.y * x.y), t[r].y - (a.x * x.y + a.y * x.x));
    }
  }
}

// The complex T = R11 \ R12 as one REAL triangular solve on the MFMA T kernel.  Row i of
// [R11 R12] is scaled by u_i = conj(R(i,i)) / |R(i,i)|, which leaves T unchanged and makes
// the diagonal real (u_i = +-1, an exact scaling, when R(i,i) is real, as on every R of the
// pivoted QR here: zlarfg leaves beta real).  With each complex entry a + ib written as the
// real 2 x 2 block [a -b; b a], the scaled R11 is an upper-triangular 2k x 2k real matrix
// (diagonal blocks |R(i,i)| I), and a complex column of T read as interleaved (re, im)
// doubles is a real column of 2k rows: the real solve is the complex one, term for term.
LRA_DEV zc zmul(zc a, zc b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
LRA_DEV zc zunit_conj(zc d) {
  const double a = hypot(d.x, d.y);
  return a == 0.0 ? make_double2(1.0, 0.0) : make_double2(d.x / a, -d.y / a);
}

// Rr (2k x 2k, column-major) from the row-scaled R11
__global__ void zrealify_R11_kernel(i64 k, const zc* R, i64 ldr, double* Rr) {
  const i64 k2 = 2 * k;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < k2 * k2;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i2 = e % k2, j2 = e / k2, i = i2 >> 1, j = j2 >> 1;
    double v = 0.0;
    if (i <= j) {
      const zc d = R[i + i * ldr];
      const zc a = i == j ? make_double2(hypot(d.x, d.y), 0.0) : zmul(zunit_conj(d), R[i + j * ldr]);
      v = (i2 & 1) == 0 ? ((j2 & 1) == 0 ? a.x : -a.y) : ((j2 & 1) == 0 ? a.y : a.x);
    }
    Rr[e] = v;
  }
}

// T(i, c) = u_i R12(i, c)
__global__ void zscale_R12_kernel(i64 k, i64 nr, const zc* R, i64 ldr, zc* T, i64 ldt) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < k * nr;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, c = e / k;
    T[i + c * ldt] = zmul(zunit_conj(R[i + i * ldr]), R[i + (k + c) * ldr]);
  }
}

// T (k x nr, ldt) = R11 \ R12 for the complex R = [R11 R12] (ldr)
int zid_T_solve(lra_ctx* c, i64 k, i64 nr, const zc* R, i64 ldr, zc* T, i64 ldt) {
  if (k <= 0 || nr <= 0) return LRA_OK;
  if (2 * k > 64 * 1184) {  // beyond the blocked real solve: one thread per right-hand side
    hipLaunchKernelGGL(ztrsm_T_kernel, dim3(grid_for(nr)), dim3(256), 0, c->stream, k, nr, R,
                       ldr, T, ldt);
    LRA_LAUNCH_CHECK();
    return LRA_OK;
  }
  double* Rr = nullptr;
  LRA_CHECK(lra_wsT(c, WS_ZT, (size_t)4 * k * k, &Rr));
  hipLaunchKernelGGL(zrealify_R11_kernel, dim3(grid_for(4 * k * k)), dim3(256), 0, c->stream, k,
                     R, ldr, Rr);
  hipLaunchKernelGGL(zscale_R12_kernel, dim3(grid_for(k * nr)), dim3(256), 0, c->stream, k, nr,
                     R, ldr, T, ldt);
  LRA_LAUNCH_CHECK();
  return lra_trsm_upper_inplace(c, 2 * k, nr, Rr, 2 * k, reinterpret_cast<double*>(T), 2 * ldt);
}

// the random objects of a complex left sketch of round `stream` on a P-row operand
struct ZDraw {
  zc* d;    // srft signs (P) / sprn coefficients (P)
  i64* idx; // srft samples (l) / subset rows (l) / sprn permutation (P)
};

int zdraw(lra_ctx* c, int sketch, uint64_t seed, uint64_t stream, i64 P, i64 l, ZDraw* z) {
  hipStream_t st = c->stream;
  switch (sketch) {
    case LRA_SKETCH_SRFT:
      hipLaunchKernelGGL(zsrft_d_kernel, dim3(grid_for(P)), dim3(256), 0, st, seed, stream, P, z->d);
      LRA_LAUNCH_CHECK();
      return lra_launch_subset_rand(st, seed, stream, P, l, z->idx);
    case LRA_SKETCH_SUB:
      return lra_launch_subset_rand(st, seed, stream, P, l, z->idx);
    case LRA_SKETCH_SPRN: {
      // the permutation of the real sparse sketch (its real coefficients are not used)
      double* sr = nullptr;
      LRA_CHECK(lra_wsT(c, WS_TMP0, (size_t)std::max<i64>(P, 1), &sr));
      LRA_CHECK(lra_launch_sprn_rand(st, seed, stream, P, sr, z->idx));
      hipLaunchKernelGGL(zsprn_s_kernel, dim3(grid_for(P)), dim3(256), 0, st, seed, stream, P, z->d);
      LRA_LAUNCH_CHECK();
      return LRA_OK;
    }
  }
  return LRA_OK;
}

// Phi (l x P complex, ld l) of the sketch kind from (seed, stream) or the injected draws
int zphi(lra_ctx* c, int sketch, uint64_t seed, uint64_t stream, i64 P, i64 l, const ZDraw* inj,
         zc* Phi) {
  hipStream_t st = c->stream;
  if (sketch == LRA_SKETCH_RANDN) {
    hipLaunchKernelGGL(zphi_randn_kernel, dim3(grid_for(l * P)), dim3(256), 0, st, seed, stream,
                       l, P, Phi);
    LRA_LAUNCH_CHECK();
    return LRA_OK;
  }
  ZDraw z{};
  if (inj) {
    z = *inj;
  } else {
    zc* d = nullptr;
    i64* idx = nullptr;
    LRA_CHECK(lra_wsT(c, WS_LP0, (size_t)std::max<i64>(P, 1), &d));
    LRA_CHECK(lra_wsT(c, WS_LP1, (size_t)std::max<i64>(std::max(P, l), 1), &idx));
    z.d = d;
    z.idx = idx;
    LRA_CHECK(zdraw(c, sketch, seed, stream, P, l, &z));
  }
  if (sketch == LRA_SKETCH_SRFT) {
    hipLaunchKernelGGL(zphi_srft_kernel, dim3(grid_for(l * P)), dim3(256), 0, st, l, P, z.d,
                       z.idx, Phi);
    LRA_LAUNCH_CHECK();
    return LRA_OK;
  }
  hipLaunchKernelGGL(zfill_zero_kernel, dim3(grid_for(l * P)), dim3(256), 0, st, l * P, Phi);
  LRA_LAUNCH_CHECK();
  if (sketch == LRA_SKETCH_SUB)
    hipLaunchKernelGGL(zphi_sub_kernel, dim3(grid_for(l)), dim3(256), 0, st, l, z.idx, Phi);
  else
    hipLaunchKernelGGL(zphi_sprn_kernel, dim3(grid_for(l)), dim3(256), 0, st, l, P, z.idx, z.d,
                       Phi);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

// B (l x N complex, ld l) = Phi * X for X the P x N complex op(A) (ld ldx)
int zsketch_apply(lra_ctx* c, i64 l, i64 P, i64 N, const zc* Phi, const zc* X, i64 ldx, zc* B) {
  if (l <= 0 || N <= 0) return LRA_OK;
  if (P <= 0) {
    hipLaunchKernelGGL(zfill_zero_kernel, dim3(grid_for(l * N)), dim3(256), 0, c->stream, l * N, B);
    LRA_LAUNCH_CHECK();
    return LRA_OK;
  }
  double* W = nullptr;
  LRA_CHECK(lra_wsT(c, WS_TMP2, (size_t)4 * l * P, &W));
  hipLaunchKernelGGL(zexpand_t_kernel, dim3(grid_for(l * P)), dim3(256), 0, c->stream, l, P, Phi,
                     W);
  LRA_LAUNCH_CHECK();
  GemmArgs g;
  g.M = 2 * l; g.N = N; g.K = 2 * P; g.alpha = 1.0; g.beta = 0.0;
  g.A = W; g.lda = 2 * P; g.transA = 1;
  g.B = reinterpret_cast<const double*>(X); g.ldb = 2 * ldx; g.transB = 0;
  g.C = reinterpret_cast<double*>(B); g.ldc = 2 * l; g.bgather = nullptr; g.prec = LRA_PREC_F64;
  return lra_gemm(c, g);
}

i64 samp(const lra_opts* o, i64 x) {
  switch (o->sketch) {
    case LRA_SKETCH_RANDN: return o->randn_samp_mul * x + o->randn_samp_add;
    case LRA_SKETCH_SRFT: return o->srft_samp_mul * x + o->srft_samp_add;
    case LRA_SKETCH_SUB: return o->sub_samp_mul * x + o->sub_samp_add;
    default: return x;
  }
}

int read_i64(lra_ctx* c, const i64* dev, i64* host) {
  LRA_HIP(hipMemcpyAsync(c->host_pinned, dev, sizeof(i64), hipMemcpyDeviceToHost, c->stream));
  LRA_HIP(hipStreamSynchronize(c->stream));
  std::memcpy(host, c->host_pinned, sizeof(i64));
  return lra_check_device_error(c);
}

}  // namespace

extern "C" int lra_zsketch_draw(lra_ctx* c, int sketch, uint64_t seed, uint64_t stream_id,
                                int64_t m, int64_t ell, double* d, int64_t* idx) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(sketch >= LRA_SKETCH_RANDN && sketch <= LRA_SKETCH_SPRN, 2);
  LRA_ARG(m >= 0, 5);
  LRA_ARG(ell >= 0, 6);
  LRA_HIP(hipSetDevice(c->device));
  if (sketch == LRA_SKETCH_RANDN) {
    LRA_ARG(d != nullptr || m * ell == 0, 7);
    if (m * ell == 0) return LRA_OK;
    return zphi(c, sketch, seed, stream_id, m, ell, nullptr, reinterpret_cast<zc*>(d));
  }
  LRA_ARG(d != nullptr || sketch == LRA_SKETCH_SUB || m == 0, 7);
  LRA_ARG(idx != nullptr || (ell == 0 && m == 0), 8);
  ZDraw z{reinterpret_cast<zc*>(d), idx};
  return zdraw(c, sketch, seed, stream_id, m, ell, &z);
}

extern "C" int lra_zsketch(lra_ctx* c, int sketch, int trans, int64_t m, int64_t n,
                           const double* A, int64_t lda, int64_t ell, const double* d,
                           const int64_t* idx, uint64_t seed, uint64_t stream_id, double* B,
                           int64_t ldb) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(sketch >= LRA_SKETCH_RANDN && sketch <= LRA_SKETCH_SPRN, 2);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 3);
  LRA_ARG(m >= 0, 4);
  LRA_ARG(n >= 0, 5);
  LRA_ARG(A != nullptr || m * n == 0, 6);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 7);
  LRA_ARG(ell >= 0, 8);
  LRA_ARG(ldb == ell || ell == 0, 14);
  LRA_HIP(hipSetDevice(c->device));
  const i64 P = trans == LRA_TRANS_N ? m : n, N = trans == LRA_TRANS_N ? n : m;
  if (P >= (1ll << 31)) {
    lra_set_error("complex sketch: %lld rows exceed 2^31", (long long)P);
    return LRA_E_UNSUPPORTED;
  }
  const zc* X = reinterpret_cast<const zc*>(A);
  i64 ldx = lda;
  if (trans == LRA_TRANS_C && m * n > 0) {
    zc* Xt = nullptr;
    LRA_CHECK(lra_wsT(c, WS_TMP1, (size_t)2 * m * n, reinterpret_cast<double**>(&Xt)));
    hipLaunchKernelGGL(zadjoint_kernel, dim3(grid_for(m * n)), dim3(256), 0, c->stream, m, n, X,
                       lda, Xt, n);
    LRA_LAUNCH_CHECK();
    X = Xt;
    ldx = n;
  }
  zc* Phi = nullptr;
  LRA_CHECK(lra_wsT(c, WS_TMP3, (size_t)2 * std::max<i64>(ell * P, 1), reinterpret_cast<double**>(&Phi)));
  ZDraw inj{const_cast<zc*>(reinterpret_cast<const zc*>(d)), const_cast<i64*>(idx)};
  const bool injected = sketch != LRA_SKETCH_RANDN && idx != nullptr;
  if (ell * P > 0)
    LRA_CHECK(zphi(c, sketch, seed, stream_id, P, ell, injected ? &inj : nullptr, Phi));
  return zsketch_apply(c, ell, P, N, Phi, X, ldx, reinterpret_cast<zc*>(B));
}

extern "C" int lra_zgeqp3_adap(lra_ctx* c, int64_t m, int64_t n, double* B, int64_t ldb,
                               int64_t rank_cap, double atol, double rtol, int64_t nb,
                               int64_t* jpvt, double* tau, int64_t* k) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0 && m < (1 << 30), 2);
  LRA_ARG(n >= 0 && n < (1 << 30), 3);
  LRA_ARG(B != nullptr || m * n == 0, 4);
  LRA_ARG(ldb >= std::max<int64_t>(1, m), 5);
  LRA_ARG(atol >= 0, 7);
  LRA_ARG(rtol >= 0, 8);
  LRA_ARG(nb > 0, 9);
  LRA_ARG(k != nullptr, 12);
  LRA_HIP(hipSetDevice(c->device));
  const i64 l = std::min(m, n);
  const i64 kmax = (rank_cap < 0 || rank_cap > l) ? l : rank_cap;
  i64* dk = nullptr;
  LRA_CHECK(lra_wsT(c, WS_SCAL, 16, &dk));
  LRA_CHECK(lra_zqrcp_global(c, B, ldb, (int)m, (int)n, (int)kmax,
                             (int)std::min<i64>(nb, std::max<i64>(kmax, 1)), atol, rtol, 0, jpvt,
                             tau, dk));
  return read_i64(c, dk, k);
}

extern "C" int lra_zid_T(lra_ctx* c, int64_t k, int64_t n, const double* R, int64_t ldr,
                         double* T, int64_t ldt) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(k >= 0, 2);
  LRA_ARG(n >= k, 3);
  LRA_ARG(ldr >= std::max<int64_t>(1, k), 5);
  LRA_ARG(ldt >= std::max<int64_t>(1, k), 7);
  if (k == 0 || n == k) return LRA_OK;
  return zid_T_solve(c, k, n - k, reinterpret_cast<const zc*>(R), ldr, reinterpret_cast<zc*>(T),
                     ldt);
}

// idfact(trans, A) for a ComplexF64 A (interleaved re/im, lda in complex elements).  T is
// complex (ldt in complex elements).  The sketch is FP64 (opts->precision is not used);
// power iterations (sketch_randn_niter > 0) and maxdet swaps return LRA_E_UNSUPPORTED.
extern "C" int lra_zidfact(lra_ctx* c, int trans, int64_t m, int64_t n, const double* A,
                           int64_t lda, const lra_opts* opts, lra_id_out* out) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(A != nullptr || m * n == 0, 5);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(out != nullptr, 8);
  if (lra_chkopts_public(opts) != LRA_OK) return -7;
  LRA_HIP(hipSetDevice(c->device));
  const lra_opts& o = *opts;
  if (o.sketch_randn_niter > 0 && o.sketch == LRA_SKETCH_RANDN) {
    lra_set_error("complex idfact: power iterations are not supported");
    return LRA_E_UNSUPPORTED;
  }
  if (o.maxdet_tol >= 0) {
    lra_set_error("complex idfact: maxdet swaps are not supported");
    return LRA_E_UNSUPPORTED;
  }
  const i64 P = trans == LRA_TRANS_N ? m : n, N = trans == LRA_TRANS_N ? n : m;
  if (P >= (1ll << 31) || N >= (1ll << 30)) {
    lra_set_error("complex idfact: %lld x %lld exceeds the supported size", (long long)P,
                  (long long)N);
    return LRA_E_UNSUPPORTED;
  }
  const zc* X = reinterpret_cast<const zc*>(A);
  i64 ldx = lda;
  if (trans == LRA_TRANS_C && m * n > 0) {
    zc* Xt = nullptr;
    LRA_CHECK(lra_wsT(c, WS_TMP1, (size_t)2 * m * n, reinterpret_cast<double**>(&Xt)));
    hipLaunchKernelGGL(zadjoint_kernel, dim3(grid_for(m * n)), dim3(256), 0, c->stream, m, n, X,
                       lda, Xt, n);
    LRA_LAUNCH_CHECK();
    X = Xt;
    ldx = n;
  }
  i64* dk = nullptr;
  LRA_CHECK(lra_wsT(c, WS_SCAL, 16, &dk));
  i64* jpvt = nullptr;
  LRA_CHECK(lra_wsT(c, WS_JPVT, (size_t)std::max<i64>(N, 1), &jpvt));
  zc* B = nullptr;
  zc* tau = nullptr;
  i64 ell = 0, k = 0, rounds = 1;
  if (o.sketch == LRA_SKETCH_NONE) {
    // pqrfact_none: the rank-adaptive QRCP of op(A) itself
    ell = P;
    LRA_CHECK(lra_wsT(c, WS_SKETCH, (size_t)2 * std::max<i64>(P * N, 1), reinterpret_cast<double**>(&B)));
    LRA_CHECK(lra_wsT(c, WS_TAU, (size_t)2 * std::max<i64>(std::min(P, N), 1), reinterpret_cast<double**>(&tau)));
    if (P * N > 0)
      hipLaunchKernelGGL(zcopy2d_kernel, dim3(grid_for(P * N)), dim3(256), 0, c->stream, P, N, X,
                         ldx, B, std::max<i64>(P, 1));
    LRA_LAUNCH_CHECK();
    const i64 l = std::min(P, N);
    const i64 kmax = (o.rank < 0 || o.rank > l) ? l : o.rank;
    LRA_CHECK(lra_zqrcp_global(c, reinterpret_cast<double*>(B), std::max<i64>(P, 1), (int)P,
                               (int)N, (int)kmax, (int)std::min<i64>(o.nb, std::max<i64>(kmax, 1)),
                               o.atol, o.rtol, 0, jpvt, reinterpret_cast<double*>(tau), dk));
    LRA_CHECK(read_i64(c, dk, &k));
  } else {
    const bool adap = o.sketchfact_adap || o.rank < 0;
    i64 target = adap ? o.nb : o.rank;
    for (i64 rnd = 0;; ++rnd, target *= 2) {
      ell = samp(&o, target);
      if (ell <= 0) { lra_set_error("sketch order %lld", (long long)ell); return LRA_E_UNSUPPORTED; }
      LRA_CHECK(lra_wsT(c, WS_SKETCH, (size_t)2 * ell * std::max<i64>(N, 1), reinterpret_cast<double**>(&B)));
      LRA_CHECK(lra_wsT(c, WS_TAU, (size_t)2 * std::max<i64>(std::min(ell, N), 1), reinterpret_cast<double**>(&tau)));
      zc* Phi = nullptr;
      LRA_CHECK(lra_wsT(c, WS_TMP3, (size_t)2 * std::max<i64>(ell * P, 1), reinterpret_cast<double**>(&Phi)));
      if (ell * P > 0) LRA_CHECK(zphi(c, o.sketch, o.seed, (uint64_t)rnd, P, ell, nullptr, Phi));
      LRA_CHECK(zsketch_apply(c, ell, P, N, Phi, X, ldx, B));
      const i64 l = std::min(ell, N);
      const i64 kmax = (o.rank < 0 || o.rank > l) ? l : o.rank;
      LRA_CHECK(lra_zqrcp_global(c, reinterpret_cast<double*>(B), ell, (int)ell, (int)N, (int)kmax,
                                 (int)std::min<i64>(o.nb, std::max<i64>(kmax, 1)), o.atol, o.rtol,
                                 adap ? (int)std::min<i64>(target, 1 << 30) : 0, jpvt,
                                 reinterpret_cast<double*>(tau), dk));
      LRA_CHECK(read_i64(c, dk, &k));
      rounds = rnd + 1;
      if (!adap || k < target) break;
    }
  }
  LRA_CHECK(lra_launch_split_ids(c->stream, N, k, jpvt, out->sk, out->rd));
  if (!out->T) {
    double* t = nullptr;
    LRA_CHECK(lra_wsT(c, WS_T, (size_t)2 * std::max<i64>(k, 1) * std::max<i64>(N - k, 1), &t));
    out->T = t;
    out->ldt = std::max<i64>(k, 1);
  }
  if (k > 0 && N > k) {
    LRA_ARG(out->ldt >= k, 8);
    LRA_CHECK(zid_T_solve(c, k, N - k, B, std::max<i64>(ell, 1), reinterpret_cast<zc*>(out->T),
                          out->ldt));
  }
  out->k = k;
  out->rounds = rounds;
  return lra_check_device_error(c);
}
