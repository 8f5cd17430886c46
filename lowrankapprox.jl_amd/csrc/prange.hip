// prange.hip -- prange (src/prange.jl:14-77): an orthonormal basis Q of the numerical range
// of op(A) (trans n / c), or of the ranges of A and A' together (trans b, square A).
//
//   randn: sketchfact(:right, trans, A) with pqrfact_retval "q" (src/prange.jl:58): round r
//          sketches B = op(A) Omega_r (P x ell, Philox stream base + r, ell = samp(target)),
//          runs the rank-adaptive QRCP over B's ell columns (geqp3_adap!, src/pqr.jl:348-418)
//          and doubles the target until k < target (src/sketch.jl:223-240, side right); Q is
//          the first k columns of B's Householder product (pqrback_postproc, src/pqr.jl:420-436).
//   sub:   prange_sub (src/prange.jl:65-78): the left sketchfact's k pivot columns of op(A),
//          orthonormalised by QR (orthcols!, src/util.jl:73-85).
//   none:  pqrfact(op(A))[:Q] (src/prange.jl:52-54).
//   b:     both right sketchfacts (A' and A; the second draws from Philox streams 2^32 + r),
//          B = [Q_r R_r, Q_c R_c] with R the k x k leading triangles, then pqrfact_backend!(B)
//          (src/prange.jl:25-46).  An exactly symmetric A takes trans n.
// srft and sprn sketches have no right-side form here (LRA_E_UNSUPPORTED).
#include "lra_internal.h"

#include <algorithm>

namespace {

// out (k x k) = upper triangle of R (ld ldr), zeros below
__global__ void prange_triu_kernel(i64 k, const double* __restrict__ R, i64 ldr,
                                   double* __restrict__ out) {
  const i64 total = k * k;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % k, j = e / k;
    out[e] = i <= j ? R[i + j * ldr] : 0.0;
  }
}

int grid_of(i64 work) { return (int)std::max<i64>(1, std::min<i64>((work + 255) / 256, 8192)); }

struct RightFact {
  double* B;   // factored sketch, P x ell (ld P)
  double* tau;
  i64 P, ell, k, rounds;
};

// sketchfact(:right, trans, A) with the Gaussian sketch; slots sb/sp/st hold B, jpvt, tau
int sketchfact_right(lra_ctx* c, int trans, i64 m, i64 n, const double* A, i64 lda,
                     const lra_opts* o, uint64_t stream_base, int sb, int sp, int st,
                     RightFact* f) {
  const i64 P = trans == LRA_TRANS_N ? m : n;
  const bool adap = o->sketchfact_adap || o->rank < 0;
  i64 target = adap ? o->nb : o->rank;
  // trans n: A Omega reads A m-contiguous, which the k-contiguous GEMM kernels cannot take as
  // the product's left operand; A' is formed once (even ld) and every round sketches it with
  // trans c -- the same product A Omega with the same Omega (it depends on its shape only)
  const double* Au = A;
  i64 ldu = lda, mu = m, nu = n;
  int tu = trans;
  if (trans == LRA_TRANS_N && m > 0 && n > 0) {
    double* At = nullptr;
    const i64 ldt = (n + 1) / 2 * 2;
    LRA_CHECK(lra_wsT(c, WS_PRT, (size_t)ldt * m, &At));
    LRA_CHECK(lra_launch_transpose(c->stream, m, n, A, lda, At, ldt));
    Au = At; ldu = ldt; mu = n; nu = m; tu = LRA_TRANS_C;
  }
  for (i64 rnd = 0;; ++rnd) {
    const i64 ell = o->randn_samp_mul * target + o->randn_samp_add;
    if (ell <= 0) {
      lra_set_error("prange: sketch order %lld", (long long)ell);
      return LRA_E_UNSUPPORTED;
    }
    double *B = nullptr, *tau = nullptr;
    i64* jp = nullptr;
    LRA_CHECK(lra_wsT(c, sb, (size_t)std::max<i64>(P * ell, 1), &B));
    LRA_CHECK(lra_wsT(c, sp, (size_t)ell, &jp));
    LRA_CHECK(lra_wsT(c, st, (size_t)std::max<i64>(std::min(P, ell), 1), &tau));
    LRA_CHECK(lra_sketch_gaussian(c, LRA_SIDE_RIGHT, tu, mu, nu, Au, ldu, ell, nullptr, o->seed,
                                  stream_base + (uint64_t)rnd, LRA_PREC_F64, B,
                                  std::max<i64>(P, 1)));
    i64 k = 0;
    LRA_CHECK(lra_geqp3_adap_f64(c, P, ell, B, std::max<i64>(P, 1), o->rank, o->atol, o->rtol,
                                 o->nb, jp, tau, &k));
    if (!adap || k < target) {
      f->B = B; f->tau = tau; f->P = P; f->ell = ell; f->k = k; f->rounds = rnd + 1;
      return LRA_OK;
    }
    target *= 2;
  }
}

// out->Q (P x k) = the first k columns of the Householder product stored in B (P x ncols)
int q_of(lra_ctx* c, i64 P, i64 k, double* B, const double* tau, int slot, lra_pqr_out* out) {
  if (k == 0) return LRA_OK;
  if (!out->Q) {
    double* q = nullptr;
    LRA_CHECK(lra_wsT(c, slot, (size_t)P * k, &q));
    out->Q = q;
    out->ldq = P;
  }
  return lra_orgqr_impl(c, P, k, B, std::max<i64>(P, 1), tau, out->Q, out->ldq);
}

}  // namespace

extern "C" int lra_prange(lra_ctx* c, int trans, int64_t m, int64_t n, const double* A,
                          int64_t lda, const lra_opts* opts, lra_pqr_out* out) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C || trans == LRA_TRANS_B, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(trans != LRA_TRANS_B || m == n, 4);  // checksquare (src/prange.jl:24)
  LRA_ARG(A != nullptr || m * n == 0, 5);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(out != nullptr, 8);
  LRA_CHECK(lra_chkopts_public(opts));
  LRA_HIP(hipSetDevice(c->device));
  lra_opts o = *opts;
  out->k = 0;
  out->rounds = 0;
  if (trans == LRA_TRANS_B && n > 0) {
    int* flag = nullptr;
    LRA_CHECK(lra_wsT(c, WS_PR9, 64, &flag));
    LRA_CHECK(lra_launch_is_symmetric(c->stream, n, A, lda, flag));
    int h = 0;
    LRA_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    LRA_HIP(hipStreamSynchronize(c->stream));
    if (h) trans = LRA_TRANS_N;  // ishermitian(A) && return prange(:n, A)
  }
  const i64 P = trans == LRA_TRANS_C ? n : m;
  if (P == 0 || m * n == 0) return LRA_OK;

  if (trans != LRA_TRANS_B) {
    if (o.sketch == LRA_SKETCH_NONE) {
      o.pqrfact_retval = LRA_RET_Q;
      return lra_pqrfact(c, trans, m, n, A, lda, &o, out);
    }
    if (o.sketch == LRA_SKETCH_SUB) {
      const i64 N = trans == LRA_TRANS_N ? n : m;
      i64* p = nullptr;
      LRA_CHECK(lra_wsT(c, WS_PR1, (size_t)N, &p));
      i64 k = 0;
      LRA_CHECK(lra_sketchfact(c, trans, m, n, A, lda, &o, 0, p, &k));
      out->k = k;
      out->rounds = 1;
      if (k == 0) return LRA_OK;
      double *B = nullptr, *tau = nullptr;
      LRA_CHECK(lra_wsT(c, WS_PR0, (size_t)P * k, &B));
      LRA_CHECK(lra_wsT(c, WS_PR2, (size_t)k + 8, &tau));
      if (trans == LRA_TRANS_N) LRA_CHECK(lra_launch_gather_cols(c->stream, m, k, A, lda, p, B, P));
      else LRA_CHECK(lra_launch_gather_rows_t(c->stream, n, k, A, lda, p, B, P));
      const i64 kq = std::min(P, k);  // orthcols! keeps min(P, k) columns
      LRA_CHECK(lra_geqrf_impl(c, P, kq, B, P, tau));
      out->k = kq;
      LRA_CHECK(q_of(c, P, kq, B, tau, WS_PR3, out));
      return lra_check_device_error(c);
    }
    if (o.sketch != LRA_SKETCH_RANDN) {
      lra_set_error("prange: the srft and sprn sketches have no right-side form here");
      return LRA_E_UNSUPPORTED;
    }
    RightFact f;
    LRA_CHECK(sketchfact_right(c, trans, m, n, A, lda, &o, 0, WS_PR0, WS_PR1, WS_PR2, &f));
    out->k = f.k;
    out->rounds = f.rounds;
    LRA_CHECK(q_of(c, P, f.k, f.B, f.tau, WS_PR3, out));
    return lra_check_device_error(c);
  }

  // ---- trans b: ranges of A' and A (square A)
  if (o.sketch != LRA_SKETCH_RANDN && o.sketch != LRA_SKETCH_NONE) {
    lra_set_error("prange(:b): only the randn and none sketches are supported here");
    return LRA_E_UNSUPPORTED;
  }
  double *Qr = nullptr, *Qc = nullptr, *Rr = nullptr, *Rc = nullptr;
  i64 kr = 0, kc = 0, rounds = 0;
  if (o.sketch == LRA_SKETCH_NONE) {
    // Fr = pqrfact!(A'), Fc = pqrfact(A), both with pqrfact_retval "qr"
    o.pqrfact_retval = LRA_RET_Q | LRA_RET_R;
    for (int side = 0; side < 2; ++side) {
      lra_pqr_out po{};
      LRA_CHECK(lra_wsT(c, side ? WS_PR5 : WS_PR3, (size_t)n * n, &po.Q));
      po.ldq = n;
      LRA_CHECK(lra_wsT(c, side ? WS_PR6 : WS_PR4, (size_t)n * n, &po.R));
      po.ldr = n;
      LRA_CHECK(lra_pqrfact(c, side ? LRA_TRANS_N : LRA_TRANS_C, m, n, A, lda, &o, &po));
      (side ? Qc : Qr) = po.Q;
      (side ? Rc : Rr) = po.R;
      (side ? kc : kr) = po.k;
    }
    // (R = [R11, R11 T] with ldr = n: its leading k x k triangle is R11)
    o.pqrfact_retval = LRA_RET_Q;
    rounds = 1;
  } else {
    RightFact fr, fc;
    LRA_CHECK(sketchfact_right(c, LRA_TRANS_C, m, n, A, lda, &o, 0, WS_PR0, WS_PR1, WS_PR2, &fr));
    LRA_CHECK(sketchfact_right(c, LRA_TRANS_N, m, n, A, lda, &o, (uint64_t)1 << 32, WS_PR5,
                               WS_PR6, WS_PR7, &fc));
    kr = fr.k;
    kc = fc.k;
    rounds = std::max(fr.rounds, fc.rounds);
    lra_pqr_out qr_{}, qc_{};
    LRA_CHECK(q_of(c, n, kr, fr.B, fr.tau, WS_PR3, &qr_));
    LRA_CHECK(q_of(c, n, kc, fc.B, fc.tau, WS_PR4, &qc_));
    Qr = qr_.Q;
    Qc = qc_.Q;
    Rr = fr.B;  // R on and above the diagonal of the factored sketches (ld n)
    Rc = fc.B;
  }
  const i64 kb = kr + kc;
  out->rounds = rounds;
  if (kb == 0) return lra_check_device_error(c);
  // B = [Q_r triu(R_r), Q_c triu(R_c)] (n x (kr + kc))
  double *B = nullptr, *Tri = nullptr;
  LRA_CHECK(lra_wsT(c, WS_PR8, (size_t)n * kb, &B));
  LRA_CHECK(lra_wsT(c, WS_PR9, (size_t)std::max<i64>(kr, kc) * std::max<i64>(kr, kc), &Tri));
  for (int side = 0; side < 2; ++side) {
    const i64 k = side ? kc : kr;
    if (k == 0) continue;
    hipLaunchKernelGGL(prange_triu_kernel, dim3(grid_of(k * k)), dim3(256), 0, c->stream, k,
                       side ? Rc : Rr, (i64)n, Tri);
    LRA_HIP(hipGetLastError());
    LRA_CHECK(lra_dgemm(c, 0, 0, n, k, k, 1.0, side ? Qc : Qr, n, Tri, k, 0.0,
                        B + (side ? kr : 0) * n, n));
  }
  // pqrfact_backend!(B, opts) with pqrfact_retval "q": the rank-adaptive QRCP of B, then Q
  i64* jp = nullptr;
  double* tau = nullptr;
  LRA_CHECK(lra_wsT(c, WS_PR10, (size_t)kb, &jp));
  LRA_CHECK(lra_wsT(c, WS_PR11, (size_t)std::min<i64>(n, kb) + 8, &tau));
  i64 k = 0;
  LRA_CHECK(lra_geqp3_adap_f64(c, n, kb, B, n, o.rank, o.atol, o.rtol, o.nb, jp, tau, &k));
  out->k = k;
  LRA_CHECK(q_of(c, n, k, B, tau, WS_PR12, out));
  return lra_check_device_error(c);
}
