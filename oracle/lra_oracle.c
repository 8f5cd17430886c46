/*
 * lra_oracle.c -- CPU restatement of LowRankApprox.jl's sketch-then-partial-factorize path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/, the smoke check
 * and bench.py's CPU-baseline leg may load it, and only as the checker.  The product path
 * (liblra_hip.so) never links, loads or calls anything in oracle/.
 *
 * Every routine states the reference code it restates (paths relative to the reference
 * repository).  All randomness (sketch matrices, sign vectors, sample indices,
 * permutations) is an explicit input, so the GPU path and this oracle see identical bits.
 *
 * Conventions: column-major storage, int64 sizes (Julia's BlasInt on ILP64), 1-based pivot
 * vectors exactly as LAPACK and the reference hand them around.
 *
 * Pinning: orc_dlaqps is checked against OpenBLAS's own dlaqps (the routine the reference
 * ccall's at src/lapack.jl:129-136) and against dgeqp3; orc_trsm_upper against
 * scipy.linalg.solve_triangular; the SRFT against numpy.fft; see tests/test_oracle.py.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

/* Fortran-ABI signature of ?laqps on ILP64 LAPACK (src/lapack.jl:117-139). */
typedef void (*orc_laqps_fn)(const i64* m, const i64* n, const i64* offset, const i64* nb,
                             i64* kb, double* A, const i64* lda, i64* jpvt, double* tau,
                             double* vn1, double* vn2, double* auxv, double* F, const i64* ldf);

#define A_(i, j) A[(i) + (j) * lda]

/* Two-norm as a plain sum of squares (the GPU kernels use the same formula). */
static double orc_nrm2(i64 n, const double* x, i64 inc) {
  double s = 0.0;
  for (i64 i = 0; i < n; ++i) s += x[i * inc] * x[i * inc];
  return sqrt(s);
}

/* sqrt(x^2 + y^2) without destructive underflow/overflow (LAPACK dlapy2). */
static double orc_lapy2(double x, double y) {
  double xa = fabs(x), ya = fabs(y);
  double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
  if (z == 0.0 || w > DBL_MAX) return w;
  double q = z / w;
  return w * sqrt(1.0 + q * q);
}

/*
 * Elementary reflector H = I - tau v v' with H [alpha; x] = [beta; 0], v(1) = 1
 * (LAPACK dlarfg, Householder generation used by dlaqps).  safmin scaling kept.
 */
void orc_dlarfg(i64 n, double* alpha, double* x, i64 incx, double* tau) {
  if (n <= 1) { *tau = 0.0; return; }
  double xnorm = orc_nrm2(n - 1, x, incx);
  if (xnorm == 0.0) { *tau = 0.0; return; }
  double beta = -copysign(orc_lapy2(*alpha, xnorm), *alpha);
  const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
  int knt = 0;
  if (fabs(beta) < safmin) {
    const double rsafmn = 1.0 / safmin;
    do {
      ++knt;
      for (i64 i = 0; i < n - 1; ++i) x[i * incx] *= rsafmn;
      beta *= rsafmn;
      *alpha *= rsafmn;
    } while (fabs(beta) < safmin && knt < 20);
    xnorm = orc_nrm2(n - 1, x, incx);
    beta = -copysign(orc_lapy2(*alpha, xnorm), *alpha);
  }
  *tau = (beta - *alpha) / beta;
  const double scal = 1.0 / (*alpha - beta);
  for (i64 i = 0; i < n - 1; ++i) x[i * incx] *= scal;
  for (int j = 0; j < knt; ++j) beta *= safmin;
  *alpha = beta;
}

/*
 * One block step of Householder QR with column pivoting, Quintana-Orti/Sun/Bischof BLAS-3
 * scheme (LAPACK dlaqps; the block step the reference calls at src/lapack.jl:129-136 from
 * geqp3_adap_main!, src/pqr.jl:397-405).  Restated from the published algorithm:
 *   per column: pivot = first argmax of vn1, swap, bring column up to date with the
 *   deferred reflectors (A -= A*F'), generate H, build column of F, update the pivot row,
 *   downdate norms (LAWN 176 formula), flag columns whose downdate lost accuracy; stop the
 *   block early at the first flagged step; then apply the block reflector to the trailing
 *   matrix and recompute the flagged norms.
 * gap (optional, may be NULL): per step, the winning norm minus the runner-up at selection
 *   time (the winning norm if there is no runner-up); used by parity tests.
 */
void orc_dlaqps_gap(i64 m, i64 n, i64 offset, i64 nb, i64* kb, double* A, i64 lda, i64* jpvt,
                    double* tau, double* vn1, double* vn2, double* auxv, double* F, i64 ldf,
                    double* gap) {
  const i64 lastrk = m < n + offset ? m : n + offset;
  const double tol3z = sqrt(DBL_EPSILON * 0.5); /* sqrt(dlamch('E')), eps = 2^-53 */
  i64 lsticc = 0; /* 1-based head of the recompute list, 0 = empty */
  i64 k = 0;      /* number of columns done in this block */
  (void)auxv;
  /* F is n x nb, column-major with leading dimension ldf. */
#define F_(i, j) F[(i) + (j) * ldf]
  while (k < nb && lsticc == 0) {
    ++k;
    const i64 rk = offset + k;        /* 1-based row of the new diagonal */
    const i64 kc = k - 1, rc = rk - 1; /* 0-based */
    /* pivot: first index of the maximum of vn1(k:n) */
    i64 pvt = kc;
    double best = vn1[kc], second = -1.0;
    for (i64 j = kc + 1; j < n; ++j) {
      if (vn1[j] > best) { second = best; best = vn1[j]; pvt = j; }
      else if (vn1[j] > second) second = vn1[j];
    }
    if (gap) gap[kc] = second < 0.0 ? best : best - second;  /* absolute norm gap */
    if (pvt != kc) {
      for (i64 i = 0; i < m; ++i) { double t = A_(i, pvt); A_(i, pvt) = A_(i, kc); A_(i, kc) = t; }
      for (i64 j = 0; j < k - 1; ++j) { double t = F_(pvt, j); F_(pvt, j) = F_(kc, j); F_(kc, j) = t; }
      i64 it = jpvt[pvt]; jpvt[pvt] = jpvt[kc]; jpvt[kc] = it;
      vn1[pvt] = vn1[kc];
      vn2[pvt] = vn2[kc];
    }
    /* A(rk:m,k) -= A(rk:m,1:k-1) * F(k,1:k-1)' */
    for (i64 j = 0; j < k - 1; ++j) {
      const double f = F_(kc, j);
      if (f != 0.0)
        for (i64 i = rc; i < m; ++i) A_(i, kc) -= A_(i, j) * f;
    }
    /* H(k) */
    if (rk < m) orc_dlarfg(m - rk + 1, &A_(rc, kc), &A_(rc + 1, kc), 1, &tau[kc]);
    else orc_dlarfg(1, &A_(rc, kc), &A_(rc, kc), 1, &tau[kc]);
    const double akk = A_(rc, kc);
    A_(rc, kc) = 1.0;
    /* F(k+1:n,k) = tau(k) * A(rk:m,k+1:n)' * A(rk:m,k) */
    for (i64 j = kc + 1; j < n; ++j) {
      double s = 0.0;
      for (i64 i = rc; i < m; ++i) s += A_(i, j) * A_(i, kc);
      F_(j, kc) = tau[kc] * s;
    }
    for (i64 j = 0; j <= kc; ++j) F_(j, kc) = 0.0;
    /* F(1:n,k) -= tau(k) * F(1:n,1:k-1) * (A(rk:m,1:k-1)' * A(rk:m,k)) */
    if (k > 1) {
      double* aux = (double*)malloc(sizeof(double) * (size_t)(k - 1));
      for (i64 j = 0; j < k - 1; ++j) {
        double s = 0.0;
        for (i64 i = rc; i < m; ++i) s += A_(i, j) * A_(i, kc);
        aux[j] = -tau[kc] * s;
      }
      for (i64 j = 0; j < k - 1; ++j)
        for (i64 i = 0; i < n; ++i) F_(i, kc) += F_(i, j) * aux[j];
      free(aux);
    }
    /* A(rk,k+1:n) -= A(rk,1:k) * F(k+1:n,1:k)' */
    for (i64 j = kc + 1; j < n; ++j) {
      double s = 0.0;
      for (i64 l = 0; l < k; ++l) s += A_(rc, l) * F_(j, l);
      A_(rc, j) -= s;
    }
    /* partial column norm downdate, LAWN 176 */
    if (rk < lastrk) {
      for (i64 j = kc + 1; j < n; ++j) {
        if (vn1[j] != 0.0) {
          double temp = fabs(A_(rc, j)) / vn1[j];
          temp = (1.0 + temp) * (1.0 - temp);
          if (temp < 0.0) temp = 0.0;
          const double r = vn1[j] / vn2[j];
          const double temp2 = temp * r * r;
          if (temp2 <= tol3z) {
            vn2[j] = (double)lsticc;
            lsticc = j + 1;
          } else {
            vn1[j] *= sqrt(temp);
          }
        }
      }
    }
    A_(rc, kc) = akk;
  }
  *kb = k;
  const i64 rk = offset + k;
  /* trailing update A(rk+1:m,kb+1:n) -= A(rk+1:m,1:kb) * F(kb+1:n,1:kb)' */
  const i64 lim = n < m - offset ? n : m - offset;
  if (k < lim) {
    for (i64 j = k; j < n; ++j)
      for (i64 l = 0; l < k; ++l) {
        const double f = F_(j, l);
        if (f != 0.0)
          for (i64 i = rk; i < m; ++i) A_(i, j) -= A_(i, l) * f;
      }
  }
  /* recompute the flagged norms */
  while (lsticc > 0) {
    const i64 j = lsticc - 1;
    const i64 next = (i64)llround(vn2[j]);
    vn1[j] = orc_nrm2(m - rk, &A_(rk, j), 1);
    vn2[j] = vn1[j];
    lsticc = next;
  }
#undef F_
}

/* Fortran-ABI wrapper with the exact dlaqps argument list (so it is interchangeable with the
 * LAPACK routine in orc_geqp3_adap). */
void orc_dlaqps_(const i64* m, const i64* n, const i64* offset, const i64* nb, i64* kb,
                 double* A, const i64* lda, i64* jpvt, double* tau, double* vn1, double* vn2,
                 double* auxv, double* F, const i64* ldf) {
  orc_dlaqps_gap(*m, *n, *offset, *nb, kb, A, *lda, jpvt, tau, vn1, vn2, auxv, F, *ldf, NULL);
}

/*
 * Rank-adaptive GEQP3: geqp3_adap! + geqp3_adap_main! (src/pqr.jl:348-418).
 *   k = min(rank < 0 ? inf : rank, min(m,n)); initial norms; ptol = max(atol, rtol*maxnrm);
 *   blocks of jb = min(nb, k-j+1) columns through laqps (offset j-1, columns j:n, ldf = n-j+1);
 *   after each block, if |A(jn-1,jn-1)| <= ptol return the first i in [j,jn) with
 *   |A(i,i)| <= ptol, minus 1.
 * fn = NULL uses orc_dlaqps_gap (and fills gap[0:k) if gap != NULL); otherwise fn is an
 * external dlaqps (e.g. OpenBLAS).  jpvt must have n entries, tau min(m,n).
 * Returns the rank k.
 */
i64 orc_geqp3_adap(i64 m, i64 n, double* A, i64 lda, i64 rank, double atol, double rtol,
                   i64 nb_opt, i64* jpvt, double* tau, orc_laqps_fn fn, double* gap) {
  for (i64 j = 0; j < n; ++j) jpvt[j] = j + 1;
  const i64 l = m < n ? m : n;
  i64 k = (rank < 0 || rank > l) ? l : rank;
  if (k <= 0) return k;
  const i64 nb = nb_opt < k ? nb_opt : k;
  const i64 lwork = 2 * n + (n + 1) * nb;
  double* work = (double*)calloc((size_t)lwork, sizeof(double));
  for (i64 j = 0; j < n; ++j) work[j] = work[n + j] = orc_nrm2(m, &A_(0, j), 1);
  double maxnrm = 0.0;
  for (i64 j = 0; j < n; ++j) if (work[j] > maxnrm) maxnrm = work[j];
  const double ptol = atol > rtol * maxnrm ? atol : rtol * maxnrm;
  i64 j = 1;
  i64 result = k;
  while (j <= k) {
    const i64 jb = nb < k - j + 1 ? nb : k - j + 1;
    const i64 off = j - 1, nn = n - j + 1, ldf = nn;
    i64 fjb = 0;
    double* vn1 = work + (j - 1);
    double* vn2 = work + n + (j - 1);
    double* auxv = work + 2 * n;
    double* F = work + 2 * n + jb;
    if (fn) fn(&m, &nn, &off, &jb, &fjb, &A_(0, j - 1), &lda, jpvt + (j - 1), tau + (j - 1),
               vn1, vn2, auxv, F, &ldf);
    else orc_dlaqps_gap(m, nn, off, jb, &fjb, &A_(0, j - 1), lda, jpvt + (j - 1),
                        tau + (j - 1), vn1, vn2, auxv, F, ldf, gap ? gap + (j - 1) : NULL);
    const i64 jn = j + fjb;
    if (fabs(A_(jn - 2, jn - 2)) <= ptol) {
      i64 i;
      for (i = j; i < jn; ++i)
        if (fabs(A_(i - 1, i - 1)) <= ptol) break;
      if (i < jn) { result = i - 1; break; }
    }
    j = jn;
  }
  free(work);
  return result;
}

/*
 * T = R11 \ R12 with R11 = triu(R(1:k,1:k)) (maxdet_t, src/pqr.jl:438-442; the ldiv! is a
 * dtrsm).  R is k x n (ldr), T is k x (n-k) (ldt), column-wise back substitution.
 */
void orc_trsm_upper(i64 k, i64 n, const double* R, i64 ldr, double* T, i64 ldt) {
  for (i64 c = 0; c < n - k; ++c) {
    double* t = T + c * ldt;
    for (i64 i = 0; i < k; ++i) t[i] = R[i + (k + c) * ldr];
    for (i64 i = k - 1; i >= 0; --i) {
      t[i] /= R[i + i * ldr];
      const double ti = t[i];
      for (i64 r = 0; r < i; ++r) t[r] -= R[r + i * ldr] * ti;
    }
  }
}

/*
 * Strong rank-revealing swaps (maxdet_swapcols! + maxdet_update!, src/pqr.jl:445-501) on the
 * T/p part: while max|T| > 1 + tol, swap skeleton i with redundant j and apply the
 * Sherman-Morrison rank-1 update of T (the ger! at :494).  R1 (k x k, ldr1) is updated as
 * in :495-500 when r1 != NULL.  p is 1-based, length n.  Returns the number of swaps.
 * findmaxabs (src/util.jl:13-23) keeps the LAST maximal entry in column-major order.
 */
i64 orc_maxdet_swapcols(i64 k, i64 n, i64* p, double* T, i64 ldt, double* r1, i64 ldr1,
                        double tol, i64 niter_max) {
  const i64 nr = n - k;
  double* work = (double*)malloc(sizeof(double) * (size_t)(n > 2 * k ? n : 2 * k));
  i64 niter = 0;
  for (;;) {
    double tmax = 0.0;
    i64 idx = -1;
    for (i64 c = 0; c < nr; ++c)
      for (i64 r = 0; r < k; ++r) {
        const double t = fabs(T[r + c * ldt]);
        if (t < tmax) continue;
        tmax = t;
        idx = r + c * k;
      }
    if (tmax <= 1.0 + tol) break;
    if (niter == niter_max) break;
    ++niter;
    const i64 i = idx % k, j = idx / k;
    i64 tmp = p[i]; p[i] = p[k + j]; p[k + j] = tmp;
    for (i64 l = 0; l < k; ++l) { work[l] = T[l + j * ldt]; T[l + j * ldt] = 0.0; }
    work[i] -= 1.0;
    T[i + j * ldt] = 1.0;
    for (i64 l = 0; l < nr; ++l) work[k + l] = T[i + l * ldt];
    const double alpha = -1.0 / (1.0 + work[i]);
    for (i64 c = 0; c < nr; ++c) {
      const double y = alpha * work[k + c];
      for (i64 r = 0; r < k; ++r) T[r + c * ldt] += work[r] * y;
    }
    if (r1) {
      for (i64 r = 0; r < k; ++r) {
        double s = 0.0;
        for (i64 l = 0; l < k; ++l) s += r1[r + l * ldr1] * work[l];
        work[k + r] = s;
      }
      for (i64 r = 0; r < k; ++r) r1[r + i * ldr1] += work[k + r];
    }
  }
  free(work);
  return niter;
}

/* Largest divisor l of n with l <= k (srft_init, src/sketch.jl:353-358). */
i64 orc_srft_l(i64 n, i64 k) {
  i64 l = k;
  while (n % l > 0) --l;
  return l;
}

/*
 * Real SRFT applied on the left: C (k x n) = SRFT(k) * B (m x n) (src/sketch.jl:474-484 with
 * srft_init :353-364, srft_reshape! :378-385, srft_apply! :396-450).  d (m) are the +-1
 * signs, idx (k) 1-based sample indices.  Per column: X(l x mm) with X[j,q] = d.*x at
 * (j-1)*mm+q, real DFT of each column of X (length l, computed here as a direct DFT), then
 * for i = 1,3,5,...: f = idx[i]-1, w = e^{-2 pi i row/mm} e^{-2 pi i col/m}, y[i] (+ y[i+1])
 * = Re (Im) of sum_j w^j (X^[col,j]); the twiddle is advanced by repeated multiplication as
 * in the reference.  The "in == 0 || in == nnyq" test of :425 compares the Base function
 * `in` with integers and is always false, so only i == k emits a single (real) entry.
 */
void orc_srft_left(i64 m, i64 n, const double* B, i64 ldb, i64 k, const double* d,
                   const i64* idx, double* C, i64 ldc) {
  const i64 l = orc_srft_l(m, k), mm = m / l;
  double* X = (double*)malloc(sizeof(double) * (size_t)(l * mm));
  double* Xr = (double*)malloc(sizeof(double) * (size_t)(l * mm));
  double* Xi = (double*)malloc(sizeof(double) * (size_t)(l * mm));
  const double two_pi = 6.283185307179586476925286766559;
  for (i64 c = 0; c < n; ++c) {
    const double* x = B + c * ldb;
    double* y = C + c * ldc;
    for (i64 j = 0; j < l; ++j)
      for (i64 q = 0; q < mm; ++q) X[j + q * l] = d[j * mm + q] * x[j * mm + q];
    /* full complex spectrum of each column of X (the halfcomplex R2HC holds the same data) */
    for (i64 q = 0; q < mm; ++q)
      for (i64 f = 0; f < l; ++f) {
        double sr = 0.0, si = 0.0;
        for (i64 j = 0; j < l; ++j) {
          const double ang = -two_pi * (double)((f * j) % l) / (double)l;
          sr += X[j + q * l] * cos(ang);
          si += X[j + q * l] * sin(ang);
        }
        Xr[f + q * l] = sr;
        Xi[f + q * l] = si;
      }
    i64 i = 0; /* 0-based slot */
    while (i < k) {
      const i64 f = idx[i] - 1;
      const i64 row = f / l, col = f % l;
      /* w = wm^row * wn^col evaluated like the reference (power of the unit roots) */
      const double wmr = cos(-two_pi / (double)mm), wmi = sin(-two_pi / (double)mm);
      const double wnr = cos(-two_pi / (double)m), wni = sin(-two_pi / (double)m);
      double ar = 1.0, ai = 0.0, br = 1.0, bi = 0.0;
      for (i64 t = 0; t < row; ++t) { double r = ar * wmr - ai * wmi; ai = ar * wmi + ai * wmr; ar = r; }
      for (i64 t = 0; t < col; ++t) { double r = br * wnr - bi * wni; bi = br * wni + bi * wnr; br = r; }
      const double wr = ar * br - ai * bi, wi = ar * bi + ai * br;
      double sr = 1.0, si = 0.0, yr = 0.0, yi = 0.0;
      for (i64 q = 0; q < mm; ++q) {
        const double a = Xr[col + q * l], b = Xi[col + q * l];
        yr += sr * a - si * b;
        yi += sr * b + si * a;
        const double t = sr * wr - si * wi;
        si = sr * wi + si * wr;
        sr = t;
      }
      y[i] = yr;
      if (i == k - 1) break;
      y[i + 1] = yi;
      i += 2;
    }
  }
  free(X); free(Xr); free(Xi);
}

/* Random row subset on the left: C[i,:] = B[r[i],:] (src/sketch.jl:248-257), r 1-based. */
void orc_subset_left(i64 m, i64 n, const double* B, i64 ldb, i64 k, const i64* r, double* C,
                     i64 ldc) {
  (void)m;
  for (i64 c = 0; c < n; ++c)
    for (i64 i = 0; i < k; ++i) C[i + c * ldc] = B[(r[i] - 1) + c * ldb];
}

/*
 * Sparse random Gaussian on the left (src/sketch.jl:571-589): row i of S has
 * p_i = fld(m-i,k)+1 nonzeros s at rows r[idx+1 .. idx+p_i] of the permutation r (1-based),
 * C[i,j] = sum_l s[idx+l] * B[r[idx+l], j], summed in l order.  s holds the m coefficients
 * back to back (row 1's p_1 values first).
 */
void orc_sprn_left(i64 m, i64 n, const double* B, i64 ldb, i64 k, const i64* r, const double* s,
                   double* C, i64 ldc) {
  for (i64 c = 0; c < n; ++c) {
    i64 idx = 0;
    for (i64 i = 1; i <= k; ++i) {
      const i64 num = m - i; /* Julia fld: floor division, also for m < k */
      const i64 p = (num >= 0 ? num / k : -((-num + k - 1) / k)) + 1;
      double acc = 0.0;
      for (i64 l = 0; l < p; ++l) acc += s[idx + l] * B[(r[idx + l] - 1) + c * ldb];
      C[(i - 1) + c * ldc] = acc;
      idx += p;
    }
  }
}
