// qrcp_global.hip -- the rank-adaptive pivoted QR (geqp3_adap!, src/pqr.jl:348-418, with the
// dlaqps steps of src/lapack.jl:117-139) for sketches too large for the register-resident
// kernel of qrcp.hip (e.g. order 1032 or 2056 on thousands of columns).
//
// The sketch stays in HBM (column-major, in place, LAPACK layout) and each pivot step is two
// launches on the context's stream, with every decision kept on the device:
//   qg_step   one workgroup: dlaqps block bookkeeping and the stop rule for the step just
//             finished, the first position of the largest partial norm (idamax), LAPACK's
//             column swap (data, jpvt, vn1, vn2), dlarfg on the pivot column (R's diagonal in
//             place, v below it, a contiguous copy of v);
//   qg_apply  one wave per trailing column (qg_applyw: two for 1025..4096 rows, four up to
//             8192): H applied, LAWN 176 downdate; a column whose downdate fails the tol3z
//             test gets its norm recomputed and ends the block.
// Once the stop rule fires, the remaining launches return at once.  The arithmetic and the
// control follow qrcp_kernel, so both give the same pivots up to rounding.
#include "lra_internal.h"
#include "lra_wave.h"

#include <algorithm>
#include <climits>
#include <cstring>

namespace {

struct GState {
  int s, j0, jb, done, k, early, pvt, flag;
  double ptol;
};

constexpr int GT = 1024;  // threads of the single-workgroup kernels

// column norms (one wave per column), jpvt = 1..n
__global__ __launch_bounds__(256) void qg_norms(int m, int n, const double* __restrict__ B,
                                                i64 ldb, double* vn1, double* vn2, i64* jpvt) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  const double* c = B + (i64)j * ldb;
  double ss = 0.0;
  for (int i = lane; i < m; i += 64) ss += c[i] * c[i];
  ss = group_sum(ss, 64);
  if (lane == 0) {
    vn1[j] = vn2[j] = sqrt(ss);
    jpvt[j] = j + 1;
  }
}

// ptol = max(atol, rtol * max_j vn1[j]); state reset
__global__ __launch_bounds__(GT) void qg_init(int n, const double* vn1, double atol, double rtol,
                                              int kmax, int nb, GState* st) {
  __shared__ double sm[GT / 64];
  double mx = 0.0;
  for (int j = threadIdx.x; j < n; j += GT) mx = fmax(mx, vn1[j]);
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m2 = sm[0];
    for (int w = 1; w < GT / 64; ++w) m2 = fmax(m2, sm[w]);
    st->ptol = fmax(atol, rtol * m2);
    st->s = 0;
    st->j0 = 0;
    st->jb = min(nb, kmax);
    st->done = kmax <= 0 ? 1 : 0;
    st->k = kmax > 0 ? kmax : 0;
    st->early = 0;
    st->pvt = 0;
    st->flag = 0;
  }
}

// step s: bookkeeping of step s-1, then the pivot of step s; returns the pivot position, or
// -1 (for every thread of the block) once the factorization has stopped
__device__ int pick_body(int s, int n, int kmax, int nb, int stop_after, const double* vn1,
                         const double* diag, GState* st) {
  __shared__ int s_done, s_pvt;
  __shared__ double sv[GT / 64];
  __shared__ int sp[GT / 64];
  if (threadIdx.x == 0) {
    int done = st->done;
    if (!done && s > 0) {
      const int last = s - 1;
      const bool blk_end = (last - st->j0 + 1 == st->jb) || (st->flag != 0) || s == kmax;
      if (blk_end) {
        const double ptol = st->ptol;
        bool stop = false;
        if (fabs(diag[last]) <= ptol) {
          int i = st->j0;
          for (; i <= last; ++i)
            if (fabs(diag[i]) <= ptol) break;
          if (i <= last) { st->k = i; stop = true; }
        }
        if (!stop && s == kmax) { st->k = kmax; stop = true; }
        if (!stop && stop_after > 0 && s >= stop_after) {
          st->k = s;
          st->early = 1;
          stop = true;
        }
        if (stop) {
          st->done = done = 1;
        } else {
          st->j0 = s;
          st->jb = min(nb, kmax - s);
        }
      }
    }
    if (!done && s == kmax) { st->k = kmax; st->done = done = 1; }
    st->flag = 0;
    s_done = done;
  }
  __syncthreads();
  if (s_done) return -1;
  // idamax over positions s..n-1: largest vn1, then lowest position
  double bv = -1.0;
  int bp = INT_MAX;
  for (int j = s + threadIdx.x; j < n; j += GT) {
    const double v = vn1[j];
    if (v > bv) { bv = v; bp = j; }  // strided order: the first max of this thread's share
  }
  double wv;
  const int wl = wave_argmax(bv, bp, 1, &wv);
  const int wp = __builtin_amdgcn_readlane(bp, wl < 0 ? 0 : wl);
  if ((threadIdx.x & 63) == 0) {
    sv[threadIdx.x >> 6] = wv;
    sp[threadIdx.x >> 6] = wl < 0 ? INT_MAX : wp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double cv = sv[0];
    int cp = sp[0];
    for (int w = 1; w < GT / 64; ++w)
      if (sv[w] > cv || (sv[w] == cv && sp[w] < cp)) { cv = sv[w]; cp = sp[w]; }
    if (cp < s || cp >= n) {
      // no candidate (NaN norms): stop here rather than pivot on a column that is not there
      st->k = s;
      st->done = 1;
      s_done = 1;
      cp = s;
    }
    st->pvt = cp;
    st->s = s;
    s_pvt = cp;
  }
  __syncthreads();
  return s_done ? -1 : s_pvt;
}

__global__ __launch_bounds__(GT) void qg_pick(int s, int n, int kmax, int nb, int stop_after,
                                              const double* vn1, const double* diag,
                                              GState* st) {
  (void)pick_body(s, n, kmax, nb, stop_after, vn1, diag, st);
}

// step s: the pivot (pick_body), then columns s and pvt swapped and H(s) built (dlarfg) in
// place, v copied
__global__ __launch_bounds__(GT) void qg_step(int s, int m, int n, int kmax, int nb,
                                              int stop_after, double* __restrict__ B, i64 ldb,
                                              i64* jpvt, double* vn1, double* vn2, double* tau,
                                              double* diag, double* v, GState* st) {
  __shared__ double sw[GT / 64];
  __shared__ double s_alpha;
  const int pvt = pick_body(s, n, kmax, nb, stop_after, vn1, diag, st);
  if (pvt < 0) return;
  double* cs = B + (i64)s * ldb;
  double* cp = B + (i64)pvt * ldb;
  // LAPACK's swap of columns s and pvt, fused with the first dlarfg pass: each thread keeps
  // its rows of the new column s (up to MAXR per thread) and sums the squares below row s
  constexpr int MAXR = 8;
  const bool regs = m <= GT * MAXR;
  double xr[MAXR];
  double ss = 0.0;
  if (regs) {
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
      const int i = (int)threadIdx.x + r * GT;
      double x = 0.0;
      if (i < m) {
        x = cp[i];
        if (pvt != s) {
          const double t = cs[i];
          cs[i] = x;
          cp[i] = t;
        }
      }
      xr[r] = x;
      if (i > s && i < m) ss += x * x;
      if (i == s) s_alpha = x;
    }
  } else {
    for (int i = threadIdx.x; i < m; i += GT) {
      const double x = cp[i];
      if (pvt != s) {
        const double t = cs[i];
        cs[i] = x;
        cp[i] = t;
      }
      if (i > s) ss += x * x;
      if (i == s) s_alpha = x;
    }
  }
  if (pvt != s && threadIdx.x == 0) {
    const i64 t = jpvt[pvt];
    jpvt[pvt] = jpvt[s];
    jpvt[s] = t;
    vn1[pvt] = vn1[s];
    vn2[pvt] = vn2[s];
  }
  ss = group_sum(ss, 64);
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = ss;
  __syncthreads();  // (also makes every thread's swapped rows visible to the others)
  double tot = 0.0;
  for (int w = 0; w < GT / 64; ++w) tot += sw[w];
  // dlarfg on cs[s..m-1]
  double alpha = s_alpha, xnorm = sqrt(tot), beta = alpha, t = 0.0, scal = 1.0, fac = 1.0;
  if (m - s > 1 && xnorm != 0.0) {
    beta = -copysign(lapy2(alpha, xnorm), alpha);
    const double safmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;
    int knt = 0;
    if (fabs(beta) < safmin) {
      const double rsafmn = 1.0 / safmin;  // a power of two: the rescaling is exact
      do {
        ++knt;
        fac *= rsafmn;
        beta *= rsafmn;
        alpha *= rsafmn;
      } while (fabs(beta) < safmin && knt < 20);
      __syncthreads();  // every thread has read sw
      double s2 = 0.0;
      for (int i = s + 1 + (int)threadIdx.x; i < m; i += GT) {
        const double x = cs[i] * fac;
        s2 += x * x;
      }
      s2 = group_sum(s2, 64);
      if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = s2;
      __syncthreads();
      double t2 = 0.0;
      for (int w = 0; w < GT / 64; ++w) t2 += sw[w];
      xnorm = sqrt(t2);
      beta = -copysign(lapy2(alpha, xnorm), alpha);
    }
    t = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
    for (int j = 0; j < knt; ++j) beta *= safmin;
  }
  if (regs) {
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
      const int i = (int)threadIdx.x + r * GT;
      if (i > s && i < m) {
        const double x = (xr[r] * fac) * scal;
        cs[i] = x;
        v[i] = x;
      }
    }
  } else {
    for (int i = s + 1 + (int)threadIdx.x; i < m; i += GT) {
      const double x = (cs[i] * fac) * scal;
      cs[i] = x;
      v[i] = x;
    }
  }
  if (threadIdx.x == 0) {
    cs[s] = beta;
    v[s] = 1.0;
    tau[s] = t;
    diag[s] = beta;
  }
}

// step s: columns s+1..n-1 (one wave each): apply H(s), downdate the partial norms.
// RPL > 0: the column's rows s..m-1 fit RPL values per lane, so they are read once into
// registers, used for the dot product and updated there (one read and one write of the
// trailing matrix per step); RPL == 0 streams the column twice.
template <int RPL>
__global__ __launch_bounds__(256) void qg_apply(int s, int m, int n, int lastrk,
                                                double* __restrict__ B, i64 ldb,
                                                const double* __restrict__ v,
                                                const double* __restrict__ tau, double* vn1,
                                                double* vn2, GState* st) {
  if (st->done) return;
  const int lane = threadIdx.x & 63;
  const int j = s + 1 + blockIdx.x * 4 + (threadIdx.x >> 6);
  const double tol3z = 1.0536712127723509e-08;  // sqrt(dlamch('E'))
  double ss = 0.0, x = 0.0;
  if constexpr (RPL > 0) {
    // v(s .. s + 64 RPL - 1) staged in LDS once per block (zero past m)
    __shared__ double sv[64 * RPL];
    for (int t = threadIdx.x; t < 64 * RPL; t += 256) sv[t] = s + t < m ? v[s + t] : 0.0;
    __syncthreads();
    if (j >= n) return;
    double* c = B + (i64)j * ldb;
    double cr[RPL];
    double w = 0.0;
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int i = s + lane + 64 * r;
      cr[r] = i < m ? c[i] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) w += sv[lane + 64 * r] * cr[r];
    w = group_sum(w, 64);
    const double tw = tau[s] * w;
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int i = s + lane + 64 * r;
      const double y = cr[r] - tw * sv[lane + 64 * r];
      if (i < m) {
        c[i] = y;
        if (i == s) x = y;
        else ss += y * y;
      }
    }
  } else {
    if (j >= n) return;
    double* c = B + (i64)j * ldb;
    double w = 0.0;
    for (int i = s + lane; i < m; i += 64) w += v[i] * c[i];
    w = group_sum(w, 64);
    const double tw = tau[s] * w;
    // update; each lane also sums the squares of its updated rows below s (the recompute
    // norm) and lane 0 keeps the new R(s, j), so nothing is read back from memory
    for (int i = s + lane; i < m; i += 64) {
      const double y = c[i] - tw * v[i];
      c[i] = y;
      if (i == s) x = y;
      else ss += y * y;
    }
  }
  x = __shfl(x, 0, 64);
  double a = vn1[j];
  if (s + 1 < lastrk && a != 0.0) {
    double temp = fabs(x) / a;
    temp = (1.0 + temp) * (1.0 - temp);
    temp = temp < 0.0 ? 0.0 : temp;
    const double q = a / vn2[j];
    const double temp2 = temp * q * q;
    if (temp2 <= tol3z) {
      ss = group_sum(ss, 64);
      if (lane == 0) {
        vn1[j] = vn2[j] = sqrt(ss);
        __hip_atomic_store(&st->flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else if (lane == 0) {
      vn1[j] = a * sqrt(temp);
    }
  }
}

// qg_apply for taller columns: NWC waves per column (rows s + 64 (NWC r + h) + lane in wave h),
// RPL values per lane, so rows up to 64 NWC RPL are read once and written once per step; the
// dot product and the recompute norm meet in LDS.
template <int NWC, int RPL>
__global__ __launch_bounds__(256) void qg_applyw(int s, int m, int n, int lastrk,
                                                 double* __restrict__ B, i64 ldb,
                                                 const double* __restrict__ v,
                                                 const double* __restrict__ tau, double* vn1,
                                                 double* vn2, GState* st) {
  if (st->done) return;
  constexpr int CPB = 4 / NWC;  // columns per block
  constexpr int ROWS = 64 * NWC * RPL;
  __shared__ double sv[ROWS];  // v(s .. s + ROWS - 1), zero past m
  __shared__ double sw[CPB][NWC], sss[CPB][NWC], sx[CPB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = wave % NWC, cl = wave / NWC;
  const int j = s + 1 + blockIdx.x * CPB + cl;
  const bool have = j < n;  // (no early exit: the barriers below involve every wave)
  const double tol3z = 1.0536712127723509e-08;  // sqrt(dlamch('E'))
  for (int t = threadIdx.x; t < ROWS; t += 256) sv[t] = s + t < m ? v[s + t] : 0.0;
  __syncthreads();
  double* c = B + (i64)(have ? j : 0) * ldb;
  double cr[RPL];
  double w = 0.0;
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int o = 64 * (NWC * r + h) + lane;
    cr[r] = (have && s + o < m) ? c[s + o] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < RPL; ++r) w += sv[64 * (NWC * r + h) + lane] * cr[r];
  w = group_sum(w, 64);
  if (lane == 0) sw[cl][h] = w;
  __syncthreads();
  double wt = 0.0;
#pragma unroll
  for (int q = 0; q < NWC; ++q) wt += sw[cl][q];
  const double tw = tau[s] * wt;
  double ss = 0.0, x = 0.0;
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int o = 64 * (NWC * r + h) + lane;
    const double y = cr[r] - tw * sv[o];
    if (have && s + o < m) {
      c[s + o] = y;
      if (o == 0) x = y;
      else ss += y * y;
    }
  }
  ss = group_sum(ss, 64);
  if (lane == 0) {
    sss[cl][h] = ss;
    if (h == 0) sx[cl] = x;  // row s is offset 0: lane 0 of the column's first wave
  }
  __syncthreads();
  if (!have || h != 0 || lane != 0) return;
  x = sx[cl];
  const double a = vn1[j];
  if (s + 1 < lastrk && a != 0.0) {
    double temp = fabs(x) / a;
    temp = (1.0 + temp) * (1.0 - temp);
    temp = temp < 0.0 ? 0.0 : temp;
    const double q = a / vn2[j];
    const double temp2 = temp * q * q;
    if (temp2 <= tol3z) {
      double ssum = 0.0;
#pragma unroll
      for (int q2 = 0; q2 < NWC; ++q2) ssum += sss[cl][q2];
      vn1[j] = vn2[j] = sqrt(ssum);
      __hip_atomic_store(&st->flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      vn1[j] = a * sqrt(temp);
    }
  }
}

__global__ void qg_finish(const GState* st, i64* k_out, int* early_out) {
  *k_out = st->k;
  if (early_out) *early_out = st->early;
}

// ---------------------------------------------------------------------------------------
// Complex (ComplexF64) steps: zlaqps' arithmetic (src/lapack.jl:117-139 calls zlaqps for
// complex element types).  Columns are interleaved (re, im) pairs; the norms, the diagonal
// and the stop rule are real.  H(s) = I - tau v v^H is applied as H^H from the left
// (A(:, j) -= conj(tau) v (v^H A(:, j))), and zlarfg makes R's diagonal real.

typedef double2 zc;
LRA_DEV zc zmul(zc a, zc b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
LRA_DEV zc zconj_mul(zc a, zc b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
LRA_DEV double zabs2(zc a) { return a.x * a.x + a.y * a.y; }
// 1 / z by Smith's scaling
LRA_DEV zc zrecip(zc z) {
  if (fabs(z.x) >= fabs(z.y)) {
    const double r = z.y / z.x, d = z.x + z.y * r;
    return make_double2(1.0 / d, -r / d);
  }
  const double r = z.x / z.y, d = z.y + z.x * r;
  return make_double2(r / d, -1.0 / d);
}
// sqrt(x^2 + y^2 + z^2) without overflow (dlapy3)
LRA_DEV double lapy3(double x, double y, double z) {
  const double xa = fabs(x), ya = fabs(y), za = fabs(z);
  const double w = fmax(xa, fmax(ya, za));
  if (w == 0.0) return xa + ya + za;
  return w * sqrt((xa / w) * (xa / w) + (ya / w) * (ya / w) + (za / w) * (za / w));
}

__global__ __launch_bounds__(256) void zq_norms(int m, int n, const zc* __restrict__ B, i64 ldb,
                                                double* vn1, double* vn2, i64* jpvt) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  const zc* c = B + (i64)j * ldb;
  double ss = 0.0;
  for (int i = lane; i < m; i += 64) ss += zabs2(c[i]);
  ss = group_sum(ss, 64);
  if (lane == 0) {
    vn1[j] = vn2[j] = sqrt(ss);
    jpvt[j] = j + 1;
  }
}

// step s: the pivot (pick_body), then columns s and pvt swapped, zlarfg on column s in
// place, v copied
__global__ __launch_bounds__(GT) void zq_step(int s, int m, int n, int kmax, int nb,
                                              int stop_after, zc* __restrict__ B, i64 ldb,
                                              i64* jpvt, double* vn1, double* vn2, zc* tau,
                                              double* diag, zc* v, GState* st) {
  __shared__ double sw[GT / 64];
  __shared__ zc s_alpha;
  const int pvt = pick_body(s, n, kmax, nb, stop_after, vn1, diag, st);
  if (pvt < 0) return;
  zc* cs = B + (i64)s * ldb;
  if (pvt != s) {
    zc* cp = B + (i64)pvt * ldb;
    for (int i = threadIdx.x; i < m; i += GT) {
      const zc t = cs[i];
      cs[i] = cp[i];
      cp[i] = t;
    }
    if (threadIdx.x == 0) {
      const i64 t = jpvt[pvt];
      jpvt[pvt] = jpvt[s];
      jpvt[s] = t;
      vn1[pvt] = vn1[s];
      vn2[pvt] = vn2[s];
    }
  }
  __syncthreads();
  double ss = 0.0;
  for (int i = s + 1 + (int)threadIdx.x; i < m; i += GT) ss += zabs2(cs[i]);
  ss = group_sum(ss, 64);
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = ss;
  if (threadIdx.x == 0) s_alpha = cs[s];
  __syncthreads();
  double tot = 0.0;
  for (int w = 0; w < GT / 64; ++w) tot += sw[w];
  double alphr = s_alpha.x, alphi = s_alpha.y, xnorm = sqrt(tot), fac = 1.0;
  zc t = make_double2(0.0, 0.0), scal = make_double2(1.0, 0.0);
  double beta = alphr;
  bool reflect = m - s > 1 ? (xnorm != 0.0 || alphi != 0.0) : (alphi != 0.0);
  if (reflect) {
    beta = -copysign(lapy3(alphr, alphi, xnorm), alphr);
    const double safmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;
    int knt = 0;
    if (fabs(beta) < safmin) {
      const double rsafmn = 1.0 / safmin;  // a power of two: the rescaling is exact
      do {
        ++knt;
        fac *= rsafmn;
        beta *= rsafmn;
        alphr *= rsafmn;
        alphi *= rsafmn;
      } while (fabs(beta) < safmin && knt < 20);
      __syncthreads();  // every thread has read sw
      double s2 = 0.0;
      for (int i = s + 1 + (int)threadIdx.x; i < m; i += GT) {
        const zc x = cs[i];
        s2 += zabs2(make_double2(x.x * fac, x.y * fac));
      }
      s2 = group_sum(s2, 64);
      if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = s2;
      __syncthreads();
      double t2 = 0.0;
      for (int w = 0; w < GT / 64; ++w) t2 += sw[w];
      xnorm = sqrt(t2);
      beta = -copysign(lapy3(alphr, alphi, xnorm), alphr);
    }
    t = make_double2((beta - alphr) / beta, -alphi / beta);
    scal = zrecip(make_double2(alphr - beta, alphi));
    for (int j = 0; j < knt; ++j) beta *= safmin;
  }
  for (int i = s + 1 + (int)threadIdx.x; i < m; i += GT) {
    const zc x0 = cs[i];
    const zc x = zmul(make_double2(x0.x * fac, x0.y * fac), scal);
    cs[i] = x;
    v[i] = x;
  }
  if (threadIdx.x == 0) {
    cs[s] = reflect ? make_double2(beta, 0.0) : s_alpha;
    v[s] = make_double2(1.0, 0.0);
    tau[s] = t;
    diag[s] = reflect ? beta : hypot(s_alpha.x, s_alpha.y);
  }
}

// step s: columns s+1..n-1 (one wave each): a_j -= conj(tau) v (v^H a_j), norm downdate
__global__ __launch_bounds__(256) void zq_apply(int s, int m, int n, int lastrk,
                                                zc* __restrict__ B, i64 ldb,
                                                const zc* __restrict__ v,
                                                const zc* __restrict__ tau, double* vn1,
                                                double* vn2, GState* st) {
  if (st->done) return;
  const int lane = threadIdx.x & 63;
  const int j = s + 1 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  const double tol3z = 1.0536712127723509e-08;  // sqrt(dlamch('E'))
  zc* c = B + (i64)j * ldb;
  double wr = 0.0, wi = 0.0;
  for (int i = s + lane; i < m; i += 64) {
    const zc p = zconj_mul(v[i], c[i]);
    wr += p.x;
    wi += p.y;
  }
  wr = group_sum(wr, 64);
  wi = group_sum(wi, 64);
  const zc ct = make_double2(tau[s].x, -tau[s].y);
  const zc tw = zmul(ct, make_double2(wr, wi));
  double ss = 0.0;
  zc x = make_double2(0.0, 0.0);
  for (int i = s + lane; i < m; i += 64) {
    const zc vv = v[i];
    const zc p = zmul(tw, vv);
    const zc y = make_double2(c[i].x - p.x, c[i].y - p.y);
    c[i] = y;
    if (i == s) x = y;
    else ss += zabs2(y);
  }
  x.x = __shfl(x.x, 0, 64);
  x.y = __shfl(x.y, 0, 64);
  double a = vn1[j];
  if (s + 1 < lastrk && a != 0.0) {
    double temp = hypot(x.x, x.y) / a;
    temp = (1.0 + temp) * (1.0 - temp);
    temp = temp < 0.0 ? 0.0 : temp;
    const double q = a / vn2[j];
    const double temp2 = temp * q * q;
    if (temp2 <= tol3z) {
      ss = group_sum(ss, 64);
      if (lane == 0) {
        vn1[j] = vn2[j] = sqrt(ss);
        __hip_atomic_store(&st->flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else if (lane == 0) {
      vn1[j] = a * sqrt(temp);
    }
  }
}

}  // namespace

// Complex rank-adaptive pivoted QR (geqp3_adap!, src/pqr.jl:348-418, with zlaqps' steps) on
// an m x n ComplexF64 matrix in HBM (interleaved re/im, leading dimension ldb in complex
// elements): jpvt (n, 1-based), tau (kmax complex), the rank into k_out.  stop_after as in
// QrcpCall.
int lra_zqrcp_global(lra_ctx* c, double* Bz, i64 ldb, int m, int n, int kmax, int nb,
                     double atol, double rtol, int stop_after, i64* jpvt, double* tau,
                     i64* k_out) {
  zc* B = reinterpret_cast<zc*>(Bz);
  if (kmax <= 0) {
    if (n > 0) LRA_CHECK(lra_launch_fill_i64(c->stream, jpvt, n, 1));
    LRA_HIP(hipMemsetAsync(k_out, 0, sizeof(i64), c->stream));
    return LRA_OK;
  }
  double* w = nullptr;
  // vn1, vn2, diag (kmax), v (m complex), state
  const size_t nd = 2 * (size_t)n + (size_t)kmax + 2 * (size_t)m + 16;
  LRA_CHECK(lra_wsT(c, WS_QG, nd, &w));
  double* vn1 = w;
  double* vn2 = vn1 + n;
  double* diag = vn2 + n;
  zc* v = reinterpret_cast<zc*>(diag + kmax + (kmax & 1));
  GState* st = reinterpret_cast<GState*>(v + m);
  const int lastrk = std::min(m, n);
  hipStream_t s0 = c->stream;
  zc* zt = reinterpret_cast<zc*>(tau);
  hipLaunchKernelGGL(zq_norms, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s0, m, n, B, ldb,
                     vn1, vn2, jpvt);
  hipLaunchKernelGGL(qg_init, dim3(1), dim3(GT), 0, s0, n, vn1, atol, rtol, kmax, nb, st);
  LRA_HIP(hipGetLastError());
  for (int s = 0; s <= kmax; ++s) {
    if (s == kmax) {
      hipLaunchKernelGGL(qg_pick, dim3(1), dim3(GT), 0, s0, s, n, kmax, nb, stop_after, vn1,
                         diag, st);
      break;
    }
    hipLaunchKernelGGL(zq_step, dim3(1), dim3(GT), 0, s0, s, m, n, kmax, nb, stop_after, B, ldb,
                       jpvt, vn1, vn2, zt, diag, v, st);
    if (s + 1 < n)
      hipLaunchKernelGGL(zq_apply, dim3((unsigned)((n - s - 1 + 3) / 4)), dim3(256), 0, s0, s,
                         m, n, lastrk, B, ldb, v, zt, vn1, vn2, st);
    LRA_HIP(hipGetLastError());
    if ((s & 63) == 63) {
      int done = 0;
      LRA_HIP(hipMemcpyAsync(c->host_pinned, &st->done, sizeof(int), hipMemcpyDeviceToHost, s0));
      LRA_HIP(hipStreamSynchronize(s0));
      std::memcpy(&done, c->host_pinned, sizeof(int));
      if (done) break;
    }
  }
  hipLaunchKernelGGL(qg_finish, dim3(1), dim3(1), 0, s0, st, k_out, nullptr);
  LRA_HIP(hipGetLastError());
  return LRA_OK;
}

// geqp3_adap (mode 0 of QrcpCall) on a sketch of any size, in HBM.
int lra_qrcp_global(lra_ctx* c, const QrcpCall& q) {
  if (q.mode != 0) {
    lra_set_error("qrcp: laqps panel %d x %d does not fit the register-resident engine", q.m,
                  q.n);
    return LRA_E_UNSUPPORTED;
  }
  if (q.offset != 0) {
    lra_set_error("qrcp: offset %d on the global-memory path", q.offset);
    return LRA_E_UNSUPPORTED;
  }
  const int m = q.m, n = q.n;
  double* w = nullptr;
  // vn1, vn2, diag (kmax), v (m), state
  const size_t nd = 2 * (size_t)n + (size_t)std::max(q.kmax, 1) + (size_t)m + 16;
  LRA_CHECK(lra_wsT(c, WS_QG, nd, &w));
  double* vn1 = w;
  double* vn2 = vn1 + n;
  double* diag = vn2 + n;
  double* v = diag + std::max(q.kmax, 1);
  GState* st = reinterpret_cast<GState*>(v + m);
  const int lastrk = std::min(m, n);
  hipStream_t s0 = c->stream;
  hipLaunchKernelGGL(qg_norms, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s0, m, n, q.A, q.lda,
                     vn1, vn2, q.jpvt);
  hipLaunchKernelGGL(qg_init, dim3(1), dim3(GT), 0, s0, n, vn1, q.atol, q.rtol, q.kmax, q.nb,
                     st);
  LRA_HIP(hipGetLastError());
  for (int s = 0; s <= q.kmax; ++s) {
    if (s == q.kmax) {  // the bookkeeping of the last step
      hipLaunchKernelGGL(qg_pick, dim3(1), dim3(GT), 0, s0, s, n, q.kmax, q.nb, q.stop_after,
                         vn1, diag, st);
      break;
    }
    hipLaunchKernelGGL(qg_step, dim3(1), dim3(GT), 0, s0, s, m, n, q.kmax, q.nb, q.stop_after,
                       q.A, q.lda, q.jpvt, vn1, vn2, q.tau, diag, v, st);
    if (s + 1 < n) {
      const dim3 grid((unsigned)((n - s - 1 + 3) / 4));
      const int rows = m - s;  // rows s..m-1 of each trailing column
      if (rows <= 64 * 8)
        hipLaunchKernelGGL(qg_apply<8>, grid, dim3(256), 0, s0, s, m, n, lastrk, q.A, q.lda, v,
                           q.tau, vn1, vn2, st);
      else if (rows <= 64 * 16)
        hipLaunchKernelGGL(qg_apply<16>, grid, dim3(256), 0, s0, s, m, n, lastrk, q.A, q.lda, v,
                           q.tau, vn1, vn2, st);
      else if (rows <= 128 * 16)
        hipLaunchKernelGGL((qg_applyw<2, 16>), dim3((unsigned)((n - s - 1 + 1) / 2)), dim3(256),
                           0, s0, s, m, n, lastrk, q.A, q.lda, v, q.tau, vn1, vn2, st);
      else if (rows <= 128 * 24)
        hipLaunchKernelGGL((qg_applyw<2, 24>), dim3((unsigned)((n - s - 1 + 1) / 2)), dim3(256),
                           0, s0, s, m, n, lastrk, q.A, q.lda, v, q.tau, vn1, vn2, st);
      else if (rows <= 128 * 32)
        hipLaunchKernelGGL((qg_applyw<2, 32>), dim3((unsigned)((n - s - 1 + 1) / 2)), dim3(256),
                           0, s0, s, m, n, lastrk, q.A, q.lda, v, q.tau, vn1, vn2, st);
      else if (rows <= 256 * 32)
        hipLaunchKernelGGL((qg_applyw<4, 32>), dim3((unsigned)(n - s - 1)), dim3(256), 0, s0, s,
                           m, n, lastrk, q.A, q.lda, v, q.tau, vn1, vn2, st);
      else
        hipLaunchKernelGGL(qg_apply<0>, grid, dim3(256), 0, s0, s, m, n, lastrk, q.A, q.lda, v,
                           q.tau, vn1, vn2, st);
    }
    LRA_HIP(hipGetLastError());
    // (every 64 steps, stop enqueueing once the device has stopped)
    if ((s & 63) == 63) {
      int done = 0;
      LRA_HIP(hipMemcpyAsync(c->host_pinned, &st->done, sizeof(int), hipMemcpyDeviceToHost, s0));
      LRA_HIP(hipStreamSynchronize(s0));
      std::memcpy(&done, c->host_pinned, sizeof(int));
      if (done) break;
    }
  }
  hipLaunchKernelGGL(qg_finish, dim3(1), dim3(1), 0, s0, st, q.k_out, q.early_out);
  LRA_HIP(hipGetLastError());
  return LRA_OK;
}
