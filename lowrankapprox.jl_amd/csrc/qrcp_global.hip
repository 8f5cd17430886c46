// qrcp_global.hip -- the rank-adaptive pivoted QR (geqp3_adap!, src/pqr.jl:348-418, with the
// dlaqps steps of src/lapack.jl:117-139) for sketches too large for the register-resident
// kernel of qrcp.hip (e.g. order 1032 or 2056 on thousands of columns).
//
// The sketch stays in HBM (column-major, in place).  Up to 8192 rows (with the uncached
// exchange area) each pivot step is ONE launch, `qg_fused` (below): every workgroup rebuilds
// the step's reflector itself, and the workgroup that finishes last picks the next pivot.
// Otherwise (or with LRA_QG_FUSED=0) each step is two launches on the context's stream, with
// every decision kept on the device:
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

// column norms (one wave per column), jpvt = 1..n (and the identity position map)
__global__ __launch_bounds__(256) void qg_norms(int m, int n, const double* __restrict__ B,
                                                i64 ldb, double* vn1, double* vn2, i64* jpvt,
                                                int* at) {
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
    if (at) at[j] = j;
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
// Fused steps: one launch per pivot step (sketches up to 8192 rows, with the uncached
// exchange area).  Columns stay where they are in memory (storage index c); the map `at` from
// LAPACK position to storage column lives on the device, and the data is permuted to LAPACK
// positions once at the end.  Launch s, every workgroup:
//   1. loads its trailing columns (positions s+1+CPB*b ..) and the pivot column at[s], and
//      rebuilds H(s) (dlarfg) in LDS -- the same reduction order in every workgroup, so all
//      copies are bitwise the same and no launch of its own is needed for it;
//   2. applies H(s) and downdates the norms (the arithmetic of qg_apply / qg_applyw);
//   3. publishes its best candidate (largest norm, then lowest position) as a tagged
//      16-byte partial in the uncached area, then takes a ticket from a two-level tree of
//      counters (64 sub-counters on their own lines, then the top one: thousands of atomics
//      on one uncached word were serialized).
// The workgroup that completes the top counter is the last of the grid: it reads every
// partial (a store still in flight is re-read until its tag arrives), runs the block
// bookkeeping and stop rule of step s (pick_body's, for step s+1), and swaps the next pivot
// into position s+1 of `at`.  No workgroup ever waits for one that has not finished its
// columns, so the launch drains beside any other kernel, grid-synchronising ones included.
// (A designated last workgroup that polled the partials instead could wait for workgroups
// that a co-resident cooperative grid kept from being dispatched.)  Workgroup 0 stores v(s), tau(s) and R(s,s); the pivot column itself is
// rewritten one launch later, when no workgroup reads it any more.

struct FState {
  int j0, jb, done, k, early, sprev, pad0, pad1;
  double ptol;
};

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// The exchange area of the fused steps: kSubCtr + 1 ticket counters kCtrStride words apart
// (the top one first), then one 16-byte partial per block from byte kPartOff.
constexpr unsigned kSubCtr = 64, kCtrStride = 16;
constexpr size_t kPartOff = 8192;

// partial tag: step tag (15 bits, never 0) | recompute flag << 15 | position << 16 |
// storage column << 40; positions and columns below 2^24 - 1
constexpr uint64_t F_NONE = (1ull << 24) - 1;
LRA_DEV uint64_t fstep_tag(int s) { return (uint64_t)(s % 32767) + 1; }
LRA_DEV uint64_t fpack(int s, int flag, int pos, int col) {
  const uint64_t pp = pos < 0 ? F_NONE : (uint64_t)pos;
  const uint64_t cc = col < 0 ? F_NONE : (uint64_t)col;
  return fstep_tag(s) | ((uint64_t)(flag != 0) << 15) | (pp << 16) | (cc << 40);
}

// ptol = max(atol, rtol * max_j vn1[j]), the pivot of step 0 swapped into position 0 of `at`
// (the identity map, from qg_norms), state reset
__global__ __launch_bounds__(GT) void qg_fstart(int n, int kmax, int nb, const double* vn1,
                                                double atol, double rtol, FState* st, int* at) {
  __shared__ double sm[GT / 64], sv[GT / 64];
  __shared__ int sp[GT / 64];
  double mx = 0.0, bv = -1.0;
  int bp = INT_MAX;
  for (int j = threadIdx.x; j < n; j += GT) {
    const double v = vn1[j];
    mx = fmax(mx, v);
    if (v > bv) { bv = v; bp = j; }  // strided order: the first max of this thread's share
  }
  mx = wave_max(mx);
  double wv;
  const int wl = wave_argmax(bv, bp, 1, &wv);
  const int wp = __builtin_amdgcn_readlane(bp, wl < 0 ? 0 : wl);
  if ((threadIdx.x & 63) == 0) {
    sm[threadIdx.x >> 6] = mx;
    sv[threadIdx.x >> 6] = wv;
    sp[threadIdx.x >> 6] = wl < 0 ? INT_MAX : wp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m2 = sm[0], cv = sv[0];
    int cp = sp[0];
    for (int w = 1; w < GT / 64; ++w) {
      m2 = fmax(m2, sm[w]);
      if (sv[w] > cv || (sv[w] == cv && sp[w] < cp)) { cv = sv[w]; cp = sp[w]; }
    }
    st->ptol = fmax(atol, rtol * m2);
    st->j0 = 0;
    st->jb = min(nb, kmax);
    st->early = 0;
    st->sprev = -1;
    int done = kmax <= 0 ? 1 : 0;
    st->k = kmax > 0 ? kmax : 0;
    if (!done && (cp < 0 || cp >= n)) {  // no candidate (NaN norms)
      st->k = 0;
      done = 1;
    }
    if (!done && cp != 0) {
      const int t = at[0];
      at[0] = at[cp];
      at[cp] = t;
    }
    st->done = done;
  }
}

// The end of a fused step s, after the block's column candidates are in LDS (cval, cpos,
// ccol, cflag): the block's best (largest norm, then lowest position) as its tagged partial;
// then the grid's last block reads every partial, runs the bookkeeping and stop rule of step
// s (pick_body's, for step s+1; rdiag = |R(s,s)| of this step) and swaps the winner into
// position s+1 of `at`.
template <int CPB>
LRA_DEV void fused_finish(int s, int n, int kmax, int nb, int stop_after, double rdiag,
                          const double* cval, const int* cpos, const int* ccol, const int* cflag,
                          int* at, const double* diag, FState* st, u64x2* part,
                          unsigned* ctr, int* err) {
  __shared__ double tv[4];
  __shared__ int tp[4], tc[4], tf[4], tok[4];
  __shared__ int s_last;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const unsigned nsub = gridDim.x < kSubCtr ? gridDim.x : kSubCtr;
  if (tid == 0) {
    double bv = cval[0];
    int bp = cpos[0], bc = ccol[0], bf = cflag[0];
#pragma unroll
    for (int q = 1; q < CPB; ++q) {
      if (cval[q] > bv || (cval[q] == bv && cpos[q] < bp)) { bv = cval[q]; bp = cpos[q]; bc = ccol[q]; }
      bf |= cflag[q];
    }
    u64x2 e;
    e.x = (uint64_t)__double_as_longlong(bv);
    e.y = bv < 0.0 ? fpack(s, bf, -1, -1) : fpack(s, bf, bp, bc);
    part[blockIdx.x] = e;  // uncached: a plain store; its tag validates it for the reader
    asm volatile("" ::: "memory");
    // a two-level ticket: sub-counter blockIdx % nsub, then the top counter; the block that
    // completes the top one is the last of the grid
    const unsigned sub = blockIdx.x % nsub;
    const unsigned cnt = gridDim.x / nsub + (sub < gridDim.x % nsub ? 1u : 0u);
    int last = 0;
    if (__hip_atomic_fetch_add(ctr + kCtrStride * (1 + sub), 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) == cnt - 1)
      last = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             nsub - 1;
    s_last = last;
  }
  __syncthreads();
  if (!s_last) return;
  // every block has taken its tickets: the counters start again from zero for the next step
  if (tid <= (int)nsub)
    __hip_atomic_store(ctr + kCtrStride * tid, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

  // ---- the last block: every partial (loaded 8 at a time per thread, then validated; a
  // partial whose store is still in flight is re-read until its tag arrives) ----
  const int G = (int)gridDim.x;
  const uint64_t want = fstep_tag(s);
  double bv = -1.0;
  int bp = INT_MAX, bc = -1, bf = 0, ok = 1;
  for (int b0 = tid; b0 < G; b0 += 256 * 8) {
    u64x2 e[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int b = b0 + 256 * q;
      e[q] = b < G ? part[b] : u64x2{0, 0};
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int b = b0 + 256 * q;
      if (b >= G || !ok) continue;
      if ((e[q].y & 0x7FFF) != want) {  // not landed yet: bounded re-read
        const volatile u64x2* src = part + b;
        unsigned spins = 0;
        do {
          __builtin_amdgcn_s_sleep(1);
          e[q] = *src;
        } while ((e[q].y & 0x7FFF) != want && ++spins < LRA_SPIN_LIMIT);
        if ((e[q].y & 0x7FFF) != want) { ok = 0; continue; }
      }
      const double v = __longlong_as_double((long long)e[q].x);
      const uint64_t pp = (e[q].y >> 16) & F_NONE;
      const int pos = pp == F_NONE ? INT_MAX : (int)pp;
      if (v > bv || (v == bv && pos < bp)) {
        bv = v;
        bp = pos;
        bc = (int)((e[q].y >> 40) & F_NONE);
      }
      bf |= (int)((e[q].y >> 15) & 1);
    }
  }
  double wv;
  const int wl = wave_argmax(bv, bp, 1, &wv);
  const int wp = __builtin_amdgcn_readlane(bp, wl < 0 ? 0 : wl);
  const int wc = __builtin_amdgcn_readlane(bc, wl < 0 ? 0 : wl);
  const int wf = __any(bf) ? 1 : 0;
  const int wok = __all(ok) ? 1 : 0;
  if (lane == 0) {
    tv[wave] = wv;
    tp[wave] = wl < 0 ? INT_MAX : wp;
    tc[wave] = wl < 0 ? -1 : wc;
    tf[wave] = wf;
    tok[wave] = wok;
  }
  __syncthreads();
  if (tid != 0) return;
  double cv = tv[0];
  int cp = tp[0], cc = tc[0], flag = tf[0], allok = tok[0];
  for (int q = 1; q < 4; ++q) {
    if (tv[q] > cv || (tv[q] == cv && tp[q] < cp)) { cv = tv[q]; cp = tp[q]; cc = tc[q]; }
    flag |= tf[q];
    allok &= tok[q];
  }
  st->sprev = s;
  if (!allok) {  // a partial never landed: stop with an error
    __hip_atomic_store(err, LRA_E_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    st->k = s + 1;
    st->done = 1;
    return;
  }
  // bookkeeping of step s (pick_body's at step s+1; R(s,s) is this launch's beta)
  const int s1 = s + 1;
  int done = 0;
  const bool blk_end = (s - st->j0 + 1 == st->jb) || flag != 0 || s1 == kmax;
  if (blk_end) {
    const double ptol = st->ptol;
    bool stop = false;
    if (fabs(rdiag) <= ptol) {
      int i = st->j0;
      for (; i < s; ++i)
        if (fabs(diag[i]) <= ptol) break;
      st->k = i;
      stop = true;
    }
    if (!stop && s1 == kmax) { st->k = kmax; stop = true; }
    if (!stop && stop_after > 0 && s1 >= stop_after) {
      st->k = s1;
      st->early = 1;
      stop = true;
    }
    if (stop) {
      done = 1;
    } else {
      st->j0 = s1;
      st->jb = min(nb, kmax - s1);
    }
  }
  if (!done && s1 == kmax) { st->k = kmax; done = 1; }
  if (!done) {
    if (cp < s1 || cp >= n || cc < 0) {  // no candidate (NaN norms)
      st->k = s1;
      done = 1;
    } else if (cp != s1) {  // LAPACK's swap of positions s+1 and the pivot's
      at[cp] = at[s1];
      at[s1] = cc;
    }
  }
  st->done = done;
}

// one fused step s; NWC waves per column, RPL rows per lane (rows s..m-1 fit 64 NWC RPL)
template <int NWC, int RPL>
__global__ __launch_bounds__(256) void qg_fused(int s, int m, int n, int kmax, int nb,
                                                int stop_after, int lastrk,
                                                double* __restrict__ B, i64 ldb, int* at,
                                                double* vn1, double* vn2, double* tau,
                                                double* diag, double* vbuf, FState* st,
                                                u64x2* part, unsigned* ctr, int* err) {
  constexpr int CPB = 4 / NWC;  // columns per block
  constexpr int ROWS = 64 * NWC * RPL;
  __shared__ double sv[ROWS];  // the pivot column, then v(s .. s + ROWS - 1), zero past m
  __shared__ double sred[4];
  __shared__ double sw[CPB][NWC], sss[CPB][NWC], sx[CPB], cval[CPB];
  __shared__ int cpos[CPB], ccol[CPB], cflag[CPB];
  if (st->done) return;  // (set by an earlier launch: the same for every thread)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int h = wave % NWC, cl = wave / NWC;
  const int p = s + 1 + (int)blockIdx.x * CPB + cl;  // trailing position of this column
  const bool have = p < n;
  // (p and so col are the same in every lane of a wave: scalar bases for the column accesses)
  const int col = __builtin_amdgcn_readfirstlane(have ? at[p] : 0);
  const int len = m - s;  // rows s..m-1 (len <= ROWS)
  const double tol3z = 1.0536712127723509e-08;  // sqrt(dlamch('E'))
  double* c = B + (i64)col * ldb + s;
  // H(s) from the pivot column (dlarfg on rows s..m-1)
  const double* pc = B + (i64)__builtin_amdgcn_readfirstlane(at[s]) * ldb + s;
  double ss = 0.0;
#pragma unroll 4
  for (int t = tid; t < ROWS; t += 256) {
    const double x = t < len ? pc[t] : 0.0;
    sv[t] = x;
    ss += t >= 1 ? x * x : 0.0;
  }
  ss = group_sum(ss, 64);
  if (lane == 0) sred[wave] = ss;
  __syncthreads();
  const double alpha0 = sv[0];
  double alpha = alpha0, xnorm = sqrt((sred[0] + sred[1]) + (sred[2] + sred[3]));
  double beta = alpha0, tu = 0.0, scal = 1.0, fac = 1.0;
  if (len > 1 && xnorm != 0.0) {
    beta = -copysign(lapy2(alpha, xnorm), alpha);
    const double safmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;
    int knt = 0;
    if (fabs(beta) < safmin) {  // (the same in every thread)
      const double rsafmn = 1.0 / safmin;  // a power of two: the rescaling is exact
      do {
        ++knt;
        fac *= rsafmn;
        beta *= rsafmn;
        alpha *= rsafmn;
      } while (fabs(beta) < safmin && knt < 20);
      __syncthreads();  // every thread has read sred
      double s2 = 0.0;
      for (int t = tid; t < ROWS; t += 256) {
        const double x = sv[t] * fac;
        s2 += (t >= 1 && t < len) ? x * x : 0.0;
      }
      s2 = group_sum(s2, 64);
      if (lane == 0) sred[wave] = s2;
      __syncthreads();
      xnorm = sqrt((sred[0] + sred[1]) + (sred[2] + sred[3]));
      beta = -copysign(lapy2(alpha, xnorm), alpha);
    }
    tu = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
    for (int j = 0; j < knt; ++j) beta *= safmin;
  }
  __syncthreads();  // every thread has read sv[0]
#pragma unroll 4
  for (int t = tid; t < ROWS; t += 256) sv[t] = t == 0 ? 1.0 : (sv[t] * fac) * scal;
  __syncthreads();
  if (blockIdx.x == 0) {
    // v(s), tau(s), R(s,s); the pivot column of step s-1 gets R(s-1,s-1) and its v
    double* vb = vbuf + (size_t)(s & 1) * m;
    for (int t = tid; t < len; t += 256) vb[s + t] = sv[t];
    if (tid == 0) {
      tau[s] = tu;
      diag[s] = beta;
    }
    if (s > 0) {
      const double* vp = vbuf + (size_t)((s - 1) & 1) * m;
      double* d = B + (i64)at[s - 1] * ldb;
      for (int t = s - 1 + tid; t < m; t += 256) d[t] = t == s - 1 ? diag[s - 1] : vp[t];
    }
  }
  // H(s) applied to this block's columns (loaded only now: holding them across the
  // reflector's reductions cost occupancy, which this HBM-bound update needs more)
  double cr[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int o = 64 * (NWC * r + h) + lane;
    cr[r] = (have && o < len) ? c[o] : 0.0;
  }
  double w = 0.0;
#pragma unroll
  for (int r = 0; r < RPL; ++r) w += sv[64 * (NWC * r + h) + lane] * cr[r];
  // re-read v from LDS below instead of keeping RPL values live across the reductions
  asm volatile("" ::: "memory");
  w = group_sum(w, 64);
  double wt = w;
  if constexpr (NWC > 1) {
    if (lane == 0) sw[cl][h] = w;
    __syncthreads();
    wt = 0.0;
#pragma unroll
    for (int q = 0; q < NWC; ++q) wt += sw[cl][q];
  }
  const double tw = tu * wt;
  double ss2 = 0.0, x = 0.0;
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int o = 64 * (NWC * r + h) + lane;
    const double y = cr[r] - tw * sv[o];
    if (have && o < len) {
      c[o] = y;
      if (o == 0) x = y;
      else ss2 += y * y;
    }
  }
  ss2 = group_sum(ss2, 64);
  if (lane == 0) {
    sss[cl][h] = ss2;
    if (h == 0) sx[cl] = x;  // row s is offset 0: lane 0 of the column's first wave
  }
  __syncthreads();
  // norm downdate (LAWN 176) or recompute, and the column's candidate entry
  if (h == 0 && lane == 0) {
    double v = -1.0;
    int fl = 0;
    if (have) {
      double a = vn1[col];
      if (s + 1 < lastrk && a != 0.0) {
        double temp = fabs(sx[cl]) / a;
        temp = (1.0 + temp) * (1.0 - temp);
        temp = temp < 0.0 ? 0.0 : temp;
        const double q = a / vn2[col];
        const double temp2 = temp * q * q;
        if (temp2 <= tol3z) {
          double ssum = 0.0;
#pragma unroll
          for (int q2 = 0; q2 < NWC; ++q2) ssum += sss[cl][q2];
          a = sqrt(ssum);
          vn1[col] = vn2[col] = a;
          fl = 1;
        } else {
          a *= sqrt(temp);
          vn1[col] = a;
        }
      }
      v = a >= 0.0 ? a : -1.0;  // (a NaN norm takes no part)
    }
    cval[cl] = v;
    cpos[cl] = have ? p : INT_MAX;
    ccol[cl] = col;
    cflag[cl] = fl;
  }
  __syncthreads();
  fused_finish<CPB>(s, n, kmax, nb, stop_after, beta, cval, cpos, ccol, cflag, at, diag, st,
                    part, ctr, err);
}

// after the last fused step: the last pivot column's R(s,s) and v, k and the early flag
__global__ __launch_bounds__(256) void qg_ffinish(int m, double* B, i64 ldb, const int* at,
                                                  const double* diag, const double* vbuf,
                                                  const FState* st, i64* k_out, int* early_out) {
  const int sp = st->sprev;
  if (sp >= 0) {
    const double* vp = vbuf + (size_t)(sp & 1) * m;
    double* d = B + (i64)at[sp] * ldb;
    for (int t = sp + (int)threadIdx.x; t < m; t += 256) d[t] = t == sp ? diag[sp] : vp[t];
  }
  if (threadIdx.x == 0) {
    *k_out = st->k;
    if (early_out) *early_out = st->early;
  }
}

// columns to LAPACK positions: B(:, p) = S(:, at[p]), jpvt[p] = at[p] + 1 (one block per p)
__global__ __launch_bounds__(256) void qg_gather(int m, const double* __restrict__ S, i64 lds,
                                                 double* __restrict__ B, i64 ldb, const int* at,
                                                 i64* jpvt) {
  const int p = blockIdx.x;
  const int c = at[p];
  const double* src = S + (i64)c * lds;
  double* dst = B + (i64)p * ldb;
  for (int i = threadIdx.x; i < m; i += 256) dst[i] = src[i];
  if (threadIdx.x == 0) jpvt[p] = (i64)c + 1;
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
                                                double* vn1, double* vn2, i64* jpvt, int* at) {
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
    if (at) at[j] = j;
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

// One fused complex step s: zq_step's zlarfg and zq_apply's update and downdate in the
// structure of qg_fused (every block rebuilds H(s) from the pivot column at[s]; its columns are
// read into registers once and written once; the grid's last block picks the next pivot).
// vbuf holds R(s,s) at row s and v below it (v(s) = 1 is implicit).
template <int NWC, int RPL>
__global__ __launch_bounds__(256) void zq_fused(int s, int m, int n, int kmax, int nb,
                                                int stop_after, int lastrk,
                                                zc* __restrict__ B, i64 ldb, int* at,
                                                double* vn1, double* vn2, zc* tau,
                                                double* diag, zc* vbuf, FState* st,
                                                u64x2* part, unsigned* ctr, int* err) {
  constexpr int CPB = 4 / NWC;  // columns per block
  constexpr int ROWS = 64 * NWC * RPL;
  __shared__ zc sv[ROWS];  // the pivot column, then v(s .. s + ROWS - 1), zero past m
  __shared__ double sred[4];
  __shared__ double swr[CPB][NWC], swi[CPB][NWC], sss[CPB][NWC], cval[CPB];
  __shared__ zc sx[CPB];
  __shared__ int cpos[CPB], ccol[CPB], cflag[CPB];
  if (st->done) return;  // (set by an earlier launch: the same for every thread)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int h = wave % NWC, cl = wave / NWC;
  const int p = s + 1 + (int)blockIdx.x * CPB + cl;  // trailing position of this column
  const bool have = p < n;
  // (p and so col are the same in every lane of a wave: scalar bases for the column accesses)
  const int col = __builtin_amdgcn_readfirstlane(have ? at[p] : 0);
  const int len = m - s;  // rows s..m-1 (len <= ROWS)
  const double tol3z = 1.0536712127723509e-08;  // sqrt(dlamch('E'))
  zc* c = B + (i64)col * ldb + s;
  // H(s) from the pivot column (zlarfg on rows s..m-1)
  const zc* pc = B + (i64)__builtin_amdgcn_readfirstlane(at[s]) * ldb + s;
  double ss = 0.0;
#pragma unroll 4
  for (int t = tid; t < ROWS; t += 256) {
    const zc x = t < len ? pc[t] : make_double2(0.0, 0.0);
    sv[t] = x;
    ss += t >= 1 ? zabs2(x) : 0.0;
  }
  ss = group_sum(ss, 64);
  if (lane == 0) sred[wave] = ss;
  __syncthreads();
  const zc alpha0 = sv[0];
  double alphr = alpha0.x, alphi = alpha0.y;
  double xnorm = sqrt((sred[0] + sred[1]) + (sred[2] + sred[3])), fac = 1.0;
  zc tu = make_double2(0.0, 0.0), scal = make_double2(1.0, 0.0);
  double beta = alphr;
  const bool reflect = len > 1 ? (xnorm != 0.0 || alphi != 0.0) : (alphi != 0.0);
  if (reflect) {
    beta = -copysign(lapy3(alphr, alphi, xnorm), alphr);
    const double safmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;
    int knt = 0;
    if (fabs(beta) < safmin) {  // (the same in every thread)
      const double rsafmn = 1.0 / safmin;  // a power of two: the rescaling is exact
      do {
        ++knt;
        fac *= rsafmn;
        beta *= rsafmn;
        alphr *= rsafmn;
        alphi *= rsafmn;
      } while (fabs(beta) < safmin && knt < 20);
      __syncthreads();  // every thread has read sred
      double s2 = 0.0;
      for (int t = tid; t < ROWS; t += 256) {
        const zc x = sv[t];
        s2 += (t >= 1 && t < len) ? zabs2(make_double2(x.x * fac, x.y * fac)) : 0.0;
      }
      s2 = group_sum(s2, 64);
      if (lane == 0) sred[wave] = s2;
      __syncthreads();
      xnorm = sqrt((sred[0] + sred[1]) + (sred[2] + sred[3]));
      beta = -copysign(lapy3(alphr, alphi, xnorm), alphr);
    }
    tu = make_double2((beta - alphr) / beta, -alphi / beta);
    scal = zrecip(make_double2(alphr - beta, alphi));
    for (int j = 0; j < knt; ++j) beta *= safmin;
  }
  const zc rss = reflect ? make_double2(beta, 0.0) : alpha0;  // R(s,s)
  const double dss = reflect ? beta : hypot(alpha0.x, alpha0.y);
  __syncthreads();  // every thread has read sv[0]
#pragma unroll 4
  for (int t = tid; t < ROWS; t += 256) {
    const zc x = sv[t];
    sv[t] = t == 0 ? make_double2(1.0, 0.0) : zmul(make_double2(x.x * fac, x.y * fac), scal);
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    // R(s,s), v(s+1..), tau(s), |R(s,s)|; the pivot column of step s-1 rewritten
    zc* vb = vbuf + (size_t)(s & 1) * m;
    for (int t = tid; t < len; t += 256) vb[s + t] = t == 0 ? rss : sv[t];
    if (tid == 0) {
      tau[s] = tu;
      diag[s] = dss;
    }
    if (s > 0) {
      const zc* vp = vbuf + (size_t)((s - 1) & 1) * m;
      zc* d = B + (i64)at[s - 1] * ldb;
      for (int t = s - 1 + tid; t < m; t += 256) d[t] = vp[t];
    }
  }
  // a_j -= conj(tau) v (v^H a_j) on this block's columns (loaded only now)
  zc cr[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int o = 64 * (NWC * r + h) + lane;
    cr[r] = (have && o < len) ? c[o] : make_double2(0.0, 0.0);
  }
  double wr = 0.0, wi = 0.0;
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const zc q = zconj_mul(sv[64 * (NWC * r + h) + lane], cr[r]);
    wr += q.x;
    wi += q.y;
  }
  // re-read v from LDS below instead of keeping RPL values live across the reductions
  asm volatile("" ::: "memory");
  wr = group_sum(wr, 64);
  wi = group_sum(wi, 64);
  if constexpr (NWC > 1) {
    if (lane == 0) {
      swr[cl][h] = wr;
      swi[cl][h] = wi;
    }
    __syncthreads();
    wr = wi = 0.0;
#pragma unroll
    for (int q = 0; q < NWC; ++q) {
      wr += swr[cl][q];
      wi += swi[cl][q];
    }
  }
  const zc tw = zmul(make_double2(tu.x, -tu.y), make_double2(wr, wi));
  double ss2 = 0.0;
  zc x = make_double2(0.0, 0.0);
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int o = 64 * (NWC * r + h) + lane;
    const zc q = zmul(tw, sv[o]);
    const zc y = make_double2(cr[r].x - q.x, cr[r].y - q.y);
    if (have && o < len) {
      c[o] = y;
      if (o == 0) x = y;
      else ss2 += zabs2(y);
    }
  }
  ss2 = group_sum(ss2, 64);
  if (lane == 0) {
    sss[cl][h] = ss2;
    if (h == 0) sx[cl] = x;  // row s is offset 0: lane 0 of the column's first wave
  }
  __syncthreads();
  // norm downdate (LAWN 176) or recompute, and the column's candidate entry
  if (h == 0 && lane == 0) {
    double v = -1.0;
    int fl = 0;
    if (have) {
      double a = vn1[col];
      if (s + 1 < lastrk && a != 0.0) {
        double temp = hypot(sx[cl].x, sx[cl].y) / a;
        temp = (1.0 + temp) * (1.0 - temp);
        temp = temp < 0.0 ? 0.0 : temp;
        const double q = a / vn2[col];
        const double temp2 = temp * q * q;
        if (temp2 <= tol3z) {
          double ssum = 0.0;
#pragma unroll
          for (int q2 = 0; q2 < NWC; ++q2) ssum += sss[cl][q2];
          a = sqrt(ssum);
          vn1[col] = vn2[col] = a;
          fl = 1;
        } else {
          a *= sqrt(temp);
          vn1[col] = a;
        }
      }
      v = a >= 0.0 ? a : -1.0;  // (a NaN norm takes no part)
    }
    cval[cl] = v;
    cpos[cl] = have ? p : INT_MAX;
    ccol[cl] = col;
    cflag[cl] = fl;
  }
  __syncthreads();
  fused_finish<CPB>(s, n, kmax, nb, stop_after, dss, cval, cpos, ccol, cflag, at, diag, st,
                    part, ctr, err);
}

// after the last fused complex step: the last pivot column's R(s,s) and v, and k
__global__ __launch_bounds__(256) void zq_ffinish(int m, zc* B, i64 ldb, const int* at,
                                                  const zc* vbuf, const FState* st,
                                                  i64* k_out) {
  const int sp = st->sprev;
  if (sp >= 0) {
    const zc* vp = vbuf + (size_t)(sp & 1) * m;
    zc* d = B + (i64)at[sp] * ldb;
    for (int t = sp + (int)threadIdx.x; t < m; t += 256) d[t] = vp[t];
  }
  if (threadIdx.x == 0) *k_out = st->k;
}

// the fused complex launch of step s: rows s..m-1 (at most 4096) pick the form
void zlaunch_fused(hipStream_t s0, int s, int m, int n, int kmax, int nb, int stop_after,
                   int lastrk, zc* B, i64 ldb, int* at, double* vn1, double* vn2, zc* tau,
                   double* diag, zc* vbuf, FState* st, u64x2* part, unsigned* ctr, int* err) {
  const int rows = m - s, tr = n - s - 1;
  auto grid = [&](int cpb) { return dim3((unsigned)std::max(1, (tr + cpb - 1) / cpb)); };
#define ZF_ARGS s, m, n, kmax, nb, stop_after, lastrk, B, ldb, at, vn1, vn2, tau, diag, vbuf, \
                st, part, ctr, err
  if (rows <= 64 * 8)
    hipLaunchKernelGGL((zq_fused<1, 8>), grid(4), dim3(256), 0, s0, ZF_ARGS);
  else if (rows <= 128 * 8)
    hipLaunchKernelGGL((zq_fused<2, 8>), grid(2), dim3(256), 0, s0, ZF_ARGS);
  else if (rows <= 256 * 8)
    hipLaunchKernelGGL((zq_fused<4, 8>), grid(1), dim3(256), 0, s0, ZF_ARGS);
  else
    hipLaunchKernelGGL((zq_fused<4, 16>), grid(1), dim3(256), 0, s0, ZF_ARGS);
#undef ZF_ARGS
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
  // vn1, vn2, diag (kmax), v (2 m complex), state, position map (n ints)
  const size_t nd = 2 * (size_t)n + (size_t)kmax + 1 + 4 * (size_t)m + 16 + ((size_t)n + 1) / 2;
  LRA_CHECK(lra_wsT(c, WS_QG, nd, &w));
  double* vn1 = w;
  double* vn2 = vn1 + n;
  double* diag = vn2 + n;
  zc* v = reinterpret_cast<zc*>(diag + kmax + (kmax & 1));
  GState* st = reinterpret_cast<GState*>(v + 2 * (size_t)m);
  int* at = reinterpret_cast<int*>(reinterpret_cast<double*>(st) + 16);
  const int lastrk = std::min(m, n);
  hipStream_t s0 = c->stream;
  zc* zt = reinterpret_cast<zc*>(tau);

  // one launch per step (rows up to 4096, with the uncached exchange area for the partials);
  // LRA_QG_FUSED=0 keeps the two launches per step below
  const char* fe = getenv("LRA_QG_FUSED");
  if (!(fe && fe[0] == '0') && m <= 4096 && n < (int)F_NONE) {
    void* area = nullptr;
    bool uc = false;
    const size_t bytes = kPartOff + 16 * (size_t)n;
    LRA_CHECK(lra_coop_area(c, bytes, &area, &uc));
    if (uc) {
      double* S = nullptr;  // the columns in storage order, for the final permutation
      LRA_CHECK(lra_wsT(c, WS_QGT, 2 * (size_t)m * n, &S));
      unsigned* ctr = reinterpret_cast<unsigned*>(area);
      u64x2* part = reinterpret_cast<u64x2*>((char*)area + kPartOff);
      FState* fs = reinterpret_cast<FState*>(st);
      LRA_CHECK(lra_launch_zero_u32(s0, (unsigned*)area, (i64)(bytes / 4)));
      hipLaunchKernelGGL(zq_norms, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s0, m, n, B,
                         ldb, vn1, vn2, jpvt, at);
      hipLaunchKernelGGL(qg_fstart, dim3(1), dim3(GT), 0, s0, n, kmax, nb, vn1, atol, rtol, fs,
                         at);
      LRA_HIP(hipGetLastError());
      for (int s = 0; s < kmax; ++s) {
        zlaunch_fused(s0, s, m, n, kmax, nb, stop_after, lastrk, B, ldb, at, vn1, vn2, zt, diag,
                      v, fs, part, ctr, c->d_err);
        LRA_HIP(hipGetLastError());
        if ((s & 63) == 63) {  // stop enqueueing once the device has stopped
          int done = 0;
          LRA_HIP(hipMemcpyAsync(c->host_pinned, &fs->done, sizeof(int), hipMemcpyDeviceToHost,
                                 s0));
          LRA_HIP(hipStreamSynchronize(s0));
          std::memcpy(&done, c->host_pinned, sizeof(int));
          if (done) break;
        }
      }
      hipLaunchKernelGGL(zq_ffinish, dim3(1), dim3(256), 0, s0, m, B, ldb, at, v, fs, k_out);
      LRA_HIP(hipMemcpy2DAsync(S, (size_t)m * sizeof(zc), B, (size_t)ldb * sizeof(zc),
                               (size_t)m * sizeof(zc), (size_t)n, hipMemcpyDeviceToDevice, s0));
      hipLaunchKernelGGL(qg_gather, dim3((unsigned)n), dim3(256), 0, s0, 2 * m, S, 2 * (i64)m,
                         Bz, 2 * ldb, at, jpvt);
      LRA_HIP(hipGetLastError());
      return LRA_OK;
    }
  }

  hipLaunchKernelGGL(zq_norms, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s0, m, n, B, ldb,
                     vn1, vn2, jpvt, nullptr);
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

namespace {

// the fused launch of step s: rows s..m-1 (at most 8192) pick the form, as for the applies
void launch_fused(hipStream_t s0, int s, const QrcpCall& q, int lastrk, int* at, double* vn1,
                  double* vn2, double* diag, double* vbuf, FState* st, u64x2* part,
                  unsigned* ctr, int* err) {
  const int m = q.m, n = q.n, rows = m - s, tr = n - s - 1;
  auto grid = [&](int cpb) { return dim3((unsigned)std::max(1, (tr + cpb - 1) / cpb)); };
#define QF_ARGS s, m, n, q.kmax, q.nb, q.stop_after, lastrk, q.A, q.lda, at, vn1, vn2, q.tau, \
                diag, vbuf, st, part, ctr, err
  if (rows <= 64 * 8)
    hipLaunchKernelGGL((qg_fused<1, 8>), grid(4), dim3(256), 0, s0, QF_ARGS);
  else if (rows <= 64 * 16)
    hipLaunchKernelGGL((qg_fused<1, 16>), grid(4), dim3(256), 0, s0, QF_ARGS);
  else if (rows <= 128 * 16)
    hipLaunchKernelGGL((qg_fused<2, 16>), grid(2), dim3(256), 0, s0, QF_ARGS);
  else if (rows <= 128 * 24)
    hipLaunchKernelGGL((qg_fused<2, 24>), grid(2), dim3(256), 0, s0, QF_ARGS);
  else if (rows <= 128 * 32)
    hipLaunchKernelGGL((qg_fused<2, 32>), grid(2), dim3(256), 0, s0, QF_ARGS);
  else
    hipLaunchKernelGGL((qg_fused<4, 32>), grid(1), dim3(256), 0, s0, QF_ARGS);
#undef QF_ARGS
}

}  // namespace

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
  // vn1, vn2, diag (kmax), v (2 m), state, position map (n ints)
  const size_t nd = 2 * (size_t)n + (size_t)std::max(q.kmax, 1) + 2 * (size_t)m + 16 +
                    ((size_t)n + 1) / 2;
  LRA_CHECK(lra_wsT(c, WS_QG, nd, &w));
  double* vn1 = w;
  double* vn2 = vn1 + n;
  double* diag = vn2 + n;
  double* v = diag + std::max(q.kmax, 1);
  GState* st = reinterpret_cast<GState*>(v + 2 * (size_t)m);
  int* at = reinterpret_cast<int*>(v + 2 * (size_t)m + 16);
  const int lastrk = std::min(m, n);
  hipStream_t s0 = c->stream;

  // one launch per step (rows up to 8192, with the uncached exchange area for the partials);
  // LRA_QG_FUSED=0 keeps the two launches per step below
  const char* fe = getenv("LRA_QG_FUSED");
  if (!(fe && fe[0] == '0') && m <= 8192 && n < (int)F_NONE && q.kmax > 0) {
    void* area = nullptr;
    bool uc = false;
    const size_t bytes = kPartOff + 16 * (size_t)n;
    LRA_CHECK(lra_coop_area(c, bytes, &area, &uc));
    if (uc) {
      double* S = nullptr;  // the columns in storage order, for the final permutation
      LRA_CHECK(lra_wsT(c, WS_QGT, (size_t)m * n, &S));
      unsigned* ctr = reinterpret_cast<unsigned*>(area);
      u64x2* part = reinterpret_cast<u64x2*>((char*)area + kPartOff);
      FState* fs = reinterpret_cast<FState*>(st);
      LRA_CHECK(lra_launch_zero_u32(s0, (unsigned*)area, (i64)(bytes / 4)));
      hipLaunchKernelGGL(qg_norms, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s0, m, n, q.A,
                         q.lda, vn1, vn2, q.jpvt, at);
      hipLaunchKernelGGL(qg_fstart, dim3(1), dim3(GT), 0, s0, n, q.kmax, q.nb, vn1, q.atol,
                         q.rtol, fs, at);
      LRA_HIP(hipGetLastError());
      for (int s = 0; s < q.kmax; ++s) {
        launch_fused(s0, s, q, lastrk, at, vn1, vn2, diag, v, fs, part, ctr, c->d_err);
        LRA_HIP(hipGetLastError());
        if ((s & 63) == 63) {  // stop enqueueing once the device has stopped
          int done = 0;
          LRA_HIP(hipMemcpyAsync(c->host_pinned, &fs->done, sizeof(int), hipMemcpyDeviceToHost,
                                 s0));
          LRA_HIP(hipStreamSynchronize(s0));
          std::memcpy(&done, c->host_pinned, sizeof(int));
          if (done) break;
        }
      }
      hipLaunchKernelGGL(qg_ffinish, dim3(1), dim3(256), 0, s0, m, q.A, q.lda, at, diag, v, fs,
                         q.k_out, q.early_out);
      LRA_HIP(hipMemcpy2DAsync(S, (size_t)m * sizeof(double), q.A, (size_t)q.lda * sizeof(double),
                               (size_t)m * sizeof(double), (size_t)n, hipMemcpyDeviceToDevice,
                               s0));
      hipLaunchKernelGGL(qg_gather, dim3((unsigned)n), dim3(256), 0, s0, m, S, (i64)m, q.A, q.lda,
                         at, q.jpvt);
      LRA_HIP(hipGetLastError());
      return LRA_OK;
    }
  }

  hipLaunchKernelGGL(qg_norms, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s0, m, n, q.A, q.lda,
                     vn1, vn2, q.jpvt, nullptr);
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
