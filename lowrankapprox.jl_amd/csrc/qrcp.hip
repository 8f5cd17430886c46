// qrcp.hip -- early-terminating Householder QR with column pivoting, register-resident.
//
// Replaces the geqp3_adap_main! loop over ?laqps blocks (src/pqr.jl:361-418 calling
// src/lapack.jl:129-136) and, in mode 1, a single ?laqps block.
//
// Design (gfx950): the ell x n sketch is spread over a co-resident grid of G workgroups;
// every column lives in the registers of `tpc` lanes (lane lic holds rows lic*RPT ..
// lic*RPT+RPT-1, so every per-row LDS offset is an immediate), so a pivot
// step never touches HBM.  Each step: every workgroup publishes its best remaining column
// (norm, position, data), one grid barrier, every workgroup picks the same global winner
// (largest norm, lowest position -- idamax's first-index rule), computes the same reflector
// redundantly in LDS, and applies it to its own columns at once (right-looking, unblocked).
// Exact arithmetic is identical to dlaqps' deferred-update scheme; the block structure of
// geqp3_adap_main! is emulated so the rank-termination test runs at the same points:
// blocks of min(nb, k-j+1) columns that end early at the first step whose norm downdate
// needed a recompute (dlaqps' lsticc), and after each block the |R(jn-1,jn-1)| <= ptol test
// returns the first small diagonal of the block.  Columns are never moved during the
// factorization; each carries its LAPACK position (jpvt order) and is written to it at the
// end.
#include "lra_internal.h"

#include <algorithm>
#include <climits>

namespace {

// Step exchange between blocks: a 16-byte slot {candidate norm, tag} per block and step parity,
// tag = step+1 | position << 28 | local column << 52 | flag << 63, written with one 16-byte
// store after the block's candidate column has been written (and acknowledged), read with one
// 16-byte load, so a reader that sees the step's tag also sees the norm and the column.
constexpr uint64_t TAG_MASK = (1ull << 28) - 1;
constexpr uint64_t POS_NONE = (1ull << 24) - 1;
constexpr uint64_t LC_NONE = (1ull << 11) - 1;
LRA_DEV uint64_t pack_tag(int step1, int pos, int lc, int flag) {
  const uint64_t pp = pos == INT_MAX ? POS_NONE : (uint64_t)pos;
  const uint64_t ll = lc < 0 ? LC_NONE : (uint64_t)lc;
  return ((uint64_t)step1 & TAG_MASK) | (pp << 28) | (ll << 52) | ((uint64_t)(flag != 0) << 63);
}

struct QrcpDev {
  double* A;
  i64 lda;
  int m, n, offset, kmax, nb, mode, stop_after, lastrk;
  double atol, rtol;
  i64* jpvt;
  double* tau;
  double* vn1g;
  double* vn2g;
  i64* k_out;
  int* early_out;
  int G, cpw, tpc, cand_stride;
  int uncached;       // exchange area bypasses the caches (no write-back/invalidate needed)
  uint32_t gen;       // launch generation, part of every chunk tag
  uint64_t* slots;    // 2G slots of 16 bytes: {norm bits, tag}
  double* cand;
  int* err;
  int* abort_flag;
};

constexpr int NB_MAX = 1024;

LRA_DEV double lapy2(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

template <int RPT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void qrcp_kernel(QrcpDev p) {
  constexpr int NW = BLOCK / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tpc = p.tpc;
  const int gpw = 64 / tpc;
  const int lic = lane & (tpc - 1);
  const int lc = wave * gpw + lane / tpc;
  const int g = blockIdx.x;
  const int col = g * p.cpw + lc;
  const bool have = (lc < p.cpw) && (col < p.n);
  const int m = p.m;
  const int rbase = lic * RPT;  // first row held by this lane
  const double tol3z = 1.0536712127723509e-08;  // sqrt(2^-53) = sqrt(dlamch('E'))

  __shared__ double s_x[64 * RPT];
  __shared__ double s_diag[NB_MAX];
  __shared__ double s_wval[NW];
  __shared__ int s_wpos[NW], s_wlc[NW], s_wflag[NW];
  __shared__ double s_tau, s_beta, s_ptol, s_winval;
  __shared__ int s_winowner, s_winlc, s_winpos, s_flagprev, s_blc;  // s_blc: exchange ok

  // s_x holds the current reflector by ABSOLUTE row: zeros above row rk and in the padding,
  // so the dot/update loops below need no per-row predicates.
  for (int t = tid; t < 64 * RPT; t += BLOCK) s_x[t] = 0.0;
  double a[RPT];
  {
    const double* src = p.A + (i64)col * p.lda + rbase;
    const int hi = have ? m - rbase : 0;
#pragma unroll
    for (int i = 0; i < RPT; ++i) a[i] = (i < hi) ? src[i] : 0.0;
  }
  int pos = col;
  bool active = have;
  i64 jp = 0;
  double vn1 = 0.0, vn2 = 0.0;
  if (p.mode == 0) {
    double ss = 0.0;
#pragma unroll
    for (int i = 0; i < RPT; ++i) ss += a[i] * a[i];
    ss = lra_group_sum(ss, tpc);
    vn1 = vn2 = sqrt(ss);
    jp = col + 1;
  } else if (have) {
    vn1 = p.vn1g[col];
    vn2 = p.vn2g[col];
    jp = p.jpvt[col];
  }
  int myflag = 0;
  int j0 = 0;
  int jb = min(p.nb, p.kmax);
  int k_result = p.kmax;
  int early = 0;
  bool aborted = false;

  int s = 0;
  for (;; ++s) {
    // ---------------- local candidate (largest norm, then lowest position) ------------
    double bv = -1.0;
    int bp = INT_MAX, blc = -1;
    if (active) { bv = vn1; bp = pos; blc = lc; }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const double ov = __shfl_xor(bv, d, 64);
      const int op = __shfl_xor(bp, d, 64);
      const int ol = __shfl_xor(blc, d, 64);
      if (ov > bv || (ov == bv && op < bp)) { bv = ov; bp = op; blc = ol; }
    }
    const int wflag = __any(myflag) ? 1 : 0;
    if (lane == 0) { s_wval[wave] = bv; s_wpos[wave] = bp; s_wlc[wave] = blc; s_wflag[wave] = wflag; }
    __syncthreads();
    // every thread reduces the NW wave candidates the same way
    double cv = s_wval[0];
    int cp = s_wpos[0], clc = s_wlc[0], cflag = s_wflag[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) {
      const double ov = s_wval[w];
      const int op = s_wpos[w];
      if (ov > cv || (ov == cv && op < cp)) { cv = ov; cp = op; clc = s_wlc[w]; }
      cflag |= s_wflag[w];
    }
    const bool done_steps = (s == p.kmax);

    if (p.G > 1) {
      const int rk = p.offset + s;
      const int len = m - rk;
      const size_t sl = (size_t)(s & 1) * p.G + g;
      // stage the block candidate's rows rk..m-1 in s_x (absolute rows; the winner's data
      // overwrites them below, rows < rk and >= m stay zero)
      if (!done_steps && have && active && lc == clc) {
        double* d = s_x + rbase;
        const int lo = rk - rbase, hi = m - rbase;
#pragma unroll
        for (int i = 0; i < RPT; ++i)
          if (i >= lo && i < hi) d[i] = a[i];
      }
      __syncthreads();
      if (wave == 0) {
        volatile u64x2* slots = reinterpret_cast<volatile u64x2*>(p.slots);
        // publish: the column as 16-byte {value, step tag} chunks (each chunk validates
        // itself, so no acknowledgement wait is needed), then the {norm, tag} slot
        const uint64_t ctag = ((uint64_t)p.gen << 32) | (uint64_t)(s + 1);
        if (!done_steps) {
          volatile u64x2* dst = reinterpret_cast<volatile u64x2*>(p.cand) + sl * p.cand_stride;
          for (int t = lane; t < len; t += 64) {
            u64x2 w;
            w.x = (uint64_t)__double_as_longlong(s_x[rk + t]);
            w.y = ctag;
            dst[t] = w;
          }
        }
        if (!p.uncached) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (lane == 0) {
          u64x2 w;
          w.x = (uint64_t)__double_as_longlong(cv);
          w.y = pack_tag(s + 1, cp, clc, cflag);
          slots[sl] = w;
        }
        // wait until every block has published this step (bounded)
        const uint64_t want = (uint64_t)(s + 1) & TAG_MASK;
        const size_t base = (size_t)(s & 1) * p.G;
        unsigned spins = 0;
        int ok = 1;
        u64x2 mine[4];  // G <= 256 slots, up to 4 per lane
        for (;;) {
          bool ready = true;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int t = lane + 64 * u;
            if (t < p.G) {
              mine[u] = slots[base + t];
              ready &= ((mine[u].y & TAG_MASK) == want);
            }
          }
          if (__all(ready)) break;
          if ((++spins & 63) == 0) {
            if (__hip_atomic_load(p.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { ok = 0; break; }
            if (spins > LRA_SPIN_LIMIT) {
              if (lane == 0) {
                __hip_atomic_store(p.err, (int)LRA_E_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(p.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              }
              ok = 0;
              break;
            }
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (!p.uncached) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        double v = -1.0;
        int vp = INT_MAX, vo = -1, vl = -1, fl = 0;
        if (ok) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int t = lane + 64 * u;
            if (t >= p.G) continue;
            const uint64_t tg = mine[u].y;
            const double sv = __longlong_as_double((long long)mine[u].x);
            const uint64_t pp = (tg >> 28) & POS_NONE;
            const int spos = pp == POS_NONE ? INT_MAX : (int)pp;
            const uint64_t ll = (tg >> 52) & LC_NONE;
            fl |= (int)(tg >> 63);
            if (sv > v || (sv == v && spos < vp)) {
              v = sv; vp = spos; vo = t; vl = ll == LC_NONE ? -1 : (int)ll;
            }
          }
#pragma unroll
          for (int d = 1; d < 64; d <<= 1) {
            const double ov = __shfl_xor(v, d, 64);
            const int op = __shfl_xor(vp, d, 64);
            const int oo = __shfl_xor(vo, d, 64);
            const int ol = __shfl_xor(vl, d, 64);
            if (ov > v || (ov == v && op < vp)) { v = ov; vp = op; vo = oo; vl = ol; }
            fl |= __shfl_xor(fl, d, 64);
          }
        }
        if (lane == 0) {
          s_winval = v; s_winpos = vp; s_winowner = vo; s_winlc = vl; s_flagprev = fl;
          s_blc = ok;
        }
      }
      __syncthreads();
      if (!s_blc) { aborted = true; break; }
    } else {
      if (tid == 0) { s_winval = cv; s_winpos = cp; s_winowner = 0; s_winlc = clc; s_flagprev = cflag; }
      __syncthreads();
    }
    myflag = 0;

    // ---------------- block bookkeeping for the step just finished (s-1) --------------
    if (s == 0) {
      if (tid == 0) s_ptol = fmax(p.atol, p.rtol * s_winval);
      __syncthreads();
    } else {
      const int last = s - 1;
      const bool size_end = (last - j0 + 1 == jb);
      const bool blk_end = size_end || (s_flagprev != 0) || done_steps;
      if (blk_end) {
        if (p.mode == 1) {
          k_result = s;
          break;
        }
        const double ptol = s_ptol;
        if (fabs(s_diag[last - j0]) <= ptol) {
          int i = j0;
          for (; i <= last; ++i)
            if (fabs(s_diag[i - j0]) <= ptol) break;
          if (i <= last) { k_result = i; break; }
        }
        if (done_steps) { k_result = p.kmax; break; }
        if (p.stop_after > 0 && s >= p.stop_after) { k_result = s; early = 1; break; }
        j0 = s;
        jb = min(p.nb, p.kmax - s);
      }
    }
    if (done_steps) { k_result = s; break; }

    // ---------------- winner column -> LDS ---------------------------------------------
    const int rk = p.offset + s;
    const int len = m - rk;
    const int win_owner = s_winowner, win_lc = s_winlc, win_pos = s_winpos;
    if (tid == 0 && rk > 0) s_x[rk - 1] = 0.0;  // previous step's v(1) = 1
    if (p.G > 1) {
      const volatile u64x2* src = reinterpret_cast<const volatile u64x2*>(p.cand) +
                                  ((size_t)(s & 1) * p.G + win_owner) * p.cand_stride;
      const uint64_t ctag = ((uint64_t)p.gen << 32) | (uint64_t)(s + 1);
      for (int t = tid; t < len; t += BLOCK) {
        u64x2 w = src[t];
        unsigned spins = 0;
        while (w.y != ctag && ++spins < LRA_SPIN_LIMIT) {  // chunk not landed yet
          __builtin_amdgcn_s_sleep(1);
          w = src[t];
        }
        if (w.y != ctag) {
          __hip_atomic_store(p.err, (int)LRA_E_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(p.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_x[rk + t] = __longlong_as_double((long long)w.x);
      }
    } else {
      if (have && active && lc == win_lc) {
        double* d = s_x + rbase;
        const int lo = rk - rbase, hi = m - rbase;
#pragma unroll
        for (int i = 0; i < RPT; ++i)
          if (i >= lo && i < hi) d[i] = a[i];
      }
    }
    __syncthreads();

    // ---------------- Householder reflector (dlarfg) in wave 0 --------------------------
    if (wave == 0) {
      double* x = s_x + rk;
      double alpha = x[0];
      double ss = 0.0;
      for (int t = 1 + lane; t < len; t += 64) ss += x[t] * x[t];
      ss = lra_wave_sum(ss);
      double xnorm = sqrt(ss);
      double tau = 0.0, beta = alpha, scal = 1.0;
      if (len > 1 && xnorm != 0.0) {
        beta = -copysign(lapy2(alpha, xnorm), alpha);
        const double safmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;
        int knt = 0;
        if (fabs(beta) < safmin) {
          const double rsafmn = 1.0 / safmin;
          do {
            ++knt;
            for (int t = 1 + lane; t < len; t += 64) x[t] *= rsafmn;
            beta *= rsafmn;
            alpha *= rsafmn;
          } while (fabs(beta) < safmin && knt < 20);
          double s2 = 0.0;
          for (int t = 1 + lane; t < len; t += 64) s2 += x[t] * x[t];
          s2 = lra_wave_sum(s2);
          xnorm = sqrt(s2);
          beta = -copysign(lapy2(alpha, xnorm), alpha);
        }
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
        for (int j = 0; j < knt; ++j) beta *= safmin;
      }
      if (lane == 0) { s_tau = tau; s_beta = beta; }
      // scale x into v (v(1) = 1 implicit, stored at s_x[0] below)
      for (int t = 1 + lane; t < len; t += 64) x[t] *= scal;
    }
    __syncthreads();
    if (tid == 0) {
      s_x[rk] = 1.0;
      if (g == 0) p.tau[s] = s_tau;
      s_diag[s - j0] = s_beta;
    }
    __syncthreads();
    const double tau = s_tau, beta = s_beta;

    // ---------------- apply H to the block's columns, position swap, norm downdate ------
    const bool is_win = have && active && (g == win_owner) && (lc == win_lc);
    if (is_win) {
      const double* vx = s_x + rbase;
      const int lo = rk - rbase, hi = m - rbase;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        if (i == lo) a[i] = beta;
        else if (i > lo && i < hi) a[i] = vx[i];
      }
      active = false;
      pos = s;
    } else {
      if (have && pos == s) pos = win_pos;  // LAPACK swaps columns k and pvt
      if (active) {
        const double* vx = s_x + rbase;
        double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
#pragma unroll
        for (int i = 0; i < RPT; i += 4) {
          w0 += vx[i] * a[i];
          if (i + 1 < RPT) w1 += vx[i + 1] * a[i + 1];
          if (i + 2 < RPT) w2 += vx[i + 2] * a[i + 2];
          if (i + 3 < RPT) w3 += vx[i + 3] * a[i + 3];
        }
        double w = lra_group_sum((w0 + w1) + (w2 + w3), tpc);
        const double tw = tau * w;
        // re-read v from LDS below instead of keeping RPT values live across the two loops
        asm volatile("" ::: "memory");
#pragma unroll
        for (int i = 0; i < RPT; ++i) a[i] -= tw * vx[i];
        if (rk + 1 < p.lastrk) {
          // new value of row rk lives in lane rk / RPT of the group, register rk % RPT
          const int irk = rk % RPT;
          double xr = 0.0;
#pragma unroll
          for (int i = 0; i < RPT; ++i) xr = (i == irk) ? a[i] : xr;
          xr = __shfl(xr, (lane & ~(tpc - 1)) | (rk / RPT), 64);
          if (vn1 != 0.0) {
            double temp = fabs(xr) / vn1;
            temp = (1.0 + temp) * (1.0 - temp);
            temp = temp < 0.0 ? 0.0 : temp;
            const double r = vn1 / vn2;
            const double temp2 = temp * r * r;
            if (temp2 <= tol3z) {
              double ss = 0.0;
              const int lo = rk - rbase;
#pragma unroll
              for (int i = 0; i < RPT; ++i) ss += (i > lo) ? a[i] * a[i] : 0.0;
              ss = lra_group_sum(ss, tpc);
              vn1 = vn2 = sqrt(ss);
              myflag = 1;
            } else {
              vn1 *= sqrt(temp);
            }
          }
        }
      }
    }
    // s_x is rewritten next step only after the next __syncthreads
  }

  if (aborted) return;
  // ---------------- write-out to LAPACK positions --------------------------------------
  // (with G > 1 every block loaded its columns before step 0's barrier, so the in-place
  // permuted write cannot overwrite a column somebody still has to read)
  if (have) {
    double* dst = p.A + (i64)pos * p.lda + rbase;
    const int hi = m - rbase;
#pragma unroll
    for (int i = 0; i < RPT; ++i)
      if (i < hi) dst[i] = a[i];
    if (lic == 0) {
      p.jpvt[pos] = jp;
      if (p.mode == 1) { p.vn1g[pos] = vn1; p.vn2g[pos] = vn2; }
    }
  }
  if (g == 0 && tid == 0) {
    *p.k_out = k_result;
    if (p.early_out) *p.early_out = early;
  }
}

template <int RPT, int BLOCK>
hipError_t launch_rpt(const QrcpDev& d, hipStream_t st, bool coop) {
  dim3 grid(d.G), block(BLOCK);
  if (coop) {
    void* args[] = {(void*)&d};
    return hipLaunchCooperativeKernel((const void*)qrcp_kernel<RPT, BLOCK>, grid, block, args,
                                      0, st);
  }
  hipLaunchKernelGGL((qrcp_kernel<RPT, BLOCK>), grid, block, 0, st, d);
  return hipGetLastError();
}

#define LRA_RPT_256(X) X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32) X(36) X(40) X(44) X(48) \
  X(52) X(56) X(60) X(64) X(68) X(72) X(76) X(80)
#define LRA_RPT_512(X) X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32) X(36) X(40) X(44) X(48)
#define LRA_RPT_1024(X) X(4) X(8) X(12) X(16) X(20)

hipError_t launch_any(int rpt, int block, const QrcpDev& d, hipStream_t st, bool coop) {
  if (block == 1024) {
    switch (rpt) {
#define CASE(R) case R: return launch_rpt<R, 1024>(d, st, coop);
      LRA_RPT_1024(CASE)
#undef CASE
    }
  } else if (block == 512) {
    switch (rpt) {
#define CASE(R) case R: return launch_rpt<R, 512>(d, st, coop);
      LRA_RPT_512(CASE)
#undef CASE
    }
  } else {
    switch (rpt) {
#define CASE(R) case R: return launch_rpt<R, 256>(d, st, coop);
      LRA_RPT_256(CASE)
#undef CASE
    }
  }
  return hipErrorInvalidValue;
}

int occupancy_any(int rpt, int block, int* nb) {
  hipError_t e = hipErrorInvalidValue;
  if (block == 1024) {
    switch (rpt) {
#define CASE(R) case R: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, qrcp_kernel<R, 1024>, 1024, 0); break;
      LRA_RPT_1024(CASE)
#undef CASE
    }
  } else if (block == 512) {
    switch (rpt) {
#define CASE(R) case R: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, qrcp_kernel<R, 512>, 512, 0); break;
      LRA_RPT_512(CASE)
#undef CASE
    }
  } else {
    switch (rpt) {
#define CASE(R) case R: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, qrcp_kernel<R, 256>, 256, 0); break;
      LRA_RPT_256(CASE)
#undef CASE
    }
  }
  return e == hipSuccess ? 0 : 1;
}

}  // namespace

// Shape selection: single 1024-thread block when the panel fits one CU's registers, else a
// cooperative grid of 256-thread blocks with `tpc` lanes per column.
struct QrcpShape {
  int tpc, rpt, block, G, cpw;
};

static bool choose_shape(lra_ctx* ctx, int m, int n, QrcpShape* out) {
  for (int tpc = 1; tpc <= 64; tpc <<= 1) {
    int rpt = (m + tpc - 1) / tpc;
    rpt = (rpt + 3) / 4 * 4;
    if (rpt < 4) rpt = 4;
    if (rpt <= 20 && n <= 1024 / tpc) {
      *out = {tpc, rpt, 1024, 1, 1024 / tpc};
      return true;
    }
    if (rpt <= 48 && n <= 512 / tpc) {
      *out = {tpc, rpt, 512, 1, 512 / tpc};
      return true;
    }
  }
  for (int tpc = 1; tpc <= 64; tpc <<= 1) {
    int rpt = (m + tpc - 1) / tpc;
    rpt = (rpt + 3) / 4 * 4;
    if (rpt < 4) rpt = 4;
    if (rpt > 80) continue;
    const int cpw = 256 / tpc;
    const int G = (n + cpw - 1) / cpw;
    int per_cu = 0;
    if (occupancy_any(rpt, 256, &per_cu) != 0 || per_cu < 1) continue;
    if (G == 1 || G <= per_cu * ctx->num_cu) {
      *out = {tpc, rpt, 256, G, cpw};
      return true;
    }
  }
  return false;
}

int lra_qrcp(lra_ctx* ctx, const QrcpCall& q) {
  if (q.kmax <= 0) {
    // nothing to factor: k = 0 (and jpvt = identity in mode 0)
    if (q.mode == 0 && q.n > 0) LRA_CHECK(lra_launch_fill_i64(ctx->stream, q.jpvt, q.n, 1));
    LRA_HIP(hipMemsetAsync(q.k_out, 0, sizeof(i64), ctx->stream));
    if (q.early_out) LRA_HIP(hipMemsetAsync(q.early_out, 0, sizeof(int), ctx->stream));
    return LRA_OK;
  }
  if (q.n >= (int)POS_NONE) {
    lra_set_error("qrcp: n=%d exceeds the position field", q.n);
    return LRA_E_UNSUPPORTED;
  }
  if (q.nb > NB_MAX) {
    lra_set_error("qrcp: nb=%d exceeds %d", q.nb, NB_MAX);
    return LRA_E_UNSUPPORTED;
  }
  QrcpShape sh;
  if (!choose_shape(ctx, q.m, q.n, &sh)) {
    lra_set_error("qrcp: panel %d x %d does not fit the register-resident engine", q.m, q.n);
    return LRA_E_UNSUPPORTED;
  }
  QrcpDev d;
  d.A = q.A; d.lda = q.lda; d.m = q.m; d.n = q.n; d.offset = q.offset; d.kmax = q.kmax;
  d.nb = q.nb; d.mode = q.mode; d.stop_after = q.stop_after;
  d.lastrk = std::min(q.m, q.n + q.offset);
  d.atol = q.atol; d.rtol = q.rtol; d.jpvt = q.jpvt; d.tau = q.tau; d.vn1g = q.vn1;
  d.vn2g = q.vn2; d.k_out = q.k_out; d.early_out = q.early_out;
  d.G = sh.G; d.cpw = sh.cpw; d.tpc = sh.tpc;
  d.cand_stride = ((q.m + 7) / 8) * 8;
  d.err = ctx->d_err;
  if (sh.G > 256) {
    lra_set_error("qrcp: %d blocks exceed the 256-slot exchange", sh.G);
    return LRA_E_UNSUPPORTED;
  }
  // exchange area: [header 256 B: abort flag][slots 2G x 16 B][cand 2G x stride doubles]
  const size_t slot_bytes = 16 * 2 * (size_t)sh.G;
  const size_t cand_bytes = 16 * 2 * (size_t)sh.G * d.cand_stride;  // {value, tag} chunks
  void* area = nullptr;
  bool uc = false;
  LRA_CHECK(lra_coop_area(ctx, 256 + slot_bytes + cand_bytes, &area, &uc));
  char* base = (char*)area;
  d.uncached = uc ? 1 : 0;
  d.abort_flag = (int*)base;
  d.slots = (uint64_t*)(base + 256);
  d.cand = (double*)(base + 256 + slot_bytes);
  // zero the header and every slot (a stale slot tag of an earlier launch must never match);
  // chunk tags carry the launch generation instead
  LRA_CHECK(lra_launch_zero_u32(ctx->stream, (unsigned*)base, (256 + slot_bytes) / 4));
  d.gen = ++ctx->qrcp_gen;
  hipError_t e = launch_any(sh.rpt, sh.block, d, ctx->stream, sh.G > 1);
  if (e != hipSuccess) {
    lra_set_error("qrcp launch (m=%d n=%d tpc=%d rpt=%d G=%d): %s", q.m, q.n, sh.tpc, sh.rpt,
                  sh.G, hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
}
