// qrcp.hip -- early-terminating Householder QR with column pivoting, register-resident.
//
// Replaces the geqp3_adap_main! loop over ?laqps blocks (src/pqr.jl:361-418 calling
// src/lapack.jl:129-136) and, in mode 1, a single ?laqps block.
//
// Design (gfx950): the ell x n sketch is spread over a co-resident grid of G workgroups;
// every column lives in the registers of `tpc` lanes (lane lic holds rows lic*RPT ..
// lic*RPT+RPT-1, so every per-row LDS offset is an immediate), so a pivot step never
// touches HBM.  One step s (pivot row rk):
//   A  each workgroup finds its best remaining column (largest norm, lowest position --
//      idamax's first-index rule) with DPP wave reductions and one barrier, and publishes
//      {norm, position} in its slot;
//   B  that column finishes the previous step's update and builds its own Householder
//      reflector (dlarfg) in registers, then publishes {tau, beta, v} as tagged chunks;
//   C  every wave polls 64 of the slots while the other columns finish the previous step's
//      rank-1 update (the update is lazy: the norm downdate only needs v'a_j and the new
//      row rk, so the axpy hides in the exchange);
//   D  the block structure of geqp3_adap_main! is emulated so the rank-termination test
//      runs at the same points: blocks of min(nb, k-j+1) columns that end early at the
//      first step whose norm downdate needed a recompute (dlaqps' lsticc), and after each
//      block the |R(jn-1,jn-1)| <= ptol test returns the first small diagonal of the block;
//   E  every workgroup copies the global winner's reflector into LDS;
//   F  dot products v'a_j, position swap and the LAWN 176 norm downdate.
// Exact arithmetic is that of dlaqps' deferred-update scheme.  Columns are never moved
// during the factorization; each carries its LAPACK position (jpvt order) and is written
// to it at the end.
#include "lra_internal.h"
#include "lra_wave.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cstdlib>

namespace {

// Step exchange between blocks: a 16-byte slot {candidate norm, tag} per block and step parity,
// tag = step+1 | position << 28 | local column << 52 | flag << 63, written with one 16-byte
// store.  The candidate's reflector is published beside it as a record of 16-byte
// {value, chunk tag} chunks: chunk 0 = tau, chunk 1 = beta, chunk 2+i = v(rk+1+i).  Every chunk
// carries (launch generation, step+1), so a reader validates each chunk by itself and no
// writer ever waits for an acknowledgement.
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
  int G, cpw, tpc, ltpc, cand_stride;  // tpc = 1 << ltpc lanes per column
  int uncached;       // exchange area bypasses the caches (no write-back/invalidate needed)
  uint32_t gen;       // launch generation, part of every chunk tag
  uint64_t* slots;    // 2G slots of 16 bytes: {norm bits, tag}
  double* cand;       // 2G records of cand_stride 16-byte chunks
  int* err;
  int* abort_flag;
};

constexpr int NB_MAX = 1024;

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

LRA_DEV void raise_error(const QrcpDev& p, int code) {
  __hip_atomic_store(p.err, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 16-byte exchange store.  Uncached memory: a plain store (no wait for its completion; the
// chunk tag validates it on the reader's side).  Cached fallback: a volatile store, made
// visible by the caller's release fence.
LRA_DEV void put_chunk(u64x2* dst, u64x2 w, int uncached) {
  if (uncached) *dst = w;
  else *(volatile u64x2*)dst = w;
}

// re-read a chunk until its tag is `tag` (bounded; on expiry or abort the error is raised)
LRA_DEV u64x2 wait_chunk(const QrcpDev& p, const u64x2* src, uint64_t tag) {
  const volatile u64x2* v = (const volatile u64x2*)src;
  u64x2 w = *v;
  unsigned spins = 0;
  while (w.y != tag) {
    if ((++spins & 63) == 0 &&
        (spins > LRA_SPIN_LIMIT ||
         __hip_atomic_load(p.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
      break;
    __builtin_amdgcn_s_sleep(1);
    if (!p.uncached) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    w = *v;
  }
  if (w.y != tag) raise_error(p, LRA_E_TIMEOUT);
  return w;
}

// One factorization on the blocks [0, p.G) of its grid; g = this block's index among them.
template <int RPT, int BLOCK, bool COOP>
__device__ __forceinline__ void qrcp_body(const QrcpDev& p, const int g) {
  constexpr int NW = BLOCK / 64;
  constexpr int XS = 64 * RPT;  // rows held by one column group (>= m)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tpc = p.tpc, ltpc = p.ltpc;
  const int lgpw = 6 - ltpc;  // columns per wave = 1 << lgpw
  const int lic = lane & (tpc - 1);
  const int lc = (wave << lgpw) + (lane >> ltpc);
  const int col = g * p.cpw + lc;
  const bool have = (lc < p.cpw) && (col < p.n);
  const int m = p.m;
  const int rbase = lic * RPT;  // first row held by this lane
  const int glead = lane & ~(tpc - 1);  // first lane of this column's group
  const double tol3z = 1.0536712127723509e-08;  // sqrt(2^-53) = sqrt(dlamch('E'))

  // s_x[par] holds the reflector of step s (par = s & 1) by ABSOLUTE row: zeros above row
  // rk and in the padding, so the dot/update loops need no per-row predicates.  Two buffers:
  // the rank-1 update of step s is finished lazily during step s+1's exchange.
  __shared__ double s_x[2][XS];
  __shared__ double s_diag[NB_MAX];
  __shared__ double s_wval[NW];
  __shared__ int s_wpos[NW], s_wlc[NW], s_wflag[NW];
  __shared__ double s_tau, s_beta, s_ptol;
  __shared__ double s_pv[NW];
  __shared__ int s_pp[NW], s_po[NW], s_pl[NW], s_pf[NW], s_pok[NW];

  for (int t = tid; t < 2 * XS; t += BLOCK) (&s_x[0][0])[t] = 0.0;
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
    ss = group_sum(ss, tpc);
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
  bool pending = false;  // rank-1 update of the previous step not applied yet
  double tw_pending = 0.0;

  for (int s = 0;; ++s) {
    const int par = s & 1;
    const int rk = p.offset + s;
    const bool done_steps = (s == p.kmax);
    // ---------------- A: block candidate (largest norm, then lowest position) -----------
    {
      double wm;
      const int wl = wave_argmax(active ? vn1 : -1.0, active ? pos : INT_MAX, tpc, &wm);
      const int wf = __any(myflag) ? 1 : 0;
      const int wposv = __builtin_amdgcn_readlane(pos, wl < 0 ? 0 : wl);
      if (lane == 0) {
        s_wval[wave] = wm;
        s_wpos[wave] = wl < 0 ? INT_MAX : wposv;
        s_wlc[wave] = wl < 0 ? -1 : (wave << lgpw) + (wl >> ltpc);
        s_wflag[wave] = wf;
      }
    }
    __syncthreads();
    double cv = s_wval[0];
    int cp = s_wpos[0], clc = s_wlc[0], cflag = s_wflag[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) {
      const double ov = s_wval[w];
      const int op = s_wpos[w];
      if (ov > cv || (ov == cv && op < cp)) { cv = ov; cp = op; clc = s_wlc[w]; }
      cflag |= s_wflag[w];
    }
    myflag = 0;
    const uint64_t ctag = ((uint64_t)p.gen << 32) | (uint64_t)(s + 1);
    const size_t sl = (size_t)par * p.G + g;
    // publish {norm, position} at once: the reflector record follows (its chunk tags let a
    // reader wait for it), so building it is off the exchange's critical path
    if (COOP && !done_steps && tid == 0) {
      if (!p.uncached) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      u64x2 w;
      w.x = (uint64_t)__double_as_longlong(cv);
      w.y = pack_tag(s + 1, cp, clc, cflag);
      put_chunk(reinterpret_cast<u64x2*>(p.slots) + sl, w, p.uncached);
      asm volatile("" ::: "memory");
    }

    // ---------------- B: the candidate column builds its reflector (dlarfg) -------------
    // The candidate first finishes the previous step's update.  G > 1: the reflector is
    // published as a record; G == 1: it goes straight to s_x[par].
    if (!done_steps && clc >= 0 && wave == (clc >> lgpw)) {
      const bool cand = have && lc == clc;
      if (cand && pending) {
        const double* vo = s_x[par ^ 1] + rbase;
#pragma unroll
        for (int i = 0; i < RPT; ++i) a[i] -= tw_pending * vo[i];
        pending = false;
      }
      const int lo = rk - rbase;
      double ss = 0.0, al = 0.0;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        ss += (i > lo) ? a[i] * a[i] : 0.0;
        al = (i == lo) ? a[i] : al;
      }
      ss = group_sum(ss, tpc);
      const double alpha0 = __shfl(al, glead | (rk / RPT), 64);
      double alpha = alpha0, xnorm = sqrt(ss), beta = alpha0, tau = 0.0, scal = 1.0, fac = 1.0;
      const int len = m - rk;
      if (len > 1 && xnorm != 0.0) {
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
          double s2 = 0.0;
#pragma unroll
          for (int i = 0; i < RPT; ++i) {
            const double xi = a[i] * fac;
            s2 += (i > lo) ? xi * xi : 0.0;
          }
          s2 = group_sum(s2, tpc);
          xnorm = sqrt(s2);
          beta = -copysign(lapy2(alpha, xnorm), alpha);
        }
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
        for (int j = 0; j < knt; ++j) beta *= safmin;
      }
      // the candidate's v goes to s_x[par] (rows <= rk get zeros; row rk becomes 1 in E),
      // unconditionally, so no per-row predicate or branch is needed
      if (cand) {
        double* d = s_x[par] + rbase;
#pragma unroll
        for (int i = 0; i < RPT; ++i) d[i] = (i > lo) ? (a[i] * fac) * scal : 0.0;
      }
      // the candidate group's first lane holds tau and beta
      const int cl = (clc & ((1 << lgpw) - 1)) << ltpc;
      const double ctau = readlane_f64(tau, cl), cbeta = readlane_f64(beta, cl);
      if (lane == 0) { s_tau = ctau; s_beta = cbeta; }
      if (COOP) {
        // the whole wave publishes {tau, beta, v(rk+1..m-1)} as coalesced tagged chunks
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        u64x2* rec = reinterpret_cast<u64x2*>(p.cand) + sl * p.cand_stride;
        const double* xs = s_x[par];
        for (int t = lane; t < len + 1; t += 64) {
          u64x2 w;
          const double val = t == 0 ? ctau : (t == 1 ? cbeta : xs[rk + t - 1]);
          w.x = (uint64_t)__double_as_longlong(val);
          w.y = ctag;
          put_chunk(rec + t, w, p.uncached);
        }
        asm volatile("" ::: "memory");  // the stores stay ahead of the exchange below
        if (!p.uncached) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      }
    }

    // ---------------- C: exchange (G > 1) while the other columns finish their update ----
    // Wave w polls slots [64w, 64w+64) and reduces them; the wave results meet in LDS.
    if (COOP && !done_steps) {
      volatile u64x2* slots = reinterpret_cast<volatile u64x2*>(p.slots);
      if (pending) {
        const double* vo = s_x[par ^ 1] + rbase;
#pragma unroll
        for (int i = 0; i < RPT; ++i) a[i] -= tw_pending * vo[i];
        pending = false;
      }
      const uint64_t want = (uint64_t)(s + 1) & TAG_MASK;
      const int t = (wave << 6) + lane;
      const bool has_slot = t < p.G;
      const volatile u64x2* src = slots + (size_t)par * p.G + t;
      u64x2 mine;
      mine.x = 0;
      mine.y = 0;
      int ok = 1;
      if ((wave << 6) < p.G) {  // wave-uniform: this wave has slots to watch (bounded wait)
        unsigned spins = 0;
        for (;;) {
          bool ready = true;
          if (has_slot) {
            mine = *src;
            ready = (mine.y & TAG_MASK) == want;
          }
          if (__all(ready)) break;
          if ((++spins & 63) == 0) {
            if (__hip_atomic_load(p.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { ok = 0; break; }
            if (spins > LRA_SPIN_LIMIT) {
              if (lane == 0) raise_error(p, LRA_E_TIMEOUT);
              ok = 0;
              break;
            }
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (!p.uncached) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      double v = -1.0;
      int vp = INT_MAX, vl = -1, fl = 0;
      if (ok && has_slot) {
        const uint64_t tg = mine.y;
        const uint64_t pp = (tg >> 28) & POS_NONE;
        const uint64_t ll = (tg >> 52) & LC_NONE;
        v = __longlong_as_double((long long)mine.x);
        vp = pp == POS_NONE ? INT_MAX : (int)pp;
        vl = ll == LC_NONE ? -1 : (int)ll;
        fl = (int)(tg >> 63);
      }
      double vmax;
      const int wl = wave_argmax(v, vp, 1, &vmax);
      const int anyfl = __any(fl) ? 1 : 0;
      const int wpos = __builtin_amdgcn_readlane(vp, wl < 0 ? 0 : wl);
      const int wlc = __builtin_amdgcn_readlane(vl, wl < 0 ? 0 : wl);
      if (lane == 0) {
        s_pv[wave] = vmax;
        s_pp[wave] = wl < 0 ? INT_MAX : wpos;
        s_po[wave] = wl < 0 ? -1 : (wave << 6) + wl;
        s_pl[wave] = wl < 0 ? -1 : wlc;
        s_pf[wave] = anyfl;
        s_pok[wave] = ok;
      }
    } else if (pending) {
      const double* vo = s_x[par ^ 1] + rbase;
#pragma unroll
      for (int i = 0; i < RPT; ++i) a[i] -= tw_pending * vo[i];
      pending = false;
    }
    __syncthreads();
    // the global winner: every thread reduces the wave results the same way
    double win_val = cv;
    int win_pos = cp, win_owner = g, win_lc = clc, flagprev = cflag, wok = 1;
    if (COOP && !done_steps) {
      win_val = s_pv[0]; win_pos = s_pp[0]; win_owner = s_po[0]; win_lc = s_pl[0];
      flagprev = s_pf[0]; wok = s_pok[0];
#pragma unroll
      for (int w = 1; w < NW; ++w) {
        const double ov = s_pv[w];
        const int op = s_pp[w];
        if (ov > win_val || (ov == win_val && op < win_pos)) {
          win_val = ov; win_pos = op; win_owner = s_po[w]; win_lc = s_pl[w];
        }
        flagprev |= s_pf[w];
        wok &= s_pok[w];
      }
    }
    if (!wok) return;  // bounded wait expired or another block failed: no write-out
    if (!done_steps && (win_owner < 0 || win_lc < 0)) {
      // no candidate anywhere (only possible with NaN norms) is an error, not a pivot
      if (tid == 0) raise_error(p, LRA_E_STATE);
      return;
    }

    // ---------------- D: block bookkeeping for the step just finished (s-1) -------------
    if (s == 0) {
      if (tid == 0) s_ptol = fmax(p.atol, p.rtol * win_val);
      // (read after the next barrier)
    } else {
      const int last = s - 1;
      const bool size_end = (last - j0 + 1 == jb);
      const bool blk_end = size_end || (flagprev != 0) || done_steps;
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

    // ---------------- E: the winner's reflector -> s_x[par] ------------------------------
    const int len = m - rk;
    double* xs = s_x[par];
    if (tid == 0) {
      if (rk >= 1) xs[rk - 1] = 0.0;  // rows of the reflector two steps back
      if (rk >= 2) xs[rk - 2] = 0.0;
      xs[rk] = 1.0;
    }
    if (COOP && win_owner != g) {  // the owner already holds its reflector in LDS
      const u64x2* rec = reinterpret_cast<const u64x2*>(p.cand) +
                         ((size_t)par * p.G + win_owner) * p.cand_stride;
      // up to 4 chunks per thread in flight at once, then validated (a chunk that has not
      // landed yet is re-read with a bounded wait)
      for (int t0 = tid; t0 < len + 1; t0 += 4 * BLOCK) {
        u64x2 w[4];
        if (p.uncached) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int t = t0 + q * BLOCK;
            w[q] = t < len + 1 ? rec[t] : u64x2{0, 0};
          }
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int t = t0 + q * BLOCK;
            if (t < len + 1) w[q] = *(const volatile u64x2*)(rec + t);
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int t = t0 + q * BLOCK;
          if (t >= len + 1) continue;
          if (w[q].y != ctag) w[q] = wait_chunk(p, rec + t, ctag);
          const double val = __longlong_as_double((long long)w[q].x);
          if (t == 0) s_tau = val;
          else if (t == 1) s_beta = val;
          else xs[rk + t - 1] = val;
        }
      }
    }
    __syncthreads();
    const double tau = s_tau, beta = s_beta;
    if (tid == 0) {
      if (g == 0) p.tau[s] = tau;
      s_diag[s - j0] = beta;  // read after the next step's barriers
    }

    // ---------------- F: apply H to the block's columns, position swap, norm downdate ----
    const bool is_win = have && active && (g == win_owner) && (lc == win_lc);
    if (is_win) {
      const double* vx = xs + rbase;  // zero in the padding rows
      const int lo = rk - rbase;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        if (i == lo) a[i] = beta;
        else if (i > lo) a[i] = vx[i];
      }
      active = false;
      pos = s;
    } else {
      if (have && pos == s) pos = win_pos;  // LAPACK swaps columns k and pvt
    }
    // every lane runs the dot (group sums need the whole wave); inactive lanes discard it
    {
      const double* vx = xs + rbase;
      double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
#pragma unroll
      for (int i = 0; i < RPT; i += 4) {
        w0 += vx[i] * a[i];
        if (i + 1 < RPT) w1 += vx[i + 1] * a[i + 1];
        if (i + 2 < RPT) w2 += vx[i + 2] * a[i + 2];
        if (i + 3 < RPT) w3 += vx[i + 3] * a[i + 3];
      }
      // re-read v from LDS below instead of keeping RPT values live across the reductions
      asm volatile("" ::: "memory");
      const double w = group_sum((w0 + w1) + (w2 + w3), tpc);
      const double tw = tau * w;
      // new value of row rk (v(rk) = 1): lives in lane rk / RPT of the group
      const int irk = rk % RPT;
      double ar = 0.0;
#pragma unroll
      for (int i = 0; i < RPT; ++i) ar = (i == irk) ? a[i] : ar;
      ar = __shfl(ar, glead | (rk / RPT), 64);
      if (active) {
        pending = true;
        tw_pending = tw;
        if (rk + 1 < p.lastrk && vn1 != 0.0) {
          const double xr = ar - tw;
          double temp = fabs(xr) / vn1;
          temp = (1.0 + temp) * (1.0 - temp);
          temp = temp < 0.0 ? 0.0 : temp;
          const double r = vn1 / vn2;
          const double temp2 = temp * r * r;
          if (temp2 <= tol3z) myflag = 1;
          else vn1 *= sqrt(temp);
        }
      }
      // columns whose norm must be recomputed finish the update now (group-uniform)
      const bool recompute = active && myflag;
      if (__any(recompute)) {
        if (recompute) {
#pragma unroll
          for (int i = 0; i < RPT; ++i) a[i] -= tw * vx[i];
          pending = false;
        }
        double ss = 0.0;
        const int lo = rk - rbase;
#pragma unroll
        for (int i = 0; i < RPT; ++i) ss += (i > lo) ? a[i] * a[i] : 0.0;
        ss = group_sum(ss, tpc);
        if (recompute) vn1 = vn2 = sqrt(ss);
      }
    }
    // s_x[par] is rewritten two steps later, after two more barriers
  }

  // (every exit of the loop comes after phase C, so no update is pending here)
  // ---------------- write-out to LAPACK positions --------------------------------------
  // (with G > 1 every block loaded its columns before step 0's exchange, so the in-place
  // permuted write cannot overwrite a column somebody still has to read)
  if (have) {
    double* dst = p.A + (i64)pos * p.lda + rbase;
    int mm = m;
    asm volatile("" : "+s"(mm));  // keep the row masks below from being computed early
    const int hi = mm - rbase;
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

template <int RPT, int BLOCK, bool COOP>
__global__ __launch_bounds__(BLOCK) void qrcp_kernel(QrcpDev p) {
  qrcp_body<RPT, BLOCK, COOP>(p, blockIdx.x);
}

// Two independent factorizations (two adaptive rounds' sketches of one matrix) in one
// co-resident grid: blocks [0, g1) factor d0, blocks [g1, g1 + d1.G) factor d1.  Each has its
// own exchange slots and records, so neither waits for the other; the shorter one ends first
// and leaves its CUs to the longer one.
template <int RPT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void qrcp_pair_kernel(QrcpDev d0, QrcpDev d1, int g1) {
  const bool second = (int)blockIdx.x >= g1;
  qrcp_body<RPT, BLOCK, true>(second ? d1 : d0, second ? (int)blockIdx.x - g1 : (int)blockIdx.x);
}

template <int RPT, int BLOCK>
hipError_t launch_pair_rpt(const QrcpDev& d0, const QrcpDev& d1, hipStream_t st) {
  dim3 grid(d0.G + d1.G), block(BLOCK);
  int g1 = d0.G;
  void* args[] = {(void*)&d0, (void*)&d1, (void*)&g1};
  return hipLaunchCooperativeKernel((const void*)qrcp_pair_kernel<RPT, BLOCK>, grid, block, args,
                                    0, st);
}

template <int RPT, int BLOCK, bool COOP>
hipError_t launch_rpt(const QrcpDev& d, hipStream_t st) {
  dim3 grid(d.G), block(BLOCK);
  if constexpr (COOP) {
    void* args[] = {(void*)&d};
    return hipLaunchCooperativeKernel((const void*)qrcp_kernel<RPT, BLOCK, true>, grid, block,
                                      args, 0, st);
  } else {
    hipLaunchKernelGGL((qrcp_kernel<RPT, BLOCK, false>), grid, block, 0, st, d);
    return hipGetLastError();
  }
}

#define LRA_RPT_256(X) X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32) X(36) X(40) X(44) X(48) \
  X(52) X(56) X(60) X(64) X(68) X(72) X(76) X(80)
#define LRA_RPT_512(X) X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32) X(36) X(40) X(44)
#define LRA_RPT_1024(X) X(4) X(8)

#define LRA_RPT_512C(X) X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32) X(36) X(40)

// Single-block shapes (512 and 1024 threads, or 256 with G == 1) run without the exchange;
// cooperative grids use 256- or 512-thread blocks with it.
hipError_t launch_any(int rpt, int block, const QrcpDev& d, hipStream_t st) {
  const bool coop = d.G > 1;
  if (block == 1024 && !coop) {
    switch (rpt) {
#define CASE(R) case R: return launch_rpt<R, 1024, false>(d, st);
      LRA_RPT_1024(CASE)
#undef CASE
    }
  } else if (block == 512 && !coop) {
    switch (rpt) {
#define CASE(R) case R: return launch_rpt<R, 512, false>(d, st);
      LRA_RPT_512(CASE)
#undef CASE
    }
  } else if (block == 512 && coop) {
    switch (rpt) {
#define CASE(R) case R: return launch_rpt<R, 512, true>(d, st);
      LRA_RPT_512C(CASE)
#undef CASE
    }
  } else if (block == 256) {
    switch (rpt) {
#define CASE(R) case R: return coop ? launch_rpt<R, 256, true>(d, st) : launch_rpt<R, 256, false>(d, st);
      LRA_RPT_256(CASE)
#undef CASE
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_pair_any(int rpt, int block, const QrcpDev& d0, const QrcpDev& d1,
                           hipStream_t st) {
  if (block == 512) {
    switch (rpt) {
#define CASE(R) case R: return launch_pair_rpt<R, 512>(d0, d1, st);
      LRA_RPT_512C(CASE)
#undef CASE
    }
  } else if (block == 256) {
    switch (rpt) {
#define CASE(R) case R: return launch_pair_rpt<R, 256>(d0, d1, st);
      LRA_RPT_256(CASE)
#undef CASE
    }
  }
  return hipErrorInvalidValue;
}

int occupancy_pair(int rpt, int block, int* nb) {
  hipError_t e = hipErrorInvalidValue;
  if (block == 256) {
    switch (rpt) {
#define CASE(R) case R: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, qrcp_pair_kernel<R, 256>, 256, 0); break;
      LRA_RPT_256(CASE)
#undef CASE
    }
  } else if (block == 512) {
    switch (rpt) {
#define CASE(R) case R: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, qrcp_pair_kernel<R, 512>, 512, 0); break;
      LRA_RPT_512C(CASE)
#undef CASE
    }
  }
  return e == hipSuccess ? 0 : 1;
}

// co-resident blocks per CU of the cooperative variant
int occupancy_coop(int rpt, int block, int* nb) {
  hipError_t e = hipErrorInvalidValue;
  if (block == 256) {
    switch (rpt) {
#define CASE(R) case R: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, qrcp_kernel<R, 256, true>, 256, 0); break;
      LRA_RPT_256(CASE)
#undef CASE
    }
  } else if (block == 512) {
    switch (rpt) {
#define CASE(R) case R: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, qrcp_kernel<R, 512, true>, 512, 0); break;
      LRA_RPT_512C(CASE)
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

// Estimated microseconds per pivot step of a cooperative shape, fitted to measured steps on
// MI355X (n = 8192, ell = 40..520): a fixed exchange + barrier cost, plus the serial work of
// a lane, which grows with the rows it holds (two waves per SIMD with 512-thread blocks hide
// part of it), plus a penalty for grids too small to spread the columns thin.
static double coop_step_cost(int rpt, int block, int G) {
  const double rows = rpt > 12 ? rpt - 12 : 0;
  double c = block == 256 ? 6.3 + 0.10 * rows : 7.2 + 0.085 * rows;
  for (int g = G; g < 256; g *= 2) c += 0.5;
  return c;
}

static bool choose_shape(lra_ctx* ctx, int m, int n, QrcpShape* out) {
  for (int tpc = 1; tpc <= 64; tpc <<= 1) {
    int rpt = (m + tpc - 1) / tpc;
    rpt = (rpt + 3) / 4 * 4;
    if (rpt < 4) rpt = 4;
    if (rpt <= 8 && n <= 1024 / tpc) {
      *out = {tpc, rpt, 1024, 1, 1024 / tpc};
      return true;
    }
    if (rpt <= 44 && n <= 512 / tpc) {
      *out = {tpc, rpt, 512, 1, 512 / tpc};
      return true;
    }
  }
  // LRA_QRCP_SHAPE=block:tpc forces a cooperative shape (when it fits)
  int fblock = 0, ftpc = 0;
  if (const char* e = getenv("LRA_QRCP_SHAPE")) {
    if (sscanf(e, "%d:%d", &fblock, &ftpc) != 2) fblock = ftpc = 0;
  }
  bool found = false;
  double best = 1e30;
  for (int block : {256, 512}) {
    const int rpt_max = block == 256 ? 80 : 40;
    for (int tpc = 1; tpc <= 64; tpc <<= 1) {
      if (fblock && (block != fblock || tpc != ftpc)) continue;
      int rpt = (m + tpc - 1) / tpc;
      rpt = (rpt + 3) / 4 * 4;
      if (rpt < 4) rpt = 4;
      if (rpt > rpt_max) continue;
      const int cpw = block / tpc;
      const int G = (n + cpw - 1) / cpw;
      if (G > 256) continue;  // at most one block per CU is co-resident with these registers
      int per_cu = 0;
      if (occupancy_coop(rpt, block, &per_cu) != 0 || per_cu < 1) continue;
      if (G > 1 && G > per_cu * ctx->num_cu) continue;
      const double cost = coop_step_cost(rpt, block, G);
      if (!found || cost < best) {
        best = cost;
        *out = {tpc, rpt, block, G, cpw};
        found = true;
      }
    }
  }
  return found;
}

static QrcpDev dev_of(lra_ctx* ctx, const QrcpCall& q, const QrcpShape& sh) {
  QrcpDev d;
  d.A = q.A; d.lda = q.lda; d.m = q.m; d.n = q.n; d.offset = q.offset; d.kmax = q.kmax;
  d.nb = q.nb; d.mode = q.mode; d.stop_after = q.stop_after;
  d.lastrk = std::min(q.m, q.n + q.offset);
  d.atol = q.atol; d.rtol = q.rtol; d.jpvt = q.jpvt; d.tau = q.tau; d.vn1g = q.vn1;
  d.vn2g = q.vn2; d.k_out = q.k_out; d.early_out = q.early_out;
  d.G = sh.G; d.cpw = sh.cpw; d.tpc = sh.tpc;
  d.ltpc = 0;
  while ((1 << d.ltpc) < sh.tpc) ++d.ltpc;
  // tau, beta and reflector rows up to the column group's padded height tpc*rpt
  d.cand_stride = ((sh.tpc * sh.rpt + 2 + 7) / 8) * 8;
  d.err = ctx->d_err;
  return d;
}

// The shapes of a pair launch: the larger sketch (rows m1 >= m0) takes the shape it would
// get alone; the smaller one takes the same rows per lane and block size with the fewest
// lanes per column that hold it.  Both must be cooperative, each within its slot exchange, and
// together co-resident.
// The smaller round's shape for a pair launched with the larger round's shape s1 (same rows
// per lane and block size: one kernel instance runs both), when both grids are co-resident.
static bool pair_fit(lra_ctx* ctx, int m0, int n, const QrcpShape& s1, QrcpShape* s0) {
  int tpc = 1;
  while (tpc * s1.rpt < m0) tpc <<= 1;
  if (tpc > 64) return false;
  const int cpw = s1.block / tpc;
  const int G = (n + cpw - 1) / cpw;
  if (G <= 1 || G > s1.block || s1.G <= 1 || s1.G > s1.block) return false;
  int per_cu = 0;
  if (occupancy_pair(s1.rpt, s1.block, &per_cu) != 0 || per_cu < 1) return false;
  if (G + s1.G > per_cu * ctx->num_cu) return false;
  *s0 = {tpc, s1.rpt, s1.block, G, cpw};
  return true;
}

// Shapes of a pair: the larger round's own best shape when the pair fits with it, else the
// cheapest shape of the larger round (by the step-cost model) with which it does.
static bool pair_shapes(lra_ctx* ctx, int m0, int m1, int n, QrcpShape* s0, QrcpShape* s1) {
  if (m0 > m1 || n <= 0) return false;
  if (!choose_shape(ctx, m1, n, s1) || s1->G <= 1) return false;
  if (pair_fit(ctx, m0, n, *s1, s0)) return true;
  bool found = false;
  double best = 1e30;
  QrcpShape b0{}, b1{};
  for (int block : {256, 512}) {
    const int rpt_max = block == 256 ? 80 : 40;
    for (int tpc = 1; tpc <= 64; tpc <<= 1) {
      int rpt = (m1 + tpc - 1) / tpc;
      rpt = (rpt + 3) / 4 * 4;
      if (rpt < 4) rpt = 4;
      if (rpt > rpt_max) continue;
      const int cpw = block / tpc;
      const int G = (n + cpw - 1) / cpw;
      if (G > 256 || G <= 1) continue;
      const QrcpShape t1 = {tpc, rpt, block, G, cpw};
      QrcpShape t0;
      if (!pair_fit(ctx, m0, n, t1, &t0)) continue;
      const double cost = coop_step_cost(rpt, block, G);
      if (!found || cost < best) {
        best = cost;
        b0 = t0;
        b1 = t1;
        found = true;
      }
    }
  }
  if (found) {
    *s0 = b0;
    *s1 = b1;
  }
  return found;
}

bool lra_qrcp_pair_fits(lra_ctx* ctx, int m0, int m1, int n) {
  const char* e = getenv("LRA_QRCP_PAIR");
  if (e && e[0] == '0') return false;
  QrcpShape s0, s1;
  return pair_shapes(ctx, m0, m1, n, &s0, &s1);
}

// Launch generations are process-wide: an exchange area passed from a destroyed context to a
// new one can then never hold a chunk tag that the new owner's launches would accept.
uint32_t lra_next_qrcp_gen() {
  static std::atomic<uint32_t> gen{0};
  uint32_t g = gen.fetch_add(1, std::memory_order_relaxed) + 1;
  return g ? g : gen.fetch_add(1, std::memory_order_relaxed) + 1;
}

int lra_qrcp_pair(lra_ctx* ctx, const QrcpCall& q0, const QrcpCall& q1) {
  QrcpShape s0, s1;
  if (q0.mode != 0 || q1.mode != 0 || q0.n != q1.n || q0.kmax <= 0 || q1.kmax <= 0 ||
      q0.nb > NB_MAX || q1.nb > NB_MAX || q0.n >= (int)POS_NONE ||
      !pair_shapes(ctx, q0.m, q1.m, q0.n, &s0, &s1)) {
    lra_set_error("qrcp pair: shapes %dx%d and %dx%d do not fit one co-resident grid", q0.m,
                  q0.n, q1.m, q1.n);
    return LRA_E_UNSUPPORTED;
  }
  QrcpDev d0 = dev_of(ctx, q0, s0), d1 = dev_of(ctx, q1, s1);
  // exchange area: [header 256 B][slots 2 G0][slots 2 G1][records 2 G0 x stride0][2 G1 x stride1]
  const size_t sb0 = 16 * 2 * (size_t)s0.G, sb1 = 16 * 2 * (size_t)s1.G;
  const size_t cb0 = sb0 * d0.cand_stride, cb1 = sb1 * d1.cand_stride;
  void* area = nullptr;
  bool uc = false;
  LRA_CHECK(lra_coop_area(ctx, 256 + sb0 + sb1 + cb0 + cb1, &area, &uc));
  char* base = (char*)area;
  d0.uncached = d1.uncached = uc ? 1 : 0;
  d0.abort_flag = d1.abort_flag = (int*)base;
  d0.slots = (uint64_t*)(base + 256);
  d1.slots = (uint64_t*)(base + 256 + sb0);
  d0.cand = (double*)(base + 256 + sb0 + sb1);
  d1.cand = (double*)(base + 256 + sb0 + sb1 + cb0);
  LRA_CHECK(lra_launch_zero_u32(ctx->stream, (unsigned*)base, (256 + sb0 + sb1) / 4));
  d0.gen = d1.gen = lra_next_qrcp_gen();
  hipError_t e;
  {
    SpinSection gate(ctx);
    e = launch_pair_any(s1.rpt, s1.block, d0, d1, ctx->stream);
  }
  if (e != hipSuccess) {
    lra_set_error("qrcp pair launch (%dx%d tpc=%d, %dx%d tpc=%d, rpt=%d G=%d+%d): %s", q0.m, q0.n,
                  s0.tpc, q1.m, q1.n, s1.tpc, s1.rpt, s0.G, s1.G, hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
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
  // LRA_QRCP_GLOBAL=1 sends every mode-0 call to the global-memory path (used by the tests)
  const char* fg = getenv("LRA_QRCP_GLOBAL");
  const bool force_global = fg && fg[0] == '1';
  if ((force_global && q.mode == 0 && q.offset == 0) || !choose_shape(ctx, q.m, q.n, &sh))
    return lra_qrcp_global(ctx, q);
  QrcpDev d = dev_of(ctx, q, sh);
  if (sh.G > sh.block) {
    lra_set_error("qrcp: %d blocks exceed the %d-slot exchange", sh.G, sh.block);
    return LRA_E_UNSUPPORTED;
  }
  // exchange area: [header 256 B: abort flag][slots 2G x 16 B][records 2G x stride x 16 B]
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
  d.gen = lra_next_qrcp_gen();
  hipError_t e;
  if (sh.G > 1) {
    SpinSection gate(ctx);
    e = launch_any(sh.rpt, sh.block, d, ctx->stream);
  } else {
    e = launch_any(sh.rpt, sh.block, d, ctx->stream);
  }
  if (e != hipSuccess) {
    lra_set_error("qrcp launch (m=%d n=%d tpc=%d rpt=%d G=%d): %s", q.m, q.n, sh.tpc, sh.rpt,
                  sh.G, hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
}
