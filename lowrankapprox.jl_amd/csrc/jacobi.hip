// jacobi.hip -- one-sided (Hestenes) Jacobi SVD of a small dense matrix: the svd!/eigen! of
// the k x k cores of psvdfact and pheigfact (src/psvd.jl:245,259, src/pheig.jl:284).
//
// X (m x n, m >= n) is orthogonalised column pair by column pair: for a pair (p, q) with
// a = |x_p|^2, b = |x_q|^2, g = x_p'x_q, the rotation that diagonalises [a g; g b] is applied
// to the two columns of X and of V.  A sweep visits every pair once in the round-robin
// ("circle") order; sweeps repeat until a sweep rotates nothing (|g| <= tol sqrt(a b)).  At
// the end the column norms of X are the singular values, X's normalised columns the left
// vectors, V the right ones (X_in = U diag(s) V'), sorted by decreasing s.  One-sided Jacobi
// computes the small singular values to high relative accuracy.
//
// Block-cyclic order (jacobi_block_kernel): the columns form NB blocks of W; one launch is a
// round of the circle order over blocks, a workgroup per block pair.  The workgroup keeps its
// 2W columns of X in LDS and visits every cross pair of its two blocks (W steps of W disjoint
// pairs; the first round of a sweep also visits each block's inner pairs), each rotation from
// dot products of the current columns, exactly as the pairwise method.  The rotations are
// accumulated in a 2W x 2W matrix Q that updates V's 2W columns once per round.  A sweep
// (every pair once) is NB - 1 launches instead of n - 1.  Matrices taller than the LDS holds
// use one launch per round of single pairs (jacobi_round_kernel).
#include "lra_internal.h"
#include "lra_wave.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

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

// Sum over the L consecutive lanes of a group (L = 16, 32 or 64); every lane of the group ends
// with the bit-identical sum (each step adds the same two operands in both partner lanes).
template <int L>
LRA_DEV double group_sum_u(double v) {
  v += dpp_f64<DPP_XOR1>(v);
  v += dpp_f64<DPP_XOR2>(v);
  v += dpp_f64<DPP_HMIRROR>(v);
  v += dpp_f64<DPP_MIRROR>(v);
  if constexpr (L == 32) v += __shfl_xor(v, 16, 64);
  if constexpr (L == 64)
    v = (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
  return v;
}

// 1/sqrt(x) for x in [1, 2]: the hardware estimate and two Newton steps
LRA_DEV double rsqrt_12(double x) {
  double r = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  r = r * (1.5 - hx * r * r);
  r = r * (1.5 - hx * r * r);
  return r;
}

// One round r of the circle order over NB blocks of W columns (a sweep's first round, first
// = 1, also visits the pairs inside each block).  A pair is handled by L lanes (L = 256 / W
// for W >= 4, so the W pairs of a step fill one wave per SIMD), each holding MR rows in
// registers; Xs columns are zero-padded to mp = L * MR >= m rows (stride mp + 16).  Dynamic
// LDS: Xs (2W columns), Qs (2W x 2W, row-major with a padded stride, so a column of Q is read
// bank-conflict free and a row contiguously).  BT threads: 512, or 256 (one wave per SIMD,
// room for 2 MR > 72 registers) for the tall blocks of W <= 2 with MR > 36.
template <int W, int MR, int BT>
__global__ __launch_bounds__(BT) void jacobi_block_kernel(int m, int n, int NB, int r, int first,
                                                           double* __restrict__ X, i64 ldx,
                                                           double* __restrict__ V, i64 ldv,
                                                           double tol, int* __restrict__ rotated,
                                                           const int* __restrict__ prev) {
  // a sweep enqueued before the host saw the previous one's count: nothing to do when the
  // previous sweep rotated nothing (it completed before this launch began)
  if (prev && *(const volatile int*)prev == 0) return;
  constexpr int C = 2 * W;
  constexpr int QS = C + 1;                 // row stride of Qs
  constexpr int NW = BT / 64;               // waves
  constexpr int L = W >= 16 ? 16 : W == 8 ? 32 : 64;  // lanes per pair
  constexpr int G = 64 / L;                 // pairs per wave
  constexpr int mp = L * MR;                // padded column length in LDS
  constexpr int mps = mp + 16;              // column stride (128 bytes of skew between columns)
  extern __shared__ double lds[];
  double* Xs = lds;
  double* Qs = lds + (size_t)C * mps;  // Qs[t * QS + j] = Q(t, j)
  __shared__ int any;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int P, Q;
  rr_pair(NB, r, blockIdx.x, &P, &Q);
  if (P > Q) { const int t = P; P = Q; Q = t; }
  // local slot c -> global column: slots [0, W) are block P, [W, 2W) block Q (P < Q, so the
  // local order is the global order)
  auto gcol = [&](int c) { return c < W ? P * W + c : Q * W + (c - W); };
  // Xs(i, c) = X(i, gcol(c)), zero for padding rows and columns; every load of a thread is
  // issued before the first store
  for (int i0 = 0; i0 < mp; i0 += BT) {
    const int i = i0 + tid;
    double tmp[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int g = gcol(c);
      const bool ok = i < m && g < n;
      const double x = X[ok ? (i64)g * ldx + i : 0];
      tmp[c] = ok ? x : 0.0;
    }
    if (i < mp) {
#pragma unroll
      for (int c = 0; c < C; ++c) Xs[c * mps + i] = tmp[c];
    }
  }
  for (int e = tid; e < C * QS; e += BT) Qs[e] = (e / QS) == (e % QS) ? 1.0 : 0.0;
  if (tid == 0) any = 0;
  __syncthreads();
  const double tol2 = tol * tol;
  int nrot = 0;
  const int grp = lane / L, gl = lane % L;
  const int pi = wave * G + grp;  // pair index in the step
  const bool act = pi < W;
  // steps: W - 1 inner steps (first round only), then W cross steps; every step holds W
  // disjoint pairs
  const int ninner = first && W > 1 ? W - 1 : 0;
  const bool busy = wave * G < W;  // waves that hold pairs (wave-uniform)
  // rotate the pair (x1 = u, x2 = w) when |g| > tol sqrt(a b); returns whether it did
  auto pair_step = [&](double (&u)[MR], double (&w)[MR], int c1, int c2, bool store1) -> bool {
    double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0, g0 = 0.0, g1 = 0.0;
#pragma unroll
    for (int q = 0; q < MR; q += 2) {
      a0 += u[q] * u[q];
      b0 += w[q] * w[q];
      g0 += u[q] * w[q];
      if (q + 1 < MR) {
        a1 += u[q + 1] * u[q + 1];
        b1 += w[q + 1] * w[q + 1];
        g1 += u[q + 1] * w[q + 1];
      }
    }
    const double a = group_sum_u<L>(a0 + a1);
    const double b = group_sum_u<L>(b0 + b1);
    const double g = group_sum_u<L>(g0 + g1);
    // |g| > tol sqrt(a b), without the square roots
    if (!(act && g != 0.0 && g * g > tol2 * (a * b))) return false;
    // t = tan(theta), the smaller root of t^2 + 2 zeta t - 1 = 0 with zeta = d / e:
    // t = sign(d) e / (|d| + sqrt(d^2 + e^2))
    const double d = b - a, e = 2.0 * g;
    const double h = sqrt(d * d + e * e);
    const double t = copysign(1.0, d) * (e / (fabs(d) + h));
    const double cs = rsqrt_12(1.0 + t * t), sn = cs * t;
    double* x1 = Xs + c1 * mps + gl;
    double* x2 = Xs + c2 * mps + gl;
#pragma unroll
    for (int q = 0; q < MR; ++q) {
      const double y = u[q], z = w[q];
      u[q] = cs * y - sn * z;
      w[q] = sn * y + cs * z;
      if (store1) x1[q * L] = u[q];
      x2[q * L] = w[q];
    }
    for (int t2 = gl; t2 < C; t2 += L) {
      double* q1 = Qs + t2 * QS + c1;
      double* q2 = Qs + t2 * QS + c2;
      const double y = *q1, z = *q2;
      *q1 = cs * y - sn * z;
      *q2 = sn * y + cs * z;
    }
    return true;
  };
  double u[MR], w[MR];
  // inner steps (a sweep's first round): W/2 pairs inside each block, circle order over W
  // players; both columns go back to LDS
  for (int step = 0; step < ninner; ++step) {
    if (busy) {
      int c1 = 0, c2 = 1;
      if (act) {
        constexpr int half = W > 1 ? W / 2 : 1;
        const int blk = pi / half, slot = pi % half;
        int a, b;
        rr_pair(W, step, slot, &a, &b);
        c1 = blk * W + min(a, b);
        c2 = blk * W + max(a, b);
      }
      const double* x1 = Xs + c1 * mps + gl;
      const double* x2 = Xs + c2 * mps + gl;
#pragma unroll
      for (int q = 0; q < MR; ++q) {
        u[q] = x1[q * L];
        w[q] = x2[q * L];
      }
      if (pair_step(u, w, c1, c2, true)) ++nrot;
    }
    __syncthreads();
  }
  // cross steps: pair pi keeps top column pi in registers for the whole round and meets
  // bottom column W + (pi + s) mod W at step s; only the bottom columns move through LDS
  const int c1 = act ? pi : 0;
  if (busy) {
    const double* x1 = Xs + c1 * mps + gl;
#pragma unroll
    for (int q = 0; q < MR; ++q) u[q] = x1[q * L];
  }
  bool moved = false;
  for (int step = 0; step < W; ++step) {
    if (busy) {
      const int c2 = act ? W + ((pi + step) & (W - 1)) : 1;
      const double* x2 = Xs + c2 * mps + gl;
#pragma unroll
      for (int q = 0; q < MR; ++q) w[q] = x2[q * L];
      if (pair_step(u, w, c1, c2, false)) {
        ++nrot;
        moved = true;
      }
    }
    __syncthreads();
  }
  if (busy && moved) {
    double* x1 = Xs + c1 * mps + gl;
#pragma unroll
    for (int q = 0; q < MR; ++q) x1[q * L] = u[q];
  }
  __syncthreads();
  if (nrot) any = 1;  // benign race: every writer stores 1
  __syncthreads();
  if (!any) return;
  // V(:, block cols) <- V(:, block cols) * Q on MFMA, as V' <- Q' V': the block of V (n x 2W)
  // is staged in LDS over Xs; lane l of a 16 x 16 output tile then holds 16 consecutive rows of
  // one column of V, so the stores are coalesced.  The first BT rows of the block are loaded
  // before X is written back, so the write-back covers their latency.
  double vt[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int g = gcol(c);
    const bool ok = tid < n && g < n;
    const double x = V[ok ? (i64)g * ldv + tid : 0];
    vt[c] = ok ? x : 0.0;
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int g = gcol(c);
    if (g >= n) continue;
    double* x = X + (i64)g * ldx;
    for (int i = tid; i < m; i += BT) x[i] = Xs[c * mps + i];
  }
  __syncthreads();  // Xs has been written out
  double* Vs = Xs;  // Vs[t * mps + i] = V(i, gcol(t))
  if (tid < n) {
#pragma unroll
    for (int c = 0; c < C; ++c) Vs[c * mps + tid] = vt[c];
  }
  for (int i0 = BT; i0 < n; i0 += BT) {
    const int i = i0 + tid;
    double tmp[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int g = gcol(c);
      const bool ok = i < n && g < n;
      const double x = V[ok ? (i64)g * ldv + i : 0];
      tmp[c] = ok ? x : 0.0;
    }
    if (i < n) {
#pragma unroll
      for (int c = 0; c < C; ++c) Vs[c * mps + i] = tmp[c];
    }
  }
  __syncthreads();
  {
    constexpr int KS = (C + 3) / 4;        // k-steps of 4
    constexpr int CT = (C + 15) / 16;      // 16-column tiles of Q
    constexpr int WPT = NW / CT;           // waves per column tile
    constexpr int TB = 2;                  // row tiles per wave whose operands load together
    const int ct = wave / WPT, rt0 = wave % WPT;
    const int rtiles = (n + 15) / 16;
    const int cq = ct * 16 + (lane & 15);  // A = Q' (row cq, k): Q(k, cq)
    double qa[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k = ks * 4 + (lane >> 4);
      qa[ks] = (k < C && cq < C) ? Qs[k * QS + cq] : 0.0;
    }
    for (int rtb = rt0; rtb < rtiles; rtb += TB * WPT) {
      double vb[TB][KS];
#pragma unroll
      for (int u = 0; u < TB; ++u) {
        const int i = (rtb + u * WPT) * 16 + (lane & 15);  // B = V' (k, i): V(i, gcol(k))
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int k = ks * 4 + (lane >> 4);
          const bool ok = k < C && i < n;
          const double x = Vs[ok ? k * mps + i : 0];
          vb[u][ks] = ok ? x : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < TB; ++u) {
        const int rt = rtb + u * WPT;
        if (rt >= rtiles) break;
        const int i = rt * 16 + (lane & 15);
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(qa[ks], vb[u][ks], acc, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int c = ct * 16 + (lane >> 4) + 4 * q;  // output (c, i) = V_new(i, gcol(c))
          if (c < C && i < n) {
            const int g = gcol(c);
            if (g < n) V[i + (i64)g * ldv] = acc[q];
          }
        }
      }
    }
  }
  if (wave == 0 && lane == 0)
    __hip_atomic_fetch_add(rotated, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// V = I and the per-sweep rotation counters rotated[0..kMaxSweeps) = 0
constexpr int kMaxSweeps = 60;
__global__ void jacobi_eye_kernel(int n, double* V, i64 ldv, int* rotated) {
  const i64 total = (i64)n * n;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x)
    V[(e % n) + (e / n) * ldv] = (e % n) == (e / n) ? 1.0 : 0.0;
  if (blockIdx.x == 0 && threadIdx.x < kMaxSweeps) rotated[threadIdx.x] = 0;
}

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

constexpr size_t kJacobiLds = 150 * 1024;

struct BlockCfg {
  int W = 0, MR = 0;
  size_t lds = 0;
};

constexpr int lanes_per_pair(int W) { return W >= 16 ? 16 : W == 8 ? 32 : 64; }

// Widest block (8, 4, 2, 1; 16 on request) whose 2W zero-padded columns of X fit in LDS with
// at most 32 rows per lane in registers (36 for W = 4: up to 2304 rows; 64 for W <= 2, in
// 256-thread workgroups beyond 36: up to 4096 rows), narrower when the matrix has few columns
// (at least 4 blocks); W = 0 (one launch per round of single pairs) when none does.  At k = 426 a sweep
// takes 0.95 ms at W = 8 and 1.06 ms at W = 16: the step cost grows with the pairs per wave
// faster than the round count falls.  LRA_JACOBI_W=<w> caps the width (tests and tuning).
BlockCfg block_cfg(i64 m, i64 n) {
  BlockCfg cfg;
  int want = 8;
  if (const char* e = getenv("LRA_JACOBI_W")) want = std::max(0, atoi(e));
  for (int W : {16, 8, 4, 2, 1}) {
    if (W > want) continue;
    if (W > 1 && (i64)4 * W > n) continue;
    const i64 L = lanes_per_pair(W);
    int MR = 0;
    for (int b : {4, 8, 16, 24, 28, 32, 36, 40, 48, 64})
      if (L * b >= m && !((W <= 2) && (b == 24 || b == 28)) &&
          (W <= 2 || b <= 32 || (W == 4 && b == 36))) {
        MR = b;
        break;
      }
    if (!MR) continue;
    const size_t lds = ((size_t)2 * W * (L * MR + 16) + (size_t)2 * W * (2 * W + 1)) * sizeof(double);
    if (lds > kJacobiLds) continue;
    cfg.W = W;
    cfg.MR = MR;
    cfg.lds = lds;
    return cfg;
  }
  return cfg;
}

template <int W, int MR>
void launch_block_wm(int NB, size_t lds, hipStream_t st, int m, int n, int r, int first,
                     double* X, i64 ldx, double* V, i64 ldv, double tol, int* rot,
                     const int* prev) {
  constexpr int BT = MR > 36 ? 256 : 512;
  static std::atomic<uint64_t> lds_set{0};
  lra_lds_limit_once(lds_set, (const void*)jacobi_block_kernel<W, MR, BT>, (int)kJacobiLds);
  hipLaunchKernelGGL((jacobi_block_kernel<W, MR, BT>), dim3((unsigned)(NB / 2)), dim3(BT), lds,
                     st, m, n, NB, r, first, X, ldx, V, ldv, tol, rot, prev);
}

template <int W>
void launch_block_w(int MR, int NB, size_t lds, hipStream_t st, int m, int n, int r, int first,
                    double* X, i64 ldx, double* V, i64 ldv, double tol, int* rot,
                    const int* prev) {
  switch (MR) {
    case 4: launch_block_wm<W, 4>(NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev); break;
    case 8: launch_block_wm<W, 8>(NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev); break;
    case 16: launch_block_wm<W, 16>(NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev); break;
    case 24:
      if constexpr (W > 2) launch_block_wm<W, 24>(NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev);
      break;
    case 28:
      if constexpr (W > 2) launch_block_wm<W, 28>(NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev);
      break;
    case 32: launch_block_wm<W, 32>(NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev); break;
    case 36:
      if constexpr (W <= 4) launch_block_wm<W, 36>(NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev);
      break;
    case 40:
      if constexpr (W <= 2) launch_block_wm<W, 40>(NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev);
      break;
    case 48:
      if constexpr (W <= 2) launch_block_wm<W, 48>(NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev);
      break;
    default:
      if constexpr (W <= 2) launch_block_wm<W, 64>(NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev);
      break;
  }
}

void launch_block(const BlockCfg& cfg, int NB, hipStream_t st, int m, int n, int r, int first,
                  double* X, i64 ldx, double* V, i64 ldv, double tol, int* rot,
                  const int* prev) {
  const int MR = cfg.MR;
  const size_t lds = cfg.lds;
  switch (cfg.W) {
    case 16: launch_block_w<16>(MR, NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev); break;
    case 8: launch_block_w<8>(MR, NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev); break;
    case 4: launch_block_w<4>(MR, NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev); break;
    case 2: launch_block_w<2>(MR, NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev); break;
    default: launch_block_w<1>(MR, NB, lds, st, m, n, r, first, X, ldx, V, ldv, tol, rot, prev); break;
  }
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
  const BlockCfg cfg = block_cfg(m, n);
  const int W = cfg.W;
  const int NB = W ? (int)((n + W - 1) / W + (((n + W - 1) / W) & 1)) : 0;
  int sw_done = 0;
  if (W) {
    // Sweep s + 1 is enqueued before the host reads sweep s's count (its kernels return at
    // once when sweep s rotated nothing), so the device does not idle while the host waits.
    // The count of sweep s goes to pinned host memory (bytes 3072.. of the context's staging
    // page) behind an event that covers sweep s only.
    int* hcnt = (int*)((char*)c->host_pinned + 3072);
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (auto& e : ev) LRA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    auto enqueue = [&](int sw) -> int {
      for (int r = 0; r < NB - 1; ++r)
        launch_block(cfg, NB, c->stream, (int)m, (int)n, r, r == 0, X, ldx, Vw, n, tol,
                     rot + sw, sw ? rot + sw - 1 : nullptr);
      LRA_HIP(hipGetLastError());
      LRA_HIP(hipMemcpyAsync(hcnt + (sw & 1), rot + sw, sizeof(int), hipMemcpyDeviceToHost,
                             c->stream));
      LRA_HIP(hipEventRecord(ev[sw & 1], c->stream));
      return LRA_OK;
    };
    int rc = enqueue(0);
    for (int sweep = 0; rc == LRA_OK && sweep < kMaxSweeps; ++sweep) {
      if (sweep + 1 < kMaxSweeps && (rc = enqueue(sweep + 1)) != LRA_OK) break;
      if (hipEventSynchronize(ev[sweep & 1]) != hipSuccess) {
        lra_set_error("gesvj: %s", hipGetErrorString(hipGetLastError()));
        rc = LRA_E_HIP;
        break;
      }
      sw_done = sweep + 1;
      if (hcnt[sweep & 1] == 0) break;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);  // released once their last record completes
    LRA_CHECK(rc);
  } else {
    for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
      for (int r = 0; r < N - 1; ++r)
        hipLaunchKernelGGL(jacobi_round_kernel, dim3(N / 2), dim3(256), 0, c->stream, (int)m,
                           (int)n, N, r, X, ldx, Vw, n, tol, rot + sweep);
      LRA_HIP(hipGetLastError());
      int h = 0;
      LRA_HIP(hipMemcpyAsync(&h, rot + sweep, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      LRA_HIP(hipStreamSynchronize(c->stream));
      sw_done = sweep + 1;
      if (h == 0) break;
    }
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
