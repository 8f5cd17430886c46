// gemm.hip -- MFMA GEMM for the sketches and the factor products (gfx950).
//
// C (M x N) = alpha * op(A) (M x K) * op(B) (K x N) + beta * C, FP64 in and out.  Replaces the
// dgemm behind mul!(C, S, A) in the Gaussian sketch (src/sketch.jl:91-110) and the products of
// the factorizations.  Arithmetic (prec):
//   F64    v_mfma_f64_16x16x4f64, FP64 accumulate (the drop-in);
//   F32    operands rounded to FP32, v_mfma_f32_16x16x4f32, FP32 accumulate;
//   BF16   operands rounded to BF16, v_mfma_f32_16x16x32_bf16, FP32 accumulate;
//   BF16X3 each operand split into three BF16 terms (hi + mid + lo ~ 24 bits), the six
//          products of order <= 2 accumulated in FP32 (~FP32 accuracy at BF16 MFMA rate).
// Tiles are staged through LDS with k contiguous ([m][k] and [n][k]), so the sketch case
// (op(A) = Omega stored K x ell, op(B) = A stored K x N) is a straight copy; the next k-tile
// is prefetched into registers while the current one is multiplied.  Blocks that share a
// B tile (the M tiles of one N tile) are mapped to the same XCD so A is fetched once into
// that XCD's L2.  Split-K writes FP64 partials that a second kernel sums in fixed order.
#include "lra_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <type_traits>
#include <condition_variable>
#include <mutex>
#include <set>
#include <tuple>
#include <vector>

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <int PREC> struct PT;
template <> struct PT<LRA_PREC_F64> { typedef double E; static const int P = 1; static const int BK = 16; static const int PAD = 1; };
template <> struct PT<LRA_PREC_F32> { typedef float E; static const int P = 1; static const int BK = 32; static const int PAD = 1; };
template <> struct PT<LRA_PREC_BF16> { typedef __bf16 E; static const int P = 1; static const int BK = 32; static const int PAD = 8; };
template <> struct PT<LRA_PREC_BF16X3> { typedef __bf16 E; static const int P = 3; static const int BK = 32; static const int PAD = 8; };

struct GemmDev {
  int M, N, K;
  double alpha, beta;
  const double* A; i64 lda; int transA;
  const double* B; i64 ldb; int transB;
  double* C; i64 ldc;
  const i64* bgather;
  double* partial;  // split-K partial sums [split][N][M]... laid out as split*(M*N) col-major M x N
  int splits, kchunk;
  int MT, NT, NG;  // M tiles, N tiles, groups (= NT * splits)
  int ct;          // store D(row, col) to C(col, row) (the transposed product, splits == 1)
};

__device__ __forceinline__ void split3(double x, __bf16* h, __bf16* m, __bf16* l) {
  const __bf16 hi = (__bf16)(float)x;
  const double r1 = x - (double)(float)hi;
  const __bf16 mi = (__bf16)(float)r1;
  const double r2 = r1 - (double)(float)mi;
  *h = hi; *m = mi; *l = (__bf16)(float)r2;
}

template <int PREC, int BM, int BN, int WAVES_M, int WAVES_N>
__global__ __launch_bounds__(256) void gemm_kernel(GemmDev p) {
  typedef typename PT<PREC>::E E;
  constexpr int P = PT<PREC>::P;
  constexpr int BK = PT<PREC>::BK;
  constexpr int LDK = BK + PT<PREC>::PAD;  // leading dimension (elements) of an LDS row
  constexpr int WTM = BM / WAVES_M, WTN = BN / WAVES_N;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int LA = BM * BK / 256;  // elements of A tile per thread
  constexpr int LB = BN * BK / 256;
  static_assert(WAVES_M * WAVES_N == 4, "4 waves");
  static_assert(LA * 256 == BM * BK && LB * 256 == BN * BK, "tile/threads");

  __shared__ E sA[P][BM * LDK];
  __shared__ E sB[P][BN * LDK];

  // ---- XCD-aware block -> (group, mt) mapping
  const int b = blockIdx.x;
  const int xcd = b & 7, idx = b >> 3;
  const int grp = (idx / p.MT) * 8 + xcd;
  const int mt = idx % p.MT;
  if (grp >= p.NG) return;
  const int nt = grp % p.NT, split = grp / p.NT;
  const int m0 = mt * BM, n0 = nt * BN;
  const int kbeg = split * p.kchunk;
  const int kend = min(p.K, kbeg + p.kchunk);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WAVES_N, wn = wave % WAVES_N;

  // ---- staging map: thread loads LA elements of the A tile and LB of the B tile.
  // k-contiguous source (trans = 1 for A, 0 for B): element e -> (row r, k c) with k fastest.
  // m/n-contiguous source: element e -> row fastest.
  double ra[LA], rb[LB];
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int t = 0; t < LA; ++t) {
      const int e = tid + 256 * t;
      int r, c;
      if (p.transA) { r = e / BK; c = e % BK; } else { r = e % BM; c = e / BM; }
      const int gm = m0 + r, gk = k0 + c;
      double v = 0.0;
      if (gm < p.M && gk < kend)
        v = p.transA ? p.A[gk + (i64)gm * p.lda] : p.A[gm + (i64)gk * p.lda];
      ra[t] = v;
    }
#pragma unroll
    for (int t = 0; t < LB; ++t) {
      const int e = tid + 256 * t;
      int r, c;
      if (!p.transB) { r = e / BK; c = e % BK; } else { r = e % BN; c = e / BN; }
      const int gn = n0 + r, gk = k0 + c;
      double v = 0.0;
      if (gn < p.N && gk < kend) {
        const i64 kk = p.bgather ? (p.bgather[gk] - 1) : gk;
        v = p.transB ? p.B[gn + kk * p.ldb] : p.B[kk + (i64)gn * p.ldb];
      }
      rb[t] = v;
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int t = 0; t < LA; ++t) {
      const int e = tid + 256 * t;
      int r, c;
      if (p.transA) { r = e / BK; c = e % BK; } else { r = e % BM; c = e / BM; }
      const int o = r * LDK + c;
      if constexpr (PREC == LRA_PREC_BF16X3) {
        __bf16 h, mi, l;
        split3(ra[t], &h, &mi, &l);
        sA[0][o] = h; sA[1][o] = mi; sA[2][o] = l;
      } else if constexpr (PREC == LRA_PREC_BF16) {
        sA[0][o] = (__bf16)(float)ra[t];
      } else {
        sA[0][o] = (E)ra[t];
      }
    }
#pragma unroll
    for (int t = 0; t < LB; ++t) {
      const int e = tid + 256 * t;
      int r, c;
      if (!p.transB) { r = e / BK; c = e % BK; } else { r = e % BN; c = e / BN; }
      const int o = r * LDK + c;
      if constexpr (PREC == LRA_PREC_BF16X3) {
        __bf16 h, mi, l;
        split3(rb[t], &h, &mi, &l);
        sB[0][o] = h; sB[1][o] = mi; sB[2][o] = l;
      } else if constexpr (PREC == LRA_PREC_BF16) {
        sB[0][o] = (__bf16)(float)rb[t];
      } else {
        sB[0][o] = (E)rb[t];
      }
    }
  };

  typedef typename std::conditional<PREC == LRA_PREC_F64, d4, f4>::type AccT;
  AccT acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = AccT{0, 0, 0, 0};

  const int l16 = lane & 15, lq = lane >> 4;
  const int am0 = wm * WTM, bn0 = wn * WTN;

  auto compute = [&]() {
    if constexpr (PREC == LRA_PREC_F64 || PREC == LRA_PREC_F32) {
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        E af[TM], bfv[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) af[i] = sA[0][(am0 + i * 16 + l16) * LDK + kk + lq];
#pragma unroll
        for (int j = 0; j < TN; ++j) bfv[j] = sB[0][(bn0 + j * 16 + l16) * LDK + kk + lq];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            if constexpr (PREC == LRA_PREC_F64)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bfv[j], acc[i][j], 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bfv[j], acc[i][j], 0, 0, 0);
          }
      }
    } else {
      // one 16x16x32 step per BK = 32
      bf8 af[P][TM], bfv[P][TN];
#pragma unroll
      for (int q = 0; q < P; ++q) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
          af[q][i] = *reinterpret_cast<const bf8*>(&sA[q][(am0 + i * 16 + l16) * LDK + lq * 8]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          bfv[q][j] = *reinterpret_cast<const bf8*>(&sB[q][(bn0 + j * 16 + l16) * LDK + lq * 8]);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          if constexpr (PREC == LRA_PREC_BF16) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][i], bfv[0][j], acc[i][j], 0, 0, 0);
          } else {
            // smallest terms first: lo*hi, hi*lo, mid*mid, mid*hi, hi*mid, hi*hi
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[2][i], bfv[0][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][i], bfv[2][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1][i], bfv[1][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1][i], bfv[0][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][i], bfv[1][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][i], bfv[0][j], acc[i][j], 0, 0, 0);
          }
        }
    }
  };

  if (kbeg < kend) {
    load_tiles(kbeg);
    store_tiles();
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
      const bool more = k0 + BK < kend;
      if (more) load_tiles(k0 + BK);
      compute();
      __syncthreads();
      if (more) {
        store_tiles();
        __syncthreads();
      }
    }
  }

  // ---- epilogue.  Output lane layout of a 16x16 tile (column l16 in every case):
  //   f32/bf16 MFMA: register r holds row 4*lq + r;  f64 MFMA: register r holds row lq + 4*r.
  // With beta != 0 the C values of one row of tiles are loaded together (unconditional loads
  // from clamped addresses), so the epilogue waits for memory once per tile row.
  const bool rd = p.splits == 1 && p.beta != 0.0;
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    double cv[TN][4];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = n0 + bn0 + j * 16 + l16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + am0 + i * 16 + (PREC == LRA_PREC_F64 ? lq + 4 * r : lq * 4 + r);
        const bool ok = rd && row < p.M && col < p.N;
        const i64 ci = p.ct ? col + (i64)row * p.ldc : row + (i64)col * p.ldc;
        const double x = p.C[ok ? ci : 0];
        cv[j][r] = ok ? x : 0.0;
      }
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = n0 + bn0 + j * 16 + l16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + am0 + i * 16 + (PREC == LRA_PREC_F64 ? lq + 4 * r : lq * 4 + r);
        if (row < p.M && col < p.N) {
          const double v = (double)acc[i][j][r];
          if (p.splits > 1) {
            p.partial[(i64)split * p.M * p.N + row + (i64)col * p.M] = v;
          } else {
            const i64 ci = p.ct ? col + (i64)row * p.ldc : row + (i64)col * p.ldc;
            p.C[ci] = rd ? p.alpha * v + p.beta * cv[j][r] : p.alpha * v;
          }
        }
      }
    }
  }
}

// FP64 GEMM for the case where both operands are k-contiguous in memory (op(A) = A' with A
// stored K x M, op(B) = B stored K x N, no gather): the sketch B = Omega * A.  Every thread
// moves 16-byte pairs (two consecutive k) from HBM into registers for the next k-tile while
// the MFMAs consume the current one from a second LDS buffer, so each k-tile costs one
// barrier.  LDS rows hold BK = 16 doubles padded to 17, which keeps the 16 lanes of an MFMA
// operand read on distinct banks.
template <int BM, int BN, int WAVES_M, int WAVES_N, int BK = 16, bool F32 = false>
__global__ __launch_bounds__(64 * WAVES_M * WAVES_N) void gemm_kc_f64_kernel(GemmDev p) {
  // F32: the same FP64 loads, rounded to FP32 on the way into LDS, v_mfma_f32_16x16x4f32
  // with FP32 accumulation (the F32 sketch precision)
  using E = typename std::conditional<F32, float, double>::type;
  typedef float f4v __attribute__((ext_vector_type(4)));
  using ACC = typename std::conditional<F32, f4v, d4>::type;
  constexpr int NT = 64 * WAVES_M * WAVES_N;
  constexpr int LDK = BK + 1, KP = BK / 2;  // KP: 16-byte k-pairs per tile row
  constexpr int WTM = BM / WAVES_M, WTN = BN / WAVES_N;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int CA = BM * KP, CB = BN * KP;  // 16-byte chunks per tile
  constexpr int LA = (CA + NT - 1) / NT, LB = (CB + NT - 1) / NT;
  __shared__ E sA[2][BM * LDK];
  __shared__ E sB[2][BN * LDK];

  const int b = blockIdx.x;
  const int xcd = b & 7, idx = b >> 3;
  const int grp = (idx / p.MT) * 8 + xcd;
  const int mt = idx % p.MT;
  if (grp >= p.NG) return;
  const int nt = grp % p.NT, split = grp / p.NT;
  const int m0 = mt * BM, n0 = nt * BN;
  const int kbeg = split * p.kchunk;
  const int kend = min(p.K, kbeg + p.kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WAVES_N, wn = wave % WAVES_N;

  typedef double d2 __attribute__((ext_vector_type(2)));
  d2 ra[LA], rb[LB];
  // chunk e of a tile -> (row e / KP, k pair e % KP)
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int t = 0; t < LA; ++t) {
      const int e = tid + NT * t;
      const int r = e / KP, c = (e % KP) * 2;
      const int gm = m0 + r, gk = k0 + c;
      d2 v = {0.0, 0.0};
      if (e < CA && gm < p.M) {
        const double* src = p.A + gk + (i64)gm * p.lda;
        if (gk + 1 < kend) v = *reinterpret_cast<const d2*>(src);
        else if (gk < kend) v.x = src[0];
      }
      ra[t] = v;
    }
#pragma unroll
    for (int t = 0; t < LB; ++t) {
      const int e = tid + NT * t;
      const int r = e / KP, c = (e % KP) * 2;
      const int gn = n0 + r, gk = k0 + c;
      d2 v = {0.0, 0.0};
      if (e < CB && gn < p.N) {
        const double* src = p.B + gk + (i64)gn * p.ldb;
        if (gk + 1 < kend) v = *reinterpret_cast<const d2*>(src);
        else if (gk < kend) v.x = src[0];
      }
      rb[t] = v;
    }
  };
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int t = 0; t < LA; ++t) {
      const int e = tid + NT * t;
      if (e < CA) {
        const int o = (e / KP) * LDK + (e % KP) * 2;
        sA[buf][o] = (E)ra[t].x;
        sA[buf][o + 1] = (E)ra[t].y;
      }
    }
#pragma unroll
    for (int t = 0; t < LB; ++t) {
      const int e = tid + NT * t;
      if (e < CB) {
        const int o = (e / KP) * LDK + (e % KP) * 2;
        sB[buf][o] = (E)rb[t].x;
        sB[buf][o + 1] = (E)rb[t].y;
      }
    }
  };

  ACC acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = ACC{0, 0, 0, 0};
  const int l16 = lane & 15, lq = lane >> 4;
  const int am0 = wm * WTM, bn0 = wn * WTN;

  // (the pointer alignment of the 16-byte loads: lda, ldb, kchunk and k0 are even)
  if (kbeg < kend) {
    load_tiles(kbeg);
    store_tiles(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
      const bool more = k0 + BK < kend;
      if (more) load_tiles(k0 + BK);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        E af[TM], bfv[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) af[i] = sA[cur][(am0 + i * 16 + l16) * LDK + kk + lq];
#pragma unroll
        for (int j = 0; j < TN; ++j) bfv[j] = sB[cur][(bn0 + j * 16 + l16) * LDK + kk + lq];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            if constexpr (F32)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bfv[j], acc[i][j], 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bfv[j], acc[i][j], 0, 0, 0);
          }
      }
      if (more) store_tiles(cur ^ 1);
      __syncthreads();
      cur ^= 1;
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = n0 + bn0 + j * 16 + l16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // f64 MFMA: register r holds row lq + 4 r; f32 MFMA: row 4 lq + r
        const int row = m0 + am0 + i * 16 + (F32 ? 4 * lq + r : lq + 4 * r);
        if (row < p.M && col < p.N) {
          const double v = (double)acc[i][j][r];
          if (p.splits > 1) {
            p.partial[(i64)split * p.M * p.N + row + (i64)col * p.M] = v;
          } else {
            double* c = p.C + row + (i64)col * p.ldc;
            *c = (p.beta == 0.0) ? p.alpha * v : p.alpha * v + p.beta * *c;
          }
        }
      }
    }
}

__global__ void splitk_reduce_kernel(int M, int N, int splits, const double* part, double alpha,
                                     double beta, double* C, i64 ldc) {
  const i64 total = (i64)M * N;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < splits; ++q) s += part[(i64)q * total + e];
    const i64 row = e % M, col = e / M;
    double* c = C + row + col * ldc;
    *c = (beta == 0.0) ? alpha * s : alpha * s + beta * *c;
  }
}

// C (M x N) = alpha A' B + beta C for a short product over a long K (M <= 32): the Gram-type
// products of the blocked QR (V'V and V'X, K = the panel's rows).  A workgroup owns a 32 x 32
// output tile and one K-range; 32-deep k-tiles of both operands are loaded coalesced (32
// consecutive k of one column per 32 lanes) into double-buffered LDS and consumed by
// v_mfma_f64_16x16x4f64, a 16 x 16 quarter per wave.  With several K-ranges the partials go
// to `part` ([split][M x N]) and splitk_reduce_kernel adds them in a fixed order.

__global__ __launch_bounds__(256) void gemm_tn_skinny_kernel(int M, int N, int K, int kchunk,
                                                             int ntiles, const double* __restrict__ A,
                                                             i64 lda, const double* __restrict__ B,
                                                             i64 ldb, double* __restrict__ part,
                                                             double alpha, double beta,
                                                             double* __restrict__ C, i64 ldc) {
  __shared__ double sA[2][32][33];  // [i][k]
  __shared__ double sB[2][32][33];  // [j][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nt = blockIdx.x % ntiles, split = blockIdx.x / ntiles;
  const int n0 = nt * 32;
  const int kbeg = split * kchunk, kend = min(K, kbeg + kchunk);
  const int kk = tid & 31, c0 = tid >> 5;  // loader: k offset, first column (+8 u)
  double ra[4], rb[4];
  auto load = [&](int k0) {
    const int k = k0 + kk;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = c0 + 8 * u;
      // unconditional loads from clamped addresses, masked afterwards
      const bool oa = k < kend && c < M, ob = k < kend && n0 + c < N;
      const double xa = A[oa ? k + (i64)c * lda : 0];
      const double xb = B[ob ? k + (i64)(n0 + c) * ldb : 0];
      ra[u] = oa ? xa : 0.0;
      rb[u] = ob ? xb : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sA[buf][c0 + 8 * u][kk] = ra[u];
      sB[buf][c0 + 8 * u][kk] = rb[u];
    }
  };
  const int wm = wave & 1, wn = wave >> 1;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  int buf = 0;
  if (kbeg < kend) {
    load(kbeg);
    store(0);
  }
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    const bool more = k0 + 32 < kend;
    if (more) load(k0 + 32);
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const double a = sA[buf][wm * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
      const double b = sB[buf][wn * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // lane holds D(row = (lane >> 4) + 4 r, col = lane & 15) of its quarter
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = wm * 16 + (lane >> 4) + 4 * r, col = n0 + wn * 16 + (lane & 15);
    if (row >= M || col >= N) continue;
    if (part) {
      part[(i64)split * M * N + row + (i64)col * M] = acc[r];
    } else {
      double* c = C + row + (i64)col * ldc;
      *c = (beta == 0.0) ? alpha * acc[r] : alpha * acc[r] + beta * *c;
    }
  }
}

// ---- configurations -------------------------------------------------------------------
struct Cfg { int id, BM, BN; };
// id -> (BM, BN, WAVES_M, WAVES_N)
static const Cfg kCfgs[] = {{0, 16, 256}, {1, 32, 256}, {2, 48, 256}, {3, 64, 256}, {4, 128, 128},
                            {5, 16, 256}, {6, 32, 256}, {7, 48, 256}, {8, 64, 256}, {9, 128, 128},
                            {10, 64, 128}, {11, 128, 256}, {12, 128, 128}, {13, 64, 256},
                            {14, 256, 128}, {15, 80, 256}, {16, 96, 256}, {17, 160, 256},
                            {18, 176, 128}, {19, 208, 128}, {20, 112, 128}, {21, 112, 256}};
constexpr int kNumCfg = 22;
// ids 5..21: the k-contiguous kernel (gemm_kc_f64_kernel, FP64 or F32 MFMA; 11, 13, 14, 16-21
// with 512 threads, 12 with 32-wide k-tiles).  15-17 hold a short sketch (ell = 72, 136) in one
// M-tile, so each k-tile of A is read from L2 once; 18 (11 MFMA rows per wave) covers
// ell = 520 in three M-tiles with 1.5% padding; 19 (13 MFMA rows per wave) covers two paired
// rounds' 400 rows (orders 136 + 264) in two M-tiles, 20 and 21 the 112 rows of orders 40 + 72
// in one.
bool kc_cfg(int id) { return id >= 5; }
bool kc_ok(const GemmArgs& g) {
  return (g.prec == LRA_PREC_F64 || g.prec == LRA_PREC_F32) && g.transA == 1 && g.transB == 0 &&
         g.bgather == nullptr &&
         (g.lda % 2) == 0 && (g.ldb % 2) == 0 && ((uintptr_t)g.A % 16) == 0 &&
         ((uintptr_t)g.B % 16) == 0;
}

template <int PREC>
hipError_t launch_cfg(int id, const GemmDev& d, int blocks, hipStream_t st) {
  switch (id) {
    case 0: hipLaunchKernelGGL((gemm_kernel<PREC, 16, 256, 1, 4>), dim3(blocks), dim3(256), 0, st, d); break;
    case 1: hipLaunchKernelGGL((gemm_kernel<PREC, 32, 256, 1, 4>), dim3(blocks), dim3(256), 0, st, d); break;
    case 2: hipLaunchKernelGGL((gemm_kernel<PREC, 48, 256, 1, 4>), dim3(blocks), dim3(256), 0, st, d); break;
    case 3: hipLaunchKernelGGL((gemm_kernel<PREC, 64, 256, 1, 4>), dim3(blocks), dim3(256), 0, st, d); break;
    case 4: hipLaunchKernelGGL((gemm_kernel<PREC, 128, 128, 2, 2>), dim3(blocks), dim3(256), 0, st, d); break;
    case 5:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<16, 256, 1, 4, 16, false>), dim3(blocks), dim3(256), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<16, 256, 1, 4, 16, true>), dim3(blocks), dim3(256), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 6:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<32, 256, 1, 4, 16, false>), dim3(blocks), dim3(256), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<32, 256, 1, 4, 16, true>), dim3(blocks), dim3(256), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 7:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<48, 256, 1, 4, 16, false>), dim3(blocks), dim3(256), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<48, 256, 1, 4, 16, true>), dim3(blocks), dim3(256), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 8:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<64, 256, 1, 4, 16, false>), dim3(blocks), dim3(256), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<64, 256, 1, 4, 16, true>), dim3(blocks), dim3(256), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 9:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<128, 128, 2, 2, 16, false>), dim3(blocks), dim3(256), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<128, 128, 2, 2, 16, true>), dim3(blocks), dim3(256), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 10:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<64, 128, 2, 2, 16, false>), dim3(blocks), dim3(256), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<64, 128, 2, 2, 16, true>), dim3(blocks), dim3(256), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 11:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<128, 256, 2, 4, 16, false>), dim3(blocks), dim3(512), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<128, 256, 2, 4, 16, true>), dim3(blocks), dim3(512), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 12:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<128, 128, 2, 2, 32, false>), dim3(blocks), dim3(256), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<128, 128, 2, 2, 32, true>), dim3(blocks), dim3(256), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 13:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<64, 256, 2, 4, 16, false>), dim3(blocks), dim3(512), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<64, 256, 2, 4, 16, true>), dim3(blocks), dim3(512), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 14:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<256, 128, 4, 2, 16, false>), dim3(blocks), dim3(512), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<256, 128, 4, 2, 16, true>), dim3(blocks), dim3(512), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 15:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<80, 256, 1, 4, 16, false>), dim3(blocks), dim3(256), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<80, 256, 1, 4, 16, true>), dim3(blocks), dim3(256), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 16:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<96, 256, 2, 4, 16, false>), dim3(blocks), dim3(512), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<96, 256, 2, 4, 16, true>), dim3(blocks), dim3(512), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 17:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<160, 256, 2, 4, 16, false>), dim3(blocks), dim3(512), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<160, 256, 2, 4, 16, true>), dim3(blocks), dim3(512), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 18:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<176, 128, 1, 8, 16, false>), dim3(blocks), dim3(512), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<176, 128, 1, 8, 16, true>), dim3(blocks), dim3(512), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 19:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<208, 128, 1, 8, 16, false>), dim3(blocks), dim3(512), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<208, 128, 1, 8, 16, true>), dim3(blocks), dim3(512), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 20:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<112, 128, 1, 8, 16, false>), dim3(blocks), dim3(512), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<112, 128, 1, 8, 16, true>), dim3(blocks), dim3(512), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    case 21:
      if constexpr (PREC == LRA_PREC_F64) hipLaunchKernelGGL((gemm_kc_f64_kernel<112, 256, 1, 8, 16, false>), dim3(blocks), dim3(512), 0, st, d);
      else if constexpr (PREC == LRA_PREC_F32) hipLaunchKernelGGL((gemm_kc_f64_kernel<112, 256, 1, 8, 16, true>), dim3(blocks), dim3(512), 0, st, d);
      else return hipErrorInvalidValue;
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_prec(int prec, int id, const GemmDev& d, int blocks, hipStream_t st) {
  if (kc_cfg(id) && prec != LRA_PREC_F64 && prec != LRA_PREC_F32) return hipErrorInvalidValue;
  switch (prec) {
    case LRA_PREC_F64: return launch_cfg<LRA_PREC_F64>(id, d, blocks, st);
    case LRA_PREC_F32: return launch_cfg<LRA_PREC_F32>(id, d, blocks, st);
    case LRA_PREC_BF16: return launch_cfg<LRA_PREC_BF16>(id, d, blocks, st);
    case LRA_PREC_BF16X3: return launch_cfg<LRA_PREC_BF16X3>(id, d, blocks, st);
  }
  return hipErrorInvalidValue;
}

int bk_of(int prec) { return prec == LRA_PREC_F64 ? 16 : 32; }

// Autotuner: the first call for a shape class times every candidate (config, splits) on the
// live operands and caches the fastest; later calls reuse the choice.  The cache is
// process-wide and filled once per class: a call that finds the class being tuned by another
// thread (another context, e.g. a serving worker) waits for that choice instead of timing on
// its own, so every product of a class in the process uses one configuration and gives the
// same bits in any order and under any overlap.
struct TuneKey {
  int prec, transA, transB, mb, nb, kb;
  bool operator<(const TuneKey& o) const {
    return std::tie(prec, transA, transB, mb, nb, kb) <
           std::tie(o.prec, o.transA, o.transB, o.mb, o.nb, o.kb);
  }
};
struct TuneVal { int cfg, splits; };
std::mutex g_tune_mu;
std::condition_variable g_tune_cv;
std::map<TuneKey, TuneVal> g_tune;
std::set<TuneKey> g_tuning;  // classes being tuned right now
int g_autotune = -1;

int bucket(i64 x) {  // power-of-two bucket
  int b = 0;
  while ((1ll << b) < x) ++b;
  return b;
}
// M's class: its 16-row MFMA tile count up to 1024 rows (a sketch order; the best M tile
// depends on how the order divides into tiles, e.g. 400 rows fill two 208-row tiles but
// 264 rows fit a 160 x 256 tile better), the power-of-two bucket beyond
int bucket_m(i64 m) { return m <= 1024 ? (int)((m + 15) / 16) : 64 + bucket(m); }

int prepare(lra_ctx* ctx, const GemmArgs& g, int cfg, int splits, GemmDev* d, int* blocks) {
  const Cfg& c = kCfgs[cfg];
  d->M = (int)g.M; d->N = (int)g.N; d->K = (int)g.K;
  d->alpha = g.alpha; d->beta = g.beta;
  d->A = g.A; d->lda = g.lda; d->transA = g.transA;
  d->B = g.B; d->ldb = g.ldb; d->transB = g.transB;
  d->C = g.C; d->ldc = g.ldc; d->bgather = g.bgather;
  const int bk = cfg == 12 ? 32 : bk_of(g.prec);
  i64 kchunk = 0;
  if (g.K <= 0) {
    splits = 1;
  } else {
    kchunk = (g.K + splits - 1) / splits;
    kchunk = (kchunk + bk - 1) / bk * bk;
    splits = (int)((g.K + kchunk - 1) / kchunk);
  }
  d->splits = splits; d->kchunk = (int)kchunk;
  d->MT = (int)((g.M + c.BM - 1) / c.BM);
  d->NT = (int)((g.N + c.BN - 1) / c.BN);
  d->NG = d->NT * splits;
  d->partial = nullptr;
  d->ct = 0;
  if (splits > 1) {
    double* part = nullptr;
    LRA_CHECK(lra_wsT(ctx, WS_SPLITK, (size_t)splits * g.M * g.N, &part));
    d->partial = part;
  }
  const i64 groups8 = ((i64)d->NG + 7) / 8;
  const i64 nblk = groups8 * 8 * d->MT;
  if (nblk > 2147483647ll) { lra_set_error("gemm: grid too large"); return LRA_E_UNSUPPORTED; }
  *blocks = (int)nblk;
  return LRA_OK;
}

int run(lra_ctx* ctx, const GemmArgs& g, int cfg, int splits) {
  GemmDev d;
  int blocks = 0;
  LRA_CHECK(prepare(ctx, g, cfg, splits, &d, &blocks));
  hipError_t e = launch_prec(g.prec, cfg, d, blocks, ctx->stream);
  if (e != hipSuccess) {
    lra_set_error("gemm launch: %s", hipGetErrorString(e));
    return LRA_E_HIP;
  }
  if (d.splits > 1) {
    const i64 total = g.M * g.N;
    int rb = (int)std::min<i64>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(rb), dim3(256), 0, ctx->stream, (int)g.M,
                       (int)g.N, d.splits, d.partial, g.alpha, g.beta, g.C, g.ldc);
    e = hipGetLastError();
    if (e != hipSuccess) {
      lra_set_error("splitk reduce launch: %s", hipGetErrorString(e));
      return LRA_E_HIP;
    }
  }
  return LRA_OK;
}

// A' B with M <= 32 and a long K on gemm_tn_skinny_kernel: 32 x 32 tiles, K split so that
// about one workgroup per CU runs (ranges of at least 256, multiples of 32).
bool tn_skinny_ok(const GemmArgs& g) {
  return g.prec == LRA_PREC_F64 && g.transA == 1 && g.transB == 0 && g.bgather == nullptr &&
         g.M <= 32 && g.K >= 256;
}

int run_tn_skinny(lra_ctx* ctx, const GemmArgs& g) {
  const i64 ntiles = (g.N + 31) / 32;
  i64 splits = std::max<i64>(1, (ctx->num_cu + ntiles - 1) / ntiles);
  splits = std::min<i64>(splits, std::max<i64>(1, g.K / 256));
  i64 kchunk = (g.K + splits - 1) / splits;
  kchunk = (kchunk + 31) / 32 * 32;
  splits = (g.K + kchunk - 1) / kchunk;
  double* part = nullptr;
  if (splits > 1) LRA_CHECK(lra_wsT(ctx, WS_SPLITK, (size_t)(splits * g.M * g.N), &part));
  hipLaunchKernelGGL(gemm_tn_skinny_kernel, dim3((unsigned)(ntiles * splits)), dim3(256), 0,
                     ctx->stream, (int)g.M, (int)g.N, (int)g.K, (int)kchunk, (int)ntiles, g.A,
                     g.lda, g.B, g.ldb, part, g.alpha, g.beta, g.C, g.ldc);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && splits > 1) {
    const i64 total = g.M * g.N;
    const int rb = (int)std::min<i64>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(rb), dim3(256), 0, ctx->stream, (int)g.M,
                       (int)g.N, (int)splits, part, g.alpha, g.beta, g.C, g.ldc);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    lra_set_error("skinny gemm launch: %s", hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
}

void candidates(const GemmArgs& g, int num_cu, std::vector<TuneVal>* out) {
  const bool kc = kc_ok(g);
  for (int cfg = 0; cfg < kNumCfg; ++cfg) {
    const Cfg& c = kCfgs[cfg];
    if (kc_cfg(cfg) && !kc) continue;
    // skip configs that waste most of the M tile
    if (c.BM > 16 && g.M <= c.BM / 2) continue;
    if ((cfg == 4 || cfg == 9 || cfg == 11 || cfg == 12) && g.M < 96) continue;
    if (cfg == 14 && g.M < 192) continue;
    const i64 tiles = ((g.M + c.BM - 1) / c.BM) * ((g.N + c.BN - 1) / c.BN);
    for (int s : {1, 2, 3, 4, 6, 8, 16, 32}) {
      if (s > 1 && tiles * s > 8ll * num_cu) break;
      if (s > 1 && g.K / s < 256) break;
      out->push_back({cfg, s});
    }
  }
  if (out->empty()) out->push_back({4, 1});
}

TuneVal heuristic(const GemmArgs& g, int num_cu) {
  int cfg = 4;
  if (g.M <= 16) cfg = 0;
  else if (g.M <= 32) cfg = 1;
  else if (g.M <= 48) cfg = 2;
  else if (g.M <= 64) cfg = 3;
  const Cfg& c = kCfgs[cfg];
  const i64 tiles = ((g.M + c.BM - 1) / c.BM) * ((g.N + c.BN - 1) / c.BN);
  int s = 1;
  while (tiles * s < 2ll * num_cu && g.K / (2 * s) >= 512 && s < 32) s *= 2;
  return {cfg, s};
}

// A short-K product with a tall C (the blocked QR's X -= V W and its like) as the transposed
// product C' = op(B)' op(A)' stored transposed: the MFMA output tile then puts 16 consecutive
// rows of one column of C in 16 consecutive lanes, so the read-modify-write of C is
// coalesced.  One K-range, the general kernel's configuration for the swapped shape.
bool ct_ok(const GemmArgs& g) {
  return g.bgather == nullptr && g.K <= 64 && g.M >= 256 && g.M >= 2 * g.N;
}

int run_ct(lra_ctx* ctx, const GemmArgs& g) {
  GemmArgs s = g;
  s.M = g.N; s.N = g.M;
  s.A = g.B; s.lda = g.ldb; s.transA = !g.transB;
  s.B = g.A; s.ldb = g.lda; s.transB = !g.transA;
  const int cfg = heuristic(s, ctx->num_cu).cfg;
  GemmDev d;
  int blocks = 0;
  LRA_CHECK(prepare(ctx, s, cfg, 1, &d, &blocks));
  d.C = g.C; d.ldc = g.ldc;
  d.ct = 1;
  const hipError_t e = launch_prec(g.prec, cfg, d, blocks, ctx->stream);
  if (e != hipSuccess) {
    lra_set_error("gemm launch: %s", hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
}

}  // namespace

int lra_gemm(lra_ctx* ctx, const GemmArgs& g) {
  if (g.M <= 0 || g.N <= 0) return LRA_OK;
  if (g.M > 2147483647ll || g.N > 2147483647ll || g.K > 2147483647ll) {
    lra_set_error("gemm: dimension exceeds int32");
    return LRA_E_UNSUPPORTED;
  }
  if (g_autotune < 0) {
    const char* env = getenv("LRA_AUTOTUNE");
    g_autotune = (env && env[0] == '0') ? 0 : 1;
  }
  // LRA_GEMM_CFG=id:splits forces a configuration (when it applies to the operands)
  if (const char* f = getenv("LRA_GEMM_CFG")) {
    int id = -1, sp = 1;
    if (sscanf(f, "%d:%d", &id, &sp) == 2 && id >= 0 && id < kNumCfg && sp >= 1 &&
        (!kc_cfg(id) || kc_ok(g)))
      return run(ctx, g, id, sp);
  }
  if (tn_skinny_ok(g)) return run_tn_skinny(ctx, g);
  if (ct_ok(g)) return run_ct(ctx, g);
  TuneKey key{g.prec, g.transA, g.transB, bucket_m(g.M), bucket(g.N), bucket(g.K)};
  TuneVal choice;
  bool have = false;
  {
    std::unique_lock<std::mutex> lk(g_tune_mu);
    for (;;) {
      auto it = g_tune.find(key);
      if (it != g_tune.end()) { choice = it->second; have = true; break; }
      if (!g_tuning.count(key)) { g_tuning.insert(key); break; }  // this call tunes it
      g_tune_cv.wait(lk);
    }
  }
  if (!have) {
    // (on an error return the class is released so that a later call tunes it)
    struct Release {
      const TuneKey& k;
      bool armed = true;
      ~Release() {
        if (!armed) return;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuning.erase(k);
        g_tune_cv.notify_all();
      }
    } release{key};
    if (!g_autotune || g.beta != 0.0 || (g.M * g.N * (g.K > 0 ? g.K : 1)) < (1ll << 24)) {
      choice = heuristic(g, ctx->num_cu);
    } else {
      std::vector<TuneVal> cands;
      candidates(g, ctx->num_cu, &cands);
      hipEvent_t e0, e1;
      LRA_HIP(hipEventCreate(&e0));
      LRA_HIP(hipEventCreate(&e1));
      auto time_of = [&](const TuneVal& c, int reps, float* ms) -> int {
        LRA_HIP(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps; ++r) LRA_CHECK(run(ctx, g, c.cfg, c.splits));
        LRA_HIP(hipEventRecord(e1, ctx->stream));
        LRA_HIP(hipEventSynchronize(e1));
        LRA_HIP(hipEventElapsedTime(ms, e0, e1));
        *ms /= (float)reps;
        return LRA_OK;
      };
      // pass 1: every candidate once warm + 2 timed; pass 2: the four fastest re-timed twice
      // more, 3 launches each, interleaved, so a single noisy timing does not pick the tile
      std::vector<std::pair<float, int>> t1;
      for (size_t i = 0; i < cands.size(); ++i) {
        LRA_CHECK(run(ctx, g, cands[i].cfg, cands[i].splits));  // warm
        float ms = 0.f;
        LRA_CHECK(time_of(cands[i], 2, &ms));
        t1.push_back({ms, (int)i});
      }
      std::sort(t1.begin(), t1.end());
      const size_t top = std::min<size_t>(4, t1.size());
      std::vector<float> best_of(top);
      for (size_t i = 0; i < top; ++i) best_of[i] = t1[i].first;
      for (int pass = 0; pass < 2; ++pass)
        for (size_t i = 0; i < top; ++i) {
          float ms = 0.f;
          LRA_CHECK(time_of(cands[t1[i].second], 3, &ms));
          best_of[i] = std::min(best_of[i], ms);
        }
      size_t win = 0;
      for (size_t i = 1; i < top; ++i)
        if (best_of[i] < best_of[win]) win = i;
      choice = cands[t1[win].second];
      hipEventDestroy(e0);
      hipEventDestroy(e1);
    }
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune[key] = choice;
      g_tuning.erase(key);
      release.armed = false;
    }
    g_tune_cv.notify_all();
  }
  return run(ctx, g, choice.cfg, choice.splits);
}

// Report the configuration chosen for a shape (for the bench's roofline bookkeeping).
extern "C" int lra_gemm_choice(int prec, int transA, int transB, int64_t M, int64_t N,
                               int64_t K, int* cfg, int* splits, int* bm, int* bn) {
  TuneKey key{prec, transA, transB, bucket_m(M), bucket(N), bucket(K)};
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(key);
  if (it == g_tune.end()) return LRA_E_STATE;
  *cfg = it->second.cfg;
  *splits = it->second.splits;
  *bm = kCfgs[it->second.cfg].BM;
  *bn = kCfgs[it->second.cfg].BN;
  return LRA_OK;
}
