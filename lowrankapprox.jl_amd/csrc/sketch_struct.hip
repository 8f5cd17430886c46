// sketch_struct.hip -- SRFT, random-subset and sparse-Gaussian sketches (left side).
//
// SRFT (src/sketch.jl:334-562, real case): with l = largest divisor of m <= ell and mm = m/l,
// output slots i = 0, 2, 4, ... of a column hold Re and Im of DFT_m(d .* x) at frequency
// f = idx[i]-1 (the last slot of an odd ell holds Re only; DC/Nyquist still emit their zero
// imaginary part, see the always-false `in == 0` test at :425; no 1/sqrt(m) scaling).  The
// reference evaluates it as an R2HC FFT of length l on each of the mm columns of the l x mm
// reshape X[p,q] = (d.*x)[p*mm+q] followed by sum_q w^q Xhat[f mod l, q] with w = e^{-2 pi i f/m}.
// Here: one workgroup per column of A streams X through LDS in slabs of 2*QC columns packed
// as QC complex sequences, transforms them with an in-place radix-2 FFT (power-of-two l) or a
// direct DFT (other l), unpacks the two real spectra and accumulates every output frequency in
// registers.  A is read exactly once (HBM-bound).
//
// Random subset (src/sketch.jl:248-257): B[i,:] = op(A)[r[i],:].
// Sparse Gaussian (src/sketch.jl:571-589): output row i sums p_i = fld(m-i,ell)+1 rows of op(A)
// taken from consecutive positions of the permutation r with coefficients s, in the
// reference's order (one thread per output row, deterministic).
#include "lra_internal.h"

#include <algorithm>

namespace {

struct cplx {
  double re, im;
};
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

constexpr int SRFT_THREADS = 256;

// grid: one block per column of A (trans n).  LDS: Z (l x qc complex) [+ W for the DFT path]
// + twiddles (l/2 complex).
__global__ __launch_bounds__(SRFT_THREADS) void srft_kernel(
    i64 m, const double* __restrict__ A, i64 lda, int l, i64 mm, int qc, int pow2, int logl,
    const double* __restrict__ d, const i64* __restrict__ idx, int ell, double* __restrict__ B,
    i64 ldb) {
  extern __shared__ double lds[];
  cplx* Z = reinterpret_cast<cplx*>(lds);               // [p][c] row-major, l x qc
  cplx* W = Z + (size_t)l * qc;                         // DFT path output (l x qc)
  cplx* tw = pow2 ? W : W + (size_t)l * qc;             // l/2 twiddles e^{-2 pi i j / l}
  const int tid = threadIdx.x;
  const i64 col = blockIdx.x;
  const double* x = A + col * lda;
  for (int j = tid; j < l / 2 + 1; j += SRFT_THREADS) {
    double s, c;
    sincospi(-2.0 * (double)j / (double)l, &s, &c);
    tw[j] = {c, s};
  }
  const int nslots = (ell + 1) / 2;
  // per-thread output slots: s = tid, tid + 256, ... (<= 4 slots per thread for ell <= 2048)
  constexpr int MAXS = 4;
  cplx acc[MAXS], wstep[MAXS];
  int fcol[MAXS], fconj[MAXS];
  i64 ff[MAXS];
#pragma unroll
  for (int u = 0; u < MAXS; ++u) {
    const int s = tid + u * SRFT_THREADS;
    acc[u] = {0.0, 0.0};
    if (s < nslots) {
      const i64 f = idx[2 * s] - 1;
      ff[u] = f;
      fcol[u] = (int)(f % l);
      fconj[u] = (int)((l - f % l) % l);
      double sn, cs;
      sincospi(-2.0 * (double)f / (double)m, &sn, &cs);
      wstep[u] = {cs, sn};
    }
  }
  const i64 nq = mm;  // real columns of X
  for (i64 q0 = 0; q0 < nq; q0 += 2 * qc) {
    __syncthreads();
    // load slab: Z[p][c] = X[p, q0+2c] + i X[p, q0+2c+1], with signs; bit-reversed p for FFT
    for (i64 e = tid; e < (i64)l * qc; e += SRFT_THREADS) {
      const int c = (int)(e % qc), p = (int)(e / qc);
      const i64 q = q0 + 2 * c;
      const i64 t0 = (i64)p * mm + q;
      const double a0 = q < nq ? d[t0] * x[t0] : 0.0;
      const double a1 = q + 1 < nq ? d[t0 + 1] * x[t0 + 1] : 0.0;
      int pp = p;
      if (pow2 && logl > 0) pp = (int)(__brev((unsigned)p) >> (32 - logl));
      Z[(size_t)pp * qc + c] = {a0, a1};
    }
    __syncthreads();
    cplx* S = Z;
    if (pow2) {
      for (int h = 1; h < l; h <<= 1) {
        const int tstep = l / (2 * h);
        for (i64 e = tid; e < (i64)(l / 2) * qc; e += SRFT_THREADS) {
          const int c = (int)(e % qc), bidx = (int)(e / qc);
          const int grp = bidx / h, k = bidx % h;
          const int i0 = grp * 2 * h + k, i1 = i0 + h;
          const cplx w = tw[k * tstep];
          const cplx a = Z[(size_t)i0 * qc + c];
          const cplx b = cmul(w, Z[(size_t)i1 * qc + c]);
          Z[(size_t)i0 * qc + c] = {a.re + b.re, a.im + b.im};
          Z[(size_t)i1 * qc + c] = {a.re - b.re, a.im - b.im};
        }
        __syncthreads();
      }
    } else {
      // direct DFT of length l: W[f][c] = sum_p Z[p][c] e^{-2 pi i f p / l}
      for (i64 e = tid; e < (i64)l * qc; e += SRFT_THREADS) {
        const int c = (int)(e % qc), f = (int)(e / qc);
        double sr = 0.0, si = 0.0;
        for (int p = 0; p < l; ++p) {
          const int j = (int)(((i64)f * p) % l);
          double sn, cs;
          if (j <= l / 2) { cs = tw[j].re; sn = tw[j].im; }
          else { cs = tw[l - j].re; sn = -tw[l - j].im; }
          const cplx z = Z[(size_t)p * qc + c];
          sr += z.re * cs - z.im * sn;
          si += z.re * sn + z.im * cs;
        }
        W[(size_t)f * qc + c] = {sr, si};
      }
      __syncthreads();
      S = W;
    }
    // unpack the two real spectra and accumulate every output frequency
#pragma unroll
    for (int u = 0; u < MAXS; ++u) {
      const int s = tid + u * SRFT_THREADS;
      if (s >= nslots) continue;
      // w^{q0} with the exponent reduced mod m (exact), then stepped by w
      const i64 e0 = (ff[u] * (q0 % m)) % m;
      double sn, cs;
      sincospi(-2.0 * (double)e0 / (double)m, &sn, &cs);
      cplx wq = {cs, sn};
      cplx a = acc[u];
      for (int c = 0; c < qc; ++c) {
        const i64 q = q0 + 2 * c;
        if (q >= nq) break;
        const cplx zc = S[(size_t)fcol[u] * qc + c];
        const cplx zn = S[(size_t)fconj[u] * qc + c];
        // X1 = (zc + conj(zn)) / 2,  X2 = (zc - conj(zn)) / (2i)
        const cplx x1 = {0.5 * (zc.re + zn.re), 0.5 * (zc.im - zn.im)};
        const cplx x2 = {0.5 * (zc.im + zn.im), -0.5 * (zc.re - zn.re)};
        cplx t = cmul(wq, x1);
        a.re += t.re; a.im += t.im;
        wq = cmul(wq, wstep[u]);
        if (q + 1 < nq) {
          t = cmul(wq, x2);
          a.re += t.re; a.im += t.im;
        }
        wq = cmul(wq, wstep[u]);
      }
      acc[u] = a;
    }
  }
#pragma unroll
  for (int u = 0; u < MAXS; ++u) {
    const int s = tid + u * SRFT_THREADS;
    if (s >= nslots) continue;
    double* y = B + col * ldb;
    y[2 * s] = acc[u].re;
    if (2 * s + 1 < ell) y[2 * s + 1] = acc[u].im;
  }
}

// ---- SRFT for power-of-two l <= 256: two register passes of length <= 16 ----------------
// e^{-2 pi i k / 16}, k = 0..15 (cos, sin)
__device__ __constant__ double kW16c[16] = {
    1.0, 0.92387953251128674, 0.70710678118654757, 0.38268343236508978, 0.0,
    -0.38268343236508978, -0.70710678118654757, -0.92387953251128674, -1.0,
    -0.92387953251128674, -0.70710678118654757, -0.38268343236508978, 0.0,
    0.38268343236508978, 0.70710678118654757, 0.92387953251128674};
__device__ __constant__ double kW16s[16] = {
    0.0, -0.38268343236508978, -0.70710678118654757, -0.92387953251128674, -1.0,
    -0.92387953251128674, -0.70710678118654757, -0.38268343236508978, 0.0,
    0.38268343236508978, 0.70710678118654757, 0.92387953251128674, 1.0,
    0.92387953251128674, 0.70710678118654757, 0.38268343236508978};

// in-place DFT of length N <= 16 (radix-2, natural order in and out): x[k] = sum_p x[p] w_N^kp
template <int N>
__device__ __forceinline__ void dft_reg(cplx* x) {
  if constexpr (N > 1) {
    cplx e[N / 2], o[N / 2];
#pragma unroll
    for (int i = 0; i < N / 2; ++i) { e[i] = x[2 * i]; o[i] = x[2 * i + 1]; }
    dft_reg<N / 2>(e);
    dft_reg<N / 2>(o);
#pragma unroll
    for (int k = 0; k < N / 2; ++k) {
      const int j = k * (16 / N);
      const cplx w = {kW16c[j], kW16s[j]};
      const cplx t = (k == 0) ? o[k] : cmul(w, o[k]);
      x[k] = {e[k].re + t.re, e[k].im + t.im};
      x[k + N / 2] = {e[k].re - t.re, e[k].im - t.im};
    }
  }
}

// One workgroup per column of A.  The column is the l x mm matrix X[p, q] = (d.*x)[p mm + q];
// slabs of 2 QC columns q are packed as QC complex sequences.  With l = L1 L2 and p = p1 L2 +
// p2, k = k1 + L1 k2: pass 1 (a thread per (sequence, p2)) loads its L1 inputs straight from
// HBM, transforms them in registers and applies w_l^(k1 p2); pass 2 (a thread per (sequence,
// k1)) transforms the L2 values from LDS; every output slot then accumulates its frequency as
// in srft_kernel.  Two workgroups share a CU (at most 255 VGPRs per lane, 69.9 KB of LDS at
// l = 256: pass 2 transforms each thread's Y values in place, so frequency k1 + L1 k2 of a
// sequence sits where pass 1 put row p2 = k2 of block k1), so one workgroup transforms while
// the other waits for its slab's loads.  MAXS output slots per thread (256 MAXS >= (ell+1)/2).
// Workgroup barrier for LDS traffic only (waits for this wave's LDS operations, not for its
// global loads).  The kernel below exchanges data between threads through LDS only.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// SIGNS: d is +-1 (drawn by the library, as srft_rand draws it for real types) and given as a
// bit mask (bit i set: d[i] = -1), staged in LDS once per workgroup; d x is then x with its
// sign flipped where the bit is set, exactly the product, without loading d per slab.
template <int L1, int L2, int MAXS, bool SIGNS>
__global__ __launch_bounds__(256, MAXS == 1 ? 2 : 1) void srft_fft_kernel(
    i64 m, const double* __restrict__ A, i64 lda, i64 mm, const double* __restrict__ d,
    const uint32_t* __restrict__ dbits, const i64* __restrict__ idx, int ell,
    double* __restrict__ B, i64 ldb) {
  constexpr int L = L1 * L2;
  constexpr int QC = 256 / (L1 > L2 ? L1 : L2);  // sequences per slab
  // LDS padding: each k1 block of Y and each row of Z is one complex longer than its data, so
  // the lanes of a wave that read different k1 blocks (pass 2) or different frequency rows
  // (the accumulation) start on different banks instead of all on one
  constexpr int YB = L2 * QC + 1;
  extern __shared__ double lds[];
  cplx* Y = reinterpret_cast<cplx*>(lds);  // [k1][p2 or k2][c], k1 blocks of YB
  cplx* tw = Y + L1 * YB;                 // w_l^j, j < l
  uint32_t* sb = reinterpret_cast<uint32_t*>(tw + L);  // SIGNS: the m-bit sign mask
  const int tid = threadIdx.x;
  const i64 col = blockIdx.x;
  const double* x = A + col * lda;
  for (int j = tid; j < L; j += 256) {
    double sn, cs;
    sincospi(-2.0 * (double)j / (double)L, &sn, &cs);
    tw[j] = {cs, sn};
  }
  if constexpr (SIGNS) {
    for (int t = tid; t < (int)((m + 31) / 32); t += 256) sb[t] = dbits[t];
  }
  const int nslots = (ell + 1) / 2;
  cplx acc[MAXS], wstep[MAXS], wstep2[MAXS];
  int fcol[MAXS], fconj[MAXS];
  i64 e0[MAXS], estep[MAXS];  // e0 = f q0 mod m, advanced by f 2 QC mod m per slab
#pragma unroll
  for (int u = 0; u < MAXS; ++u) {
    const int s = tid + u * 256;
    acc[u] = {0.0, 0.0};
    fcol[u] = fconj[u] = 0;
    e0[u] = estep[u] = 0;
    wstep[u] = wstep2[u] = {1.0, 0.0};
    if (s < nslots) {
      const i64 f = idx[2 * s] - 1;
      estep[u] = (f * (2 * QC)) % m;
      // Y offsets of frequency f mod l and of its conjugate partner (-f) mod l
      const int kc = (int)(f % L), kn = (L - kc) % L;
      fcol[u] = (kc % L1) * YB + (kc / L1) * QC;
      fconj[u] = (kn % L1) * YB + (kn / L1) * QC;
      double sn, cs;
      sincospi(-2.0 * (double)f / (double)m, &sn, &cs);
      wstep[u] = {cs, sn};
      sincospi(-4.0 * (double)f / (double)m, &sn, &cs);
      wstep2[u] = {cs, sn};
    }
  }
  // pass-1 role: sequence c1, column p2 of the reshape
  const bool p1act = tid < QC * L2;
  const int c1 = tid % QC, p2 = tid / QC;
  // pass-2 role: sequence c2, frequency k1
  const bool p2act = tid < QC * L1;
  const int c2 = tid % QC, k1 = tid / QC;
  const i64 nq = mm;
  // 16-byte loads of column pairs when every pair is 16-byte aligned in both x and d
  const bool vec = (mm & 1) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0 &&
                   (SIGNS || (reinterpret_cast<uintptr_t>(d) & 15) == 0);
  lds_barrier();  // the twiddles are written by other waves than the ones reading them
  for (i64 q0 = 0; q0 < nq; q0 += 2 * QC) {
    // pass 1.  Every load is unconditional, from a clamped address, with the mask applied
    // afterwards: loads under branches would each wait for the previous one.
    if (p1act) {
      const i64 q = q0 + 2 * c1;
      const bool ma = q < nq, mb = q + 1 < nq;
      cplx v[L1];
      if (vec) {
        // columns q, q + 1 as one 16-byte load (q even, mm even; q clamped to nq - 2)
        const i64 qv = min(q, nq - 2);
        d2v rx[L1];
#pragma unroll
        for (int a = 0; a < L1; ++a)
          rx[a] = *reinterpret_cast<const d2v*>(x + (i64)(a * L2 + p2) * mm + qv);
        if constexpr (SIGNS) {
          // rows pos, pos + 1 (pos even) share a mask word
#pragma unroll
          for (int a = 0; a < L1; ++a) {
            const int pos = (int)((i64)(a * L2 + p2) * mm + qv);
            const uint32_t w = sb[pos >> 5] >> (pos & 31);
            const double x0 = (w & 1u) ? -rx[a].x : rx[a].x;
            const double x1 = (w & 2u) ? -rx[a].y : rx[a].y;
            v[a] = {ma ? x0 : 0.0, mb ? x1 : 0.0};
          }
        } else {
#pragma unroll
          for (int a = 0; a < L1; ++a) {
            const d2v dv = *reinterpret_cast<const d2v*>(d + (i64)(a * L2 + p2) * mm + qv);
            v[a] = {ma ? dv.x * rx[a].x : 0.0, mb ? dv.y * rx[a].y : 0.0};
          }
        }
      } else {
        const i64 qa = min(q, nq - 1), qb = min(q + 1, nq - 1);
        double rx[L1][2];
#pragma unroll
        for (int a = 0; a < L1; ++a) {
          const i64 base = (i64)(a * L2 + p2) * mm;
          rx[a][0] = x[base + qa];
          rx[a][1] = x[base + qb];
        }
#pragma unroll
        for (int a = 0; a < L1; ++a) {
          const i64 base = (i64)(a * L2 + p2) * mm;
          if constexpr (SIGNS) {
            const int pa = (int)(base + qa), pb = (int)(base + qb);
            const double x0 = ((sb[pa >> 5] >> (pa & 31)) & 1u) ? -rx[a][0] : rx[a][0];
            const double x1 = ((sb[pb >> 5] >> (pb & 31)) & 1u) ? -rx[a][1] : rx[a][1];
            v[a] = {ma ? x0 : 0.0, mb ? x1 : 0.0};
          } else {
            v[a] = {ma ? d[base + qa] * rx[a][0] : 0.0, mb ? d[base + qb] * rx[a][1] : 0.0};
          }
        }
      }
      dft_reg<L1>(v);
#pragma unroll
      for (int a = 0; a < L1; ++a) {
        const cplx t = (a * p2 == 0) ? v[a] : cmul(tw[(a * p2) % L], v[a]);
        Y[(size_t)a * YB + p2 * QC + c1] = t;
      }
    }
    lds_barrier();
    // pass 2, in place: frequency k1 + L1 k2 goes where row p2 = k2 was
    if (p2act) {
      cplx v[L2];
#pragma unroll
      for (int b = 0; b < L2; ++b) v[b] = Y[(size_t)k1 * YB + b * QC + c2];
      dft_reg<L2>(v);
#pragma unroll
      for (int b = 0; b < L2; ++b) Y[(size_t)k1 * YB + b * QC + c2] = v[b];
    }
    lds_barrier();
    // accumulate every output frequency (unpacking the two real spectra of a sequence)
#pragma unroll
    for (int u = 0; u < MAXS; ++u) {
      const int s = tid + u * 256;
      if (s >= nslots) continue;
      double sn, cs;
      sincospi(-2.0 * (double)e0[u] / (double)m, &sn, &cs);
      e0[u] += estep[u];
      if (e0[u] >= m) e0[u] -= m;
      cplx wq = {cs, sn};
      cplx a = acc[u];
#pragma unroll 2
      for (int c = 0; c < QC; ++c) {
        const i64 q = q0 + 2 * c;
        if (q >= nq) break;
        const cplx zc = Y[fcol[u] + c];
        const cplx zn = Y[fconj[u] + c];
        // twice the two real spectra: 2 X1 = zc + conj(zn), 2 X2 = (zc - conj(zn)) / i; the
        // factor 1/2 is applied once at the end.  Columns q and q + 1 enter as
        // w^q (X1 + w X2), and w^q advances by w^2.
        cplx x = {zc.re + zn.re, zc.im - zn.im};
        if (q + 1 < nq) {
          const cplx x2 = {zc.im + zn.im, zn.re - zc.re};
          const cplx t = cmul(wstep[u], x2);
          x.re += t.re; x.im += t.im;
        }
        const cplx t = cmul(wq, x);
        a.re += t.re; a.im += t.im;
        wq = cmul(wq, wstep2[u]);
      }
      acc[u] = a;
    }
    lds_barrier();  // Y read by every thread before the next slab's pass 1 writes it
  }
#pragma unroll
  for (int u = 0; u < MAXS; ++u) {
    const int s = tid + u * 256;
    if (s >= nslots) continue;
    double* y = B + col * ldb;
    y[2 * s] = 0.5 * acc[u].re;
    if (2 * s + 1 < ell) y[2 * s + 1] = 0.5 * acc[u].im;
  }
}

template <int L1, int L2, int MAXS, bool SIGNS>
int launch_srft_fft_s(hipStream_t st, i64 Q, i64 P, const double* X, i64 ldx, i64 mm,
                      const double* d, const uint32_t* dbits, const i64* idx, int ell,
                      double* B, i64 ldb) {
  constexpr int L = L1 * L2;
  constexpr int QC = 256 / (L1 > L2 ? L1 : L2);
  // Y (L1 blocks of L2 QC + 1) and the L twiddles, 16 bytes each, then the sign mask
  const size_t bytes = ((size_t)L1 * (L2 * QC + 1) + L) * 16 +
                       (SIGNS ? (size_t)((P + 31) / 32) * 4 : 0);
  if (bytes > 64 * 1024) {
    static std::atomic<uint64_t> lds_set{0};  // per device, set once
    lra_lds_limit_once(lds_set, (const void*)srft_fft_kernel<L1, L2, MAXS, SIGNS>,
                       (int)std::max<size_t>(bytes, 96 * 1024));
  }
  hipLaunchKernelGGL((srft_fft_kernel<L1, L2, MAXS, SIGNS>), dim3((unsigned)Q), dim3(256),
                     bytes, st, P, X, ldx, mm, d, dbits, idx, ell, B, ldb);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    lra_set_error("srft launch: %s", hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
}

// dbits: the sign mask of a library-drawn d (P <= kSignRows), or null (d as given)
template <int L1, int L2>
int launch_srft_fft(hipStream_t st, i64 Q, i64 P, const double* X, i64 ldx, i64 mm,
                    const double* d, const uint32_t* dbits, const i64* idx, int ell, double* B,
                    i64 ldb) {
  // output slots per thread: (ell + 1) / 2 slots over 256 threads
  if (ell <= 512 && dbits)
    return launch_srft_fft_s<L1, L2, 1, true>(st, Q, P, X, ldx, mm, d, dbits, idx, ell, B, ldb);
  if (ell <= 512)
    return launch_srft_fft_s<L1, L2, 1, false>(st, Q, P, X, ldx, mm, d, dbits, idx, ell, B, ldb);
  return launch_srft_fft_s<L1, L2, 4, false>(st, Q, P, X, ldx, mm, d, dbits, idx, ell, B, ldb);
}

// bits[w] bit j set when d[32 w + j] < 0
__global__ void srft_sign_bits_kernel(i64 P, const double* __restrict__ d, uint32_t* bits) {
  const i64 w = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (w >= (P + 31) / 32) return;
  uint32_t b = 0;
  for (int j = 0; j < 32; ++j) {
    const i64 i = 32 * w + j;
    if (i < P && d[i] < 0.0) b |= 1u << j;
  }
  bits[w] = b;
}

// sign masks up to 65536 rows (8 KB): the kernel's LDS stays within two workgroups per CU
constexpr i64 kSignRows = 65536;

// B[i, j] = A[r[i]-1, j]
__global__ void subset_kernel(i64 n, const double* A, i64 lda, int ell, const i64* r, double* B,
                              i64 ldb) {
  const i64 total = (i64)ell * n;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    const int i = (int)(e % ell);
    const i64 j = e / ell;
    B[i + j * ldb] = A[(r[i] - 1) + j * lda];
  }
}

// One block per column, the column first staged in LDS with coalesced loads (the output rows
// gather it in permutation order, which read from HBM was one scattered 8-byte load per
// term); then as sprn_kernel: thread i sums output row i over its group of the permutation,
// the same products in the same order.  m <= kSprnLdsRows.
constexpr i64 kSprnLdsRows = 16384;  // 128 KB of LDS
__global__ __launch_bounds__(256) void sprn_lds_kernel(i64 m, const double* __restrict__ A,
                                                       i64 lda, int ell,
                                                       const i64* __restrict__ r,
                                                       const double* __restrict__ s,
                                                       double* __restrict__ B, i64 ldb) {
  extern __shared__ double xs[];
  const i64 col = blockIdx.x;
  const double* x = A + col * lda;
  for (i64 t = threadIdx.x; t < m; t += 256) xs[t] = x[t];
  __syncthreads();
  const i64 qq = m / ell, rr = m % ell;
  for (int i = threadIdx.x; i < ell; i += 256) {
    const i64 t = i;
    const i64 start = t <= rr ? t * (qq + 1) : rr * (qq + 1) + (t - rr) * qq;
    const i64 p = (i + 1) <= rr ? qq + 1 : qq;
    double acc = 0.0;
    for (i64 l = 0; l < p; ++l) acc += s[start + l] * xs[r[start + l] - 1];
    B[i + col * ldb] = acc;
  }
}

// one block per column; thread i sums output row i over its group of the permutation
__global__ void sprn_kernel(i64 m, const double* __restrict__ A, i64 lda, int ell,
                            const i64* __restrict__ r, const double* __restrict__ s,
                            double* __restrict__ B, i64 ldb) {
  const i64 col = blockIdx.x;
  const double* x = A + col * lda;
  for (int i = threadIdx.x; i < ell; i += blockDim.x) {
    // group sizes p_t = fld(m-t, ell) + 1 (t = 1..ell): qq+1 for t <= rr, qq after,
    // with m = qq*ell + rr (also right for m < ell: qq = 0, empty groups beyond m)
    const i64 qq = m / ell, rr = m % ell;
    const i64 t = i;  // groups before this one (1-based row i+1)
    const i64 start = t <= rr ? t * (qq + 1) : rr * (qq + 1) + (t - rr) * qq;
    const i64 p = (i + 1) <= rr ? qq + 1 : qq;
    double acc = 0.0;
    for (i64 l = 0; l < p; ++l) acc += s[start + l] * x[r[start + l] - 1];
    B[i + col * ldb] = acc;
  }
}

}  // namespace

#define LRA_LAUNCH_CHECK(what)                                                          \
  do {                                                                                  \
    hipError_t _e = hipGetLastError();                                                  \
    if (_e != hipSuccess) {                                                             \
      lra_set_error("%s launch: %s", what, hipGetErrorString(_e));                      \
      return LRA_E_HIP;                                                                 \
    }                                                                                   \
  } while (0)

// op(A) as a column-major P x Q matrix: trans n uses A directly, trans c a transposed copy.
static int op_matrix(lra_ctx* c, int trans, i64 m, i64 n, const double* A, i64 lda,
                     const double** X, i64* ldx, i64* P, i64* Q) {
  if (trans == LRA_TRANS_N) {
    *X = A; *ldx = lda; *P = m; *Q = n;
    return LRA_OK;
  }
  double* t = nullptr;
  LRA_CHECK(lra_wsT(c, WS_TMP0, (size_t)m * n, &t));
  LRA_CHECK(lra_launch_transpose(c->stream, m, n, A, lda, t, std::max<i64>(n, 1)));
  *X = t; *ldx = std::max<i64>(n, 1); *P = n; *Q = m;
  return LRA_OK;
}

int lra_sketch_srft_impl(lra_ctx* c, int trans, i64 m, i64 n, const double* A, i64 lda,
                         i64 ell, const double* d, const i64* idx, uint64_t seed,
                         uint64_t stream_id, double* B, i64 ldb) {
  const double* X;
  i64 ldx, P, Q;
  LRA_CHECK(op_matrix(c, trans, m, n, A, lda, &X, &ldx, &P, &Q));
  if (ell <= 0 || Q <= 0) return LRA_OK;
  if (P <= 0) { lra_set_error("srft: empty operand"); return LRA_E_UNSUPPORTED; }
  if (ell > 2048) { lra_set_error("srft: ell=%lld > 2048", (long long)ell); return LRA_E_UNSUPPORTED; }
  if (Q > 2147483647ll) { lra_set_error("srft: too many columns"); return LRA_E_UNSUPPORTED; }
  const uint32_t* dbits = nullptr;
  if (!d) {
    double* dd = nullptr;
    i64* ii = nullptr;
    LRA_CHECK(lra_wsT(c, WS_LP0, (size_t)P, &dd));
    LRA_CHECK(lra_wsT(c, WS_LP1, (size_t)ell, &ii));
    LRA_CHECK(lra_launch_srft_rand(c->stream, seed, stream_id, P, ell, dd, ii));
    d = dd;
    idx = ii;
    if (P <= kSignRows) {  // d is +-1 by construction: pass it as a sign mask
      uint32_t* bits = nullptr;
      const i64 nw = (P + 31) / 32;
      LRA_CHECK(lra_wsT(c, WS_LP2, (size_t)nw, &bits));
      hipLaunchKernelGGL(srft_sign_bits_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0,
                         c->stream, P, d, bits);
      LRA_LAUNCH_CHECK("srft signs");
      dbits = bits;
    }
  }
  i64 l = std::min<i64>(ell, P);
  while (P % l) --l;
  const i64 mm = P / l;
  const bool pow2 = (l & (l - 1)) == 0;
  if (pow2 && ell <= 2048) {
    // two register passes for power-of-two l <= 256
    switch (l) {
      case 256: return launch_srft_fft<16, 16>(c->stream, Q, P, X, ldx, mm, d, dbits, idx, (int)ell, B, ldb);
      case 128: return launch_srft_fft<16, 8>(c->stream, Q, P, X, ldx, mm, d, dbits, idx, (int)ell, B, ldb);
      case 64: return launch_srft_fft<8, 8>(c->stream, Q, P, X, ldx, mm, d, dbits, idx, (int)ell, B, ldb);
      case 32: return launch_srft_fft<8, 4>(c->stream, Q, P, X, ldx, mm, d, dbits, idx, (int)ell, B, ldb);
      case 16: return launch_srft_fft<4, 4>(c->stream, Q, P, X, ldx, mm, d, dbits, idx, (int)ell, B, ldb);
      default: break;
    }
  }
  int logl = 0;
  while ((1ll << logl) < l) ++logl;
  // packed complex columns per slab: LDS = l*qc*16 (+ l*qc*16 for the DFT path) + twiddles
  const size_t lds_cap = 120 * 1024;
  int qc = 16;
  auto need = [&](int q) {
    return (size_t)l * q * 16 * (pow2 ? 1 : 2) + (size_t)(l / 2 + 1) * 16;
  };
  while (qc > 1 && need(qc) > lds_cap) qc >>= 1;
  if (need(qc) > lds_cap) { lra_set_error("srft: l=%lld too large", (long long)l); return LRA_E_UNSUPPORTED; }
  if (need(qc) > 64 * 1024) {
    static std::atomic<uint64_t> lds_set{0};  // per device, set once to the cap need(qc) obeys
    lra_lds_limit_once(lds_set, (const void*)srft_kernel, (int)lds_cap);
  }
  hipLaunchKernelGGL(srft_kernel, dim3((unsigned)Q), dim3(SRFT_THREADS), need(qc), c->stream, P,
                     X, ldx, (int)l, mm, qc, pow2 ? 1 : 0, logl, d, idx, (int)ell, B, ldb);
  LRA_LAUNCH_CHECK("srft");
  return LRA_OK;
}

int lra_sketch_subset_impl(lra_ctx* c, int trans, i64 m, i64 n, const double* A, i64 lda,
                           i64 ell, const i64* r, uint64_t seed, uint64_t stream_id, double* B,
                           i64 ldb) {
  const double* X;
  i64 ldx, P, Q;
  LRA_CHECK(op_matrix(c, trans, m, n, A, lda, &X, &ldx, &P, &Q));
  if (ell <= 0 || Q <= 0) return LRA_OK;
  if (!r) {
    i64* rr = nullptr;
    LRA_CHECK(lra_wsT(c, WS_LP1, (size_t)ell, &rr));
    LRA_CHECK(lra_launch_subset_rand(c->stream, seed, stream_id, P, ell, rr));
    r = rr;
  }
  const i64 total = ell * Q;
  const int blocks = (int)std::min<i64>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(subset_kernel, dim3(blocks), dim3(256), 0, c->stream, Q, X, ldx, (int)ell,
                     r, B, ldb);
  LRA_LAUNCH_CHECK("subset");
  return LRA_OK;
}

int lra_sketch_sprn_impl(lra_ctx* c, int trans, i64 m, i64 n, const double* A, i64 lda,
                         i64 ell, const i64* r, const double* s, uint64_t seed,
                         uint64_t stream_id, double* B, i64 ldb) {
  const double* X;
  i64 ldx, P, Q;
  LRA_CHECK(op_matrix(c, trans, m, n, A, lda, &X, &ldx, &P, &Q));
  if (ell <= 0 || Q <= 0) return LRA_OK;
  if (!r) {
    double* ss = nullptr;
    i64* rr = nullptr;
    LRA_CHECK(lra_wsT(c, WS_LP0, (size_t)P, &ss));
    LRA_CHECK(lra_wsT(c, WS_LP1, (size_t)P, &rr));
    LRA_CHECK(lra_launch_sprn_rand(c->stream, seed, stream_id, P, ss, rr));
    r = rr;
    s = ss;
  }
  if (Q > 2147483647ll) { lra_set_error("sprn: too many columns"); return LRA_E_UNSUPPORTED; }
  if (P <= kSprnLdsRows) {
    static std::atomic<uint64_t> lds_set{0};
    lra_lds_limit_once(lds_set, (const void*)sprn_lds_kernel,
                       (int)(kSprnLdsRows * sizeof(double)));
    hipLaunchKernelGGL(sprn_lds_kernel, dim3((unsigned)Q), dim3(256),
                       (size_t)std::max<i64>(P, 1) * sizeof(double), c->stream, P, X, ldx,
                       (int)ell, r, s, B, ldb);
  } else {
    hipLaunchKernelGGL(sprn_kernel, dim3((unsigned)Q), dim3(256), 0, c->stream, P, X, ldx,
                       (int)ell, r, s, B, ldb);
  }
  LRA_LAUNCH_CHECK("sprn");
  return LRA_OK;
}

extern "C" int lra_sketch_srft(lra_ctx* c, int trans, int64_t m, int64_t n, const double* A,
                               int64_t lda, int64_t ell, const double* d, const int64_t* idx,
                               uint64_t seed, uint64_t stream_id, double* B, int64_t ldb) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 6);
  LRA_ARG(ell >= 0, 7);
  LRA_ARG((d == nullptr) == (idx == nullptr), 9);
  LRA_ARG(ldb >= std::max<int64_t>(1, ell), 13);
  return lra_sketch_srft_impl(c, trans, m, n, A, lda, ell, d, idx, seed, stream_id, B, ldb);
}

extern "C" int lra_sketch_subset(lra_ctx* c, int trans, int64_t m, int64_t n, const double* A,
                                 int64_t lda, int64_t ell, const int64_t* r, uint64_t seed,
                                 uint64_t stream_id, double* B, int64_t ldb) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 6);
  LRA_ARG(ell >= 0, 7);
  LRA_ARG(ldb >= std::max<int64_t>(1, ell), 12);
  return lra_sketch_subset_impl(c, trans, m, n, A, lda, ell, r, seed, stream_id, B, ldb);
}

extern "C" int lra_sketch_sprn(lra_ctx* c, int trans, int64_t m, int64_t n, const double* A,
                               int64_t lda, int64_t ell, const int64_t* r, const double* s,
                               uint64_t seed, uint64_t stream_id, double* B, int64_t ldb) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 6);
  LRA_ARG(ell >= 0, 7);
  LRA_ARG((r == nullptr) == (s == nullptr), 9);
  LRA_ARG(ldb >= std::max<int64_t>(1, ell), 13);
  return lra_sketch_sprn_impl(c, trans, m, n, A, lda, ell, r, s, seed, stream_id, B, ldb);
}
