// maxdet.hip -- strong rank-revealing swaps that maximise |det(R11)|
// (maxdet_swapcols! / maxdet_update!, src/pqr.jl:445-501; findmaxabs, src/util.jl:13-23).
//
// One iteration: (i, j) = position of the largest |T| (the LAST maximum in column-major
// order, as findmaxabs keeps replacing on ties); stop when it is <= 1 + tol or the iteration
// budget is spent; else swap p[i] <-> p[k+j] and apply the Sherman-Morrison update
//   u = T[:, j] - e_i,  T[:, j] = e_i,  v = T[i, :],  T += (-1 / (1 + u_i)) u v'
// (the ger of maxdet_update!, with BLAS's y*alpha-first order, in separately rounded
// multiply and add so the result matches dger bit for bit).  Three launches per iteration
// (argmax partials, decide + prepare, rank-1 update) keep every state word on the device;
// the host reads the state back once per batch of iterations.
#include "lra_internal.h"

#include <algorithm>

namespace {

struct MdState {
  double tol;
  i64 niter, niter_max;
  int done, active;
  double alpha;
  i64 i, j;
};

constexpr int MD_THREADS = 256;
constexpr int MD_MAX_PARTIALS = 1024;

// larger |value| wins; on a tie the larger (later) linear index wins
LRA_DEV bool md_better(double v, i64 ix, double bv, i64 bix) {
  return v > bv || (v == bv && ix > bix);
}

__global__ __launch_bounds__(MD_THREADS) void md_argmax_partial(
    i64 k, i64 nr, const double* __restrict__ T, i64 ldt, i64 cols_per_block,
    double* __restrict__ pv, i64* __restrict__ pix, const MdState* __restrict__ st) {
  if (st->done) return;
  const i64 c0 = (i64)blockIdx.x * cols_per_block;
  const i64 c1 = std::min<i64>(nr, c0 + cols_per_block);
  double bv = -1.0;
  i64 bix = -1;
  for (i64 c = c0; c < c1; ++c)
    for (i64 r = threadIdx.x; r < k; r += MD_THREADS) {
      const double t = fabs(T[r + c * ldt]);
      const i64 ix = r + c * k;
      if (md_better(t, ix, bv, bix)) { bv = t; bix = ix; }
    }
  __shared__ double sv[MD_THREADS];
  __shared__ i64 si[MD_THREADS];
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bix;
  __syncthreads();
  for (int w = MD_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double ov = sv[threadIdx.x + w];
      const i64 oi = si[threadIdx.x + w];
      if (md_better(ov, oi, sv[threadIdx.x], si[threadIdx.x])) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pix[blockIdx.x] = si[0];
  }
}

// single block: global argmax, the stop test, the swap, u, T[:, j] = e_i and v = T[i, :]
__global__ __launch_bounds__(MD_THREADS) void md_prepare(
    i64 k, i64 nr, double* __restrict__ T, i64 ldt, i64* __restrict__ p, int nparts,
    const double* __restrict__ pv, const i64* __restrict__ pix, double* __restrict__ u,
    double* __restrict__ v, MdState* __restrict__ st) {
  __shared__ double sv[MD_THREADS];
  __shared__ i64 si[MD_THREADS];
  __shared__ int s_go;
  __shared__ i64 s_i, s_j;
  if (st->done) return;
  double bv = -1.0;
  i64 bix = -1;
  for (int b = threadIdx.x; b < nparts; b += MD_THREADS)
    if (md_better(pv[b], pix[b], bv, bix)) { bv = pv[b]; bix = pix[b]; }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bix;
  __syncthreads();
  for (int w = MD_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double ov = sv[threadIdx.x + w];
      const i64 oi = si[threadIdx.x + w];
      if (md_better(ov, oi, sv[threadIdx.x], si[threadIdx.x])) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double tmax = sv[0];
    int go = 1;
    st->active = 0;
    if (!(tmax > 1.0 + st->tol) || si[0] < 0) go = 0;  // converged
    else if (st->niter == st->niter_max) go = 0;      // budget spent (-1: unlimited)
    if (!go) {
      st->done = 1;
    } else {
      st->niter += 1;
      const i64 i = si[0] % k, j = si[0] / k;
      const i64 t = p[i];
      p[i] = p[k + j];
      p[k + j] = t;
      s_i = i;
      s_j = j;
    }
    s_go = go;
  }
  __syncthreads();
  if (!s_go) return;
  const i64 i = s_i, j = s_j;
  double* Tj = T + j * ldt;
  for (i64 l = threadIdx.x; l < k; l += MD_THREADS) {
    u[l] = Tj[l];
    Tj[l] = 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    u[i] = u[i] - 1.0;
    Tj[i] = 1.0;
    st->alpha = -1.0 / (1.0 + u[i]);
    st->i = i;
    st->j = j;
    st->active = 1;
  }
  __syncthreads();
  for (i64 c = threadIdx.x; c < nr; c += MD_THREADS) v[c] = T[i + c * ldt];
}

// T += u * (alpha * v)'
__global__ __launch_bounds__(MD_THREADS) void md_update(i64 k, i64 nr, int rblocks,
                                                       double* __restrict__ T,
                                                       i64 ldt, const double* __restrict__ u,
                                                       const double* __restrict__ v,
                                                       const MdState* __restrict__ st) {
#pragma clang fp contract(off)  // dger's separately rounded multiply and add, no FMA
  if (!st->active) return;
  const double alpha = st->alpha;
  const i64 c = blockIdx.x / rblocks;  // one column per group of rblocks blocks
  if (c >= nr) return;
  const double y = alpha * v[c];
  for (i64 r = (i64)(blockIdx.x % rblocks) * MD_THREADS + threadIdx.x; r < k;
       r += (i64)rblocks * MD_THREADS)
    T[r + c * ldt] = T[r + c * ldt] + u[r] * y;
}

__global__ void md_init(MdState* st, double tol, i64 niter_max) {
  st->tol = tol;
  st->niter = 0;
  st->niter_max = niter_max;
  st->done = 0;
  st->active = 0;
  st->alpha = 0.0;
  st->i = st->j = 0;
}

}  // namespace

// p (n, device, 1-based positions: first k = skeleton, rest = redundant), T k x (n-k).
extern "C" int lra_maxdet_swapcols_f64(lra_ctx* c, int64_t k, int64_t n, int64_t* p, double* T,
                                       int64_t ldt, double tol, int64_t niter_max,
                                       int64_t* niter) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(k >= 0, 2);
  LRA_ARG(n >= k, 3);
  LRA_ARG(p != nullptr || n == 0, 4);
  LRA_ARG(T != nullptr || k * (n - k) == 0, 5);
  LRA_ARG(ldt >= std::max<int64_t>(1, k), 6);
  LRA_ARG(niter != nullptr, 9);
  LRA_ARG((n - k) * 8 < (1ll << 31), 3);  // grid of the rank-1 update
  LRA_HIP(hipSetDevice(c->device));
  *niter = 0;
  const i64 nr = n - k;
  if (k == 0 || nr == 0) return LRA_OK;
  // partial argmax blocks: whole columns per block, at most MD_MAX_PARTIALS blocks
  const i64 cpb = std::max<i64>(1, std::max<i64>((nr + MD_MAX_PARTIALS - 1) / MD_MAX_PARTIALS,
                                                 (i64)(4096 / std::max<i64>(k, 1))));
  const int nparts = (int)((nr + cpb - 1) / cpb);
  char* ws = nullptr;
  const size_t state_bytes = 256;
  LRA_CHECK(lra_wsT(c, WS_MD0, state_bytes + (size_t)nparts * 16, &ws));
  MdState* st = (MdState*)ws;
  double* pv = (double*)(ws + state_bytes);
  i64* pix = (i64*)(ws + state_bytes + (size_t)nparts * 8);
  double* uv = nullptr;
  LRA_CHECK(lra_wsT(c, WS_MD1, (size_t)(k + nr), &uv));
  double* u = uv;
  double* v = uv + k;
  hipLaunchKernelGGL(md_init, dim3(1), dim3(1), 0, c->stream, st, tol, (i64)niter_max);
  LRA_HIP(hipGetLastError());
  const int ublocks = (int)std::min<i64>(8, (k + MD_THREADS - 1) / MD_THREADS);
  const dim3 ugrid((unsigned)(nr * ublocks));
  MdState hs{};
  const int kBatch = 8;
  for (;;) {
    for (int it = 0; it < kBatch; ++it) {
      hipLaunchKernelGGL(md_argmax_partial, dim3(nparts), dim3(MD_THREADS), 0, c->stream, k, nr,
                         T, ldt, cpb, pv, pix, st);
      hipLaunchKernelGGL(md_prepare, dim3(1), dim3(MD_THREADS), 0, c->stream, k, nr, T, ldt, p,
                         nparts, pv, pix, u, v, st);
      hipLaunchKernelGGL(md_update, ugrid, dim3(MD_THREADS), 0, c->stream, k, nr, ublocks, T,
                         ldt, u, v, st);
    }
    LRA_HIP(hipGetLastError());
    LRA_HIP(hipMemcpyAsync(&hs, st, sizeof(MdState), hipMemcpyDeviceToHost, c->stream));
    LRA_HIP(hipStreamSynchronize(c->stream));
    if (hs.done) break;
  }
  *niter = hs.niter;
  return LRA_OK;
}
