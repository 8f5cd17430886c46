// batched.hip -- the ID of many independent blocks in persistent launches: every block runs the
// adaptive rounds of idfact (src/id.jl:434-447 -> sketchfact_randn, src/sketch.jl:223-240 ->
// geqp3_adap!, src/pqr.jl:348-418 -> T = R11 \ R12, src/pqr.jl:438-442) inside one
// workgroup of 1024 threads.
//
// A launch is one kernel per adaptive round ("stage"): stage r sketches with order
// ell_r = samp(nb 2^r) and Philox stream r, like idfact's round r.  Stage 0 takes every block
// (sorted by decreasing m*n, so the large ones start first); a block whose rank reaches the
// round's target is appended on the device to the next stage's list, and a block that no
// stage can take is finished by lra_idfact afterwards.  The stages are separate kernels on one
// stream (a counter reset + one kernel per stage), so the launch can be captured in a hipGraph.
//
// Per block, in one workgroup that takes the next block from the stage's work queue:
//   1. sketch  B (ell x n) = Omega(:, 1:m) * A on FP64 MFMA.  Omega_r is the ell_r x mmax
//      matrix of Philox stream r shared by every block: its first m columns are exactly the
//      Omega that idfact draws for an m-row block, so a block's result is the result of idfact
//      on that block alone.  Omega is staged through LDS in row chunks; each wave computes
//      16-column strips of B, with the next 16 rows of A in flight while the current ones are
//      multiplied, and stages them in a per-workgroup scratch slab;
//   2. pivoted QR of B with TPC threads per column (ell/TPC rows each in registers): per step
//      every wave publishes its best column (largest partial norm, then lowest position, as
//      idamax) together with that column's values; after a barrier every wave picks the pivot
//      (a DPP argmax over the wave candidates) and runs the block bookkeeping, wave 0 builds the
//      Householder reflector into LDS, and after a second barrier every wave applies it to its
//      columns.  Norm downdating, the recompute criterion, the block boundaries of dlaqps and
//      the stopping rule of geqp3_adap! are those of the single-matrix kernel in qrcp.hip;
//   3. sk / rd from the final column positions and T(:, j) = R11 \ R(1:k, j) by back
//      substitution in each column's registers.
#include "lra_internal.h"
#include "lra_wave.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <numeric>
#include <type_traits>
#include <vector>

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));  // 8-byte aligned pair
// global-memory views of A (the block pointers are read from the descriptor table, so the
// compiler cannot infer their address space): global loads count against vmcnt only, and an
// LDS wait does not wait for them
typedef __attribute__((address_space(1))) d2u gd2u;
typedef __attribute__((address_space(1))) double gdouble;
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

constexpr int SLAB = 40 * 1024;  // scratch doubles per workgroup (>= LM * columns, every variant)
constexpr int NSTAGE = 4;
constexpr int NKIND = 3;  // kernels per stage: small / one workgroup / workgroup pair

struct BDesc {
  const double* A;
  i64 lda;
  i64* sk;
  i64* rd;
  double* T;
  i64 ldt;
  int m, n;
};

struct BStage {
  const BDesc* desc;
  const int* list;          // this stage's blocks
  const unsigned* list_n;   // device count of `list` (stages >= 1), or null: nq
  int nq;
  unsigned* head;           // work-queue head, zeroed before the launch
  const double* omega;      // ell x mmax, row r at omega + r * ldo (k-contiguous)
  i64 ldo;
  double* scratch;          // SLAB doubles per workgroup
  i64* k_out;               // per block: rank, or -1 (to lra_idfact)
  int* next_list[NKIND];    // rejected blocks for the next stage's kernels (by width)
  unsigned* next_n[NKIND];
  int next_maxn[NKIND];
  int ell, rank, nb, target;
  double atol, rtol;
  // workgroup pairs: the exchange slots, launch generation (tags) and error word
  u64x2* xchg;
  int uncached;
  const unsigned* gen;
  int* err;
};

// queue heads and list sizes to zero; the launch generation (part of every exchange tag) + 1
__global__ void reset_kernel(unsigned* c, int n, unsigned* gen) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
  if (threadIdx.x == 0) *gen = *gen + 1u;
}

// barrier for LDS hand-offs only: outstanding global loads into registers stay in flight
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// ---- exchange between the two workgroups of a pair (one column block split in halves) ----
// A slot holds LM + 2 tagged 16-byte chunks: {norm}, {position | flag << 32}, then the
// candidate column's rows.  Each chunk carries the tag in its second half, so a reader
// validates every chunk on its own; with uncached memory the stores need no fence.
LRA_DEV uint64_t xtag(unsigned gen, int it, int st) {
  return (1ull << 63) | ((uint64_t)(gen & 0x7FFFFFu) << 40) |
         ((uint64_t)(unsigned)it << 8) | (uint64_t)(st & 0xFF);
}

template <int LM>
LRA_DEV void xpublish(u64x2* slot, uint64_t tag, double val, int pos, int flag,
                      const double* col, int uncached) {
  const int lane = threadIdx.x & 63;
  for (int c = lane; c < LM + 2; c += 64) {
    u64x2 w;
    w.y = tag;
    if (c == 0) w.x = (uint64_t)__double_as_longlong(val);
    else if (c == 1) w.x = (uint64_t)(uint32_t)pos | ((uint64_t)(uint32_t)flag << 32);
    else w.x = col ? (uint64_t)__double_as_longlong(col[c - 2]) : 0ull;
    if (uncached) slot[c] = w;
    else *(volatile u64x2*)(slot + c) = w;
  }
  if (!uncached) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

// Wait (bounded) for the partner's slot of tag `tag`; the rows go to dst (this wave's LDS).
// Returns false when the wait expired or another wait already did (the error word is set).
template <int LM>
LRA_DEV bool xpoll(const u64x2* slot, uint64_t tag, int uncached, int* err, double* val,
                   int* pos, int* flag, double* dst) {
  constexpr int S = LM + 2;
  const int lane = threadIdx.x & 63;
  const volatile u64x2* v = (const volatile u64x2*)slot;
  u64x2 w0 = {0ull, 0ull}, w1 = {0ull, 0ull};
  unsigned spins = 0;
  bool ok = true;
  for (;;) {
    if (lane < S) w0 = v[lane];
    if (lane + 64 < S) w1 = v[lane + 64];
    const bool mine = (lane >= S || w0.y == tag) && (lane + 64 >= S || w1.y == tag);
    if (__all(mine)) break;
    if ((++spins & 255) == 0 &&
        (spins > LRA_SPIN_LIMIT / 64 ||
         __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
      ok = false;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
    if (!uncached) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  if (!ok) {
    if (lane == 0)
      __hip_atomic_store(err, (int)LRA_E_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
  }
  const uint64_t h0 = __builtin_amdgcn_readlane((int)(w0.x & 0xffffffffu), 0) & 0xffffffffull;
  const uint64_t h1 = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(w0.x >> 32), 0);
  *val = __longlong_as_double((long long)((h1 << 32) | h0));
  *pos = __builtin_amdgcn_readlane((int)(w0.x & 0xffffffffu), 1);
  *flag = __builtin_amdgcn_readlane((int)(w0.x >> 32), 1);
  if (lane >= 2 && lane < S) dst[lane - 2] = __longlong_as_double((long long)w0.x);
  if (lane + 64 < S) dst[lane + 62] = __longlong_as_double((long long)w1.x);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return true;
}

// rows k0 + 4 lq .. +3 of the columns of the wave's first NV strips: A is the block, voff[i]
// the byte offset of row 4 lq of this lane's column in strip i; FULL: the whole group
// k0 .. k0+15 lies within the m rows, else rows past m read as zero
template <int SPW, int NV, bool FULL>
LRA_DEV void batched_load_rows(const double* A, const i64 (&voff)[SPW], int k0, int m,
                               d2 (&x)[SPW][2]) {
  const char* base = (const char*)(A + k0);
  if constexpr (FULL) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const gd2u* src = (const gd2u*)(base + voff[i]);
      x[i][0] = src[0];
      x[i][1] = src[1];
    }
  } else {
    const int kr = k0 + 4 * (((int)threadIdx.x & 63) >> 4);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const gdouble* src = (const gdouble*)(base + voff[i]);
      x[i][0].x = kr < m ? src[0] : 0.0;
      x[i][0].y = kr + 1 < m ? src[1] : 0.0;
      x[i][1].x = kr + 2 < m ? src[2] : 0.0;
      x[i][1].y = kr + 3 < m ? src[3] : 0.0;
    }
  }
}

// acc[i] += Omega(:, kg .. kg+15 of the staged chunk) * x[i] for the wave's first NV strips
// (rows ell .. LM-1 of the staged chunk are zero, so every row is read unconditionally)
template <int LM, int SPW, int NV, int NTL, int OLD>
LRA_DEV void batched_mfma_group(const double* sO, int kg, int lq, int l16,
                                const d2 (&x)[SPW][2], d4 (&acc)[SPW][NTL]) {
  const int kr = kg + 4 * lq;
  double av[NTL][4];
#pragma unroll
  for (int t = 0; t < NTL; ++t) {
    // (rows past LM, in the last tile, read row LM-1: they only reach output rows past LM)
    const int row = 16 * t + l16 < LM ? 16 * t + l16 : LM - 1;
    const d2 y0 = *reinterpret_cast<const d2*>(sO + row * OLD + kr);
    const d2 y1 = *reinterpret_cast<const d2*>(sO + row * OLD + kr + 2);
    av[t][0] = y0.x; av[t][1] = y0.y; av[t][2] = y1.x; av[t][3] = y1.y;
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double bv[4] = {x[i][0].x, x[i][0].y, x[i][1].x, x[i][1].y};
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int t = 0; t < NTL; ++t)
        acc[i][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t][u], bv[u], acc[i][t], 0, 0, 0);
  }
}

// LM: largest sketch order; TPC: threads per column (RPT = LM / TPC rows each); CL: workgroups
// per block (2: each takes half the columns and the pair exchanges its pivot candidates);
// BT: threads per workgroup (512: two workgroups per CU, so one's sketch overlaps the
// other's pivot steps; 1024: one).
template <int LM, int TPC, int CL, int BT>
__global__ __launch_bounds__(BT, 4) void id_batched_kernel(BStage p) {
  constexpr int BW = BT / 64;                // waves per workgroup
  constexpr int RPT = LM / TPC;
  constexpr int CPW = BT / TPC;              // largest n
  constexpr int NTL = (LM + 15) / 16;        // 16-row MFMA tiles of the sketch
  constexpr int SPW = NTL <= 3 ? 2 : 1;      // strips per wave per pass
  // rows of A per Omega chunk (512-thread workgroups: small enough for two per CU)
  constexpr int MC = BT == 1024 ? (NTL <= 3 ? 256 : 128) : (NTL <= 3 ? 128 : 32);
  constexpr int OLD = MC + 2;                // LDS row stride of the chunk (16-byte rows)
  constexpr int KM = LM < 64 ? LM : 64;      // R11 capacity (k < the round's target <= 64)
  constexpr int NR = (LM + 63) / 64;         // rows per lane in the reflector build
  static_assert(RPT * TPC == LM && LM * CPW <= SLAB && (RPT % 2) == 0, "shape");

  __shared__ __align__(16) double sO[LM * OLD];  // Omega chunk, row-major
  __shared__ double sCand[BW][LM];               // each wave's best column
  __shared__ double sX[LM];                      // the step's reflector, by absolute row
  __shared__ double sR[KM * KM];                 // R11, column-major
  __shared__ double sDiag[LM];
  __shared__ double sVal[BW];
  __shared__ int sPos[BW], sLane[BW], sFlag[BW];
  __shared__ double sTau;                        // the step's tau (written by wave 0)
  __shared__ double sMax[BW];
  __shared__ double sPc[LM];   // the partner's candidate column (pairs; wave 0 fetches it)
  __shared__ double sXv;       // ... its norm, position and recompute flag
  __shared__ int sXp, sXf;
  __shared__ int sBlk, sAbort;

  const int tid = threadIdx.x, lane = tid & 63;
  // wave-uniform: the LDS addresses derived from it are scalar
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l16 = lane & 15, lq = lane >> 4;
  const int lc = tid / TPC, lic = tid % TPC;   // column, row part
  const int rbase = lic * RPT;
  double* Bs = p.scratch + (i64)blockIdx.x * SLAB;
  const double tol3z = 1.0536712127723509e-08;  // sqrt(dlamch('E'))
  const int nq = p.list_n ? (int)*p.list_n : p.nq;
  // pairs take the blocks it = pair, pair + npairs, ... (both halves walk the same list)
  const int part = CL == 2 ? (int)(blockIdx.x & 1) : 0;
  const int pair = blockIdx.x >> 1, npairs = gridDim.x >> 1;
  const unsigned gen = CL == 2 ? *p.gen : 0u;
  // slots: [pair][half][block parity][step parity].  Consecutive blocks use different
  // areas, so a half that has moved on to the next block never overwrites a slot its partner
  // may still be reading; within a block, step s+2 reuses step s's slot only after the
  // partner has published step s+1, i.e. after it finished reading step s.
  u64x2* mybase = nullptr;
  const u64x2* peerbase = nullptr;
  if (CL == 2) {
    mybase = p.xchg + (size_t)(pair * 2 + part) * 4 * (LM + 2);
    peerbase = p.xchg + (size_t)(pair * 2 + (part ^ 1)) * 4 * (LM + 2);
  }
  if (tid == 0) sAbort = 0;
  int it = CL == 2 ? pair - npairs : -1;

  for (;;) {
    __syncthreads();  // the previous block is done with every LDS array
    int qi;
    if constexpr (CL == 2) {
      it += npairs;
      qi = it;
      if (sAbort) return;
    } else {
      if (tid == 0)
        sBlk = (int)__hip_atomic_fetch_add(p.head, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      qi = sBlk;
    }
    if (qi >= nq) break;
    const int b = p.list[qi];
    const BDesc d = p.desc[b];
    const int m = d.m, n = d.n;
    u64x2* myslot = nullptr;
    const u64x2* peerslot = nullptr;
    if (CL == 2) {
      const int bpar = (qi / npairs) & 1;
      myslot = mybase + bpar * 2 * (LM + 2);
      peerslot = peerbase + bpar * 2 * (LM + 2);
    }
    const int c0col = part * CPW;               // first column of this workgroup's half
    const int nloc = min(CPW, n - c0col);       // its columns (CL == 1: all n)
    // (opaque per block, so row masks derived from ell are not hoisted out of the block loop
    // into registers that stay live across it)
    int ell = p.ell;
    asm volatile("" : "+s"(ell));

    // ---------------- 1. sketch: Bs(0:ell, 0:n) = Omega(:, 0:m) * A ----------------------
    // Each wave owns up to SPW 16-column strips per pass.  The rows of A are read in groups
    // of 16 (lane: rows kr .. kr+3 of its column l16), one group ahead of the MFMAs that use
    // them, so a group's loads are in flight while the previous group is multiplied; the
    // Omega chunk (MC rows) is staged in LDS and shared by every strip.
    {
      const int nstrips = (max(nloc, 0) + 15) >> 4;
      const int spw = nstrips > BW ? SPW : 1;
      const int npass = (nstrips + BW * spw - 1) / (BW * spw);
      for (int pass = 0; pass < npass; ++pass) {
        d4 acc[SPW][NTL];
#pragma unroll
        for (int i = 0; i < SPW; ++i)
#pragma unroll
          for (int t = 0; t < NTL; ++t) acc[i][t] = d4{0.0, 0.0, 0.0, 0.0};
        const int q0 = pass * BW * spw + wave * spw;
        int nvs = min(spw, nstrips - q0);  // this wave's strips in this pass (wave-uniform)
        nvs = nvs < 0 ? 0 : nvs;
        // column offsets (row 0); columns past nloc read column 0 of the half: their products
        // land in output columns that are never stored
        i64 voff[SPW];
#pragma unroll
        for (int i = 0; i < SPW; ++i) {
          const int col = (q0 + i) * 16 + l16;
          voff[i] = ((i64)(c0col + (col < nloc ? col : 0)) * d.lda + 4 * lq) * 8;
        }
        // NV: the wave's strips in this pass (a compile-time count, so every load below is
        // unconditional and the wait for a group's rows leaves the next group's loads in flight)
        auto run = [&](auto nv_c) {
          constexpr int NV = decltype(nv_c)::value;
          const int mfull = m & ~15;  // rows in whole 16-row groups
          d2 xa[SPW][2], xb[SPW][2];  // two row groups in flight (ping-pong)
          if (NV > 0 && mfull > 0) batched_load_rows<SPW, NV, true>(d.A, voff, 0, m, xa);
          for (int c0 = 0; c0 < m; c0 += MC) {
            const int mc = min(MC, m - c0);
            lds_barrier();  // every wave is done with the previous chunk
            for (int e = tid; e < LM * (MC / 2); e += BT) {
              const int r = e / (MC / 2), cc = (e % (MC / 2)) * 2;
              d2 v = {0.0, 0.0};
              if (r < ell) {
                const double* src = p.omega + (i64)r * p.ldo + c0 + cc;
                if (cc + 1 < mc) v = *reinterpret_cast<const d2*>(src);
                else if (cc < mc) v.x = src[0];
              }
              *reinterpret_cast<d2*>(sO + r * OLD + cc) = v;
            }
            lds_barrier();
            if constexpr (NV > 0) {
              // whole groups of this chunk, two per iteration, every load and MFMA
              // unconditional, so waiting for one group leaves the next one's loads in flight.
              // Every chunk but the last holds an even number of whole groups and starts on
              // xa; the last iteration's xa load is the next chunk's first group, or the last
              // whole group again (clamped, not used).
              const int gend = min(mc, mfull - c0);
              int kg = 0;
              for (; kg + 32 <= gend; kg += 32) {
                batched_load_rows<SPW, NV, true>(d.A, voff, c0 + kg + 16, m, xb);
                batched_mfma_group<LM, SPW, NV, NTL, OLD>(sO, kg, lq, l16, xa, acc);
                batched_load_rows<SPW, NV, true>(d.A, voff, min(c0 + kg + 32, mfull - 16), m, xa);
                batched_mfma_group<LM, SPW, NV, NTL, OLD>(sO, kg + 16, lq, l16, xb, acc);
              }
              if (kg < gend)  // an odd last whole group (last chunk only), already in xa
                batched_mfma_group<LM, SPW, NV, NTL, OLD>(sO, kg, lq, l16, xa, acc);
              if (c0 + mc == m && mfull < m) {  // the last, partial group
                batched_load_rows<SPW, NV, false>(d.A, voff, mfull, m, xb);
                batched_mfma_group<LM, SPW, NV, NTL, OLD>(sO, mfull - c0, lq, l16, xb, acc);
              }
            }
          }
        };
        if (nvs >= SPW) run(std::integral_constant<int, SPW>{});
        else if (nvs == 1) run(std::integral_constant<int, 1>{});
        else run(std::integral_constant<int, 0>{});
        // output element (row 16t + lq + 4r, column l16) of each 16 x 16 tile; rows ell .. LM-1
        // are zero (so are the staged Omega rows), rows past LM are not stored
#pragma unroll
        for (int i = 0; i < SPW; ++i) {
          const int col = (q0 + i) * 16 + l16;
          if (i < nvs && col < nloc) {
            gdouble* dst = (gdouble*)Bs + (i64)col * LM + lq;
#pragma unroll
            for (int t = 0; t < NTL; ++t)
#pragma unroll
              for (int r = 0; r < 4; ++r)
                if (16 * t + 4 * r < LM) dst[16 * t + 4 * r] = acc[i][t][r];
          }
        }
      }
    }
    __syncthreads();  // Bs complete (same workgroup: visible after the barrier)

    // ---------------- 2. pivoted QR of B: column lc, rows rbase .. rbase+RPT-1 ------------
    const bool have = lc < nloc;
    double a[RPT];
    {
      const double* src = Bs + (i64)lc * LM + rbase;
      // (row tests as i < ell - rbase: a per-lane bound against compile-time i, instead of
      // RPT per-lane row numbers the compiler would keep live)
      const int el = ell - rbase;
#pragma unroll
      for (int i = 0; i < RPT; i += 2) {
        d2 v = {0.0, 0.0};
        if (have) v = *reinterpret_cast<const d2*>(src + i);
        a[i] = (i < el) ? v.x : 0.0;
        a[i + 1] = (i + 1 < el) ? v.y : 0.0;
      }
    }
    double vn1, vn2;
    {
      double ss = 0.0;
#pragma unroll
      for (int i = 0; i < RPT; ++i) ss += a[i] * a[i];
      vn1 = vn2 = sqrt(group_sum(ss, TPC));
      const double wm = wave_max(have ? vn1 : 0.0);
      if (lane == 0) sMax[wave] = wm;
    }
    __syncthreads();
    double maxnrm = sMax[0];
#pragma unroll
    for (int w = 1; w < BW; ++w) maxnrm = fmax(maxnrm, sMax[w]);
    if constexpr (CL == 2) {
      // both halves need the largest column norm of the whole block (exchange 0, parity 1).
      // Only wave 0 exchanges (one poller per workgroup keeps the uncached traffic of the
      // waits low); the others take the result from LDS after the barrier.
      if (wave == 0) {
        xpublish<LM>(myslot + (LM + 2), xtag(gen, qi, 0), maxnrm, 0, 0, nullptr, p.uncached);
        double pv;
        int pp, pf;
        const bool ok = xpoll<LM>(peerslot + (LM + 2), xtag(gen, qi, 0), p.uncached, p.err, &pv,
                                  &pp, &pf, sPc);
        if (lane == 0) {
          sXv = pv;
          if (!ok) sAbort = 1;
        }
      }
      __syncthreads();
      if (sAbort) return;
      maxnrm = fmax(maxnrm, sXv);
    }
    const double ptol = fmax(p.atol, p.rtol * maxnrm);
    int kcap = min(ell, n);
    if (p.rank >= 0) kcap = min(kcap, p.rank);
    const int nbe = min(p.nb, max(kcap, 1));
    const int lastrk = min(ell, n);
    int pos = c0col + lc;
    bool active = have;
    int myflag = 0;
    int j0 = 0, jb = min(nbe, kcap);  // dlaqps block bookkeeping (wave 0's copy is the one used)
    int k_res = kcap;
    int reject = 0;
    for (int s = 0;; ++s) {
      const int par = s & 1;
      // lane-derived values recomputed each step from an opaque copy of the thread index: kept
      // live across the loop, the addresses built from them were spilled to scratch memory and
      // reloaded (one memory round trip each) on the step's critical path
      int tid_o = tid;
      asm volatile("" : "+v"(tid_o));
      const int lane = tid_o & 63, lic = tid_o % TPC, rbase = lic * RPT;
      // A: this wave's best column publishes its values and its (norm, position)
      {
        double wv;
        const int wl = wave_argmax(active ? vn1 : -1.0, active ? pos : INT_MAX, TPC, &wv);
        const int wf = __any(myflag) ? 1 : 0;
        if (wl >= 0 && (lane & ~(TPC - 1)) == wl) {
#pragma unroll
          for (int i = 0; i < RPT; ++i) sCand[wave][rbase + i] = a[i];
        }
        const int wpos = __builtin_amdgcn_readlane(pos, wl < 0 ? 0 : wl);
        if (lane == 0) {
          sVal[wave] = wv;
          sPos[wave] = wl < 0 ? INT_MAX : wpos;
          sLane[wave] = wl;
          sFlag[wave] = wf;
        }
      }
      myflag = 0;
      __syncthreads();
      if (CL == 2 && sAbort) return;

      // B: the pivot (largest norm, then lowest position; lane w holds wave w's candidate),
      // computed by every wave; pairs swap it with the other half (wave 0 exchanges, the
      // others read the outcome from LDS): the larger norm wins, then the lower position
      int cw, cp, cwl, cflag;
      double cv;
      {
        const bool lw = lane < BW;
        const double ov = lw ? sVal[lane] : -1.0;
        const int op = lw ? sPos[lane] : INT_MAX;
        const int wl = wave_argmax(ov, op, 1, &cv);
        cw = wl < 0 ? 0 : wl;
        cp = wl < 0 ? INT_MAX : __builtin_amdgcn_readlane(op, wl);
        cwl = __builtin_amdgcn_readlane(lw ? sLane[lane] : -1, cw);
        cflag = __any(lw && sFlag[lane] != 0) ? 1 : 0;
      }
      if constexpr (CL == 2) {
        // (sPc, the partner's column, is written and read by wave 0 only)
        if (wave == 0) {
          xpublish<LM>(myslot + par * (LM + 2), xtag(gen, qi, s + 1), cv, cp, cflag,
                       cwl >= 0 ? sCand[cw] : nullptr, p.uncached);
          double pv;
          int pp, pf;
          const bool ok = xpoll<LM>(peerslot + par * (LM + 2), xtag(gen, qi, s + 1),
                                    p.uncached, p.err, &pv, &pp, &pf, sPc);
          if (lane == 0) {
            sXv = pv; sXp = pp; sXf = pf;
            if (!ok) sAbort = 1;
          }
        }
        __syncthreads();
        if (sAbort) return;
        const double pv = sXv;
        const int pp = __builtin_amdgcn_readfirstlane(sXp);
        cflag |= __builtin_amdgcn_readfirstlane(sXf);
        if (pv > cv || (pv == cv && pp < cp)) { cv = pv; cp = pp; cw = -1; cwl = -1; }
      }
      // D: dlaqps block bookkeeping for the step just finished (s-1) and the stop rule
      const bool done_steps = (s == kcap);
      if (s > 0) {
        const int last = s - 1;
        const bool blk_end = (last - j0 + 1 == jb) || (cflag != 0) || done_steps;
        if (blk_end) {
          if (fabs(sDiag[last]) <= ptol) {
            int i = j0;
            for (; i <= last; ++i)
              if (fabs(sDiag[i]) <= ptol) break;
            if (i <= last) { k_res = i; break; }
          }
          if (done_steps) { k_res = kcap; break; }
          if (p.target > 0 && s >= p.target) { k_res = s; reject = 1; break; }
          j0 = s;
          jb = min(nbe, kcap - s);
        }
      }
      if (done_steps) { k_res = s; break; }

      // E (wave 0): the reflector (dlarfg) of the pivot column into sX by absolute row (zero
      // above row rk, 1 at rk) and R11's column rk; the other waves wait at the barrier
      {
        if (wave == 0) {
          const int rk = s;
          const double* xc = cw >= 0 ? sCand[cw] : sPc;
          double xr[NR];
          double ss = 0.0;
#pragma unroll
          for (int h = 0; h < NR; ++h) {
            const int r = lane + 64 * h;
            xr[h] = r < LM ? xc[r] : 0.0;
            ss += (r > rk && r < ell) ? xr[h] * xr[h] : 0.0;
          }
          ss = group_sum(ss, 64);
          const double alpha0 = xc[rk];
          double alpha = alpha0, xnorm = sqrt(ss), beta = alpha0, scal = 1.0, fac = 1.0;
          double tau = 0.0;
          if (ell - rk > 1 && xnorm != 0.0) {
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
              for (int h = 0; h < NR; ++h) {
                const int r = lane + 64 * h;
                const double xs = xr[h] * fac;
                s2 += (r > rk && r < ell) ? xs * xs : 0.0;
              }
              xnorm = sqrt(group_sum(s2, 64));
              beta = -copysign(lapy2(alpha, xnorm), alpha);
            }
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
            for (int j = 0; j < knt; ++j) beta *= safmin;
          }
#pragma unroll
          for (int h = 0; h < NR; ++h) {
            const int r = lane + 64 * h;
            const double vr =
                (r > rk && r < ell) ? (xr[h] * fac) * scal : (r == rk ? 1.0 : 0.0);
            if (r < LM) sX[r] = vr;
            if (rk < KM) {
              if (r < rk) sR[r + rk * KM] = xr[h];
              if (r == rk) sR[rk + rk * KM] = beta;
            }
          }
          if (lane == 0) {
            sDiag[rk] = beta;
            sTau = tau;
          }
        }
      }
      __syncthreads();

      // F: position swap (LAPACK swaps columns rk and pvt), apply H, downdate the norms
      const int rk = s;
      const bool is_win = active && cw >= 0 && wave == cw && (lane & ~(TPC - 1)) == cwl;
      if (is_win) {
        active = false;
        pos = s;
      } else if (have && pos == s) {
        pos = cp;
      }
      // (the TPC lanes of a column share `active`, so the group sums stay within active lanes)
      int flag = 0;
      if (active) {
        const double* v = sX + rbase;
        double w0 = 0.0, w1 = 0.0;
#pragma unroll
        for (int i = 0; i < RPT; i += 2) {
          w0 += v[i] * a[i];
          w1 += v[i + 1] * a[i + 1];
        }
        const double tw = sTau * group_sum(w0 + w1, TPC);
        // re-read v from LDS below rather than keeping its values live beside the column
        asm volatile("" ::: "memory");
        double xs = 0.0;
        const int rl = rk - rbase;  // row rk within this lane's part
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          a[i] -= tw * v[i];
          xs = (i == rl) ? a[i] : xs;
        }
        xs = group_sum(xs, TPC);  // new R(rk, j), from the lane that holds row rk
        if (rk + 1 < lastrk && vn1 != 0.0) {
          double temp = fabs(xs) / vn1;
          temp = (1.0 + temp) * (1.0 - temp);
          temp = temp < 0.0 ? 0.0 : temp;
          const double q = vn1 / vn2;
          const double temp2 = temp * q * q;
          if (temp2 <= tol3z) flag = 1;
          else vn1 *= sqrt(temp);
        }
        if (flag) {
          // the block ends here; the flagged norm is recomputed from the updated rows
          double s3 = 0.0;
#pragma unroll
          for (int i = 0; i < RPT; ++i) s3 += (i > rl) ? a[i] * a[i] : 0.0;
          vn1 = vn2 = sqrt(group_sum(s3, TPC));
          myflag = 1;
        }
      }
    }
    if (p.target > 0 && k_res >= p.target) reject = 1;
    __syncthreads();  // R11 complete

    // ---------------- 3. sk, rd, T = R11 \ R12 ----------------------------------------
    const int k = k_res;
    if (reject) {
      if (tid == 0 && part == 0) {
        p.k_out[b] = -1;
        for (int kk = 0; kk < NKIND; ++kk) {
          if (p.next_list[kk] && n <= p.next_maxn[kk]) {
            const unsigned slot = __hip_atomic_fetch_add(p.next_n[kk], 1u, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
            p.next_list[kk][slot] = b;
            break;
          }
        }
      }
      continue;
    }
    if (tid == 0 && part == 0) p.k_out[b] = k;
    const bool redund = have && pos >= k;
    if (have && lic == 0 && pos < k) d.sk[pos] = c0col + lc + 1;
    if (!redund) continue;
    if (lic == 0) d.rd[pos - k] = c0col + lc + 1;
    // back substitution; row i lives in part i / RPT of the column, at index i % RPT, so
    // every index below is a compile-time constant within its part's branch
#pragma unroll
    for (int i = KM - 1; i >= 0; --i) {
      if (i < k) {
        const int owner = i / RPT;
        double ti = lic == owner ? a[i % RPT] / sR[i + i * KM] : 0.0;
        ti = group_sum(ti, TPC);
        if (lic == owner) a[i % RPT] = ti;
#pragma unroll
        for (int part = 0; part < TPC; ++part) {
          const int lim = min(RPT, i - part * RPT);  // rows of this part above row i
          if (lic == part) {
#pragma unroll
            for (int ii = 0; ii < lim; ++ii) a[ii] -= sR[part * RPT + ii + i * KM] * ti;
          }
        }
        asm volatile("" ::: "memory");  // one column of R11 in flight at a time
      }
    }
    double* tcol = d.T + (i64)(pos - k) * d.ldt + rbase;
#pragma unroll
    for (int ii = 0; ii < RPT; ++ii)
      if (rbase + ii < k) tcol[ii] = a[ii];
  }
}

// the kernel variants: per order, up to three kernels by block width.  Order 40 (one thread
// per column): 512-thread workgroups, two per CU, for n <= 512 and one 1024-thread workgroup
// for 512 < n <= 1024.  Order 72 (two threads per column): 512-thread workgroups for
// n <= 256, 1024-thread ones for n <= 512 and workgroup pairs for 512 < n <= 1024.
struct Variant {
  int lmax;
  int maxn[NKIND];  // widest block of each kernel (0: none)
};
constexpr Variant kVariants[] = {{40, {512, 1024, 0}}, {72, {256, 512, 1024}}};
// kernel kind: 0 = 512-thread workgroups (two per CU), 1 = one 1024-thread workgroup,
// 2 = workgroup pair of 1024-thread workgroups
constexpr int kPerCu[NKIND] = {2, 1, 0};

int variant_of(i64 ell) {
  for (int v = 0; v < 2; ++v)
    if (ell <= kVariants[v].lmax) return v;
  return -1;
}

hipError_t launch_variant(int v, int kind, const BStage& s, int grid, hipStream_t st) {
  if (v == 0 && kind == 0)
    hipLaunchKernelGGL((id_batched_kernel<40, 1, 1, 512>), dim3(grid), dim3(512), 0, st, s);
  else if (v == 0)
    hipLaunchKernelGGL((id_batched_kernel<40, 1, 1, 1024>), dim3(grid), dim3(1024), 0, st, s);
  else if (kind == 0)
    hipLaunchKernelGGL((id_batched_kernel<72, 2, 1, 512>), dim3(grid), dim3(512), 0, st, s);
  else if (kind == 1)
    hipLaunchKernelGGL((id_batched_kernel<72, 2, 1, 1024>), dim3(grid), dim3(1024), 0, st, s);
  else
    hipLaunchKernelGGL((id_batched_kernel<72, 2, 2, 1024>), dim3(grid), dim3(1024), 0, st, s);
  return hipGetLastError();
}

}  // namespace

// ---------------------------------------------------------------------------------------
// host side

struct lra_batch_plan {
  lra_ctx* c = nullptr;
  lra_opts o{};
  i64 nblk = 0;
  std::vector<i64> m, n, lda, ldt;
  std::vector<const double*> A;
  std::vector<i64*> sk, rd;
  std::vector<double*> T;
  std::vector<int> direct;  // blocks that go to lra_idfact without the batched kernels
  int nq[NKIND] = {};       // stage-0 blocks of each kernel
  int w0[NSTAGE] = {};      // per stage: widest block of the 512-thread kernel (0: none)
  std::vector<int> queued;  // blocks the kernels take (stage 0), in block order
  int grid = 0, npairs = 0, nstage = 0;  // grid: workgroups the scratch slabs cover
  int ell[NSTAGE] = {}, target[NSTAGE] = {}, variant[NSTAGE] = {};
  BDesc* d_desc = nullptr;
  int* d_lists = nullptr;        // [stage][kind] lists of nblk entries
  unsigned* d_counts = nullptr;  // [0, NK NSTAGE): queue heads, [NK NSTAGE, 2 NK NSTAGE): list
                                 // sizes, [2 NK NSTAGE]: launch generation
  double* d_omega[NSTAGE] = {};
  i64 ldo = 0;
  double* d_scratch = nullptr;
  i64* d_k = nullptr;
  u64x2* xchg = nullptr;
  bool uncached = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

static void plan_free(lra_batch_plan* P) {
  if (!P) return;
  if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  void* bufs[] = {P->d_desc, P->d_lists, P->d_counts, P->d_scratch, P->d_k};
  for (void* q : bufs)
    if (q) (void)hipFree(q);
  for (double* q : P->d_omega)
    if (q) (void)hipFree(q);
  delete P;
}

// widest block of stage r's kernel `kind` (0: the kernel is not launched)
static int maxn_of(const lra_batch_plan* P, int r, int kind) {
  return kind == 0 ? P->w0[r] : kVariants[P->variant[r]].maxn[kind];
}

// Stage 0's work lists: each queued block goes to the first kernel whose width limit holds
// it, each list by decreasing m*n (the large blocks start first); uploaded to the device and
// finished (the host vectors are temporaries).
static int upload_stage0(lra_batch_plan* P) {
  std::vector<int> q[NKIND];
  for (int b : P->queued) {
    int kind = 0;
    while (kind < NKIND && (maxn_of(P, 0, kind) == 0 || P->n[b] > maxn_of(P, 0, kind))) ++kind;
    if (kind < NKIND) q[kind].push_back(b);  // (queued blocks fit stage 0's widest kernel)
  }
  hipError_t e = hipSuccess;
  for (int kind = 0; kind < NKIND; ++kind) {
    std::stable_sort(q[kind].begin(), q[kind].end(),
                     [&](int x, int y) { return P->m[x] * P->n[x] > P->m[y] * P->n[y]; });
    P->nq[kind] = (int)q[kind].size();
    if (P->nq[kind] > 0 && e == hipSuccess)
      e = hipMemcpyAsync(P->d_lists + (size_t)kind * P->nblk, q[kind].data(),
                         sizeof(int) * P->nq[kind], hipMemcpyHostToDevice, P->c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(P->c->stream);
  if (e != hipSuccess) {
    lra_set_error("batched plan: %s", hipGetErrorString(e));
    return LRA_E_HIP;
  }
  return LRA_OK;
}

static int enqueue(lra_batch_plan* P) {
  lra_ctx* c = P->c;
  if (P->nq[0] + P->nq[1] + P->nq[2] == 0) return LRA_OK;
  constexpr int NC = NKIND * NSTAGE;  // queue heads, then list sizes
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, c->stream, P->d_counts, 2 * NC,
                     P->d_counts + 2 * NC);
  LRA_HIP(hipGetLastError());
  auto list = [&](int r, int kind) {
    return P->d_lists + (size_t)(NKIND * r + kind) * P->nblk;
  };
  for (int r = 0; r < P->nstage; ++r) {
    for (int kind = 0; kind < NKIND; ++kind) {
      if (maxn_of(P, r, kind) == 0) continue;
      if (r == 0 && P->nq[kind] == 0) continue;
      BStage s;
      s.desc = P->d_desc;
      s.list = list(r, kind);
      s.list_n = r == 0 ? nullptr : P->d_counts + NC + NKIND * r + kind;
      s.nq = r == 0 ? P->nq[kind] : (int)P->nblk;
      s.head = P->d_counts + NKIND * r + kind;
      s.omega = P->d_omega[r];
      s.ldo = P->ldo;
      s.scratch = P->d_scratch;
      s.k_out = P->d_k;
      const bool more = r + 1 < P->nstage;
      for (int kk = 0; kk < NKIND; ++kk) {
        const bool has = more && maxn_of(P, r + 1, kk) > 0;
        s.next_list[kk] = has ? list(r + 1, kk) : nullptr;
        s.next_n[kk] = has ? P->d_counts + NC + NKIND * (r + 1) + kk : nullptr;
        s.next_maxn[kk] = has ? maxn_of(P, r + 1, kk) : 0;
      }
      s.ell = P->ell[r];
      s.rank = (int)std::min<i64>(P->o.rank, INT_MAX);
      s.nb = (int)P->o.nb;
      s.target = P->target[r];
      s.atol = P->o.atol;
      s.rtol = P->o.rtol;
      s.xchg = P->xchg;
      s.uncached = P->uncached ? 1 : 0;
      s.gen = P->d_counts + 2 * NC;
      s.err = c->d_err;
      const int work = r == 0 ? P->nq[kind] : (int)P->nblk;
      const bool pair_kernel = kind == 2;
      const int grid = pair_kernel ? 2 * P->npairs
                                   : std::max(1, std::min(work, kPerCu[kind] * c->num_cu));
      hipError_t e;
      if (pair_kernel) {
        SpinSection gate(c);
        e = launch_variant(P->variant[r], kind, s, grid, c->stream);
      } else {
        e = launch_variant(P->variant[r], kind, s, grid, c->stream);
      }
      if (e != hipSuccess) {
        lra_set_error("batched launch: %s", hipGetErrorString(e));
        return LRA_E_HIP;
      }
    }
  }
  return LRA_OK;
}

// The batch's autotuner: which blocks of each round run in the 512-thread kernel (two
// workgroups per CU) and which in the 1024-thread one is a width limit per round.  The plan
// times the candidate limits on its own blocks -- the compiled limit, half of it, none, per
// round -- with one warm launch and the best of two timed launches each, and keeps the
// fastest.  Every block's result is the same under any limit (each block is factored by one
// workgroup or pair exactly as idfact would), so the choice changes time only.
// LRA_BATCHED_W0=w0,w1 forces the limits of rounds 0 and 1; LRA_AUTOTUNE=0, or fewer than 256
// blocks, keeps the compiled ones.
static int tune_widths(lra_batch_plan* P) {
  const int lim0 = kVariants[P->variant[0]].maxn[0];
  const int lim1 = P->nstage > 1 ? kVariants[P->variant[1]].maxn[0] : 0;
  if (const char* f = getenv("LRA_BATCHED_W0")) {
    int a = -1, b = -1;
    const int got = sscanf(f, "%d,%d", &a, &b);
    if (got >= 1) P->w0[0] = std::max(0, std::min(a, lim0));
    if (got >= 2 && P->nstage > 1) P->w0[1] = std::max(0, std::min(b, lim1));
    return upload_stage0(P);
  }
  const char* at = getenv("LRA_AUTOTUNE");
  if ((at && at[0] == '0') || P->queued.size() < 256) return LRA_OK;
  lra_ctx* c = P->c;
  std::vector<std::pair<int, int>> cands;
  for (int a : {lim0, lim0 / 2, 0})
    for (int b : {lim1, lim1 / 2, 0}) {
      cands.push_back({a, b});
      if (P->nstage < 2) break;
    }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  LRA_HIP(hipEventCreate(&e0));
  LRA_HIP(hipEventCreate(&e1));
  int st = LRA_OK;
  float best = 0.0f;
  std::pair<int, int> pick = cands[0];
  for (size_t i = 0; i < cands.size() && st == LRA_OK; ++i) {
    P->w0[0] = cands[i].first;
    if (P->nstage > 1) P->w0[1] = cands[i].second;
    st = upload_stage0(P);
    if (st == LRA_OK && i == 0) st = enqueue(P);  // warm: first launch of each kernel
    float t = 0.0f;
    for (int rep = 0; rep < 2 && st == LRA_OK; ++rep) {
      float ms = 0.0f;
      if (hipEventRecord(e0, c->stream) != hipSuccess) st = LRA_E_HIP;
      if (st == LRA_OK) st = enqueue(P);
      if (st == LRA_OK && (hipEventRecord(e1, c->stream) != hipSuccess ||
                           hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) {
        lra_set_error("batched autotune: event timing failed");
        st = LRA_E_HIP;
      }
      if (rep == 0 || ms < t) t = ms;
    }
    if (st == LRA_OK && (i == 0 || t < best)) { best = t; pick = cands[i]; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st == LRA_OK && P->xchg) st = lra_check_device_error(c);  // a pair's wait expired
  if (st != LRA_OK) return st;
  P->w0[0] = pick.first;
  if (P->nstage > 1) P->w0[1] = pick.second;
  return upload_stage0(P);
}

int lra_chkopts_public(const lra_opts* o);

extern "C" int lra_idfact_batched_plan(lra_ctx* c, int64_t nblk, const int64_t* m,
                                       const int64_t* n, const double* const* A,
                                       const int64_t* lda, const lra_opts* opts,
                                       int64_t* const* sk, int64_t* const* rd,
                                       double* const* T, const int64_t* ldt,
                                       lra_batch_plan** plan) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(nblk >= 0 && nblk <= INT_MAX, 2);
  LRA_ARG(m != nullptr || nblk == 0, 3);
  LRA_ARG(n != nullptr || nblk == 0, 4);
  LRA_ARG(A != nullptr || nblk == 0, 5);
  LRA_ARG(lda != nullptr || nblk == 0, 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(sk != nullptr || nblk == 0, 8);
  LRA_ARG(rd != nullptr || nblk == 0, 9);
  LRA_ARG(T != nullptr || nblk == 0, 10);
  LRA_ARG(ldt != nullptr || nblk == 0, 11);
  LRA_ARG(plan != nullptr, 12);
  *plan = nullptr;
  if (lra_chkopts_public(opts) != LRA_OK) return -7;
  for (i64 b = 0; b < nblk; ++b) {
    if (m[b] < 0 || n[b] < 0 || lda[b] < std::max<i64>(1, m[b]) || ldt[b] < 1 ||
        (A[b] == nullptr && m[b] * n[b] != 0)) {
      lra_set_error("batched: block %lld: bad shape (m %lld, n %lld, lda %lld, ldt %lld)",
                    (long long)b, (long long)m[b], (long long)n[b], (long long)lda[b],
                    (long long)ldt[b]);
      return -3;
    }
  }
  LRA_HIP(hipSetDevice(c->device));
  lra_batch_plan* P = new lra_batch_plan();
  P->c = c;
  P->o = *opts;
  P->nblk = nblk;
  P->m.assign(m, m + nblk); P->n.assign(n, n + nblk); P->lda.assign(lda, lda + nblk);
  P->ldt.assign(ldt, ldt + nblk); P->A.assign(A, A + nblk);
  P->sk.assign(sk, sk + nblk); P->rd.assign(rd, rd + nblk); P->T.assign(T, T + nblk);
  const lra_opts& o = P->o;
  // What the kernels implement: the Gaussian sketch without power iterations or maxdet
  // swaps, rounds of order <= 72.  Low-precision sketches are not used: A is read once in
  // FP64 either way, and below rtol 1e-6 idfact recomputes a low-precision round in FP64.
  const bool kernel_ok = o.sketch == LRA_SKETCH_RANDN && o.sketch_randn_niter == 0 &&
                         o.maxdet_tol < 0 && (o.precision == LRA_PREC_F64 || o.rtol < 1e-6);
  const bool adap = o.sketchfact_adap || o.rank < 0;
  i64 tg = adap ? o.nb : o.rank;
  for (int r = 0; r < NSTAGE && kernel_ok; ++r) {
    const i64 ell = o.randn_samp_mul * tg + o.randn_samp_add;
    const int v = variant_of(ell);
    if (ell < 1 || v < 0) break;
    // the ranks a stage returns (below its target, or the fixed rank) must fit R11
    const i64 kbound = adap ? tg : std::min<i64>(o.rank, ell);
    if (kbound > std::min(kVariants[v].lmax, 64)) break;
    P->ell[r] = (int)ell;
    P->target[r] = adap ? (int)tg : 0;
    P->variant[r] = v;
    P->nstage = r + 1;
    if (!adap) break;
    tg *= 2;
  }
  for (int r = 0; r < P->nstage; ++r) P->w0[r] = kVariants[P->variant[r]].maxn[0];
  i64 mmax = 0;
  int nwide = 0;  // stage-0 blocks of the workgroup-pair kernel
  const Variant& V0 = kVariants[P->variant[0]];
  for (i64 b = 0; b < nblk; ++b) {
    // stage 0 needs the block's width within its variant, and T rows for its largest rank
    int kind = 0;
    while (kind < NKIND && (V0.maxn[kind] == 0 || n[b] > V0.maxn[kind])) ++kind;
    const bool ok = P->nstage > 0 && n[b] >= 1 && kind < NKIND && m[b] <= INT_MAX &&
                    ldt[b] >= std::max<i64>(1, std::min<i64>(P->ell[0], n[b]));
    if (ok) {
      P->queued.push_back((int)b);
      nwide += kind == 2;
      mmax = std::max(mmax, m[b]);
    } else {
      P->direct.push_back((int)b);
    }
  }
  // a later stage writes T for ranks up to its own order: stages beyond what every queued
  // block's ldt allows are dropped (forwarding is decided on the device from n alone)
  for (int r = 1; r < P->nstage; ++r) {
    bool all = true;
    for (int b : P->queued) all = all && ldt[b] >= std::min<i64>(P->ell[r], n[b]);
    if (!all) { P->nstage = r; break; }
  }
  bool pairs = V0.maxn[2] > 0 && nwide > 0;
  for (int r = 1; r < P->nstage; ++r) pairs = pairs || kVariants[P->variant[r]].maxn[2] > 0;
  const int nqueued = (int)P->queued.size();
  // scratch slabs: the most workgroups any launch has (two per CU for the 512-thread kernel)
  P->grid = std::max(1, std::min(std::max(nqueued, P->nstage > 1 ? (int)nblk : 0),
                                 2 * c->num_cu));
  P->npairs = std::max(1, c->num_cu / 2);
  int st = LRA_OK;
  if (pairs) {
    void* xp = nullptr;
    st = lra_coop_area(c, sizeof(u64x2) * 4 * 74 * 2 * (size_t)P->npairs, &xp, &P->uncached);
    P->xchg = (u64x2*)xp;
  }
  if (st == LRA_OK && nqueued > 0) {
    std::vector<BDesc> hd(nblk);
    for (i64 b = 0; b < nblk; ++b) {
      hd[b].A = A[b]; hd[b].lda = lda[b]; hd[b].sk = sk[b]; hd[b].rd = rd[b]; hd[b].T = T[b];
      hd[b].ldt = ldt[b]; hd[b].m = (int)std::min<i64>(m[b], INT_MAX);
      hd[b].n = (int)std::min<i64>(n[b], INT_MAX);
    }
    P->ldo = std::max<i64>(2, (mmax + 1) & ~(i64)1);
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc((void**)&P->d_desc, sizeof(BDesc) * nblk);
    if (e == hipSuccess)
      e = hipMalloc((void**)&P->d_lists, sizeof(int) * nblk * NKIND * NSTAGE);
    if (e == hipSuccess)
      e = hipMalloc((void**)&P->d_counts, sizeof(unsigned) * (2 * NKIND * NSTAGE + 1));
    if (e == hipSuccess)
      e = hipMemsetAsync(P->d_counts, 0, sizeof(unsigned) * (2 * NKIND * NSTAGE + 1), c->stream);
    for (int r = 0; r < P->nstage && e == hipSuccess; ++r)
      e = hipMalloc((void**)&P->d_omega[r], sizeof(double) * P->ell[r] * P->ldo);
    if (e == hipSuccess)
      e = hipMalloc((void**)&P->d_scratch,
                    sizeof(double) * SLAB * (size_t)std::max(P->grid, 2 * P->npairs));
    if (e == hipSuccess) e = hipMalloc((void**)&P->d_k, sizeof(i64) * std::max<i64>(nblk, 1));
    if (e == hipSuccess)
      e = hipMemcpyAsync(P->d_desc, hd.data(), sizeof(BDesc) * nblk, hipMemcpyHostToDevice,
                         c->stream);
    if (e != hipSuccess) {
      lra_set_error("batched plan: %s", hipGetErrorString(e));
      st = LRA_E_HIP;
    }
    // Omega of round r (Philox stream r), the first m columns of which are every block's
    // round-r sketch matrix; stored k-contiguous (ell_r rows of ldo)
    for (int r = 0; r < P->nstage && st == LRA_OK; ++r)
      st = lra_launch_gaussian_fill_t(c->stream, o.seed, (uint64_t)r, P->ell[r], mmax,
                                      P->d_omega[r], P->ldo);
    // the host copy above reads `hd`: finish it before that vector goes away
    if (st == LRA_OK && hipStreamSynchronize(c->stream) != hipSuccess) {
      lra_set_error("batched plan: stream synchronize failed");
      st = LRA_E_HIP;
    }
    if (st == LRA_OK) st = upload_stage0(P);
    if (st == LRA_OK) st = tune_widths(P);
  }
  if (st != LRA_OK) {
    plan_free(P);
    return st;
  }
  *plan = P;
  return LRA_OK;
}

// Enqueue the batched kernels on the context's stream (a graph launch once captured).
extern "C" int lra_idfact_batched_plan_widths(const lra_batch_plan* P, int* widths, int count) {
  LRA_ARG(P != nullptr, 1);
  LRA_ARG(widths != nullptr || count <= 0, 2);
  for (int r = 0; r < count; ++r) widths[r] = r < P->nstage ? P->w0[r] : -1;
  return LRA_OK;
}

extern "C" int lra_idfact_batched_launch(lra_batch_plan* P) {
  LRA_ARG(P != nullptr, 1);
  LRA_HIP(hipSetDevice(P->c->device));
  if (P->gexec) {
    LRA_HIP(hipGraphLaunch(P->gexec, P->c->stream));
    return LRA_OK;
  }
  return enqueue(P);
}

// Capture the launch in a hipGraph (later launches replay it).
extern "C" int lra_idfact_batched_capture(lra_batch_plan* P) {
  LRA_ARG(P != nullptr, 1);
  lra_ctx* c = P->c;
  LRA_HIP(hipSetDevice(c->device));
  if (P->gexec) return LRA_OK;
  LRA_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int st = enqueue(P);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (st != LRA_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) {
    lra_set_error("batched capture: %s", hipGetErrorString(e));
    return LRA_E_HIP;
  }
  P->graph = g;
  LRA_HIP(hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0));
  return LRA_OK;
}

// Wait for the launch, then finish the blocks the kernels left (k = -1) or did not take
// through lra_idfact.  k_out (host, nblk) receives every block's rank.  A block finished there
// whose rank exceeds its ldt keeps sk / rd but not T; the call then reports LRA_E_UNSUPPORTED
// after finishing every other block (k_out tells which).
extern "C" int lra_idfact_batched_finish(lra_batch_plan* P, int64_t* k_out,
                                         int64_t* n_fallback) {
  LRA_ARG(P != nullptr, 1);
  lra_ctx* c = P->c;
  LRA_HIP(hipSetDevice(c->device));
  std::vector<i64> hk(P->nblk, -1);
  if (P->nq[0] + P->nq[1] + P->nq[2] > 0) {
    LRA_HIP(hipMemcpyAsync(hk.data(), P->d_k, sizeof(i64) * P->nblk, hipMemcpyDeviceToHost,
                           c->stream));
  }
  LRA_HIP(hipStreamSynchronize(c->stream));
  if (P->xchg) LRA_CHECK(lra_check_device_error(c));  // a pair's bounded wait expired
  for (int b : P->direct) hk[b] = -1;
  i64 nf = 0, over = -1;
  for (i64 b = 0; b < P->nblk; ++b) {
    if (hk[b] >= 0) continue;
    ++nf;
    lra_id_out out{};
    out.sk = P->sk[b]; out.rd = P->rd[b]; out.T = nullptr; out.ldt = 0;  // library-owned T
    const int st = lra_idfact(c, LRA_TRANS_N, P->m[b], P->n[b], P->A[b], P->lda[b], &P->o, &out);
    if (st != LRA_OK) return st;
    const i64 k = out.k, nr = P->n[b] - k;
    hk[b] = k;
    if (k > P->ldt[b]) {
      if (over < 0) over = b;
      continue;
    }
    if (k > 0 && nr > 0)
      LRA_HIP(hipMemcpy2DAsync(P->T[b], sizeof(double) * P->ldt[b], out.T,
                               sizeof(double) * out.ldt, sizeof(double) * k, nr,
                               hipMemcpyDeviceToDevice, c->stream));
    // the next idfact reuses the library-owned T
    LRA_HIP(hipStreamSynchronize(c->stream));
  }
  if (k_out) std::copy(hk.begin(), hk.end(), k_out);
  if (n_fallback) *n_fallback = nf;
  if (over >= 0) {
    lra_set_error("batched: block %lld has rank %lld > ldt %lld (T not written)",
                  (long long)over, (long long)hk[over], (long long)P->ldt[over]);
    return LRA_E_UNSUPPORTED;
  }
  return LRA_OK;
}

extern "C" int lra_idfact_batched_plan_destroy(lra_batch_plan* P) {
  if (!P) return LRA_OK;
  (void)hipSetDevice(P->c->device);
  (void)hipStreamSynchronize(P->c->stream);
  plan_free(P);
  return LRA_OK;
}

extern "C" int lra_idfact_batched(lra_ctx* c, int64_t nblk, const int64_t* m, const int64_t* n,
                                  const double* const* A, const int64_t* lda,
                                  const lra_opts* opts, int64_t* const* sk, int64_t* const* rd,
                                  double* const* T, const int64_t* ldt, int64_t* k_out,
                                  int64_t* n_fallback) {
  lra_batch_plan* P = nullptr;
  LRA_CHECK(lra_idfact_batched_plan(c, nblk, m, n, A, lda, opts, sk, rd, T, ldt, &P));
  int st = lra_idfact_batched_launch(P);
  if (st == LRA_OK) st = lra_idfact_batched_finish(P, k_out, n_fallback);
  lra_idfact_batched_plan_destroy(P);
  return st;
}
