// lra_internal.h -- shared declarations of liblra_hip.so (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/lra.h"

#include <atomic>

typedef int64_t i64;

// ---------------------------------------------------------------------------------------
// host-side error plumbing

void lra_set_error(const char* fmt, ...);
#define LRA_HIP(expr)                                                                     \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess) {                                                               \
      lra_set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(_e));   \
      return LRA_E_HIP;                                                                   \
    }                                                                                     \
  } while (0)
#define LRA_CHECK(expr)               \
  do {                                \
    int _s = (expr);                  \
    if (_s != LRA_OK) return _s;      \
  } while (0)
#define LRA_ARG(cond, idx)                                               \
  do {                                                                   \
    if (!(cond)) {                                                       \
      lra_set_error("argument %d invalid: %s", (int)(idx), #cond);       \
      return -(int)(idx);                                                \
    }                                                                    \
  } while (0)

// ---------------------------------------------------------------------------------------
// context: one stream, a caching workspace, the cooperative-kernel exchange buffers

struct lra_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  size_t hbm_budget = 0;
  int num_cu = 0;
  int lds_bytes = 0;
  // workspace slots: grown on demand, never shrunk until destroy
  static const int kSlots = 80;
  void* ws[kSlots] = {};
  size_t ws_size[kSlots] = {};
  // small pinned host staging area for scalars read back synchronously
  void* host_pinned = nullptr;
  // per-context cooperative exchange area (uncached device memory when available: allocated
  // once, never freed before the process ends)
  void* coop = nullptr;       // ordinary memory, grown on demand
  size_t coop_size = 0;
  void* coop_uc = nullptr;    // uncached, fixed size, never freed (pooled per device)
  size_t coop_uc_size = 0;
  bool coop_uc_tried = false;
  int* d_err = nullptr;  // device error word (timeout etc.)
  hipEvent_t spin_ev = nullptr;  // recorded after this context's grid-synchronising kernels
};

// Kernels whose workgroups wait for each other (the cooperative QRCP, the geqrf panel, the
// batched workgroup pairs) need all their workgroups resident at once.  Two of them from
// different contexts running together could each hold part of the device while waiting for
// their remaining workgroups, so on one device they run one at a time, in launch order: a
// SpinSection around such a launch makes the context's stream wait for the previous one.
// Other kernels overlap freely.  (Skipped while the stream is being captured.)
struct SpinSection {
  explicit SpinSection(lra_ctx* c);
  ~SpinSection();
  lra_ctx* c;
  bool held = false;
};
void lra_spin_forget(lra_ctx* c);  // at context destruction

// The matrix a factorization works on: its entries in device memory (A != nullptr), or a
// matrix-free operator reached through the caller's apply callback (LinearOperator,
// src/linop.jl:9-40).  herm: the operator is Hermitian (HermitianLinearOperator); for a dense
// source the drivers test exact symmetry themselves where the reference calls ishermitian.
struct LraSrc {
  i64 m = 0, n = 0;
  const double* A = nullptr;
  i64 lda = 1;
  lra_apply_fn apply = nullptr;
  void* user = nullptr;
  bool herm = false;
  bool op() const { return apply != nullptr; }
};
inline LraSrc lra_src_dense(i64 m, i64 n, const double* A, i64 lda) {
  LraSrc s;
  s.m = m; s.n = n; s.A = A; s.lda = lda;
  return s;
}
inline LraSrc lra_src_op(i64 m, i64 n, lra_apply_fn apply, void* user, bool herm) {
  LraSrc s;
  s.m = m; s.n = n; s.apply = apply; s.user = user; s.herm = herm;
  return s;
}
// C (P x k) = op(A)[:, idx] for op(A) = A (trans n) or A' (trans c): a gather for a dense
// source, one application to k one-hot vectors for an operator (getcols, src/util.jl:25-34)
int lra_src_getcols(lra_ctx* c, const LraSrc& s, int trans, const i64* idx, i64 k, double* C,
                    i64 ldc);
// D (k x k) = A[idx, idx] (getindex(L, rows, cols), src/linop.jl:70-74)
int lra_src_principal(lra_ctx* c, const LraSrc& s, const i64* idx, i64 k, double* D, i64 ldd);
// idfact(trans, A, opts) on either kind of source (src/id.jl:434-447)
int lra_idfact_src(lra_ctx* c, int trans, const LraSrc& s, const lra_opts* opts,
                   lra_id_out* out);
// the reference's chkopts!(opts, A) for an operator (src/LowRankApprox.jl:142-148): a sketch
// other than randn is replaced by randn with a warning (verb != 0)
void lra_opts_for_operator(lra_opts* o);

// named workspace slots (each call path uses fixed slots so they can overlap safely)
enum WsSlot {
  WS_OMEGA = 0, WS_SKETCH, WS_SKETCH2, WS_SPLITK, WS_JPVT, WS_TAU, WS_SCAL, WS_R,
  WS_T, WS_TMP0, WS_TMP1, WS_TMP2, WS_TMP3, WS_LP0, WS_LP1, WS_LP2, WS_BATCH0, WS_BATCH1,
  WS_BATCH2, WS_QR0, WS_QR1, WS_QR2, WS_X0, WS_MD0, WS_MD1, WS_OP0, WS_OP1, WS_CUR0,
  WS_F0, WS_F1, WS_F2, WS_F3, WS_F4, WS_F5, WS_F6, WS_F7, WS_F8, WS_F9, WS_J0, WS_J1,
  WS_PI0, WS_PI1, WS_PI2, WS_QG, WS_PR0, WS_PR1, WS_PR2, WS_PR3, WS_PR4, WS_PR5, WS_PR6,
  WS_PR7, WS_PR8, WS_PR9, WS_PR10, WS_PR11, WS_PR12, WS_CU0, WS_CUR1, WS_OP2, WS_OP3,
  WS_SKETCH3, WS_JPVT2, WS_TAU2, WS_SKETCHP, WS_QGT, WS_ZT, WS_PRT, WS_COUNT
};
static_assert(WS_COUNT <= lra_ctx::kSlots, "workspace slots");

// Raise the dynamic-LDS limit of kernel `fn` to `bytes` once per device.  `done` holds one
// bit per device; a bit is set only after the attribute call returned, so a caller that sees
// it set may launch, and concurrent first callers at worst set the attribute twice.
inline void lra_lds_limit_once(std::atomic<uint64_t>& done, const void* fn, int bytes) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  const uint64_t bit = 1ull << (dev & 63);
  if (done.load(std::memory_order_acquire) & bit) return;
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipGetLastError();
  done.fetch_or(bit, std::memory_order_release);
}
int lra_chkopts_public(const lra_opts* o);  // chkopts! (0 or -(bad field))
int lra_ws(lra_ctx* ctx, int slot, size_t bytes, void** out);
template <class T>
static inline int lra_wsT(lra_ctx* ctx, int slot, size_t count, T** out) {
  return lra_ws(ctx, slot, count * sizeof(T), (void**)out);
}
int lra_coop_area(lra_ctx* ctx, size_t bytes, void** out, bool* uncached);
// Next launch generation of the tagged exchanges (process-wide, never 0).
uint32_t lra_next_qrcp_gen();
int lra_check_device_error(lra_ctx* ctx);

// ---------------------------------------------------------------------------------------
// device helpers

#ifdef __HIP_DEVICE_COMPILE__
#define LRA_DEV __device__ __forceinline__
#else
#define LRA_DEV __host__ __device__ inline
#endif

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): counter-based generator, 4x32-bit counter, 2x32-bit key, 10 rounds.
struct philox4 {
  uint32_t x[4];
};
__host__ __device__ inline philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                 uint32_t c3, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c0;
    const uint64_t p1 = (uint64_t)M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0; k1 += W1;
  }
  philox4 r;
  r.x[0] = c0; r.x[1] = c1; r.x[2] = c2; r.x[3] = c3;
  return r;
}

// Two standard normals #(2*pair) and #(2*pair+1) of stream `stream` under `seed`
// (Box-Muller on two 53-bit uniforms).
__device__ inline void lra_normal2(uint64_t seed, uint64_t stream, uint64_t pair, double* z0,
                                   double* z1) {
  philox4 r = philox4x32_10((uint32_t)pair, (uint32_t)(pair >> 32), (uint32_t)stream,
                            (uint32_t)(stream >> 32), (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t a = ((uint64_t)r.x[0] << 32) | r.x[1];
  const uint64_t b = ((uint64_t)r.x[2] << 32) | r.x[3];
  const double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
  const double u2 = (double)(b >> 11) * 0x1.0p-53;          // [0, 1)
  const double rad = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  *z0 = rad * c;
  *z1 = rad * s;
}
__device__ inline double lra_normal1(uint64_t seed, uint64_t stream, uint64_t idx) {
  double z0, z1;
  lra_normal2(seed, stream, idx >> 1, &z0, &z1);
  return (idx & 1) ? z1 : z0;
}
// 64 random bits #idx of stream `stream` (for signs, sample indices, permutations)
__device__ inline uint64_t lra_bits64(uint64_t seed, uint64_t stream, uint64_t idx) {
  philox4 r = philox4x32_10((uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)stream,
                            (uint32_t)(stream >> 32) ^ 0x5bd1e995u, (uint32_t)seed,
                            (uint32_t)(seed >> 32));
  return ((uint64_t)r.x[0] << 32) | r.x[1];
}

// wave-level sum over lanes that share `width` (power of two <= 64) consecutive lanes
__device__ __forceinline__ double lra_group_sum(double v, int width) {
  for (int d = 1; d < width; d <<= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ double lra_wave_sum(double v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// ---------------------------------------------------------------------------------------
// bounded grid barrier for co-resident (cooperatively launched) grids.
// counter increases monotonically; the caller passes target = (phase+1) * nblocks.
// Every wait is bounded: on expiry the device error word is set and the abort flag raised,
// so every block leaves its loop and the grid drains.

#define LRA_SPIN_LIMIT (1u << 26)

__device__ __forceinline__ bool lra_grid_barrier(unsigned* counter, unsigned target,
                                                 int* err, int* abort_flag) {
  __shared__ int s_ok;
  // make this wave's global writes visible device-wide (L2 write-back), then arrive
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    // relaxed polling: no cache invalidation per poll
    while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = 0;
        break;
      }
      if (++spins > LRA_SPIN_LIMIT) {
        __hip_atomic_store(err, (int)LRA_E_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    // one acquire for the block: invalidates this CU's L1 and the XCD's L2 lines so the
    // block's subsequent loads see the other blocks' data
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// ---------------------------------------------------------------------------------------
// kernel launch wrappers implemented in the .hip files

// rng.hip
int lra_launch_gaussian_fill(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 rows,
                             i64 cols, double* out, i64 ld);
// the rows x cols Philox matrix stored transposed (cols x rows, leading dimension ld)
int lra_launch_gaussian_fill_t(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 rows,
                               i64 cols, double* out, i64 ld);
int lra_launch_gaussian_fill_f32(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 rows,
                                 i64 cols, float* out, i64 ld);
int lra_launch_srft_rand(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 m, i64 ell,
                         double* d, i64* idx);
int lra_launch_subset_rand(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 m, i64 ell,
                           i64* r);
int lra_launch_sprn_rand(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 m, double* s,
                         i64* r);
int lra_launch_fill_i64(hipStream_t st, i64* p, i64 n, i64 start);  // p[i] = start + i
int lra_launch_zero_u32(hipStream_t st, unsigned* p, i64 n);

// gemm.hip: C (M x N) = alpha * op(A) (M x K) * op(B) (K x N) + beta * C, FP64 in/out.
// transA/transB: 0 = as stored (column-major), 1 = transposed.  bgather (optional, 1-based)
// selects the rows of op(B): op(B)[k, j] = B'[bgather[k]-1, j].  prec selects the MFMA
// arithmetic.  omega_seed != ~0 makes op(A) the Gaussian generated on the fly is NOT used;
// sketches pass a materialised Omega.
struct GemmArgs {
  i64 M, N, K;
  double alpha, beta;
  const double* A; i64 lda; int transA;
  const double* B; i64 ldb; int transB;
  double* C; i64 ldc;
  const i64* bgather;
  int prec;
};
int lra_gemm(lra_ctx* ctx, const GemmArgs& g);

// qr.hip: blocked Householder QR (geqrf layout) and the explicit thin Q
int lra_geqrf_impl(lra_ctx* ctx, i64 m, i64 n, double* A, i64 lda, double* tau);
int lra_orgqr_impl(lra_ctx* ctx, i64 m, i64 k, const double* A, i64 lda, const double* tau,
                   double* Q, i64 ldq);

// qrcp.hip
struct QrcpCall {
  double* A; i64 lda;
  int m, n, offset;
  int kmax, nb;
  int mode;  // 0 = geqp3_adap, 1 = laqps block
  double atol, rtol;
  i64* jpvt;   // device, n (in/out; mode 0 initialised to 1..n by the kernel)
  double* tau; // device, kmax
  double* vn1; double* vn2;  // device, n (mode 1 in/out)
  int stop_after;            // mode 0: >0 stops at the first block end with steps >= value
  i64* k_out;                // device scalar: rank (mode 0) / kb (mode 1)
  int* early_out;            // device scalar: 1 when stop_after fired
};
int lra_qrcp(lra_ctx* ctx, const QrcpCall& q);
// Two mode-0 factorizations of equal n (q0.m <= q1.m) in one co-resident launch;
// LRA_E_UNSUPPORTED (nothing launched) when the pair does not fit.  lra_qrcp_pair_fits tells
// beforehand (and is false when LRA_QRCP_PAIR=0).
bool lra_qrcp_pair_fits(lra_ctx* ctx, int m0, int m1, int n);
int lra_qrcp_pair(lra_ctx* ctx, const QrcpCall& q0, const QrcpCall& q1);
// qrcp_global.hip: the same (mode 0) with the sketch in HBM, for sketches beyond the
// register-resident kernel
int lra_qrcp_global(lra_ctx* ctx, const QrcpCall& q);

// trsm.hip: T = R11 \ R12; in-place op(U) X = B with U the upper k x k triangle of R
int lra_launch_id_T(hipStream_t st, i64 k, i64 n, const double* R, i64 ldr, double* T, i64 ldt);
// the same for any k: row blocks of 1184 (GEMM updates between MFMA block solves) beyond one slab
int lra_id_T(lra_ctx* c, i64 k, i64 n, const double* R, i64 ldr, double* T, i64 ldt);
// X = R \ X in place, R k x k upper triangular, X k x nr (trsm.hip)
int lra_trsm_upper_inplace(lra_ctx* c, i64 k, i64 nr, const double* R, i64 ldr, double* T,
                           i64 ldt);
int lra_launch_trsm_upper(hipStream_t st, i64 k, i64 nrhs, const double* R, i64 ldr, double* B,
                          i64 ldb, int trans);
int lra_launch_axpy2d(hipStream_t st, i64 m, i64 n, double alpha, const double* X, i64 ldx,
                      double* Y, i64 ldy);

// misc.hip
int lra_launch_gather_cols(hipStream_t st, i64 m, i64 k, const double* A, i64 lda,
                           const i64* cols, double* out, i64 ldo);           // out = A[:, cols]
int lra_launch_gather_rows_t(hipStream_t st, i64 n, i64 k, const double* A, i64 lda,
                             const i64* rows, double* out, i64 ldo);         // out = A[rows,:]'
int lra_launch_transpose(hipStream_t st, i64 m, i64 n, const double* A, i64 lda, double* out,
                         i64 ldo);
int lra_launch_copy2d(hipStream_t st, i64 m, i64 n, const double* A, i64 lda, double* out,
                      i64 ldo);
int lra_launch_split_ids(hipStream_t st, i64 n, i64 k, const i64* jpvt, i64* sk, i64* rd);
// S (rows x k) = one-hot columns e_idx[j] (idx 1-based)
int lra_launch_onehot(hipStream_t st, i64 rows, i64 k, const i64* idx, double* S, i64 lds);
// out (k x n) = A[rows, :] (rows 1-based, k entries)
int lra_launch_gather_rows(hipStream_t st, i64 k, i64 n, const double* A, i64 lda,
                           const i64* rows, double* out, i64 ldo);
// *flag = 1 when the n x n matrix A is exactly symmetric, else 0 (device int)
int lra_launch_is_symmetric(hipStream_t st, i64 n, const double* A, i64 lda, int* flag);
