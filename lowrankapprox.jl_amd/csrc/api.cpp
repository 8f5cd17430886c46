// api.cpp -- C ABI of liblra_hip.so: contexts, workspace, and the host drivers that mirror
// the reference's control flow (adaptive sketch loop, postprocessing) on top of the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lra_internal.h"

#include <mutex>

#include <functional>

static thread_local char g_err[1024] = "";

void lra_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char* lra_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------------------
// options

extern "C" void lra_opts_default(lra_opts* o) {
  // LRAOptions(Float64) defaults, src/LowRankApprox.jl:96-119
  std::memset(o, 0, sizeof(*o));
  o->atol = 0.0;
  o->maxdet_niter = -1;
  o->maxdet_tol = -1.0;
  o->nb = 32;
  o->pheig_orthtol = std::sqrt(2.220446049250313e-16);
  o->pqrfact_retval = LRA_RET_Q | LRA_RET_R;
  o->rank = -1;
  o->rtol = 5 * 2.220446049250313e-16;
  o->sketch = LRA_SKETCH_RANDN;
  o->sketch_randn_niter = 0;
  o->sketchfact_adap = 1;
  o->randn_samp_mul = 1; o->randn_samp_add = 8;
  o->srft_samp_mul = 1; o->srft_samp_add = 8;
  o->sub_samp_mul = 4; o->sub_samp_add = 8;
  o->snorm_niter = 32;
  o->verb = 1;
  o->precision = LRA_PREC_F64;
  o->seed = 0;
}

// chkopts! (src/LowRankApprox.jl:133-141)
static int chkopts(const lra_opts* o) {
  if (!(o->atol >= 0)) { lra_set_error("atol"); return -1; }
  if (!(o->nb > 0)) { lra_set_error("nb"); return -1; }
  if (!(o->pheig_orthtol >= 0)) { lra_set_error("pheig_orthtol"); return -1; }
  if (!(o->rtol >= 0)) { lra_set_error("rtol"); return -1; }
  if (o->sketch < LRA_SKETCH_NONE || o->sketch > LRA_SKETCH_SPRN) { lra_set_error("sketch"); return -1; }
  if (o->precision < LRA_PREC_F64 || o->precision > LRA_PREC_BF16X3) { lra_set_error("precision"); return -1; }
  return LRA_OK;
}
int lra_chkopts_public(const lra_opts* o) { return chkopts(o); }

// ---------------------------------------------------------------------------------------
// device-wide order of the grid-synchronising kernels

namespace {
struct SpinGate {
  std::mutex mu;
  hipEvent_t last = nullptr;
};
SpinGate g_spin[64];
}  // namespace

SpinSection::SpinSection(lra_ctx* ctx) : c(ctx) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  if (cs != hipStreamCaptureStatusNone || c->device < 0 || c->device >= 64) return;
  SpinGate& g = g_spin[c->device];
  g.mu.lock();
  held = true;
  // (this context's own last event is already ordered by its stream)
  if (g.last && g.last != c->spin_ev) (void)hipStreamWaitEvent(c->stream, g.last, 0);
}

SpinSection::~SpinSection() {
  if (!held) return;
  SpinGate& g = g_spin[c->device];
  if (!c->spin_ev && hipEventCreateWithFlags(&c->spin_ev, hipEventDisableTiming) != hipSuccess) {
    c->spin_ev = nullptr;
    (void)hipGetLastError();
  }
  if (c->spin_ev && hipEventRecord(c->spin_ev, c->stream) == hipSuccess) g.last = c->spin_ev;
  g.mu.unlock();
}

void lra_spin_forget(lra_ctx* c) {
  if (c->device < 0 || c->device >= 64 || !c->spin_ev) return;
  SpinGate& g = g_spin[c->device];
  std::lock_guard<std::mutex> lk(g.mu);
  if (g.last == c->spin_ev) g.last = nullptr;
  (void)hipEventDestroy(c->spin_ev);
  c->spin_ev = nullptr;
}

// ---------------------------------------------------------------------------------------
// context

extern "C" int lra_ctx_create(int device, size_t hbm_budget, lra_ctx** out) {
  LRA_ARG(out != nullptr, 3);
  int ndev = 0;
  LRA_HIP(hipGetDeviceCount(&ndev));
  LRA_ARG(device >= 0 && device < ndev, 1);
  LRA_HIP(hipSetDevice(device));
  lra_ctx* c = new lra_ctx();
  c->device = device;
  c->hbm_budget = hbm_budget;
  hipDeviceProp_t prop;
  LRA_HIP(hipGetDeviceProperties(&prop, device));
  c->num_cu = prop.multiProcessorCount;
  c->lds_bytes = (int)prop.sharedMemPerMultiprocessor;
  LRA_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
  c->stream = c->own_stream;
  LRA_HIP(hipHostMalloc(&c->host_pinned, 4096, hipHostMallocDefault));
  LRA_HIP(hipMalloc(&c->d_err, 64));
  LRA_HIP(hipMemset(c->d_err, 0, 64));
  *out = c;
  return LRA_OK;
}

// Uncached exchange areas of destroyed contexts, per device.  Uncached memory that kernels
// have read is never freed before the process ends; a new context takes a pooled area before
// allocating one, so the areas held are bounded by the contexts alive at once.  A pooled area
// is zeroed when it is returned and launch generations are process-wide, so no tag written by
// its earlier owner can match a later reader.
namespace {
std::mutex g_uc_mu;
std::vector<std::pair<void*, size_t>> g_uc_pool[64];
}  // namespace

extern "C" int lra_ctx_destroy(lra_ctx* c) {
  if (!c) return LRA_OK;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  lra_spin_forget(c);
  for (int i = 0; i < lra_ctx::kSlots; ++i)
    if (c->ws[i]) hipFree(c->ws[i]);
  if (c->coop) hipFree(c->coop);
  if (c->coop_uc && c->device >= 0 && c->device < 64 &&
      hipMemsetAsync(c->coop_uc, 0, c->coop_uc_size, c->stream) == hipSuccess &&
      hipStreamSynchronize(c->stream) == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_uc_mu);
    g_uc_pool[c->device].push_back({c->coop_uc, c->coop_uc_size});
  } else {
    (void)hipGetLastError();  // (an area that could not be cleared stays unused)
  }
  if (c->d_err) hipFree(c->d_err);
  if (c->host_pinned) hipHostFree(c->host_pinned);
  if (c->own_stream) hipStreamDestroy(c->own_stream);
  delete c;
  return LRA_OK;
}

extern "C" int lra_ctx_set_stream(lra_ctx* c, void* stream) {
  LRA_ARG(c != nullptr, 1);
  c->stream = (hipStream_t)stream;  // NULL is the device's null (legacy default) stream
  return LRA_OK;
}

extern "C" int lra_sync(lra_ctx* c) {
  LRA_ARG(c != nullptr, 1);
  LRA_HIP(hipStreamSynchronize(c->stream));
  return lra_check_device_error(c);
}

extern "C" int lra_device_info(lra_ctx* c, int64_t* num_cu, int64_t* lds) {
  LRA_ARG(c != nullptr, 1);
  if (num_cu) *num_cu = c->num_cu;
  if (lds) *lds = c->lds_bytes;
  return LRA_OK;
}

static size_t total_ws(lra_ctx* c) {
  size_t s = c->coop_size;
  for (int i = 0; i < lra_ctx::kSlots; ++i) s += c->ws_size[i];
  return s;
}

int lra_ws(lra_ctx* c, int slot, size_t bytes, void** out) {
  if (bytes == 0) bytes = 256;
  if (c->ws_size[slot] < bytes) {
    const size_t want = std::max(bytes, c->ws_size[slot] + c->ws_size[slot] / 4);
    if (c->hbm_budget && total_ws(c) - c->ws_size[slot] + want > c->hbm_budget) {
      lra_set_error("workspace: %zu bytes exceed the HBM budget %zu", want, c->hbm_budget);
      return LRA_E_ALLOC;
    }
    if (c->ws[slot]) {
      LRA_HIP(hipStreamSynchronize(c->stream));
      LRA_HIP(hipFree(c->ws[slot]));
      c->ws[slot] = nullptr;
      c->ws_size[slot] = 0;
    }
    hipError_t e = hipMalloc(&c->ws[slot], want);
    if (e != hipSuccess) {
      lra_set_error("workspace hipMalloc(%zu): %s", want, hipGetErrorString(e));
      return LRA_E_ALLOC;
    }
    c->ws_size[slot] = want;
  }
  *out = c->ws[slot];
  return LRA_OK;
}

// Exchange area of the cooperative kernels.  First choice: a fixed 64 MiB of uncached device
// memory (reads and writes bypass the per-XCD L2, so blocks exchange data without cache
// write-back/invalidate); it is never freed before the process ends.  Larger requests fall
// back to ordinary memory, with which the kernels use agent-scope fences.
int lra_coop_area(lra_ctx* c, size_t bytes, void** out, bool* uncached) {
  static const size_t kUncached = (size_t)64 << 20;
  if (!c->coop_uc_tried) {
    c->coop_uc_tried = true;
    void* p = nullptr;
    if (c->device >= 0 && c->device < 64) {
      std::lock_guard<std::mutex> lk(g_uc_mu);
      auto& pool = g_uc_pool[c->device];
      if (!pool.empty()) {
        c->coop_uc = pool.back().first;
        c->coop_uc_size = pool.back().second;
        pool.pop_back();
      }
    }
    if (c->coop_uc) {
      // (taken from the pool, already zeroed)
    } else if (hipExtMallocWithFlags(&p, kUncached, hipDeviceMallocUncached) == hipSuccess) {
      c->coop_uc = p;
      c->coop_uc_size = kUncached;
    } else {
      (void)hipGetLastError();
    }
  }
  if (c->coop_uc && bytes <= c->coop_uc_size) {
    *out = c->coop_uc;
    *uncached = true;
    return LRA_OK;
  }
  if (c->coop_size < bytes) {
    if (c->coop) {
      LRA_HIP(hipStreamSynchronize(c->stream));
      LRA_HIP(hipFree(c->coop));
      c->coop = nullptr;
    }
    const size_t want = std::max(bytes, (size_t)1 << 20);
    hipError_t e = hipMalloc(&c->coop, want);
    if (e != hipSuccess) {
      lra_set_error("coop hipMalloc(%zu): %s", want, hipGetErrorString(e));
      return LRA_E_ALLOC;
    }
    c->coop_size = want;
  }
  *out = c->coop;
  *uncached = false;
  return LRA_OK;
}

int lra_check_device_error(lra_ctx* c) {
  int h = 0;
  LRA_HIP(hipMemcpy(&h, c->d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (h != 0) {
    hipMemset(c->d_err, 0, sizeof(int));
    lra_set_error("device error %d (a bounded wait expired in a cooperative kernel)", h);
    return h;
  }
  return LRA_OK;
}

// synchronous read of device scalars through the pinned staging area
// (the device error word comes back in the same synchronisation)
template <class T>
static int read_scalars(lra_ctx* c, const T* dev, T* host, int count) {
  char* hp = static_cast<char*>(c->host_pinned);
  if (sizeof(T) * count > 2048) {
    lra_set_error("read_scalars: %d values exceed the staging area", count);
    return LRA_E_UNSUPPORTED;
  }
  LRA_HIP(hipMemcpyAsync(hp, dev, sizeof(T) * count, hipMemcpyDeviceToHost, c->stream));
  LRA_HIP(hipMemcpyAsync(hp + 2048, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  LRA_HIP(hipStreamSynchronize(c->stream));
  std::memcpy(host, hp, sizeof(T) * count);
  int h = 0;
  std::memcpy(&h, hp + 2048, sizeof(int));
  if (h != 0) {
    LRA_HIP(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
    LRA_HIP(hipStreamSynchronize(c->stream));
    lra_set_error("device error %d (a bounded wait expired in a cooperative kernel)", h);
    return h;
  }
  return LRA_OK;
}

// ---------------------------------------------------------------------------------------
// primitives

extern "C" int lra_gaussian_fill(lra_ctx* c, uint64_t seed, uint64_t stream_id, int64_t rows,
                                 int64_t cols, double* out, int64_t ld) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(rows >= 0, 4);
  LRA_ARG(cols >= 0, 5);
  LRA_ARG(out != nullptr || rows * cols == 0, 6);
  LRA_ARG(ld >= std::max<int64_t>(1, rows), 7);
  return lra_launch_gaussian_fill(c->stream, seed, stream_id, rows, cols, out, ld);
}

// Gaussian sketch into B; Omega materialised in WS_OMEGA when not injected.
static int sketch_gaussian_impl(lra_ctx* c, int side, int trans, i64 m, i64 n, const double* A,
                                i64 lda, i64 ell, const double* omega, uint64_t seed,
                                uint64_t stream_id, int prec, double* B, i64 ldb) {
  // op(A) is p x q
  const i64 p = trans == LRA_TRANS_N ? m : n;
  const i64 q = trans == LRA_TRANS_N ? n : m;
  GemmArgs g;
  g.alpha = 1.0; g.beta = 0.0; g.C = B; g.ldc = ldb; g.bgather = nullptr; g.prec = prec;
  if (side == LRA_SIDE_LEFT) {
    // B (ell x q) = Omega (ell x p) * op(A)
    g.M = ell; g.N = q; g.K = p;
    if (omega) {
      g.A = omega; g.lda = ell; g.transA = 0;
    } else {
      // generated Omega is stored transposed (p x ell): both GEMM operands k-contiguous
      double* w = nullptr;
      const i64 ldw = (p + 1) / 2 * 2;  // even: 16-byte aligned k pairs
      LRA_CHECK(lra_wsT(c, WS_OMEGA, (size_t)ldw * ell, &w));
      LRA_CHECK(lra_launch_gaussian_fill_t(c->stream, seed, stream_id, ell, p, w, ldw));
      g.A = w; g.lda = ldw; g.transA = 1;
    }
    g.B = A; g.ldb = lda; g.transB = trans == LRA_TRANS_N ? 0 : 1;
  } else {
    // B (p x ell) = op(A) (p x q) * Omega (q x ell)
    const double* om = omega;
    if (!om) {
      double* w = nullptr;
      LRA_CHECK(lra_wsT(c, WS_OMEGA, (size_t)q * ell, &w));
      LRA_CHECK(lra_launch_gaussian_fill(c->stream, seed, stream_id, q, ell, w, q));
      om = w;
    }
    g.M = p; g.N = ell; g.K = q;
    g.A = A; g.lda = lda; g.transA = trans == LRA_TRANS_N ? 0 : 1;
    g.B = om; g.ldb = q; g.transB = 0;
  }
  return lra_gemm(c, g);
}

// The Gaussian sketches of two rounds in one GEMM: C ((ell0 + ell1) x q, ld ell0 + ell1) =
// [Omega_s0; Omega_s1] * op(A), rows 0..ell0-1 the round of Philox stream s0 and the rest
// the round of stream s1 (each Omega is the one its round draws alone).  One tall product
// instead of two short ones, whose M tiles leave the MFMA units part-idle.
static int sketch_gaussian_left2(lra_ctx* c, int trans, i64 m, i64 n, const double* A, i64 lda,
                                 i64 ell0, i64 ell1, uint64_t seed, uint64_t s0, uint64_t s1,
                                 double* C) {
  const i64 p = trans == LRA_TRANS_N ? m : n;
  const i64 q = trans == LRA_TRANS_N ? n : m;
  const i64 L = ell0 + ell1, ldw = (p + 1) / 2 * 2;
  double* w = nullptr;
  LRA_CHECK(lra_wsT(c, WS_OMEGA, (size_t)ldw * L, &w));
  LRA_CHECK(lra_launch_gaussian_fill_t(c->stream, seed, s0, ell0, p, w, ldw));
  LRA_CHECK(lra_launch_gaussian_fill_t(c->stream, seed, s1, ell1, p, w + ell0 * ldw, ldw));
  GemmArgs g;
  g.alpha = 1.0; g.beta = 0.0; g.C = C; g.ldc = L; g.bgather = nullptr; g.prec = LRA_PREC_F64;
  g.M = L; g.N = q; g.K = p;
  g.A = w; g.lda = ldw; g.transA = 1;
  g.B = A; g.ldb = lda; g.transB = trans == LRA_TRANS_N ? 0 : 1;
  return lra_gemm(c, g);
}

extern "C" int lra_sketch_draw(lra_ctx* c, int sketch, uint64_t seed, uint64_t stream_id,
                               int64_t m, int64_t ell, double* d, int64_t* idx) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(sketch >= LRA_SKETCH_RANDN && sketch <= LRA_SKETCH_SPRN, 2);
  LRA_ARG(m >= 0, 5);
  LRA_ARG(ell >= 0, 6);
  LRA_HIP(hipSetDevice(c->device));
  switch (sketch) {
    case LRA_SKETCH_RANDN:
      LRA_ARG(d != nullptr || m * ell == 0, 7);
      return lra_launch_gaussian_fill(c->stream, seed, stream_id, ell, m, d, std::max<i64>(ell, 1));
    case LRA_SKETCH_SRFT:
      LRA_ARG(d != nullptr || m == 0, 7);
      LRA_ARG(idx != nullptr || ell == 0, 8);
      return lra_launch_srft_rand(c->stream, seed, stream_id, m, ell, d, idx);
    case LRA_SKETCH_SUB:
      LRA_ARG(idx != nullptr || ell == 0, 8);
      return lra_launch_subset_rand(c->stream, seed, stream_id, m, ell, idx);
    default:
      LRA_ARG(d != nullptr || m == 0, 7);
      LRA_ARG(idx != nullptr || m == 0, 8);
      return lra_launch_sprn_rand(c->stream, seed, stream_id, m, d, idx);
  }
}

extern "C" int lra_sketch_gaussian(lra_ctx* c, int side, int trans, int64_t m, int64_t n,
                                   const double* A, int64_t lda, int64_t ell,
                                   const double* omega, uint64_t seed, uint64_t stream_id,
                                   int prec, double* B, int64_t ldb) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(side == LRA_SIDE_LEFT || side == LRA_SIDE_RIGHT, 2);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 3);
  LRA_ARG(m >= 0, 4);
  LRA_ARG(n >= 0, 5);
  LRA_ARG(A != nullptr || m * n == 0, 6);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 7);
  LRA_ARG(ell >= 0, 8);
  LRA_ARG(prec >= LRA_PREC_F64 && prec <= LRA_PREC_BF16X3, 12);
  const i64 rows = side == LRA_SIDE_LEFT ? ell : (trans == LRA_TRANS_N ? m : n);
  LRA_ARG(ldb >= std::max<int64_t>(1, rows), 14);
  return sketch_gaussian_impl(c, side, trans, m, n, A, lda, ell, omega, seed, stream_id, prec,
                              B, ldb);
}

extern "C" int lra_geqp3_adap_f64(lra_ctx* c, int64_t m, int64_t n, double* B, int64_t ldb,
                                  int64_t rank_cap, double atol, double rtol, int64_t nb,
                                  int64_t* jpvt, double* tau, int64_t* k) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0 && m < (1 << 20), 2);
  LRA_ARG(n >= 0 && n < (1 << 30), 3);
  LRA_ARG(B != nullptr || m * n == 0, 4);
  LRA_ARG(ldb >= std::max<int64_t>(1, m), 5);
  LRA_ARG(atol >= 0, 7);
  LRA_ARG(rtol >= 0, 8);
  LRA_ARG(nb > 0, 9);
  LRA_ARG(k != nullptr, 12);
  const i64 l = std::min(m, n);
  const i64 kmax = (rank_cap < 0 || rank_cap > l) ? l : rank_cap;
  i64* dk = nullptr;
  LRA_CHECK(lra_wsT(c, WS_SCAL, 16, &dk));
  QrcpCall q{};
  q.A = B; q.lda = ldb; q.m = (int)m; q.n = (int)n; q.offset = 0; q.kmax = (int)kmax;
  q.nb = (int)std::min<i64>(nb, std::max<i64>(kmax, 1));
  q.mode = 0; q.atol = atol; q.rtol = rtol; q.jpvt = jpvt; q.tau = tau; q.vn1 = nullptr;
  q.vn2 = nullptr; q.stop_after = 0; q.k_out = dk; q.early_out = nullptr;
  LRA_CHECK(lra_qrcp(c, q));
  i64 hk = 0;
  LRA_CHECK(read_scalars(c, dk, &hk, 1));
  *k = hk;
  return LRA_OK;
}

extern "C" int lra_laqps_f64(lra_ctx* c, int64_t m, int64_t n, int64_t offset, int64_t nb,
                             int64_t* kb, double* A, int64_t lda, int64_t* jpvt, double* tau,
                             double* vn1, double* vn2, double* auxv, double* F, int64_t ldf) {
  (void)auxv; (void)F; (void)ldf;
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0 && m < (1 << 20), 2);
  LRA_ARG(n >= 0 && n < (1 << 30), 3);
  LRA_ARG(offset >= 0 && offset <= m, 4);
  LRA_ARG(nb >= 0, 5);
  LRA_ARG(kb != nullptr, 6);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 8);
  const i64 kmax = std::min<i64>(nb, std::min<i64>(n, m - offset));
  i64* dk = nullptr;
  LRA_CHECK(lra_wsT(c, WS_SCAL, 16, &dk));
  QrcpCall q{};
  q.A = A; q.lda = lda; q.m = (int)m; q.n = (int)n; q.offset = (int)offset;
  q.kmax = (int)kmax; q.nb = (int)std::max<i64>(kmax, 1); q.mode = 1; q.atol = 0; q.rtol = 0;
  q.jpvt = jpvt; q.tau = tau; q.vn1 = vn1; q.vn2 = vn2; q.stop_after = 0; q.k_out = dk;
  q.early_out = nullptr;
  if (kmax <= 0) { *kb = 0; return LRA_OK; }
  LRA_CHECK(lra_qrcp(c, q));
  i64 hk = 0;
  LRA_CHECK(read_scalars(c, dk, &hk, 1));
  *kb = hk;
  return LRA_OK;
}

extern "C" int lra_id_T_f64(lra_ctx* c, int64_t k, int64_t n, const double* R, int64_t ldr,
                            double* T, int64_t ldt) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(k >= 0, 2);
  LRA_ARG(n >= k, 3);
  LRA_ARG(ldr >= std::max<int64_t>(1, k), 5);
  LRA_ARG(ldt >= std::max<int64_t>(1, k), 7);
  return lra_id_T(c, k, n, R, ldr, T, ldt);
}

// ---------------------------------------------------------------------------------------
// sketchfact: the adaptive loop (src/sketch.jl:223-240 and the srft/sub/sprn siblings)

struct SketchFactOut {
  double* B;  // factored sketch (ell x N, ld = ell)
  double* Bsave;  // the sketch before factorization (mixed form only, else nullptr)
  i64 ell, N;
  i64* jpvt;
  double* tau;
  i64 k;
  i64 rounds;
  int prec;      // arithmetic of the accepted round
  i64 stream_id; // Philox stream of the accepted round's sketch
};

int lra_sketch_any(lra_ctx* c, int sketch, int trans, i64 m, i64 n, const double* A, i64 lda,
                   i64 ell, uint64_t seed, uint64_t stream_id, int prec, double* B, i64 ldb);

static i64 samp_of(const lra_opts* o, i64 x) {
  switch (o->sketch) {
    case LRA_SKETCH_RANDN: return o->randn_samp_mul * x + o->randn_samp_add;
    case LRA_SKETCH_SRFT: return o->srft_samp_mul * x + o->srft_samp_add;
    case LRA_SKETCH_SUB: return o->sub_samp_mul * x + o->sub_samp_add;
    default: return x;  // sprn: order = n (src/sketch.jl:680)
  }
}

// Relative noise floor of a sketch computed with `prec` (FP32 accumulation): diagonals of
// its pivoted QR below floor * max column norm carry no information.
static double prec_floor(int prec) {
  switch (prec) {
    case LRA_PREC_F32: return 1e-6;
    case LRA_PREC_BF16X3: return 1e-6;
    case LRA_PREC_BF16: return 5e-2;
    default: return 0.0;
  }
}

// sketchfact(:left, trans, A) (src/sketch.jl:223-240 and the srft/sub/sprn drivers).
// Mixed precision (o->precision != F64, Gaussian sketch): a round on low-precision MFMA is
// trusted when its answer does not depend on diagonals below the precision's noise floor:
//   - rtol >= floor: the low-precision round is final (selection + rank); T is refined in FP64
//     afterwards (lra_mixed_refine_T);
//   - rtol <  floor: a round that certifies rank >= target above the floor doubles as usual;
//     the first round that cannot is recomputed in FP64 with the same Omega, and every later
//     round stays FP64, so the accepted factorization is exactly the FP64 drop-in's.
// `make(ell, stream, prec, B)` writes the ell x N sketch of round `stream` into B (ld = ell).
// `make2(ell0, ell1, stream0, stream1, C)`, when given, writes two rounds' FP64 sketches as
// the row blocks of one (ell0 + ell1) x N matrix C (ld ell0 + ell1) for a pair of rounds.
using SketchFn = std::function<int(i64, uint64_t, int, double*)>;
using SketchFn2 = std::function<int(i64, i64, uint64_t, uint64_t, double*)>;

static int sketchfact_run(lra_ctx* c, i64 N, const SketchFn& make, const lra_opts* o,
                          uint64_t stream_base, SketchFactOut* out, bool pair_ok = true,
                          const SketchFn2& make2 = nullptr) {
  const bool gauss = o->sketch == LRA_SKETCH_RANDN;
  int prec = gauss ? o->precision : LRA_PREC_F64;
  const bool adap = o->sketchfact_adap || o->rank < 0;
  i64 target = adap ? o->nb : o->rank;
  i64* dk = nullptr;
  LRA_CHECK(lra_wsT(c, WS_SCAL, 16, &dk));
  int* dearly = (int*)(dk + 4);
  for (i64 rnd = 0;;) {
    const i64 ell = samp_of(o, target);
    if (ell <= 0) { lra_set_error("sketch order %lld", (long long)ell); return LRA_E_UNSUPPORTED; }
    // Rounds r and r+1 of an FP64 adaptive loop run as a pair whenever both fit one
    // co-resident grid (0+1, 2+3, ...; a round that does not pair runs alone and the next
    // one tries again): both sketches are formed by one GEMM and factored in one launch,
    // round r+1 on the larger share of the grid, so round r's pivoted QR costs no time of its
    // own (dense matrices only: a caller's operator is applied exactly once per round, as
    // the reference does).  Each round's factorization is exactly what it is alone (its own
    // sketch, stop rule and stop_after); the first accepted round is taken, as the
    // sequential loop would.  Round 0 pairs only for wide sketches (N >= 4096): on small
    // problems the pair's larger launch and GEMM cost more than round 0's QR saves, and
    // small ranks are accepted in round 0.
    if (pair_ok && adap && prec == LRA_PREC_F64 && (rnd >= 1 || N >= 4096)) {
      const i64 ell2 = samp_of(o, 2 * target);
      if (ell2 > ell && ell2 < (1 << 20) && N < (1 << 30) &&
          lra_qrcp_pair_fits(c, (int)std::min(ell, N), (int)std::min(ell2, N), (int)N) &&
          ell <= N && ell2 <= N) {
        const i64 ells[2] = {ell, ell2}, targets[2] = {target, 2 * target};
        double* Bs[2] = {nullptr, nullptr};
        i64* jp[2] = {nullptr, nullptr};
        double* tu[2] = {nullptr, nullptr};
        QrcpCall qs[2];
        for (int h = 0; h < 2; ++h) {
          LRA_CHECK(lra_wsT(c, h ? WS_SKETCH3 : WS_SKETCH, (size_t)ells[h] * std::max<i64>(N, 1), &Bs[h]));
          LRA_CHECK(lra_wsT(c, h ? WS_JPVT2 : WS_JPVT, (size_t)std::max<i64>(N, 1), &jp[h]));
          LRA_CHECK(lra_wsT(c, h ? WS_TAU2 : WS_TAU, (size_t)std::max<i64>(std::min(ells[h], N), 1), &tu[h]));
        }
        // merged: both sketches as the row blocks of one matrix (one GEMM), factored in place
        double* Cp = nullptr;
        const i64 L = ell + ell2;
        if (make2) {
          LRA_CHECK(lra_wsT(c, WS_SKETCHP, (size_t)L * std::max<i64>(N, 1), &Cp));
          LRA_CHECK(make2(ell, ell2, stream_base + (uint64_t)rnd, stream_base + (uint64_t)rnd + 1,
                          Cp));
        }
        for (int h = 0; h < 2; ++h) {
          if (!Cp) LRA_CHECK(make(ells[h], stream_base + (uint64_t)(rnd + h), prec, Bs[h]));
          const i64 l = std::min(ells[h], N);
          const i64 kmax = (o->rank < 0 || o->rank > l) ? l : o->rank;
          QrcpCall& q = qs[h];
          q = QrcpCall{};
          q.A = Cp ? Cp + (h ? ell : 0) : Bs[h];
          q.lda = Cp ? L : ells[h];
          q.m = (int)ells[h]; q.n = (int)N; q.offset = 0;
          q.kmax = (int)kmax; q.nb = (int)std::min<i64>(o->nb, std::max<i64>(kmax, 1));
          q.mode = 0; q.atol = o->atol; q.rtol = o->rtol; q.jpvt = jp[h]; q.tau = tu[h];
          q.stop_after = (int)std::min<i64>(targets[h], 1 << 30);
          q.k_out = dk + h; q.early_out = nullptr;
        }
        const int st = lra_qrcp_pair(c, qs[0], qs[1]);
        if (st == LRA_OK) {
          i64 hv[2] = {0, 0};
          LRA_CHECK(read_scalars(c, dk, hv, 2));
          for (int h = 0; h < 2; ++h) {
            if (hv[h] < targets[h]) {
              if (Cp)  // the accepted round's factored rows, compact (ld = its order)
                LRA_CHECK(lra_launch_copy2d(c->stream, ells[h], N, Cp + (h ? ell : 0), L, Bs[h],
                                            ells[h]));
              out->B = Bs[h]; out->Bsave = nullptr; out->ell = ells[h]; out->N = N;
              out->jpvt = jp[h]; out->tau = tu[h]; out->k = hv[h]; out->rounds = rnd + h + 1;
              out->prec = prec; out->stream_id = stream_base + rnd + h;
              return LRA_OK;
            }
          }
          target *= 4;
          rnd += 2;
          continue;
        }
        if (st != LRA_E_UNSUPPORTED) return st;
        (void)hipGetLastError();  // not launched: factor round by round below
      }
    }
    double* B = nullptr;
    LRA_CHECK(lra_wsT(c, WS_SKETCH, (size_t)ell * std::max<i64>(N, 1), &B));
    LRA_CHECK(make(ell, stream_base + (uint64_t)rnd, prec, B));
    const double floor = prec_floor(prec);
    const bool lowp = prec != LRA_PREC_F64;
    double* Bsave = nullptr;
    if (lowp && o->rtol >= floor) {
      LRA_CHECK(lra_wsT(c, WS_SKETCH2, (size_t)ell * std::max<i64>(N, 1), &Bsave));
      LRA_CHECK(lra_launch_copy2d(c->stream, ell, N, B, ell, Bsave, ell));
    }
    i64* jpvt = nullptr;
    double* tau = nullptr;
    LRA_CHECK(lra_wsT(c, WS_JPVT, (size_t)std::max<i64>(N, 1), &jpvt));
    LRA_CHECK(lra_wsT(c, WS_TAU, (size_t)std::max<i64>(std::min(ell, N), 1), &tau));
    const i64 l = std::min(ell, N);
    const i64 kmax = (o->rank < 0 || o->rank > l) ? l : o->rank;
    QrcpCall q{};
    q.A = B; q.lda = ell; q.m = (int)ell; q.n = (int)N; q.offset = 0; q.kmax = (int)kmax;
    q.nb = (int)std::min<i64>(o->nb, std::max<i64>(kmax, 1)); q.mode = 0;
    q.atol = o->atol; q.rtol = lowp ? std::max(o->rtol, floor) : o->rtol;
    q.jpvt = jpvt; q.tau = tau; q.vn1 = nullptr;
    q.vn2 = nullptr; q.stop_after = adap ? (int)std::min<i64>(target, 1 << 30) : 0;
    q.k_out = dk; q.early_out = dearly;
    LRA_CHECK(lra_qrcp(c, q));
    // with rtol below the floor, a rejected low-precision round also reports |R(0,0)| (the
    // largest column norm) and |R(t-1,t-1)|, t = target: R is written back on every exit
    const bool look = adap && lowp && o->rtol < floor && target >= 1 && target <= l;
    if (look) {
      LRA_HIP(hipMemcpyAsync(dk + 2, B, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      LRA_HIP(hipMemcpyAsync(dk + 3, B + (target - 1) * (ell + 1), sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
    }
    i64 hv[4] = {0, 0, 0, 0};
    LRA_CHECK(read_scalars(c, dk, hv, look ? 4 : 1));
    const i64 hk = hv[0];
    const bool accept = !adap || hk < target;
    if (accept && lowp && o->rtol < floor) {
      // the low-precision sketch cannot resolve this round: redo it in FP64 (same Omega)
      prec = LRA_PREC_F64;
      continue;
    }
    if (!accept && look) {
      // the next round (twice the target) is run in FP64 from the start when this round's
      // last pivot is already within kFloorAhead of the floor: a low-precision attempt would
      // most likely stop on the floor and be recomputed.  Only the precision of a round
      // changes; an FP64 round decides exactly as the FP64 drop-in does.
      constexpr double kFloorAhead = 1e4;
      double r0 = 0.0, rl = 0.0;
      std::memcpy(&r0, &hv[2], sizeof(double));
      std::memcpy(&rl, &hv[3], sizeof(double));
      if (fabs(rl) < kFloorAhead * floor * fabs(r0)) prec = LRA_PREC_F64;
    }
    if (accept) {
      out->B = B; out->Bsave = Bsave; out->ell = ell; out->N = N; out->jpvt = jpvt;
      out->tau = tau; out->k = hk; out->rounds = rnd + 1; out->prec = prec;
      out->stream_id = stream_base + rnd;
      return LRA_OK;
    }
    target *= 2;
    ++rnd;
  }
}

// orthrows! (src/util.jl:86-97, thin = false) of B (ell x N, ld ell): the rows become the
// orthonormal Q of B's LQ factorization, computed as the QR of B' (the same dlarfg steps as
// gelqf on B); rows beyond min(ell, N) are zeroed.
static int orthrows(lra_ctx* c, i64 ell, i64 N, double* B) {
  const i64 k = std::min(ell, N);
  if (k == 0) return LRA_OK;
  double *X = nullptr, *Q = nullptr;
  LRA_CHECK(lra_wsT(c, WS_PI0, (size_t)N * ell + 2 * (size_t)ell + 16, &X));
  double* tau = X + (size_t)N * ell;
  LRA_CHECK(lra_wsT(c, WS_PI1, (size_t)N * k, &Q));
  LRA_CHECK(lra_launch_transpose(c->stream, ell, N, B, ell, X, N));  // X = B' (N x ell)
  LRA_CHECK(lra_geqrf_impl(c, N, k, X, N, tau));
  LRA_CHECK(lra_orgqr_impl(c, N, k, X, N, tau, Q, N));
  if (k < ell) LRA_HIP(hipMemsetAsync(B, 0, sizeof(double) * ell * N, c->stream));
  // B(0:k, :) = Q'
  if (k == ell) return lra_launch_transpose(c->stream, N, k, Q, N, B, ell);
  double* Qt = nullptr;
  LRA_CHECK(lra_wsT(c, WS_PI0, (size_t)N * ell + 2 * (size_t)ell + 16, &Qt));
  LRA_CHECK(lra_launch_transpose(c->stream, N, k, Q, N, Qt, k));
  return lra_launch_copy2d(c->stream, k, N, Qt, k, B, ell);
}

// Y (ell x Q) = X (ell x R) * M with M = A (which 0: R = m, Q = n) or A' (which 1: R = n,
// Q = m).  A dense source is one GEMM; an operator is applied to X' and the result transposed
// back (mul!(C, X, L), src/linop.jl:105-113, as (L' X')').
static int src_rmul(lra_ctx* c, const LraSrc& s, int which, i64 ell, const double* X, i64 ldx,
                    double* Y, i64 ldy) {
  const i64 R = which == 0 ? s.m : s.n, Q = which == 0 ? s.n : s.m;
  if (!s.op()) {
    GemmArgs g;
    g.bgather = nullptr; g.prec = LRA_PREC_F64; g.alpha = 1.0; g.beta = 0.0;
    g.M = ell; g.N = Q; g.K = R; g.A = X; g.lda = ldx; g.transA = 0;
    g.B = s.A; g.ldb = s.lda; g.transB = which; g.C = Y; g.ldc = ldy;
    return lra_gemm(c, g);
  }
  double *Xt = nullptr, *Yt = nullptr;
  LRA_CHECK(lra_wsT(c, WS_OP2, (size_t)std::max<i64>(R * ell, 1), &Xt));
  LRA_CHECK(lra_wsT(c, WS_OP3, (size_t)std::max<i64>(Q * ell, 1), &Yt));
  LRA_CHECK(lra_launch_transpose(c->stream, ell, R, X, ldx, Xt, std::max<i64>(R, 1)));
  const int st = s.apply(s.user, which == 0 ? 1 : 0, ell, Xt, std::max<i64>(R, 1), Yt,
                         std::max<i64>(Q, 1), (void*)c->stream);
  if (st != 0) {
    lra_set_error("operator apply returned %d", st);
    return LRA_E_STATE;
  }
  return lra_launch_transpose(c->stream, Q, ell, Yt, std::max<i64>(Q, 1), Y, ldy);
}

// sketch_randn_ln / _lc power iterations (src/sketch.jl:129-174): Bp (ell x N) alternately
// orthonormalised and multiplied by op(A)' and op(A); a Hermitian A skips the second
// orthonormalisation (the two sketches swap).  A is a dense matrix or an operator.
static int power_iterate(lra_ctx* c, int trans, const LraSrc& s, i64 ell, double* B, i64 niter,
                         bool herm) {
  const i64 N = trans == LRA_TRANS_N ? s.n : s.m, P = trans == LRA_TRANS_N ? s.m : s.n;
  // op(A)' is A' (which 1) for trans n and A (which 0) for trans c; op(A) the other one
  const int w_opt = trans == LRA_TRANS_N ? 1 : 0, w_op = 1 - w_opt;
  double* Bq = nullptr;
  LRA_CHECK(lra_wsT(c, WS_PI2, (size_t)ell * std::max(N, P), &Bq));
  for (i64 it = 0; it < niter; ++it) {
    LRA_CHECK(orthrows(c, ell, N, B));
    LRA_CHECK(src_rmul(c, s, w_opt, ell, B, ell, Bq, ell));  // Bq (ell x P) = Bp op(A)'
    if (herm) {  // Bp, Bq = Bq, Bp (P == N)
      LRA_CHECK(lra_launch_copy2d(c->stream, ell, N, Bq, ell, B, ell));
      continue;
    }
    LRA_CHECK(orthrows(c, ell, P, Bq));
    LRA_CHECK(src_rmul(c, s, w_op, ell, Bq, ell, B, ell));  // Bp (ell x N) = Bq op(A)
  }
  return LRA_OK;
}

void lra_opts_for_operator(lra_opts* o) {
  if (o->sketch != LRA_SKETCH_RANDN) {
    if (o->verb) fprintf(stderr, "lra: invalid sketch method for an operator; using randn\n");
    o->sketch = LRA_SKETCH_RANDN;
  }
  o->precision = LRA_PREC_F64;  // the operator is applied in FP64; no entries to refine from
}

// Gaussian sketch B (ell x N) = Omega * op(L) of an operator, formed as (op(L)' Omega')':
// one application of the caller's operator to ell vectors (Omega the ell x P Philox matrix of
// the round, generated transposed).
static int op_sketch(lra_ctx* c, int trans, const LraSrc& s, i64 ell, uint64_t seed,
                     uint64_t stream, double* B) {
  const i64 N = trans == LRA_TRANS_N ? s.n : s.m, P = trans == LRA_TRANS_N ? s.m : s.n;
  double *X = nullptr, *Y = nullptr;
  LRA_CHECK(lra_wsT(c, WS_OP0, (size_t)std::max<i64>(P * ell, 1), &X));
  LRA_CHECK(lra_wsT(c, WS_OP1, (size_t)std::max<i64>(N * ell, 1), &Y));
  LRA_CHECK(lra_launch_gaussian_fill_t(c->stream, seed, stream, ell, P, X, std::max<i64>(P, 1)));
  const int st = s.apply(s.user, trans == LRA_TRANS_N ? 1 : 0, ell, X, std::max<i64>(P, 1), Y,
                         std::max<i64>(N, 1), (void*)c->stream);
  if (st != 0) {
    lra_set_error("operator apply returned %d", st);
    return LRA_E_STATE;
  }
  return lra_launch_transpose(c->stream, N, ell, Y, std::max<i64>(N, 1), B, ell);
}

// sketchfact(:left, trans, A) on a dense matrix (any sketch) or an operator (Gaussian only,
// the caller has applied lra_opts_for_operator), with sketch_randn_niter power iterations.
static int sketchfact_left(lra_ctx* c, int trans, const LraSrc& s, const lra_opts* o,
                           SketchFactOut* out, uint64_t stream_base = 0) {
  const i64 N = trans == LRA_TRANS_N ? s.n : s.m;  // columns of op(A)
  const i64 niter = o->sketch == LRA_SKETCH_RANDN ? o->sketch_randn_niter : 0;
  bool herm = false;
  if (niter > 0 && s.m == s.n) {  // ishermitian(A) (src/sketch.jl:132)
    if (s.op()) {
      herm = s.herm;
    } else {
      int* flag = nullptr;
      LRA_CHECK(lra_wsT(c, WS_PI2, 64, &flag));
      LRA_CHECK(lra_launch_is_symmetric(c->stream, s.n, s.A, s.lda, flag));
      int h = 0;
      LRA_CHECK(read_scalars(c, flag, &h, 1));
      herm = h != 0;
    }
  }
  SketchFn make = [=](i64 ell, uint64_t stream, int prec, double* B) -> int {
    if (s.op()) LRA_CHECK(op_sketch(c, trans, s, ell, o->seed, stream, B));
    else LRA_CHECK(lra_sketch_any(c, o->sketch, trans, s.m, s.n, s.A, s.lda, ell, o->seed,
                                  stream, prec, B, ell));
    if (niter > 0) LRA_CHECK(power_iterate(c, trans, s, ell, B, niter, herm));
    return LRA_OK;
  };
  SketchFn2 make2 = nullptr;
  if (!s.op() && o->sketch == LRA_SKETCH_RANDN && niter == 0)
    make2 = [=](i64 ell0, i64 ell1, uint64_t s0, uint64_t s1, double* C) -> int {
      return sketch_gaussian_left2(c, trans, s.m, s.n, s.A, s.lda, ell0, ell1, o->seed, s0, s1,
                                   C);
    };
  return sketchfact_run(c, N, make, o, stream_base, out, !s.op(), make2);
}

int lra_src_getcols(lra_ctx* c, const LraSrc& s, int trans, const i64* idx, i64 k, double* C,
                    i64 ldc) {
  if (k <= 0) return LRA_OK;
  if (!s.op()) {
    if (trans == LRA_TRANS_N) return lra_launch_gather_cols(c->stream, s.m, k, s.A, s.lda, idx, C, ldc);
    return lra_launch_gather_rows_t(c->stream, s.n, k, s.A, s.lda, idx, C, ldc);
  }
  // op(L)[:, idx] = op(L) * S for the one-hot S (R x k) (src/linop.jl:54-67)
  const i64 R = trans == LRA_TRANS_N ? s.n : s.m;
  double* S = nullptr;
  LRA_CHECK(lra_wsT(c, WS_OP2, (size_t)std::max<i64>(R * k, 1), &S));
  LRA_CHECK(lra_launch_onehot(c->stream, R, k, idx, S, std::max<i64>(R, 1)));
  const int st = s.apply(s.user, trans == LRA_TRANS_N ? 0 : 1, k, S, std::max<i64>(R, 1), C, ldc,
                         (void*)c->stream);
  if (st != 0) {
    lra_set_error("operator apply returned %d", st);
    return LRA_E_STATE;
  }
  return LRA_OK;
}

// Low-precision sketch selected the skeleton: recompute T in FP64 from A's skeleton columns
// with one step of iterative refinement.  T0 = R11 \ R12 of the low-precision sketch; the
// residual E = op(A)[:, rd] - op(A)[:, sk] * T0 is formed in FP64 (the one full-size FP64
// GEMM), sketched with the same Omega in the low precision (E is small, so its rounding is
// second order), and the correction solves the sketched least-squares problem through the
// semi-normal equations with the sketch's R11:  R11' R11 dT = (S*op(A)[:, sk])' (S*E).
static int mixed_refine_T(lra_ctx* c, int trans, i64 m, i64 n, const double* A, i64 lda,
                          const SketchFactOut& f, const i64* sk, const i64* rd, double* T,
                          i64 ldt, const lra_opts* o) {
  const i64 k = f.k, N = f.N, nr = N - k, ell = f.ell;
  const i64 P = trans == LRA_TRANS_N ? m : n;  // rows of op(A)
  double *C = nullptr, *E = nullptr, *SE = nullptr, *Bsk = nullptr, *X = nullptr;
  LRA_CHECK(lra_wsT(c, WS_TMP0, (size_t)P * k, &C));
  LRA_CHECK(lra_wsT(c, WS_TMP1, (size_t)P * nr, &E));
  LRA_CHECK(lra_wsT(c, WS_TMP2, (size_t)ell * nr, &SE));
  LRA_CHECK(lra_wsT(c, WS_TMP3, (size_t)ell * k, &Bsk));
  LRA_CHECK(lra_wsT(c, WS_R, (size_t)k * nr, &X));
  if (trans == LRA_TRANS_N) {
    LRA_CHECK(lra_launch_gather_cols(c->stream, m, k, A, lda, sk, C, P));
    LRA_CHECK(lra_launch_gather_cols(c->stream, m, nr, A, lda, rd, E, P));
  } else {
    LRA_CHECK(lra_launch_gather_rows_t(c->stream, n, k, A, lda, sk, C, P));
    LRA_CHECK(lra_launch_gather_rows_t(c->stream, n, nr, A, lda, rd, E, P));
  }
  // E -= C * T0   (FP64 MFMA)
  GemmArgs g;
  g.M = P; g.N = nr; g.K = k; g.alpha = -1.0; g.beta = 1.0;
  g.A = C; g.lda = P; g.transA = 0; g.B = T; g.ldb = ldt; g.transB = 0;
  g.C = E; g.ldc = P; g.bgather = nullptr; g.prec = LRA_PREC_F64;
  LRA_CHECK(lra_gemm(c, g));
  // S*E with the accepted round's Omega, low precision
  LRA_CHECK(lra_sketch_any(c, LRA_SKETCH_RANDN, LRA_TRANS_N, P, nr, E, P, ell, o->seed,
                           (uint64_t)f.stream_id, f.prec, SE, ell));
  // X = (S*op(A)[:, sk])' * (S*E), S*op(A)[:, sk] = saved sketch columns sk
  LRA_CHECK(lra_launch_gather_cols(c->stream, ell, k, f.Bsave, ell, sk, Bsk, ell));
  g.M = k; g.N = nr; g.K = ell; g.alpha = 1.0; g.beta = 0.0;
  g.A = Bsk; g.lda = ell; g.transA = 1; g.B = SE; g.ldb = ell; g.transB = 0;
  g.C = X; g.ldc = k; g.prec = LRA_PREC_F64;
  LRA_CHECK(lra_gemm(c, g));
  // dT = R11^{-1} R11^{-T} X, then T += dT
  LRA_CHECK(lra_launch_trsm_upper(c->stream, k, nr, f.B, ell, X, k, 1));
  LRA_CHECK(lra_launch_trsm_upper(c->stream, k, nr, f.B, ell, X, k, 0));
  LRA_CHECK(lra_launch_axpy2d(c->stream, k, nr, 1.0, X, k, T, ldt));
  return LRA_OK;
}

// ---------------------------------------------------------------------------------------
// idfact (src/id.jl:434-447)

using RefineFn = std::function<int(const i64*, const i64*, double*, i64)>;
static int finish_id(lra_ctx* c, const SketchFactOut& f, i64 N, const lra_opts* op,
                     lra_id_out* out, const RefineFn& refine);

int lra_idfact_src(lra_ctx* c, int trans, const LraSrc& s, const lra_opts* opts,
                   lra_id_out* out) {
  if (chkopts(opts) != LRA_OK) return -7;
  lra_opts o = *opts;
  o.pqrfact_retval = LRA_RET_T;
  if (s.op()) lra_opts_for_operator(&o);
  const i64 m = s.m, n = s.n;
  const i64 N = trans == LRA_TRANS_N ? n : m;
  SketchFactOut f{};
  if (o.sketch == LRA_SKETCH_NONE) {
    // pqrfact_none: rank-adaptive QRCP of op(A) itself (src/pqr.jl:323-327,343-346)
    const i64 P = trans == LRA_TRANS_N ? m : n;
    double* B = nullptr;
    LRA_CHECK(lra_wsT(c, WS_SKETCH, (size_t)std::max<i64>(P * N, 1), &B));
    if (trans == LRA_TRANS_N) LRA_CHECK(lra_launch_copy2d(c->stream, m, n, s.A, s.lda, B, m));
    else LRA_CHECK(lra_launch_transpose(c->stream, m, n, s.A, s.lda, B, n));
    i64* jpvt = nullptr;
    double* tau = nullptr;
    LRA_CHECK(lra_wsT(c, WS_JPVT, (size_t)std::max<i64>(N, 1), &jpvt));
    LRA_CHECK(lra_wsT(c, WS_TAU, (size_t)std::max<i64>(std::min(P, N), 1), &tau));
    i64 k = 0;
    LRA_CHECK(lra_geqp3_adap_f64(c, P, N, B, std::max<i64>(P, 1), o.rank, o.atol, o.rtol, o.nb,
                                 jpvt, tau, &k));
    f.B = B; f.Bsave = nullptr; f.ell = P; f.N = N; f.jpvt = jpvt; f.tau = tau; f.k = k;
    f.rounds = 1; f.prec = LRA_PREC_F64; f.stream_id = 0;
  } else {
    LRA_CHECK(sketchfact_left(c, trans, s, &o, &f));
  }
  return finish_id(c, f, N, &o, out, [&](const i64* sk, const i64* rd, double* T, i64 ldt) {
    if (s.op()) return (int)LRA_OK;  // operators are sketched in FP64 (no refinement)
    return mixed_refine_T(c, trans, m, n, s.A, s.lda, f, sk, rd, T, ldt, &o);
  });
}

extern "C" int lra_idfact(lra_ctx* c, int trans, int64_t m, int64_t n, const double* A,
                          int64_t lda, const lra_opts* opts, lra_id_out* out) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(A != nullptr || m * n == 0, 5);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(out != nullptr, 8);
  return lra_idfact_src(c, trans, lra_src_dense(m, n, A, lda), opts, out);
}

// The tail of idfact shared by the dense and matrix-free drivers: sk/rd from the pivots,
// T = R11 \ R12 (maxdet_t), the mixed-precision refinement when the accepted round was
// low precision, and the determinant-maximising swaps when maxdet_tol >= 0
// (pqrback_postproc, src/pqr.jl:420-436).
static int finish_id(lra_ctx* c, const SketchFactOut& f, i64 N, const lra_opts* op,
                     lra_id_out* out, const RefineFn& refine) {
  const lra_opts& o = *op;
  const i64 k = f.k;
  LRA_CHECK(lra_launch_split_ids(c->stream, N, k, f.jpvt, out->sk, out->rd));
  if (!out->T) {
    // library-owned T (valid until the next call on this context), compact ldt = k
    double* t = nullptr;
    LRA_CHECK(lra_wsT(c, WS_T, (size_t)std::max<i64>(k, 1) * std::max<i64>(N - k, 1), &t));
    out->T = t;
    out->ldt = std::max<i64>(k, 1);
  }
  if (k > 0 && N > k) {
    LRA_ARG(out->ldt >= k, 8);
    LRA_CHECK(lra_id_T(c, k, N, f.B, f.ell, out->T, out->ldt));
    if (f.prec != LRA_PREC_F64 && f.Bsave) LRA_CHECK(refine(out->sk, out->rd, out->T, out->ldt));
    if (o.maxdet_tol >= 0) {
      // strong-RRQR swaps (pqrback_postproc with maxdet, src/pqr.jl:426-433)
      i64* p = nullptr;
      LRA_CHECK(lra_wsT(c, WS_TMP3, (size_t)N, &p));
      LRA_HIP(hipMemcpyAsync(p, out->sk, sizeof(i64) * k, hipMemcpyDeviceToDevice, c->stream));
      LRA_HIP(hipMemcpyAsync(p + k, out->rd, sizeof(i64) * (N - k), hipMemcpyDeviceToDevice, c->stream));
      i64 it = 0;
      LRA_CHECK(lra_maxdet_swapcols_f64(c, k, N, p, out->T, out->ldt, o.maxdet_tol, o.maxdet_niter, &it));
      LRA_HIP(hipMemcpyAsync(out->sk, p, sizeof(i64) * k, hipMemcpyDeviceToDevice, c->stream));
      LRA_HIP(hipMemcpyAsync(out->rd, p + k, sizeof(i64) * (N - k), hipMemcpyDeviceToDevice, c->stream));
    }
  }
  out->k = k;
  out->rounds = f.rounds;
  return LRA_OK;
}

// ---------------------------------------------------------------------------------------
// matrix-free idfact (idfact(L::LinearOperator), src/id.jl:434-447 with the LinOp products
// of src/linop.jl:105-113): the Gaussian sketch Omega*op(L) is formed as (op(L)' Omega')',
// one application of the caller's operator to ell vectors per adaptive round, and each power
// iteration applies op(L) and op(L)' once more (src/sketch.jl:129-174).

extern "C" int lra_idfact_op(lra_ctx* c, int trans, int64_t m, int64_t n, lra_apply_fn apply,
                             void* user, int isherm, const lra_opts* opts, lra_id_out* out) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(apply != nullptr, 5);
  LRA_ARG(!isherm || m == n, 7);
  LRA_ARG(opts != nullptr, 8);
  LRA_ARG(out != nullptr, 9);
  LRA_HIP(hipSetDevice(c->device));
  return lra_idfact_src(c, trans, lra_src_op(m, n, apply, user, isherm != 0), opts, out);
}

// ---------------------------------------------------------------------------------------
// sketchfact (src/sketch.jl:52-66) with pqrfact_retval = "": column pivots and rank of
// op(A); sketch = none is pqrfact_backend! on op(A) itself (src/pqr.jl:323-327,343-346).

static int pivots_of(lra_ctx* c, int trans, const LraSrc& s, const lra_opts* o,
                     uint64_t stream_base, SketchFactOut* f) {
  const i64 m = s.m, n = s.n;
  const i64 N = trans == LRA_TRANS_N ? n : m;
  if (o->sketch != LRA_SKETCH_NONE) return sketchfact_left(c, trans, s, o, f, stream_base);
  const i64 P = trans == LRA_TRANS_N ? m : n;
  double* B = nullptr;
  LRA_CHECK(lra_wsT(c, WS_SKETCH, (size_t)std::max<i64>(P * N, 1), &B));
  if (trans == LRA_TRANS_N) LRA_CHECK(lra_launch_copy2d(c->stream, m, n, s.A, s.lda, B, std::max<i64>(m, 1)));
  else LRA_CHECK(lra_launch_transpose(c->stream, m, n, s.A, s.lda, B, std::max<i64>(n, 1)));
  i64* jpvt = nullptr;
  double* tau = nullptr;
  LRA_CHECK(lra_wsT(c, WS_JPVT, (size_t)std::max<i64>(N, 1), &jpvt));
  LRA_CHECK(lra_wsT(c, WS_TAU, (size_t)std::max<i64>(std::min(P, N), 1), &tau));
  i64 k = 0;
  LRA_CHECK(lra_geqp3_adap_f64(c, P, N, B, std::max<i64>(P, 1), o->rank, o->atol, o->rtol, o->nb,
                               jpvt, tau, &k));
  *f = SketchFactOut{};
  f->B = B; f->ell = P; f->N = N; f->jpvt = jpvt; f->tau = tau; f->k = k; f->rounds = 1;
  f->prec = LRA_PREC_F64;
  return LRA_OK;
}

extern "C" int lra_sketchfact(lra_ctx* c, int trans, int64_t m, int64_t n, const double* A,
                              int64_t lda, const lra_opts* opts, uint64_t stream_base,
                              int64_t* p, int64_t* k) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(trans == LRA_TRANS_N || trans == LRA_TRANS_C, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(A != nullptr || m * n == 0, 5);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(p != nullptr, 9);
  LRA_ARG(k != nullptr, 10);
  if (chkopts(opts) != LRA_OK) return -7;
  LRA_HIP(hipSetDevice(c->device));
  lra_opts o = *opts;
  o.precision = LRA_PREC_F64;
  SketchFactOut f{};
  LRA_CHECK(pivots_of(c, trans, lra_src_dense(m, n, A, lda), &o, stream_base, &f));
  const i64 N = trans == LRA_TRANS_N ? n : m;
  if (N > 0)
    LRA_HIP(hipMemcpyAsync(p, f.jpvt, sizeof(i64) * N, hipMemcpyDeviceToDevice, c->stream));
  *k = f.k;
  return LRA_OK;
}

// ---------------------------------------------------------------------------------------
// curfact (src/cur.jl:532-564): row and column skeletons from two chained sketchfacts.  The
// first skeleton (of length kr or kc, which may exceed min(m,n) when the sketch shows more
// rank than the matrix has, e.g. at rtol = 0) stays in a library workspace; only the final
// min(kr, kc) indices are written to the caller's buffers.

static int curfact_src(lra_ctx* c, const LraSrc& s, const lra_opts* opts, int64_t* rows,
                       int64_t* cols, int64_t* k) {
  if (chkopts(opts) != LRA_OK) return -6;
  LRA_HIP(hipSetDevice(c->device));
  lra_opts o = *opts;
  o.pqrfact_retval = 0;
  o.precision = LRA_PREC_F64;
  if (s.op()) lra_opts_for_operator(&o);
  const i64 m = s.m, n = s.n;
  *k = 0;
  if (m == 0 || n == 0) return LRA_OK;
  // ishermitian(A) (exactly symmetric for real A): one column sketchfact, rows = cols
  bool herm = false;
  if (m == n) {
    if (s.op()) {
      herm = s.herm;
    } else {
      int* flag = nullptr;
      LRA_CHECK(lra_wsT(c, WS_CUR0, 64, &flag));
      LRA_CHECK(lra_launch_is_symmetric(c->stream, n, s.A, s.lda, flag));
      int h = 0;
      LRA_CHECK(read_scalars(c, flag, &h, 1));
      herm = h != 0;
    }
  }
  SketchFactOut f{};
  if (herm) {
    LRA_CHECK(pivots_of(c, LRA_TRANS_N, s, &o, 0, &f));
    LRA_HIP(hipMemcpyAsync(cols, f.jpvt, sizeof(i64) * f.k, hipMemcpyDeviceToDevice, c->stream));
    LRA_HIP(hipMemcpyAsync(rows, f.jpvt, sizeof(i64) * f.k, hipMemcpyDeviceToDevice, c->stream));
    *k = f.k;
    LRA_HIP(hipStreamSynchronize(c->stream));
    return LRA_OK;
  }
  // the second sketchfact draws from Philox streams 2^32 + round
  const uint64_t base2 = 1ull << 32;
  i64 kr = 0, kc = 0;
  double* X = nullptr;
  i64* first = nullptr;  // the first stage's skeleton (up to m or n entries)
  LRA_CHECK(lra_wsT(c, WS_CUR1, (size_t)std::max<i64>(std::max(m, n), 1), &first));
  if (m >= n) {
    LRA_CHECK(pivots_of(c, LRA_TRANS_C, s, &o, 0, &f));
    kr = f.k;
    if (kr > 0)
      LRA_HIP(hipMemcpyAsync(first, f.jpvt, sizeof(i64) * kr, hipMemcpyDeviceToDevice, c->stream));
    // A[rows, :] as the transpose of X = A[rows, :]' (n x kr), factored with trans = c
    LRA_CHECK(lra_wsT(c, WS_CUR0, (size_t)std::max<i64>(n * kr, 1) + 8, &X));
    LRA_CHECK(lra_src_getcols(c, s, LRA_TRANS_C, first, kr, X, n));
    LRA_CHECK(pivots_of(c, LRA_TRANS_C, lra_src_dense(n, kr, X, std::max<i64>(n, 1)), &o, base2,
                        &f));
    kc = f.k;
    const i64 kk = std::min(kr, kc);
    if (kk > 0) {
      LRA_HIP(hipMemcpyAsync(rows, first, sizeof(i64) * kk, hipMemcpyDeviceToDevice, c->stream));
      LRA_HIP(hipMemcpyAsync(cols, f.jpvt, sizeof(i64) * kk, hipMemcpyDeviceToDevice, c->stream));
    }
  } else {
    LRA_CHECK(pivots_of(c, LRA_TRANS_N, s, &o, 0, &f));
    kc = f.k;
    if (kc > 0)
      LRA_HIP(hipMemcpyAsync(first, f.jpvt, sizeof(i64) * kc, hipMemcpyDeviceToDevice, c->stream));
    LRA_CHECK(lra_wsT(c, WS_CUR0, (size_t)std::max<i64>(m * kc, 1) + 8, &X));
    LRA_CHECK(lra_src_getcols(c, s, LRA_TRANS_N, first, kc, X, m));
    LRA_CHECK(pivots_of(c, LRA_TRANS_C, lra_src_dense(m, kc, X, std::max<i64>(m, 1)), &o, base2,
                        &f));
    kr = f.k;
    const i64 kk = std::min(kr, kc);
    if (kk > 0) {
      LRA_HIP(hipMemcpyAsync(rows, f.jpvt, sizeof(i64) * kk, hipMemcpyDeviceToDevice, c->stream));
      LRA_HIP(hipMemcpyAsync(cols, first, sizeof(i64) * kk, hipMemcpyDeviceToDevice, c->stream));
    }
  }
  *k = std::min(kr, kc);
  LRA_HIP(hipStreamSynchronize(c->stream));
  return LRA_OK;
}

extern "C" int lra_curfact(lra_ctx* c, int64_t m, int64_t n, const double* A, int64_t lda,
                           const lra_opts* opts, int64_t* rows, int64_t* cols, int64_t* k) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(A != nullptr || m * n == 0, 4);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 5);
  LRA_ARG(opts != nullptr, 6);
  LRA_ARG(rows != nullptr || std::min(m, n) == 0, 7);
  LRA_ARG(cols != nullptr || std::min(m, n) == 0, 8);
  LRA_ARG(k != nullptr, 9);
  return curfact_src(c, lra_src_dense(m, n, A, lda), opts, rows, cols, k);
}

// curfact on an operator (src/cur.jl:527-564 with A a LinearOperator): the first sketchfact
// applies the operator, A[rows, :] / A[:, cols] come from one application to one-hot vectors
// (src/linop.jl:54-67), and the second sketchfact runs on that dense block.
extern "C" int lra_curfact_op(lra_ctx* c, int64_t m, int64_t n, lra_apply_fn apply, void* user,
                              int isherm, const lra_opts* opts, int64_t* rows, int64_t* cols,
                              int64_t* k) {
  LRA_ARG(c != nullptr, 1);
  LRA_ARG(m >= 0, 2);
  LRA_ARG(n >= 0, 3);
  LRA_ARG(apply != nullptr, 4);
  LRA_ARG(!isherm || m == n, 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(rows != nullptr || std::min(m, n) == 0, 8);
  LRA_ARG(cols != nullptr || std::min(m, n) == 0, 9);
  LRA_ARG(k != nullptr, 10);
  return curfact_src(c, lra_src_op(m, n, apply, user, isherm != 0), opts, rows, cols, k);
}

// ---------------------------------------------------------------------------------------
// sketch dispatch (sketch(), src/sketch.jl:35-50) -- left side, op(A) = A or A'

int lra_sketch_srft_impl(lra_ctx* c, int trans, i64 m, i64 n, const double* A, i64 lda,
                         i64 ell, const double* d, const i64* idx, uint64_t seed,
                         uint64_t stream_id, double* B, i64 ldb);
int lra_sketch_subset_impl(lra_ctx* c, int trans, i64 m, i64 n, const double* A, i64 lda,
                           i64 ell, const i64* r, uint64_t seed, uint64_t stream_id, double* B,
                           i64 ldb);
int lra_sketch_sprn_impl(lra_ctx* c, int trans, i64 m, i64 n, const double* A, i64 lda,
                         i64 ell, const i64* r, const double* s, uint64_t seed,
                         uint64_t stream_id, double* B, i64 ldb);

int lra_sketch_any(lra_ctx* c, int sketch, int trans, i64 m, i64 n, const double* A, i64 lda,
                   i64 ell, uint64_t seed, uint64_t stream_id, int prec, double* B, i64 ldb) {
  switch (sketch) {
    case LRA_SKETCH_RANDN:
      return sketch_gaussian_impl(c, LRA_SIDE_LEFT, trans, m, n, A, lda, ell, nullptr, seed,
                                  stream_id, prec, B, ldb);
    case LRA_SKETCH_SRFT:
      return lra_sketch_srft_impl(c, trans, m, n, A, lda, ell, nullptr, nullptr, seed,
                                  stream_id, B, ldb);
    case LRA_SKETCH_SUB:
      return lra_sketch_subset_impl(c, trans, m, n, A, lda, ell, nullptr, seed, stream_id, B,
                                    ldb);
    case LRA_SKETCH_SPRN:
      return lra_sketch_sprn_impl(c, trans, m, n, A, lda, ell, nullptr, nullptr, seed,
                                  stream_id, B, ldb);
  }
  lra_set_error("unknown sketch %d", sketch);
  return LRA_E_UNSUPPORTED;
}
