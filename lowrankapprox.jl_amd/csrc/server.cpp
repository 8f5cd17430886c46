// server.cpp -- a serving loop for idfact / psvdfact requests on one GPU (MI355X extension; the
// reference has no server: each call factors one matrix, src/id.jl:434-447,
// src/psvd.jl:238-272).
//
// Requests name a matrix that stays resident in device memory (the caller keeps it alive until
// the request is done).  They are admitted in arrival order as soon as a worker is free and
// their memory estimate fits the HBM budget next to the memory already held; the factors stay
// in device memory (owned by the server) until the caller releases them.  Each worker is a
// host thread with its own context, i.e. its own stream and workspace, so the kernels of
// admitted requests overlap on the device (the grid-synchronising ones are ordered device-wide,
// see SpinSection).  Every result is the one lra_idfact / lra_psvdfact gives for the request
// alone: the workers run exactly those entry points.
#include "lra_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Request {
  int64_t id = 0;
  int kind = LRA_REQ_IDFACT;
  int64_t m = 0, n = 0, lda = 0;
  const double* A = nullptr;
  lra_opts o{};
  size_t est = 0;   // admission estimate (bytes), held while running
  size_t held = 0;  // bytes of the stored factors, held until release
  int status = LRA_REQ_QUEUED;
  int err = LRA_OK;
  std::string msg;
  double t_submit = 0, t_start = 0, t_done = 0;
  // factors (device, server-owned)
  int64_t k = 0;
  int64_t* sk = nullptr;
  int64_t* rd = nullptr;
  double* T = nullptr;
  double* U = nullptr;
  double* S = nullptr;
  double* Vt = nullptr;
};

size_t idfact_estimate(int64_t m, int64_t n) {
  // sketch rounds up to order ~ min(m,n): Omega' (ell x m), the sketch and its copy
  // (ell x n each), split-K partials, T (<= ell x n), plus sk/rd
  const size_t ell = (size_t)std::min(m, n) + 8;
  return 8 * (ell * ((size_t)m + 4 * (size_t)n)) + 16 * (size_t)n;
}

size_t psvd_estimate(int64_t m, int64_t n) {
  const size_t k = (size_t)std::min(m, n);
  // the ID above, then the skeleton's QR (m x k), the k x n core, its SVD and U (m x k)
  return idfact_estimate(m, n) + 8 * (3 * (size_t)m * k + 3 * k * (size_t)n + 2 * k * k);
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

}  // namespace

struct lra_server {
  int device = 0;
  size_t budget = 0, in_use = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::map<int64_t, Request*> reqs;
  std::vector<std::thread> workers;
  std::vector<lra_ctx*> ctxs;
  int64_t next_id = 1;
  int running = 0;
  bool stop = false;
  std::chrono::steady_clock::time_point t0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

// Run one request on a worker's context; copies the library-owned factors into server-owned
// device memory.  Returns LRA_OK or an error status (message in r->msg).
int serve_one(lra_server* S, lra_ctx* c, Request* r) {
  (void)S;
  LRA_HIP(hipSetDevice(c->device));
  if (r->kind == LRA_REQ_IDFACT) {
    const int64_t n = r->n;
    LRA_HIP(hipMalloc((void**)&r->sk, sizeof(int64_t) * std::max<int64_t>(n, 1)));
    LRA_HIP(hipMalloc((void**)&r->rd, sizeof(int64_t) * std::max<int64_t>(n, 1)));
    lra_id_out out{};
    out.sk = r->sk; out.rd = r->rd; out.T = nullptr; out.ldt = 0;
    LRA_CHECK(lra_idfact(c, LRA_TRANS_N, r->m, n, r->A, r->lda, &r->o, &out));
    r->k = out.k;
    const int64_t nr = n - out.k;
    if (out.k > 0 && nr > 0) {
      LRA_HIP(hipMalloc((void**)&r->T, sizeof(double) * out.k * nr));
      LRA_HIP(hipMemcpy2DAsync(r->T, sizeof(double) * out.k, out.T, sizeof(double) * out.ldt,
                               sizeof(double) * out.k, nr, hipMemcpyDeviceToDevice, c->stream));
    }
    r->held = 16 * (size_t)std::max<int64_t>(n, 1) + 8 * (size_t)out.k * (size_t)nr;
  } else {
    lra_psvd_out out{};
    LRA_CHECK(lra_psvdfact(c, r->m, r->n, r->A, r->lda, &r->o, &out));
    const int64_t k = out.k;
    r->k = k;
    if (k > 0) {
      LRA_HIP(hipMalloc((void**)&r->U, sizeof(double) * r->m * k));
      LRA_HIP(hipMalloc((void**)&r->S, sizeof(double) * k));
      LRA_HIP(hipMalloc((void**)&r->Vt, sizeof(double) * k * r->n));
      LRA_HIP(hipMemcpy2DAsync(r->U, sizeof(double) * r->m, out.U, sizeof(double) * out.ldu,
                               sizeof(double) * r->m, k, hipMemcpyDeviceToDevice, c->stream));
      LRA_HIP(hipMemcpyAsync(r->S, out.S, sizeof(double) * k, hipMemcpyDeviceToDevice,
                             c->stream));
      LRA_HIP(hipMemcpy2DAsync(r->Vt, sizeof(double) * k, out.Vt, sizeof(double) * out.ldvt,
                               sizeof(double) * k, r->n, hipMemcpyDeviceToDevice, c->stream));
    }
    r->held = 8 * (size_t)k * ((size_t)r->m + (size_t)r->n + 1);
  }
  LRA_HIP(hipStreamSynchronize(c->stream));
  return lra_check_device_error(c);
}

void worker_main(lra_server* S, int w) {
  lra_ctx* c = S->ctxs[w];
  (void)hipSetDevice(S->device);
  for (;;) {
    Request* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      // admission: the oldest request, once its estimate fits next to what is held
      S->cv_work.wait(lk, [&] {
        return S->stop ||
               (!S->queue.empty() && S->in_use + S->queue.front()->est <= S->budget);
      });
      if (S->stop) return;
      r = S->queue.front();
      S->queue.pop_front();
      S->in_use += r->est;
      ++S->running;
      r->status = LRA_REQ_RUNNING;
      r->t_start = S->now();
    }
    const int st = serve_one(S, c, r);
    std::string msg = st == LRA_OK ? std::string() : std::string(lra_last_error());
    {
      std::lock_guard<std::mutex> lk(S->mu);
      r->t_done = S->now();
      S->in_use -= r->est;
      --S->running;
      if (st == LRA_OK) {
        S->in_use += r->held;
        r->status = LRA_REQ_DONE;
      } else {
        r->held = 0;
        r->err = st;
        r->msg = msg;
        r->status = LRA_REQ_FAILED;
      }
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
}

void free_factors(Request* r) {
  dev_free(r->sk); dev_free(r->rd); dev_free(r->T);
  dev_free(r->U); dev_free(r->S); dev_free(r->Vt);
  r->sk = r->rd = nullptr;
  r->T = r->U = r->S = r->Vt = nullptr;
}

}  // namespace

extern "C" int lra_server_create(int device, size_t hbm_budget, int nworkers, lra_server** out) {
  LRA_ARG(hbm_budget > 0, 2);
  LRA_ARG(nworkers >= 1 && nworkers <= 64, 3);
  LRA_ARG(out != nullptr, 4);
  *out = nullptr;
  lra_server* S = new lra_server();
  S->device = device;
  S->budget = hbm_budget;
  S->t0 = std::chrono::steady_clock::now();
  for (int w = 0; w < nworkers; ++w) {
    lra_ctx* c = nullptr;
    const int st = lra_ctx_create(device, 0, &c);
    if (st != LRA_OK) {
      for (lra_ctx* x : S->ctxs) lra_ctx_destroy(x);
      delete S;
      return st;
    }
    S->ctxs.push_back(c);
  }
  for (int w = 0; w < nworkers; ++w) S->workers.emplace_back(worker_main, S, w);
  *out = S;
  return LRA_OK;
}

// Queue a request on the matrix A (device, m x n, leading dimension lda), which must stay
// valid until the request is done.  A request whose estimate exceeds the whole budget is
// refused (LRA_E_ALLOC).
extern "C" int lra_server_submit(lra_server* S, int kind, int64_t m, int64_t n, const double* A,
                                 int64_t lda, const lra_opts* opts, int64_t* handle) {
  LRA_ARG(S != nullptr, 1);
  LRA_ARG(kind == LRA_REQ_IDFACT || kind == LRA_REQ_PSVDFACT, 2);
  LRA_ARG(m >= 0, 3);
  LRA_ARG(n >= 0, 4);
  LRA_ARG(A != nullptr || m * n == 0, 5);
  LRA_ARG(lda >= std::max<int64_t>(1, m), 6);
  LRA_ARG(opts != nullptr, 7);
  LRA_ARG(handle != nullptr, 8);
  Request* r = new Request();
  r->kind = kind; r->m = m; r->n = n; r->A = A; r->lda = lda; r->o = *opts;
  r->est = kind == LRA_REQ_IDFACT ? idfact_estimate(m, n) : psvd_estimate(m, n);
  if (r->est > S->budget) {
    lra_set_error("server: request %lld x %lld needs ~%zu bytes, over the %zu-byte budget",
                  (long long)m, (long long)n, r->est, S->budget);
    delete r;
    return LRA_E_ALLOC;
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    r->t_submit = S->now();
    S->reqs[r->id] = r;
    S->queue.push_back(r);
    *handle = r->id;
  }
  S->cv_work.notify_all();
  return LRA_OK;
}

extern "C" int lra_server_poll(lra_server* S, int64_t handle, int32_t* status) {
  LRA_ARG(S != nullptr, 1);
  LRA_ARG(status != nullptr, 3);
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(handle);
  if (it == S->reqs.end()) {
    lra_set_error("server: unknown handle %lld", (long long)handle);
    return LRA_E_STATE;
  }
  *status = it->second->status;
  return LRA_OK;
}

// Wait until the request is done or failed (timeout_s < 0: no limit).
extern "C" int lra_server_wait(lra_server* S, int64_t handle, double timeout_s, int32_t* status) {
  LRA_ARG(S != nullptr, 1);
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(handle);
  if (it == S->reqs.end()) {
    lra_set_error("server: unknown handle %lld", (long long)handle);
    return LRA_E_STATE;
  }
  Request* r = it->second;
  auto done = [&] { return r->status == LRA_REQ_DONE || r->status == LRA_REQ_FAILED; };
  if (timeout_s < 0) S->cv_done.wait(lk, done);
  else S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  if (status) *status = r->status;
  return LRA_OK;
}

extern "C" int lra_server_result(lra_server* S, int64_t handle, lra_server_out* out) {
  LRA_ARG(S != nullptr, 1);
  LRA_ARG(out != nullptr, 3);
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(handle);
  if (it == S->reqs.end()) {
    lra_set_error("server: unknown handle %lld", (long long)handle);
    return LRA_E_STATE;
  }
  const Request* r = it->second;
  *out = lra_server_out{};
  out->kind = r->kind;
  out->status = r->status;
  out->m = r->m;
  out->n = r->n;
  out->k = r->k;
  out->sk = r->sk; out->rd = r->rd; out->T = r->T; out->ldt = std::max<int64_t>(r->k, 1);
  out->U = r->U; out->ldu = std::max<int64_t>(r->m, 1); out->S = r->S; out->Vt = r->Vt;
  out->ldvt = std::max<int64_t>(r->k, 1);
  out->t_submit = r->t_submit; out->t_start = r->t_start; out->t_done = r->t_done;
  if (r->status == LRA_REQ_FAILED) {
    lra_set_error("server: request %lld failed with status %d: %s", (long long)handle, r->err,
                  r->msg.c_str());
    return r->err;
  }
  return LRA_OK;
}

// Free a finished request's factors (their memory leaves the budget) and forget the handle.
extern "C" int lra_server_release(lra_server* S, int64_t handle) {
  LRA_ARG(S != nullptr, 1);
  Request* r = nullptr;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(handle);
    if (it == S->reqs.end()) {
      lra_set_error("server: unknown handle %lld", (long long)handle);
      return LRA_E_STATE;
    }
    r = it->second;
    if (r->status != LRA_REQ_DONE && r->status != LRA_REQ_FAILED) {
      lra_set_error("server: request %lld is not finished", (long long)handle);
      return LRA_E_STATE;
    }
    S->reqs.erase(it);
    S->in_use -= r->held;
  }
  LRA_HIP(hipSetDevice(S->device));
  free_factors(r);
  delete r;
  S->cv_work.notify_all();
  return LRA_OK;
}

extern "C" int lra_server_stats(lra_server* S, int64_t* queued, int64_t* running,
                                size_t* bytes_held) {
  LRA_ARG(S != nullptr, 1);
  std::lock_guard<std::mutex> lk(S->mu);
  if (queued) *queued = (int64_t)S->queue.size();
  if (running) *running = S->running;
  if (bytes_held) *bytes_held = S->in_use;
  return LRA_OK;
}

// Stops the workers after the running requests (queued ones are dropped), frees everything.
extern "C" int lra_server_destroy(lra_server* S) {
  if (!S) return LRA_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (auto& t : S->workers) t.join();
  (void)hipSetDevice(S->device);
  for (auto& kv : S->reqs) {
    free_factors(kv.second);
    delete kv.second;
  }
  for (lra_ctx* c : S->ctxs) lra_ctx_destroy(c);
  delete S;
  return LRA_OK;
}
