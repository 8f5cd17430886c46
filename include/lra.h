/*
 * lra.h -- C ABI of liblra_hip.so, the MI355X (gfx950) implementation of LowRankApprox.jl's
 * sketch-then-partial-factorize path.
 *
 * Conventions (shared by every entry point):
 *   - column-major storage, int64_t sizes and leading dimensions (Julia's BlasInt on ILP64);
 *   - every matrix/vector pointer is DEVICE memory unless the parameter says "host";
 *   - pivot / index vectors are 1-based, exactly as the reference hands them around;
 *   - calls are asynchronous on the context's stream unless documented otherwise; lra_sync
 *     waits; a context is used by one host thread at a time (contexts are independent);
 *   - status: 0 = ok, < 0 = -(index of the first invalid argument, 1-based) like LAPACK info,
 *     > 0 = runtime error (LRA_E_*); lra_last_error() returns a message for the calling thread.
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference repository, version 0.5.5).
 */
#ifndef LRA_H
#define LRA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LRA_OK = 0,
  LRA_E_HIP = 1,       /* HIP runtime error */
  LRA_E_ALLOC = 2,     /* device allocation failed / HBM budget exceeded */
  LRA_E_UNSUPPORTED = 3, /* shape outside what the kernels were built for */
  LRA_E_TIMEOUT = 4,   /* a cooperative kernel's bounded wait expired */
  LRA_E_STATE = 5      /* invalid handle / state */
};

/* sketch types (LRAOptions.sketch, src/LowRankApprox.jl:86,106) */
enum { LRA_SKETCH_NONE = 0, LRA_SKETCH_RANDN = 1, LRA_SKETCH_SRFT = 2, LRA_SKETCH_SUB = 3,
       LRA_SKETCH_SPRN = 4 };
/* arithmetic of the sketch GEMM (MI355X extension; F64 is the reference behaviour) */
enum { LRA_PREC_F64 = 0, LRA_PREC_F32 = 1, LRA_PREC_BF16 = 2, LRA_PREC_BF16X3 = 3 };
/* pqrfact_retval bits ("q", "r", "t"; src/pqr.jl:299-301) */
enum { LRA_RET_Q = 1, LRA_RET_R = 2, LRA_RET_T = 4 };
/* trans / side symbols (:n / :c, :left / :right) */
enum { LRA_TRANS_N = 0, LRA_TRANS_C = 1, LRA_TRANS_B = 2 /* prange only */ };
enum { LRA_SIDE_LEFT = 0, LRA_SIDE_RIGHT = 1 };

/*
 * Mirror of LRAOptions (src/LowRankApprox.jl:77-94), same field names and, through
 * lra_opts_default, the same Float64 defaults (:96-119).  The sketchfact_*_samp functions
 * are affine maps n -> mul*n + add (defaults n+8, n+8, 4n+8).  The fields after `verb` are
 * MI355X extensions.
 */
typedef struct lra_opts {
  double atol;
  int64_t maxdet_niter;
  double maxdet_tol;
  int64_t nb;
  double pheig_orthtol;
  int32_t pqrfact_retval; /* LRA_RET_* bits */
  int64_t rank;
  double rtol;
  int32_t sketch; /* LRA_SKETCH_* */
  int64_t sketch_randn_niter;
  int32_t sketchfact_adap;
  int64_t randn_samp_mul, randn_samp_add;
  int64_t srft_samp_mul, srft_samp_add;
  int64_t sub_samp_mul, sub_samp_add;
  int64_t snorm_niter;
  int32_t verb;
  /* --- extensions --- */
  int32_t precision; /* LRA_PREC_*: arithmetic of the sketch GEMM */
  uint64_t seed;     /* key of the counter-based generator (Philox4x32-10) */
} lra_opts;

typedef struct lra_ctx lra_ctx;

/* ID result (IDPackedV, src/id.jl:15-19).  sk and rd are caller device buffers of N entries
 * each (N = columns of op(A)); the first k of sk and N-k of rd are written.  T: when the
 * caller passes T = NULL the library stores T (k x (N-k)) in a buffer it owns, valid until the
 * next call on the same context, and sets T and ldt = max(k,1); otherwise T is a caller buffer
 * with ldt >= k and room for ldt*(N-k) doubles.  k and the adaptive round count are written
 * on the host. */
typedef struct lra_id_out {
  int64_t* sk;
  int64_t* rd;
  double* T;
  int64_t ldt;
  int64_t k;      /* host: rank */
  int64_t rounds; /* host: adaptive rounds used */
} lra_id_out;

/* Partial QR (PQRFactors / PartialQR, src/pqr.jl:25-40, 86-100): op(A)[:, p] ~ Q R with
 * Q (P x k), R = [R11, R11 T] (k x N), p (N, 1-based) and T (k x (N-k)).  Buffers left NULL
 * are library-owned (valid until the next call on the context) and filled in. */
typedef struct lra_pqr_out {
  double* Q;
  int64_t ldq;
  double* R;
  int64_t ldr;
  int64_t* p;
  double* T;
  int64_t ldt;
  int64_t k;      /* host: rank */
  int64_t rounds; /* host: adaptive rounds of the ID */
} lra_pqr_out;

/* Partial SVD (PartialSVD, src/psvd.jl:16-24): A ~ U diag(S) Vt with U (m x k), S (k),
 * Vt (k x n).  NULL buffers are library-owned. */
typedef struct lra_psvd_out {
  double* U;
  int64_t ldu;
  double* S;
  double* Vt;
  int64_t ldvt;
  int64_t k; /* host */
} lra_psvd_out;

/* Partial Hermitian eigendecomposition (PartialHermEigen, src/pheig.jl:16-22): A ~ V diag(w) V'
 * with w ascending (kn negative, then kp positive), V (n x k).  NULL buffers are
 * library-owned. */
typedef struct lra_pheig_out {
  double* values;
  double* vectors;
  int64_t ldv;
  int64_t k, kn, kp; /* host */
} lra_pheig_out;

/* Matrix-free operator (LinearOperator mul!/mulc!, src/linop.jl:9-40,105-113):
 * Y (rows x nrhs, ldy) = op(L) * X (cols x nrhs, ldx) with op = L (trans 0) or L' (trans 1),
 * device buffers, work enqueued on `stream` (a hipStream_t). Return 0 on success. */
typedef int (*lra_apply_fn)(void* user, int trans, int64_t nrhs, const double* X, int64_t ldx,
                            double* Y, int64_t ldy, void* stream);

/* ---- context ------------------------------------------------------------------------ */
void lra_opts_default(lra_opts* opts);
int lra_ctx_create(int device, size_t hbm_budget, lra_ctx** ctx);
int lra_ctx_destroy(lra_ctx* ctx);
/* Enqueue all work of ctx on `stream` (a hipStream_t; NULL = the null stream).  A new
 * context starts on a non-blocking stream of its own. */
int lra_ctx_set_stream(lra_ctx* ctx, void* stream);
int lra_sync(lra_ctx* ctx);
const char* lra_last_error(void);
/* Device properties the host side needs (CUs, LDS per CU). */
int lra_device_info(lra_ctx* ctx, int64_t* num_cu, int64_t* lds_bytes);

/* ---- primitives --------------------------------------------------------------------- */

/* Standard-normal fill of a rows x cols column-major matrix from the counter-based
 * generator: element (i,j) is normal #(j*rows+i) of stream `stream_id` under `seed`.
 * Replaces crandn (src/util.jl:4) behind RandomGaussian (src/sketch.jl:91-110). */
int lra_gaussian_fill(lra_ctx* ctx, uint64_t seed, uint64_t stream_id, int64_t rows,
                      int64_t cols, double* out, int64_t ld);

/* The random objects a left sketch of an m-row operand draws from (seed, stream_id), exactly
 * those the sketch kernels and the adaptive loop (round r uses stream_base + r) generate
 * internally, so that a caller can inject the same bits into another implementation:
 *   randn: d = Omega (ell x m, column-major), idx unused (crandn, src/util.jl:4-11);
 *   srft:  d = signs (m), idx = frequency samples (ell, 1-based) (srft_rand, src/sketch.jl:338-361);
 *   sub:   idx = sampled rows (ell, 1-based), d unused (rand(1:m, k), src/sketch.jl:252);
 *   sprn:  d = coefficients (m), idx = permutation (m, 1-based) (src/sketch.jl:575-579).
 * Device buffers. */
int lra_sketch_draw(lra_ctx* ctx, int sketch, uint64_t seed, uint64_t stream_id, int64_t m,
                    int64_t ell, double* d, int64_t* idx);

/* Gaussian sketch B = Omega*op(A) (side left) or op(A)*Omega (side right) with op = A or
 * A' (trans), Omega either injected (omega != NULL: ell x K left, K x ell right,
 * column-major, ld = rows) or generated from (seed, stream_id) as lra_gaussian_fill does.
 * prec selects the MFMA arithmetic (F32/BF16/BF16X3 accumulate in FP32).
 * Replaces mul!(C, RandomGaussian, A) -> crandn + dgemm (src/sketch.jl:91-110, 129-220). */
int lra_sketch_gaussian(lra_ctx* ctx, int side, int trans, int64_t m, int64_t n,
                        const double* A, int64_t lda, int64_t ell, const double* omega,
                        uint64_t seed, uint64_t stream_id, int prec, double* B, int64_t ldb);

/* SRFT sketch on the left, B (ell x n) = SRFT(ell) * op(A) (real case, reference quirks
 * kept).  d: +-1 signs (K entries), idx: 1-based samples (ell entries), both device; when
 * d == NULL both are generated from (seed, stream_id).
 * Replaces mul!(C, SRFT, A) (src/sketch.jl:474-484, 499-510) incl. FFTW r2r. */
int lra_sketch_srft(lra_ctx* ctx, int trans, int64_t m, int64_t n, const double* A,
                    int64_t lda, int64_t ell, const double* d, const int64_t* idx,
                    uint64_t seed, uint64_t stream_id, double* B, int64_t ldb);

/* Random row subset on the left, B[i,:] = op(A)[r[i],:] (r 1-based, device, or generated
 * when NULL).  Replaces mul!(C, RandomSubset, A) (src/sketch.jl:248-257, 271-281). */
int lra_sketch_subset(lra_ctx* ctx, int trans, int64_t m, int64_t n, const double* A,
                      int64_t lda, int64_t ell, const int64_t* r, uint64_t seed,
                      uint64_t stream_id, double* B, int64_t ldb);

/* Sparse random Gaussian on the left (perm r 1-based and coefficients s, K entries each,
 * device, or generated when NULL).  Replaces mul!(C, SparseRandomGaussian, A)
 * (src/sketch.jl:571-589, 614-633). */
int lra_sketch_sprn(lra_ctx* ctx, int trans, int64_t m, int64_t n, const double* A,
                    int64_t lda, int64_t ell, const int64_t* r, const double* s,
                    uint64_t seed, uint64_t stream_id, double* B, int64_t ldb);

/* Exact ?laqps drop-in for the ccall at src/lapack.jl:129-136 (argument list of dlaqps by
 * value instead of by reference): one block of at most nb pivoted Householder steps on the
 * m x n panel A (rows offset+1..m active), jpvt/vn1/vn2 in/out, tau out, *kb (host) out.
 * auxv and F are accepted for signature parity and not used (they are dlaqps workspace). */
int lra_laqps_f64(lra_ctx* ctx, int64_t m, int64_t n, int64_t offset, int64_t nb,
                  int64_t* kb, double* A, int64_t lda, int64_t* jpvt, double* tau,
                  double* vn1, double* vn2, double* auxv, double* F, int64_t ldf);

/* Rank-adaptive GEQP3 (geqp3_adap! + geqp3_adap_main!, src/pqr.jl:348-418) in one launch:
 * B (m x n) overwritten by the LAPACK-layout factorization, jpvt (n, 1-based) and tau
 * (min(m,n)) out, *k (host) = rank.  rank_cap < 0 means min(m,n). */
int lra_geqp3_adap_f64(lra_ctx* ctx, int64_t m, int64_t n, double* B, int64_t ldb,
                       int64_t rank_cap, double atol, double rtol, int64_t nb, int64_t* jpvt,
                       double* tau, int64_t* k);

/* T = R11 \ R12 for R = triu(R(1:k,:)) (maxdet_t, src/pqr.jl:438-442; the dtrsm behind
 * ldiv!(UpperTriangular, .)).  R is k x n (ldr), T k x (n-k) (ldt). */
int lra_id_T_f64(lra_ctx* ctx, int64_t k, int64_t n, const double* R, int64_t ldr, double* T,
                 int64_t ldt);

/* Strong-RRQR determinant-maximising swaps on (p, T) (maxdet_swapcols!/maxdet_update!,
 * src/pqr.jl:445-501); p device 1-based (n), T k x (n-k).  *niter (host) out. */
int lra_maxdet_swapcols_f64(lra_ctx* ctx, int64_t k, int64_t n, int64_t* p, double* T,
                            int64_t ldt, double tol, int64_t niter_max, int64_t* niter);

/* Spectral norm by power iteration (snorm, src/snorm.jl:14-45) of a dense A (isherm != 0:
 * iterate on A itself).  x0 (n, device) is the start vector or NULL (generated). */
int lra_snorm(lra_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t lda, int isherm,
              const lra_opts* opts, const double* x0, double* result);

/* snormdiff(A, B) = snorm(A - B) for two dense m x n matrices (src/snorm.jl:50-53; the
 * difference is applied as A x - B x, src/linop.jl:155-187). */
int lra_snormdiff(lra_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t lda,
                  const double* B, int64_t ldb, int isherm, const lra_opts* opts,
                  const double* x0, double* result);

/* snormdiff(A, ID) = snorm(A - A[:,sk]*[I T]P') (src/snorm.jl:50-53 with the ID as a
 * LinearOperator): the difference is applied matrix-free.  sk/rd/T as in lra_id_out. */
int lra_snormdiff_id(lra_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t lda,
                     int64_t k, const int64_t* sk, const int64_t* rd, const double* T,
                     int64_t ldt, const lra_opts* opts, const double* x0, double* result);

/* Allocation-free lra_snormdiff_id, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments and result as lra_snormdiff_id; its vectors
 * and partial sums live in a fixed 32 MiB device array per GPU (held until the process
 * ends), so it needs n + m + k + 64 + chunks*m <= 4 Mi doubles (chunks <= 4 * CUs * 256 / m
 * + 1; every shape up to about 10^6 rows fits) and returns LRA_E_UNSUPPORTED beyond.  It
 * never calls hipMalloc/hipFree.  Calls on one GPU share that array: calls that overlap on
 * one GPU (from different contexts or threads) race. */
int lra_snormdiff_id_noalloc(lra_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t lda,
                             int64_t k, const int64_t* sk, const int64_t* rd, const double* T,
                             int64_t ldt, const lra_opts* opts, const double* x0,
                             double* result);

/* snorm of a matrix-free operator (snorm(L::LinearOperator), src/snorm.jl:14-45 with
 * src/linop.jl:105-113): `apply` is called with nrhs = 1 for L x (trans 0) and L' y
 * (trans 1); isherm != 0 iterates on L itself. */
int lra_snorm_op(lra_ctx* ctx, int64_t m, int64_t n, lra_apply_fn apply, void* user,
                 int isherm, const lra_opts* opts, const double* x0, double* result);

/* Householder QR of a tall m x n matrix (m >= n) in place, LAPACK geqrf layout (R on and
 * above the diagonal, reflectors below, tau[n]) -- the qr! of pqrfact/psvdfact/pheigfact
 * (src/pqr.jl:298, src/psvd.jl:244,258, src/pheig.jl:282).  Reflectors follow dlarfg, so the
 * signs of R's diagonal are dgeqrf's. */
int lra_geqrf_f64(lra_ctx* ctx, int64_t m, int64_t n, double* A, int64_t lda, double* tau);

/* Q (m x k, ldq) = first k columns of H_1...H_k from lra_geqrf_f64's output (LAPACK.orgqr!,
 * src/pqr.jl:427). */
int lra_orgqr_f64(lra_ctx* ctx, int64_t m, int64_t k, const double* A, int64_t lda,
                  const double* tau, double* Q, int64_t ldq);

/* C = alpha op(A) op(B) + beta C in FP64 on MFMA (trans 0 = N, 1 = T; BLAS argument
 * order): the dgemm behind mul! on the dense factors (src/pqr.jl:86-175, src/psvd.jl:61-124,
 * src/pheig.jl:54-100). */
int lra_dgemm(lra_ctx* ctx, int transa, int transb, int64_t m, int64_t n, int64_t k,
              double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
              double beta, double* C, int64_t ldc);

/* Y = V X (trans 0: X n x nrhs, Y k x nrhs) or V' X (trans 1: X k x nrhs, Y n x nrhs) for the
 * interpolation matrix V = [I T] P' of an ID (mul!/mulc! of IDPackedV, src/id.jl:94-200). */
int lra_id_apply(lra_ctx* ctx, int trans, int64_t k, int64_t n, const int64_t* sk,
                 const int64_t* rd, const double* T, int64_t ldt, int64_t nrhs,
                 const double* X, int64_t ldx, double* Y, int64_t ldy);

/* X (m x n) = diag(d) X (the diagonal factor of PartialSVD / PartialHermEigen products). */
int lra_scale_rows(lra_ctx* ctx, int64_t m, int64_t n, const double* d, double* X,
                   int64_t ldx);

/* ---- whole-pipeline entry points (mirror the reference API) -------------------------- */

/* idfact(trans, A, opts) (src/id.jl:434-447): sketchfact(:left, trans, A) (adaptive loop,
 * src/sketch.jl:223-240 and siblings) -> QRCP -> T.  With opts->precision != F64 the sketch
 * GEMM runs on low-precision MFMA and only selects the skeleton; T is recomputed in FP64
 * from A's skeleton columns with one step of iterative refinement. */
int lra_idfact(lra_ctx* ctx, int trans, int64_t m, int64_t n, const double* A, int64_t lda,
               const lra_opts* opts, lra_id_out* out);

/* ---- matrix-free entry points (AbstractLinearOperator arguments) ---------------------
 * The operator L (m x n) is reached only through `apply` (see lra_apply_fn).  isherm != 0
 * declares a Hermitian operator (HermitianLinearOperator, src/linop.jl:29-40; m == n): the
 * reference's ishermitian(L) is true for it and false for a plain LinearOperator.  As
 * chkopts!(opts, L) does (src/LowRankApprox.jl:142-148), a sketch other than randn is
 * replaced by randn (with a warning on stderr when opts->verb != 0); the operator is applied
 * in FP64, so opts->precision is ignored. */

/* idfact(trans, L, opts) (src/id.jl:434-447 through src/linop.jl:105-113): each adaptive
 * round applies op(L)' to ell Gaussian vectors; sketch_randn_niter power iterations apply
 * op(L) and op(L)' once more each (src/sketch.jl:129-174). */
int lra_idfact_op(lra_ctx* ctx, int trans, int64_t m, int64_t n, lra_apply_fn apply,
                  void* user, int isherm, const lra_opts* opts, lra_id_out* out);

/* pqrfact(trans, L, opts) (src/pqr.jl:285-320): the ID of op(L), then the skeleton columns
 * op(L)[:, sk] from one application to k one-hot vectors (getcols, src/util.jl:30-34 and
 * src/linop.jl:54-67) and their thin QR. */
int lra_pqrfact_op(lra_ctx* ctx, int trans, int64_t m, int64_t n, lra_apply_fn apply,
                   void* user, int isherm, const lra_opts* opts, lra_pqr_out* out);

/* psvdfact(L, opts) / psvdvals (src/psvd.jl:238-290) on an operator. */
int lra_psvdfact_op(lra_ctx* ctx, int64_t m, int64_t n, lra_apply_fn apply, void* user,
                    int isherm, const lra_opts* opts, lra_psvd_out* out);
int lra_psvdvals_op(lra_ctx* ctx, int64_t m, int64_t n, lra_apply_fn apply, void* user,
                    int isherm, const lra_opts* opts, double* S, int64_t* k);

/* pheigfact(L, opts) / pheigvals (src/pheig.jl:276-314) on a Hermitian operator (n x n; the
 * caller guarantees L' = L).  L[sk, sk] comes from one application to k one-hot vectors
 * (src/linop.jl:70-74). */
int lra_pheigfact_op(lra_ctx* ctx, int64_t n, lra_apply_fn apply, void* user,
                     const lra_opts* opts, lra_pheig_out* out);
int lra_pheigvals_op(lra_ctx* ctx, int64_t n, lra_apply_fn apply, void* user,
                     const lra_opts* opts, double* values, int64_t* k);

/* curfact(L, opts) (src/cur.jl:527-564): isherm != 0 takes one column sketchfact and
 * rows = cols; otherwise the first sketchfact applies the operator, the selected rows
 * (columns) of L come from one application to one-hot vectors, and the second sketchfact runs
 * on that dense block.  rows, cols as in lra_curfact. */
int lra_curfact_op(lra_ctx* ctx, int64_t m, int64_t n, lra_apply_fn apply, void* user,
                   int isherm, const lra_opts* opts, int64_t* rows, int64_t* cols,
                   int64_t* k);

/* sketchfact(:left, trans, A, opts) with pqrfact_retval = "" (src/sketch.jl:52-66 and the
 * adaptive drivers :223-240, :313-330, :545-562, :674-690; sketch = none is
 * pqrfact_backend! on op(A), src/pqr.jl:323-327): column pivots of op(A) into p (device,
 * N = columns of op(A) entries, 1-based) and the rank into *k (host).  Round r of the
 * adaptive loop draws its random sketch from Philox stream stream_base + r. */
int lra_sketchfact(lra_ctx* ctx, int trans, int64_t m, int64_t n, const double* A,
                   int64_t lda, const lra_opts* opts, uint64_t stream_base, int64_t* p,
                   int64_t* k);

/* curfact(A, opts) (src/cur.jl:532-564; sketch = none: curfact_none, :574-606): row and
 * column skeletons.  rows and cols (device, min(m,n) entries each) receive k 1-based
 * indices, *k (host); k <= min(m,n) always (the first stage's longer skeleton stays in the
 * library's workspace and is cut to min(kr, kc) as the reference does).  An exactly symmetric A (ishermitian for real A) takes one column
 * sketchfact and rows = cols; otherwise two chained sketchfacts, the second drawing from
 * Philox streams 2^32 + round. */
int lra_curfact(lra_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t lda,
                const lra_opts* opts, int64_t* rows, int64_t* cols, int64_t* k);

/* CUR(A, CURPackedU) (src/cur.jl:85-93; HermCUR / SymCUR :102-116 invert the same
 * intersection): from a row and column skeleton (rows, cols: device, k entries, 1-based),
 * C = A[:, cols] (m x k), R = A[rows, :] (k x n), and U = pinv(A[rows, cols]) as
 * PartialSVD(V, 1 ./ sigma, U') of the intersection's SVD W = U diag(sigma) V':
 * V (k x k), sinv = 1 ./ sigma (k, 1/0 = Inf as in the reference), Ut = U' (k x k).
 * C or R may be NULL (not formed). */
int lra_cur_build(lra_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t lda,
                  const int64_t* rows, const int64_t* cols, int64_t k, double* C, int64_t ldc,
                  double* R, int64_t ldr, double* V, int64_t ldv, double* sinv, double* Ut,
                  int64_t ldut);

/* pqrfact(trans, A, opts) (src/pqr.jl:285-320): the ID of op(A), then the thin QR of its
 * skeleton columns; pqrfact_retval bits choose Q, R, T.  sketch = none is the rank-adaptive
 * QRCP of op(A) itself (pqrfact_none, :323-327). */
int lra_pqrfact(lra_ctx* ctx, int trans, int64_t m, int64_t n, const double* A, int64_t lda,
                const lra_opts* opts, lra_pqr_out* out);

/* prange(trans, A, opts) (src/prange.jl:14-77): Q (P x k, orthonormal columns) spanning the
 * numerical range of op(A) (trans n / c; P = rows of op(A)), or of A and A' together
 * (trans LRA_TRANS_B, square A; an exactly symmetric A takes trans n).  sketch randn: the
 * right-side adaptive sketchfact (src/sketch.jl:223-240) and its Householder Q; sub:
 * prange_sub (:65-78); none: pqrfact(op(A))[:Q]; trans b with randn or none only; srft and
 * sprn return LRA_E_UNSUPPORTED.  Only out->Q, ldq (NULL Q: library-owned, valid until the
 * next call on the context), k and rounds are set. */
int lra_prange(lra_ctx* ctx, int trans, int64_t m, int64_t n, const double* A, int64_t lda,
               const lra_opts* opts, lra_pqr_out* out);

/* psvdfact(A, opts) / psvdvals (src/psvd.jl:238-300): the ID, the thin QR of the skeleton,
 * the SVD of the k x n core through its QR and a one-sided Jacobi SVD, truncated by
 * psvdrank (:302-308).  psvdvals writes the k singular values to S (device). */
int lra_psvdfact(lra_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t lda,
                 const lra_opts* opts, lra_psvd_out* out);
int lra_psvdvals(lra_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t lda,
                 const lra_opts* opts, double* S, int64_t* k);

/* pheigfact(A, opts) / pheigvals for a Hermitian (symmetric) A (src/pheig.jl:276-314). */
int lra_pheigfact(lra_ctx* ctx, int64_t n, const double* A, int64_t lda, const lra_opts* opts,
                  lra_pheig_out* out);
int lra_pheigvals(lra_ctx* ctx, int64_t n, const double* A, int64_t lda, const lra_opts* opts,
                  double* values, int64_t* k);

/* One-sided Jacobi SVD of X (m x n, m >= n), the core of psvdfact/pheigfact: X = U diag(s) V'
 * with s decreasing; X is overwritten (MI355X extension; replaces svd!/eigen! on the small
 * core, src/psvd.jl:245, src/pheig.jl:284). */
int lra_gesvj_f64(lra_ctx* ctx, int64_t m, int64_t n, double* X, int64_t ldx, double* U,
                  int64_t ldu, double* s, double* V, int64_t ldv);

/* ---- complex element type (ComplexF64) --------------------------------------------------
 * Complex matrices are column-major with interleaved (re, im) doubles; leading dimensions
 * count complex elements.  These replace the reference's paths for T <: Complex
 * (src/id.jl:434-447 with crandn, src/util.jl:4-11; the complex SRFT, src/sketch.jl:345-376,
 * 452-472; zlaqps through src/lapack.jl:117-139).  The QRCP runs in HBM (two launches per
 * pivot step); complex power iterations, maxdet swaps and the complex partial QR / SVD /
 * eigen factorizations are not provided. */

/* The random objects of a complex left sketch of round stream_id on an m-row operand:
 *   randn: d = Phi (ell x m complex, 2*ell*m doubles), Phi[p,i] = G[2p,i] + i G[2p+1,i] for the
 *          real 2ell x m Gaussian G that lra_sketch_draw(randn) gives (crandn's layout);
 *   srft:  d = unit-modulus signs (m complex), idx = frequency samples (ell, 1-based);
 *   sub:   idx = rows (ell, 1-based);
 *   sprn:  d = coefficients (m complex), idx = permutation (m, 1-based). */
int lra_zsketch_draw(lra_ctx* ctx, int sketch, uint64_t seed, uint64_t stream_id, int64_t m,
                     int64_t ell, double* d, int64_t* idx);

/* Complex left sketch B (ell x N, ldb = ell) = Phi * op(A), op = A or A^H, Phi of the sketch
 * kind drawn from (seed, stream_id) or injected (d, idx as lra_zsketch_draw gives them; idx
 * NULL draws them; randn always draws).  SRFT: B[p, j] = DFT(d .* op(A)[:, j])[idx[p]]. */
int lra_zsketch(lra_ctx* ctx, int sketch, int trans, int64_t m, int64_t n, const double* A,
                int64_t lda, int64_t ell, const double* d, const int64_t* idx, uint64_t seed,
                uint64_t stream_id, double* B, int64_t ldb);

/* Complex rank-adaptive GEQP3 (geqp3_adap!, src/pqr.jl:348-418, zlaqps arithmetic) in place:
 * jpvt (n, 1-based), tau (min(m,n) complex), *k (host). */
int lra_zgeqp3_adap(lra_ctx* ctx, int64_t m, int64_t n, double* B, int64_t ldb,
                    int64_t rank_cap, double atol, double rtol, int64_t nb, int64_t* jpvt,
                    double* tau, int64_t* k);

/* Complex T = R11 \ R12 (maxdet_t, src/pqr.jl:438-442). */
int lra_zid_T(lra_ctx* ctx, int64_t k, int64_t n, const double* R, int64_t ldr, double* T,
              int64_t ldt);

/* idfact(trans, A, opts) for a ComplexF64 A: sk, rd as lra_idfact, T complex (ldt in complex
 * elements; T = NULL: library-owned).  Gaussian, SRFT, subset and sparse sketches and
 * sketch = none; LRA_E_UNSUPPORTED for sketch_randn_niter > 0 or maxdet_tol >= 0. */
int lra_zidfact(lra_ctx* ctx, int trans, int64_t m, int64_t n, const double* A, int64_t lda,
                const lra_opts* opts, lra_id_out* out);

/* Batched ID of many independent blocks (MI355X extension; no counterpart in the reference,
 * which factors blocks one idfact call at a time, src/id.jl:434-447).  Block b (m[b] x n[b],
 * A[b] with lda[b]) gets exactly the result of lra_idfact on it alone with the same opts: its
 * first adaptive round runs inside one persistent fused launch (sketch GEMM + pivoted QR +
 * T), and a block whose rank reaches that round's target is completed by lra_idfact.  All
 * per-block arrays are HOST arrays; the pointers in them are device addresses.  sk[b] and
 * rd[b] hold n[b] entries, T[b] holds ldt[b] * n[b] doubles with ldt[b] >= the block's rank.
 * k_out (host, nblk) receives the ranks, *n_fallback the number of blocks that went through
 * lra_idfact. */
int lra_idfact_batched(lra_ctx* ctx, int64_t nblk, const int64_t* m, const int64_t* n,
                       const double* const* A, const int64_t* lda, const lra_opts* opts,
                       int64_t* const* sk, int64_t* const* rd, double* const* T,
                       const int64_t* ldt, int64_t* k_out, int64_t* n_fallback);

/* The same in three phases, for callers that replay one batch (MI355X extension):
 * _plan validates the blocks, uploads their descriptors and draws Omega (allocates);
 * _launch enqueues the counter reset + persistent kernel on the context's stream (or replays
 * the graph made by _capture), allocating nothing; _finish waits and completes the blocks left
 * to lra_idfact. */
typedef struct lra_batch_plan lra_batch_plan;
int lra_idfact_batched_plan(lra_ctx* ctx, int64_t nblk, const int64_t* m, const int64_t* n,
                            const double* const* A, const int64_t* lda, const lra_opts* opts,
                            int64_t* const* sk, int64_t* const* rd, double* const* T,
                            const int64_t* ldt, lra_batch_plan** plan);
int lra_idfact_batched_capture(lra_batch_plan* plan);
int lra_idfact_batched_launch(lra_batch_plan* plan);
int lra_idfact_batched_finish(lra_batch_plan* plan, int64_t* k_out, int64_t* n_fallback);
int lra_idfact_batched_plan_destroy(lra_batch_plan* plan);
/* The plan's kernel split (MI355X extension, no counterpart in the reference): widths[r] is
 * the widest block of round r that runs in a 512-thread workgroup (two per CU) rather than a
 * 1024-thread one or a workgroup pair, 0 for none, -1 past the plan's rounds; count entries.
 * _plan picks them by timing the candidates on the blocks when LRA_AUTOTUNE is not 0 and the
 * batch has at least 256 blocks (LRA_BATCHED_W0=w0,w1 forces them); results do not depend on
 * them. */
int lra_idfact_batched_plan_widths(const lra_batch_plan* plan, int* widths, int count);

/*
 * Serving (MI355X extension; the reference has no server, each call factors one matrix).
 * Requests on device-resident matrices are admitted in arrival order when a worker is free
 * and their memory estimate fits the HBM budget next to the factors already held; workers
 * (host threads with a context each) overlap their kernels on the device; factors stay in
 * device memory until released.  Each result is what lra_idfact / lra_psvdfact returns for
 * the request alone (the workers call them).  Times are seconds on the server's clock.
 */
enum { LRA_REQ_IDFACT = 0, LRA_REQ_PSVDFACT = 1 };
enum { LRA_REQ_QUEUED = 0, LRA_REQ_RUNNING = 1, LRA_REQ_DONE = 2, LRA_REQ_FAILED = 3 };
typedef struct lra_server lra_server;
typedef struct lra_server_out {
  int32_t kind, status;
  int64_t m, n, k;
  int64_t* sk; /* idfact: sk (k), rd (n-k), T (k x (n-k), ldt) */
  int64_t* rd;
  double* T;
  int64_t ldt;
  double* U; /* psvdfact: U (m x k, ldu), S (k), Vt (k x n, ldvt) */
  int64_t ldu;
  double* S;
  double* Vt;
  int64_t ldvt;
  double t_submit, t_start, t_done;
} lra_server_out;
int lra_server_create(int device, size_t hbm_budget, int nworkers, lra_server** srv);
int lra_server_submit(lra_server* srv, int kind, int64_t m, int64_t n, const double* A,
                      int64_t lda, const lra_opts* opts, int64_t* handle);
int lra_server_poll(lra_server* srv, int64_t handle, int32_t* status);
int lra_server_wait(lra_server* srv, int64_t handle, double timeout_s, int32_t* status);
int lra_server_result(lra_server* srv, int64_t handle, lra_server_out* out);
int lra_server_release(lra_server* srv, int64_t handle);
int lra_server_stats(lra_server* srv, int64_t* queued, int64_t* running, size_t* bytes_held);
int lra_server_destroy(lra_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* LRA_H */
