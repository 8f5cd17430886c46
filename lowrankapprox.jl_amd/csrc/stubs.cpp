// stubs.cpp -- entry points not built yet; each reports LRA_E_UNSUPPORTED.
#include "lra_internal.h"
struct SketchFactOut;



extern "C" int lra_maxdet_swapcols_f64(lra_ctx*, int64_t, int64_t, int64_t*, double*, int64_t,
                                       double, int64_t, int64_t*) {
  lra_set_error("maxdet: not built"); return LRA_E_UNSUPPORTED; }


extern "C" int lra_idfact_op(lra_ctx*, int, int64_t, int64_t, lra_apply_fn, void*,
                             const lra_opts*, lra_id_out*) {
  lra_set_error("idfact_op: not built"); return LRA_E_UNSUPPORTED; }
extern "C" int lra_idfact_batched(lra_ctx*, int64_t, const int64_t*, const int64_t*,
                                  const double* const*, const int64_t*, const lra_opts*,
                                  int64_t* const*, int64_t* const*, double* const*,
                                  const int64_t*, int64_t*, int64_t*) {
  lra_set_error("batched: not built"); return LRA_E_UNSUPPORTED; }



