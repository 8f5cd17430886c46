// stubs.cpp -- entry points not built yet; each reports LRA_E_UNSUPPORTED.
#include "lra_internal.h"
struct SketchFactOut;





extern "C" int lra_idfact_batched(lra_ctx*, int64_t, const int64_t*, const int64_t*,
                                  const double* const*, const int64_t*, const lra_opts*,
                                  int64_t* const*, int64_t* const*, double* const*,
                                  const int64_t*, int64_t*, int64_t*) {
  lra_set_error("batched: not built"); return LRA_E_UNSUPPORTED; }



