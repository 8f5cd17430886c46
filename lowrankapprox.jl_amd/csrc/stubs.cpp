// stubs.cpp -- entry points not built yet; each reports LRA_E_UNSUPPORTED.
#include "lra_internal.h"
