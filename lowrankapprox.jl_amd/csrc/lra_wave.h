// lra_wave.h -- wave-level (64-lane) reductions on DPP lane permutations and the small
// scalar helpers shared by the pivoted-QR kernels.
#pragma once
#include "lra_internal.h"

#include <climits>

LRA_DEV double lapy2(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

// ---- wave-level helpers on DPP lane permutations (VALU latency instead of LDS permutes) ----
// quad_perm [1,0,3,2] = lane^1, quad_perm [2,3,0,1] = lane^2, row_half_mirror pairs the two
// quads of each 8 lanes, row_mirror the two halves of each 16 lanes.  Every pattern is an
// involution, so a+b is formed as b+a in the partner lane: all lanes of a group end with
// bit-identical results.
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HMIRROR = 0x141, DPP_MIRROR = 0x140;

template <int CTRL>
LRA_DEV double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL>
LRA_DEV int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
LRA_DEV double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// sum over the `width` consecutive lanes of a group (width a power of two <= 64); call with
// the whole wave active
LRA_DEV double group_sum(double v, int width) {
  if (width >= 2) v += dpp_f64<DPP_XOR1>(v);
  if (width >= 4) v += dpp_f64<DPP_XOR2>(v);
  if (width >= 8) v += dpp_f64<DPP_HMIRROR>(v);
  if (width >= 16) v += dpp_f64<DPP_MIRROR>(v);
  if (width >= 32) v += __shfl_xor(v, 16, 64);
  if (width >= 64) v += __shfl_xor(v, 32, 64);
  return v;
}
// wave-uniform maximum (NaN-free inputs)
LRA_DEV double wave_max(double v) {
  v = fmax(v, dpp_f64<DPP_XOR1>(v));
  v = fmax(v, dpp_f64<DPP_XOR2>(v));
  v = fmax(v, dpp_f64<DPP_HMIRROR>(v));
  v = fmax(v, dpp_f64<DPP_MIRROR>(v));
  return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)),
              fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}
LRA_DEV int wave_min_i32(int v) {
  v = min(v, dpp_i32<DPP_XOR1>(v));
  v = min(v, dpp_i32<DPP_XOR2>(v));
  v = min(v, dpp_i32<DPP_HMIRROR>(v));
  v = min(v, dpp_i32<DPP_MIRROR>(v));
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
// Wave argmax with idamax's rule (largest value, then lowest key); lanes with value < 0 take
// no part.  Returns the winning lane or -1; `width` lanes may share one (value, key) pair.
LRA_DEV int wave_argmax(double v, int key, int width, double* vmax) {
  const double mx = wave_max(v);
  *vmax = mx;
  if (mx < 0.0) return -1;
  uint64_t msk = __ballot(v == mx);
  if (__popcll(msk) > (unsigned)width) {  // a tie between different keys: lowest key wins
    const int kmin = wave_min_i32(v == mx ? key : INT_MAX);
    msk = __ballot(v == mx && key == kmin);
  }
  return __builtin_amdgcn_readfirstlane(__ffsll((long long)msk) - 1);
}
