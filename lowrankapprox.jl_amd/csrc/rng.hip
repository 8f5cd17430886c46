// rng.hip -- on-device counter-based randomness for the sketches.
//
// Replaces the host RNG calls of the reference: crandn (src/util.jl:4) behind RandomGaussian
// (src/sketch.jl:91-110), srft_rand / rand(1:n,k) (src/sketch.jl:338-361), rand(1:m,k) of
// RandomSubset (:252) and randperm + crandn of SparseRandomGaussian (:575-579).  Every value
// is a pure function of (seed, stream, index), so any kernel can regenerate any element.
#include "lra_internal.h"

namespace {

__global__ void gaussian_fill_kernel(uint64_t seed, uint64_t stream, i64 rows, i64 cols,
                                     double* out, i64 ld) {
  const i64 total = rows * cols;
  const i64 npairs = (total + 1) / 2;
  for (i64 pr = blockIdx.x * (i64)blockDim.x + threadIdx.x; pr < npairs;
       pr += (i64)gridDim.x * blockDim.x) {
    double z0, z1;
    lra_normal2(seed, stream, (uint64_t)pr, &z0, &z1);
    const i64 e0 = 2 * pr, e1 = 2 * pr + 1;
    out[(e0 % rows) + (e0 / rows) * ld] = z0;
    if (e1 < total) out[(e1 % rows) + (e1 / rows) * ld] = z1;
  }
}

// The same matrix stored transposed: out[j + i*ld] = element (i, j) = normal #(j*rows + i).
// One thread per output element (consecutive threads write consecutive addresses); each
// evaluates the pair that holds its normal.
__global__ void gaussian_fill_t_kernel(uint64_t seed, uint64_t stream, i64 rows, i64 cols,
                                       double* out, i64 ld) {
  const i64 total = rows * cols;
  for (i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x; t < total;
       t += (i64)gridDim.x * blockDim.x) {
    const i64 j = t % cols, i = t / cols;
    const i64 e = j * rows + i;
    double z0, z1;
    lra_normal2(seed, stream, (uint64_t)(e >> 1), &z0, &z1);
    out[j + i * ld] = (e & 1) ? z1 : z0;
  }
}

__global__ void gaussian_fill_f32_kernel(uint64_t seed, uint64_t stream, i64 rows, i64 cols,
                                         float* out, i64 ld) {
  const i64 total = rows * cols;
  const i64 npairs = (total + 1) / 2;
  for (i64 pr = blockIdx.x * (i64)blockDim.x + threadIdx.x; pr < npairs;
       pr += (i64)gridDim.x * blockDim.x) {
    double z0, z1;
    lra_normal2(seed, stream, (uint64_t)pr, &z0, &z1);
    const i64 e0 = 2 * pr, e1 = 2 * pr + 1;
    out[(e0 % rows) + (e0 / rows) * ld] = (float)z0;
    if (e1 < total) out[(e1 % rows) + (e1 / rows) * ld] = (float)z1;
  }
}

// unbiased integer in [0, n) from 64 random bits (multiply-shift; bias < n / 2^64)
__device__ inline i64 bounded(uint64_t bits, i64 n) {
  return (i64)__umul64hi(bits, (uint64_t)n);
}

// SRFT: signs d[i] = +-1 with probability 1/2 (srft_rand, src/sketch.jl:338-344) and
// idx[j] uniform in 1..m with replacement (src/sketch.jl:361).
__global__ void srft_rand_kernel(uint64_t seed, uint64_t stream, i64 m, i64 ell, double* d,
                                 i64* idx) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < m + ell;
       i += (i64)gridDim.x * blockDim.x) {
    const uint64_t b = lra_bits64(seed, stream, (uint64_t)i);
    if (i < m) d[i] = (b >> 63) ? 1.0 : -1.0;
    else idx[i - m] = 1 + bounded(b, m);
  }
}

__global__ void subset_rand_kernel(uint64_t seed, uint64_t stream, i64 m, i64 ell, i64* r) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < ell;
       i += (i64)gridDim.x * blockDim.x)
    r[i] = 1 + bounded(lra_bits64(seed, stream, (uint64_t)i), m);
}

// Random permutation of 1..m as a bijective keyed hash: a 4-round Feistel network on the
// smallest power-of-four domain >= m with cycle walking (Black & Rogaway, "Ciphers with
// arbitrary finite domains", CT-RSA 2002).  r[i] = perm(i) + 1; s[i] standard normals.
__device__ inline uint32_t feistel_round(uint32_t x, uint32_t key, int r, uint32_t mask) {
  uint32_t h = x * 0x9E3779B1u ^ (key + 0x7F4A7C15u * (uint32_t)(r + 1));
  h ^= h >> 15; h *= 0x2C1B3C6Du; h ^= h >> 12; h *= 0x297A2D39u; h ^= h >> 15;
  return h & mask;
}
__device__ inline i64 feistel_perm(i64 i, i64 m, int half_bits, uint64_t seed, uint64_t stream) {
  const uint32_t mask = (1u << half_bits) - 1u;
  const uint32_t k0 = (uint32_t)(seed ^ (stream * 0x9E3779B97F4A7C15ull));
  const uint32_t k1 = (uint32_t)((seed >> 32) ^ (stream >> 7));
  uint64_t x = (uint64_t)i;
  do {
    uint32_t L = (uint32_t)(x >> half_bits) & mask, R = (uint32_t)x & mask;
    for (int r = 0; r < 4; ++r) {
      const uint32_t nl = R;
      R = L ^ feistel_round(R, (r & 1) ? k1 : k0, r, mask);
      L = nl;
    }
    x = ((uint64_t)L << half_bits) | R;
  } while ((i64)x >= m);
  return (i64)x;
}

__global__ void sprn_rand_kernel(uint64_t seed, uint64_t stream, i64 m, int half_bits,
                                 double* s, i64* r) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < m;
       i += (i64)gridDim.x * blockDim.x) {
    r[i] = feistel_perm(i, m, half_bits, seed, stream) + 1;
    s[i] = lra_normal1(seed, stream ^ 0xA5A5A5A5ull, (uint64_t)i);
  }
}

__global__ void fill_i64_kernel(i64* p, i64 n, i64 start) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n;
       i += (i64)gridDim.x * blockDim.x)
    p[i] = start + i;
}

__global__ void zero_u32_kernel(unsigned* p, i64 n) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n;
       i += (i64)gridDim.x * blockDim.x)
    p[i] = 0u;
}

inline int grid_for(i64 work, int block = 256) {
  i64 g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

}  // namespace

#define LRA_LAUNCH_CHECK()                                                   \
  do {                                                                       \
    hipError_t _e = hipGetLastError();                                       \
    if (_e != hipSuccess) {                                                  \
      lra_set_error("%s:%d launch: %s", __FILE__, __LINE__, hipGetErrorString(_e)); \
      return LRA_E_HIP;                                                      \
    }                                                                        \
  } while (0)

int lra_launch_gaussian_fill(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 rows,
                             i64 cols, double* out, i64 ld) {
  if (rows <= 0 || cols <= 0) return LRA_OK;
  hipLaunchKernelGGL(gaussian_fill_kernel, dim3(grid_for((rows * cols + 1) / 2)), dim3(256),
                     0, st, seed, stream_id, rows, cols, out, ld);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_gaussian_fill_t(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 rows,
                               i64 cols, double* out, i64 ld) {
  if (rows <= 0 || cols <= 0) return LRA_OK;
  hipLaunchKernelGGL(gaussian_fill_t_kernel, dim3(grid_for(rows * cols)), dim3(256), 0, st, seed,
                     stream_id, rows, cols, out, ld);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_gaussian_fill_f32(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 rows,
                                 i64 cols, float* out, i64 ld) {
  if (rows <= 0 || cols <= 0) return LRA_OK;
  hipLaunchKernelGGL(gaussian_fill_f32_kernel, dim3(grid_for((rows * cols + 1) / 2)),
                     dim3(256), 0, st, seed, stream_id, rows, cols, out, ld);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_srft_rand(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 m, i64 ell,
                         double* d, i64* idx) {
  hipLaunchKernelGGL(srft_rand_kernel, dim3(grid_for(m + ell)), dim3(256), 0, st, seed,
                     stream_id, m, ell, d, idx);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_subset_rand(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 m, i64 ell,
                           i64* r) {
  hipLaunchKernelGGL(subset_rand_kernel, dim3(grid_for(ell)), dim3(256), 0, st, seed,
                     stream_id, m, ell, r);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_sprn_rand(hipStream_t st, uint64_t seed, uint64_t stream_id, i64 m, double* s,
                         i64* r) {
  int bits = 1;
  while ((1ll << (2 * bits)) < m) ++bits;
  if (bits > 31) {
    lra_set_error("sprn: m=%lld too large for the permutation generator", (long long)m);
    return LRA_E_UNSUPPORTED;
  }
  hipLaunchKernelGGL(sprn_rand_kernel, dim3(grid_for(m)), dim3(256), 0, st, seed, stream_id,
                     m, bits, s, r);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_fill_i64(hipStream_t st, i64* p, i64 n, i64 start) {
  if (n <= 0) return LRA_OK;
  hipLaunchKernelGGL(fill_i64_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, n, start);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}

int lra_launch_zero_u32(hipStream_t st, unsigned* p, i64 n) {
  if (n <= 0) return LRA_OK;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, n);
  LRA_LAUNCH_CHECK();
  return LRA_OK;
}
