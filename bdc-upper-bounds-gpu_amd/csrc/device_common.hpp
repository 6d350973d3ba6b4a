// Device helpers shared by the kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace bdc {

constexpr int kPartStrideDev = 4;

// ---- symmetry group {id, reverse, complement, reverse∘complement} ----
// Same group and representative rule as backend/BAA.cu:195-258; written with
// the hardware bit-reverse instruction instead of the mask ladder.
__host__ __device__ inline uint32_t rev_bits(uint32_t y, int k) {
#if defined(__HIP_DEVICE_COMPILE__)
  return k == 0 ? 0u : (__builtin_bitreverse32(y) >> (32 - k));
#else
  uint32_t r = 0;
  for (int i = 0; i < k; ++i) r |= ((y >> i) & 1u) << (k - 1 - i);
  return r;
#endif
}
__host__ __device__ inline uint32_t comp_bits(uint32_t y, int k) {
  return k >= 32 ? ~y : (y ^ ((1u << k) - 1u));
}
__host__ __device__ inline uint32_t rep_of(uint32_t x, int n) {
  uint32_t r = rev_bits(x, n), c = comp_bits(x, n), rc = rev_bits(c, n);
  uint32_t m1 = r < x ? r : x, m2 = c < rc ? c : rc;
  return m2 < m1 ? m2 : m1;
}
__host__ __device__ inline int orbit_size_of(uint32_t x, int n) {
  uint32_t r = rev_bits(x, n), rc = rev_bits(comp_bits(x, n), n);
  return (x == r || x == rc) ? 2 : 4;
}

// Storage index of input x when the columns of P are kept in bit-reversed
// order (the low cr bits of x reversed; cr = 0: natural order).  An
// involution: it also maps a storage index back to its input.
__host__ __device__ inline uint64_t store_index(uint64_t x, int cr) {
  if (cr == 0) return x;
  const uint64_t m = (1ull << cr) - 1ull;
  return (x & ~m) | rev_bits((uint32_t)(x & m), cr);
}

// rank of a representative r among all representatives (ascending), from the
// representative bitmask and its per-word exclusive popcount prefix.
__device__ inline uint32_t rank_of_rep(uint32_t r, const uint64_t *mask, const uint32_t *wordbase) {
  uint32_t w = r >> 6, b = r & 63u;
  uint64_t below = b ? (mask[w] & ((~0ull) >> (64 - b))) : 0ull;
  return wordbase[w] + (uint32_t)__popcll(below);
}

// ---- wave / block reductions (wave64) ----
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Block reduction of three values (sum, max, sum) for blockDim.x <= 1024; the
// result is valid in thread 0.  `sh` must hold 3 * 16 doubles.
__device__ inline void block_reduce3(double &s0, double &m1, double &s2, double *sh) {
  s0 = wave_sum(s0);
  m1 = wave_max(m1);
  s2 = wave_sum(s2);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) {
    sh[w] = s0;
    sh[16 + w] = m1;
    sh[32 + w] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = -INFINITY, c = 0.0;
    for (int i = 0; i < nw; ++i) {
      a += sh[i];
      b = fmax(b, sh[16 + i]);
      c += sh[32 + i];
    }
    s0 = a;
    m1 = b;
    s2 = c;
  }
}

__device__ inline double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int i = 0; i < nw; ++i) r += sh[i];
  __syncthreads();
  return r;  // valid in every thread
}

__device__ inline double block_max(double v, double *sh) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = -INFINITY;
  for (int i = 0; i < nw; ++i) r = fmax(r, sh[i]);
  __syncthreads();
  return r;
}

// Device scalar slots (ctx->scal, FP64).
enum : int {
  kScalPScale = 0,   // p = pscale * P
  kScalDRef = 1,     // exponent offset of the fused update
  kScalMu = 2,       // centring of ln q
  kScalIter = 3,     // iteration counter (as double)
  kScalLastMx = 4,   // last mx (bits)
  kScalVScale = 5,   // split: exponent (as double) of the V/W limb scale
  kScalDMax = 6,     // scratch: max D of the last D pass
  kScalSumW = 7,     // scratch
  kScalDualScale = 8, // 1 / sum_x exp(beta (D - Dmax)) of the last dual pass
  kScalOne = 9,      // constant 1.0
  kScalDualU = 10,   // U_beta of the last bdc_dual call
  kScalCount = 16,
};

}  // namespace bdc
