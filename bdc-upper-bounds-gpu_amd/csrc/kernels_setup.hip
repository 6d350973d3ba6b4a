// Setup kernels: subsequence-count caches, stacked count operands,
// complement/reversal orbits and ranks, initial distributions, and the hoisted
// entropy term H(x)/C - ln C.  All run once per context.
#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "device_common.hpp"
#include "fast_log.hpp"
#include "kernels.hpp"

namespace bdc {

// ---------------------------------------------------------------- caches
// c(x,y) by the O(nk) prefix DP of backend/generate_bit_transition_cache.cpp:23-44
// (dp[j] = ways to form y[0..j) from the prefix of x read so far), one thread
// per (x,y) pair, the DP row held in registers.  Output index (x<<k)|y
// (cache_io.h layout).  n <= 20.
template <typename T>
__global__ void __launch_bounds__(256) k_count_cache(int n, int k, T *__restrict__ out) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (1ull << (n + k))) return;
  const uint32_t x = (uint32_t)(idx >> k), y = (uint32_t)(idx & ((1ull << k) - 1));
  uint32_t dp[21];
#pragma unroll
  for (int j = 0; j <= 20; ++j) dp[j] = (j == 0);
  for (int i = 1; i <= n; ++i) {
    const uint32_t xi = (x >> (n - i)) & 1u;
#pragma unroll
    for (int j = 20; j >= 1; --j) {
      if (j <= k && j <= i) {
        const uint32_t yj = (y >> (k - j)) & 1u;
        dp[j] += (yj == xi) ? dp[j - 1] : 0u;
      }
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int j = 0; j <= 20; ++j)
    if (j == k) r = dp[j];
  out[idx] = (T)r;
}

int launch_count_cache_u32(int n, int k, uint32_t *out, hipStream_t s) {
  const uint64_t tot = 1ull << (n + k);
  hipLaunchKernelGGL(k_count_cache<uint32_t>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, k, out);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}
int launch_count_cache_u16(int n, int k, uint16_t *out, hipStream_t s) {
  const uint64_t tot = 1ull << (n + k);
  hipLaunchKernelGGL(k_count_cache<uint16_t>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, k, out);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// Stacked operand: dst[r * ldd + off + c] = src[r * (1<<m) + c] for c < 2^m,
// r < R (a count cache copied into its column block).
__global__ void k_stack_cols(const uint16_t *__restrict__ src, int m, uint64_t R, uint16_t *__restrict__ dst,
                             int ldd, int off) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t w = 1ull << m;
  if (i >= R * w) return;
  const uint64_t r = i >> m, c = i & (w - 1);
  dst[r * ldd + off + c] = src[i];
}
int launch_stack_cols(const uint16_t *src, int m, uint64_t R, uint16_t *dst, int ldd, int off, hipStream_t s) {
  const uint64_t tot = R << m;
  hipLaunchKernelGGL(k_stack_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, src, m, R, dst, ldd, off);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// fp16 copy of a u16 count matrix (R x C row-major), optionally transposed
// (dst[c * R + r]).  Counts above 2048 that fp16 cannot hold are rounded to
// nearest-even here; the exact difference is applied by the residual lists.
__global__ void k_counts_to_half(const uint16_t *__restrict__ src, uint64_t R, uint64_t C, int transpose,
                                 _Float16 *__restrict__ dst) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R * C) return;
  const uint64_t r = i / C, c = i % C;
  const _Float16 v = (_Float16)(float)src[i];
  if (transpose)
    dst[c * R + r] = v;
  else
    dst[i] = v;
}
int launch_counts_to_half(const uint16_t *src, uint64_t R, uint64_t C, int transpose, _Float16 *dst, hipStream_t s) {
  const uint64_t tot = R * C;
  hipLaunchKernelGGL(k_counts_to_half, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, src, R, C, transpose, dst);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// ---------------------------------------------------------------- orbits
// Bitmask of representatives (rep(x) == x, backend/main_BAA.cu:46-50) and the
// popcount of every 64-bit word.
__global__ void k_repmask(int n, uint64_t nwords, uint64_t *__restrict__ mask, uint32_t *__restrict__ pop) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwords) return;
  const uint64_t nx = 1ull << n;
  uint64_t m = 0;
  for (int b = 0; b < 64; ++b) {
    const uint64_t x = w * 64 + b;
    if (x < nx && rep_of((uint32_t)x, n) == (uint32_t)x) m |= 1ull << b;
  }
  mask[w] = m;
  pop[w] = (uint32_t)__popcll(m);
}

// Exclusive scan of uint32 (totals < 2^32): three phases over blocks of 1024
// elements (256 threads x 4).
__global__ void __launch_bounds__(256) k_scan_partial(const uint32_t *__restrict__ in, uint64_t len,
                                                      uint32_t *__restrict__ out, uint32_t *__restrict__ bsum) {
  __shared__ uint32_t sh[256];
  const uint64_t base = (uint64_t)blockIdx.x * 1024 + threadIdx.x * 4;
  uint32_t v[4], s = 0;
  for (int i = 0; i < 4; ++i) {
    v[i] = (base + i < len) ? in[base + i] : 0u;
    s += v[i];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    uint32_t t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = sh[threadIdx.x] - s;  // exclusive prefix of this thread
  for (int i = 0; i < 4; ++i) {
    if (base + i < len) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = sh[255];
}
__global__ void __launch_bounds__(256) k_scan_blocks(uint32_t *__restrict__ bsum, uint64_t nb, uint32_t *__restrict__ total) {
  // single block, sequential over chunks of 256
  __shared__ uint32_t sh[256];
  uint32_t carry = 0;
  for (uint64_t c = 0; c < nb; c += 256) {
    const uint64_t i = c + threadIdx.x;
    const uint32_t v = i < nb ? bsum[i] : 0u;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      uint32_t t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + sh[threadIdx.x] - v;
    const uint32_t blk = sh[255];
    __syncthreads();
    carry += blk;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}
__global__ void k_scan_add(uint32_t *__restrict__ out, uint64_t len, const uint32_t *__restrict__ bsum) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) out[i] += bsum[i / 1024];
}

int launch_exclusive_scan(const uint32_t *in, uint64_t len, uint32_t *out, uint32_t *tmp_bsum, uint32_t *total,
                          hipStream_t s) {
  const uint64_t nb = (len + 1023) / 1024;
  hipLaunchKernelGGL(k_scan_partial, dim3((unsigned)nb), dim3(256), 0, s, in, len, out, tmp_bsum);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(256), 0, s, tmp_bsum, nb, total);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, out, len, tmp_bsum);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

int launch_repmask(int n, uint64_t nwords, uint64_t *mask, uint32_t *pop, hipStream_t s) {
  hipLaunchKernelGGL(k_repmask, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, n, nwords, mask, pop);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// ascending representative list: word w writes its set bits at wordbase[w] + i
__global__ void k_orbit_list(uint64_t nwords, const uint64_t *__restrict__ mask, const uint32_t *__restrict__ wordbase,
                             uint32_t *__restrict__ orbits) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwords) return;
  uint64_t m = mask[w];
  uint32_t o = wordbase[w];
  while (m) {
    const int b = __ffsll((unsigned long long)m) - 1;
    orbits[o++] = (uint32_t)(w * 64 + b);
    m &= m - 1;
  }
}
int launch_orbit_list(uint64_t nwords, const uint64_t *mask, const uint32_t *wordbase, uint32_t *orbits, hipStream_t s) {
  hipLaunchKernelGGL(k_orbit_list, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, nwords, mask, wordbase,
                     orbits);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// ---------------------------------------------------------------- distributions
// 1-D grids: at most 2^20 blocks of 256 (a dispatch holds fewer than 2^32
// work-items per dimension, and 2^n = 2^32 inputs at n = 32); kernels loop.
static unsigned grid_for(uint64_t len) { return (unsigned)std::min<uint64_t>(1u << 20, (len + 255) / 256); }

// P[x] = val (uniform, backend/main_BAA.cu:64)
__global__ void k_fill(double *__restrict__ P, uint64_t len, double val) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x)
    P[i] = val;
}
int launch_fill(double *P, uint64_t len, double val, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(len)), dim3(256), 0, s, P, len, val);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// full-space expansion of an orbit-indexed array: P[s] = scale * Q[rank(rep(x))] + add
// for the input x stored at s (store_index with cr column bits reversed)
__global__ void k_expand_orbits(int n, const double *__restrict__ Q, const uint64_t *__restrict__ mask,
                                const uint32_t *__restrict__ wordbase, double scale, double add, int cr,
                                double *__restrict__ P) {
  const uint64_t nx = 1ull << n;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nx; s += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = rep_of((uint32_t)store_index(s, cr), n);
    P[s] = scale * Q[rank_of_rep(r, mask, wordbase)] + add;
  }
}
int launch_expand_orbits(int n, const double *Q, const uint64_t *mask, const uint32_t *wordbase, double scale,
                         double add, int cr, double *P, hipStream_t s) {
  hipLaunchKernelGGL(k_expand_orbits, dim3(grid_for(1ull << n)), dim3(256), 0, s, n, Q, mask, wordbase, scale, add, cr,
                     P);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// Q[i] = scale * P[store(orbits[i])]  (scale read from a device scalar)
__global__ void k_gather_orbits(uint64_t norb, const uint32_t *__restrict__ orbits, const double *__restrict__ P,
                                const double *__restrict__ scal, int cr, double *__restrict__ Q) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < norb) Q[i] = scal[kScalPScale] * P[store_index(orbits[i], cr)];
}
int launch_gather_orbits(uint64_t norb, const uint32_t *orbits, const double *P, const double *scal, int cr, double *Q,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_gather_orbits, dim3((unsigned)((norb + 255) / 256)), dim3(256), 0, s, norb, orbits, P, scal,
                     cr, Q);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// out[i] = scale * src[store(idx[i])], scale = scal[slot] (1 when slot < 0)
__global__ void k_gather_at(const double *__restrict__ src, const uint32_t *__restrict__ idx, uint64_t len,
                            const double *__restrict__ scal, int slot, int cr, double *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) out[i] = (slot >= 0 ? scal[slot] : 1.0) * src[store_index(idx[i], cr)];
}
int launch_gather_at(const double *src, const uint32_t *idx, uint64_t len, const double *scal, int slot, int cr,
                     double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_at, dim3(grid_for(len)), dim3(256), 0, s, src, idx, len, scal, slot, cr, out);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// out[r][c] = in[r][rev(c)] over rows of 2^h doubles (h <= 14): one row per
// block, staged in LDS so both the reads and the writes are contiguous.  With
// rev an involution this both applies and undoes the column order.
__global__ void __launch_bounds__(1024) k_permute_cols(const double *__restrict__ in, double *__restrict__ out, int h) {
  extern __shared__ double row[];
  const uint32_t w = 1u << h;
  const double *src = in + (uint64_t)blockIdx.x * w;
  double *dst = out + (uint64_t)blockIdx.x * w;
  for (uint32_t c = threadIdx.x; c < w; c += blockDim.x) row[c] = src[c];
  __syncthreads();
  for (uint32_t c = threadIdx.x; c < w; c += blockDim.x) dst[c] = row[rev_bits(c, h)];
}
int launch_permute_cols(const double *in, double *out, uint64_t rows, int h, hipStream_t s) {
  if (h < 1 || h > 14 || rows == 0 || rows > 0x7fffffffull) return BDC_ERR_ARG;
  hipLaunchKernelGGL(k_permute_cols, dim3((unsigned)rows), dim3(1024), (size_t)8 << h, s, in, out, h);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// out[r][:] = in[rev(r)][:] for 2^h rows of `width` u16
__global__ void k_permute_rows_u16(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int h, int width) {
  const uint64_t tot = ((uint64_t)1 << h) * width;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = (uint32_t)(i / width), c = (uint32_t)(i % width);
    out[i] = in[(uint64_t)rev_bits(r, h) * width + c];
  }
}
int launch_permute_rows_u16(const uint16_t *in, uint16_t *out, int h, int width, hipStream_t s) {
  hipLaunchKernelGGL(k_permute_rows_u16, dim3(grid_for(((uint64_t)1 << h) * width)), dim3(256), 0, s, in, out, h, width);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// Counter-based generator: splitmix64 of (seed, stream, counter).
__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double u01(uint64_t seed, uint64_t i, uint64_t c) {
  const uint64_t z = mix64(mix64(seed ^ mix64(i)) + c);
  return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);  // (0,1)
}
// Gamma(alpha) by Marsaglia–Tsang (alpha >= 1) with the U^{1/alpha} boost for
// alpha < 1; normals by Box–Muller.  The rejection loop is bounded (128 tries;
// acceptance is > 95% per try for every alpha).
__device__ double gamma_sample(double alpha, uint64_t seed, uint64_t i) {
  double boost = 1.0;
  uint64_t ctr = 0;
  if (alpha < 1.0) {
    boost = pow(u01(seed, i, ctr++), 1.0 / alpha);
    alpha += 1.0;
  }
  const double d = alpha - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
  for (int t = 0; t < 128; ++t) {
    const double u1 = u01(seed, i, ctr++), u2 = u01(seed, i, ctr++);
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    double v = 1.0 + c * z;
    if (v <= 0.0) continue;
    v = v * v * v;
    const double u = u01(seed, i, ctr++);
    if (log(u) < 0.5 * z * z + d - d * v + d * log(v)) return boost * d * v;
  }
  return boost * d;
}
// Q[i] = Gamma_i / |O_i| (normalised by the caller)
__global__ void k_dirichlet_orbits(int n, uint64_t norb, const uint32_t *__restrict__ orbits, double alpha,
                                   uint64_t seed, double *__restrict__ Q, double *__restrict__ mass) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= norb) return;
  const double g = gamma_sample(alpha, seed, i);
  mass[i] = g;
  Q[i] = g / (double)orbit_size_of(orbits[i], n);
}
int launch_dirichlet_orbits(int n, uint64_t norb, const uint32_t *orbits, double alpha, uint64_t seed, double *Q,
                            double *mass, hipStream_t s) {
  hipLaunchKernelGGL(k_dirichlet_orbits, dim3((unsigned)((norb + 255) / 256)), dim3(256), 0, s, n, norb, orbits, alpha,
                     seed, Q, mass);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// deterministic block-partial sums then one block (for setup-time totals)
__global__ void __launch_bounds__(256) k_sum_partial(const double *__restrict__ v, uint64_t len,
                                                     double *__restrict__ part) {
  __shared__ double sh[16];
  double s = 0.0;
  const uint64_t per = (len + gridDim.x - 1) / gridDim.x;
  const uint64_t b = (uint64_t)blockIdx.x * per, e = b + per < len ? b + per : len;
  for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) s += v[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void __launch_bounds__(256) k_sum_final(const double *__restrict__ part, int np, double *__restrict__ out) {
  __shared__ double sh[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) *out = s;
}
int launch_sum(const double *v, uint64_t len, double *part, int np, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_partial, dim3(np), dim3(256), 0, s, v, len, part);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, part, np, out);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// ---------------------------------------------------------------- entropy term
// H(x) = sum_y c(x,y) ln c(x,y) for every representative x (the per-pair
// W ln W of backend/BAA.cu:401-411, hoisted out of the iteration), with
//   c(x, y) = sum_j c(x1, y_hi(j)) c(x2, y_lo(k-j))      (backend/BAA.cu:355-392).
// The block for x1 holds the "row set" of x1, A[j][z] = c(x1, z) for |z| = j,
// and, for each representative x = (x1, x2) in turn, the row set B[m][z] =
// c(x2, z) of x2; then it sweeps the outputs y.  Row-set entries are built
// independently (no level-by-level pass) from two per-string half tables:
//   c(s, z) = sum_i c(s_hi, first i symbols of z) c(s_lo, last |z| - i symbols).
// When both row sets do not fit LDS (half length 16 with k >= 16, or 16/15
// with large k), each launch covers one quadrant of the outputs: the top bit
// of y fixed to t (A keeps only the entries z whose top bit is t) and the bottom
// bit fixed to b (B keeps only the entries whose bottom bit is b); the four
// launches add their sums in a fixed order.  A thread sweeps 4 outputs that
// differ in two adjacent bits, so for all but the last splits the A entry is
// shared and the 4 B entries are consecutive (one 8-byte LDS load).
struct EntropyArgs {
  int n, k, n1, n2;
  int rt, rb;        // restrict the A row set to y's top bit t / the B row set to y's bottom bit b
  int t, b;          // quadrant of this launch
  int jlo, jhi;      // split range (backend/BAA.cu:355-357)
  int aLo, aHi;      // stored levels of A (non-identity) and of B
  int bLo, bHi;
  int gLo;           // B levels [gLo, bHi] are built 8 entries per thread (none when gLo > bHi)
  int aWords, bWords;  // u16 entries of the A and B regions (multiples of 4)
  int groups;        // 4-output groups per quadrant
  uint32_t x1base;
  int accumulate;    // 0: write H, 1: add to H
  int htoff[9];      // byte offset of the half tables of strings of length l (l <= 8) in htab, -1: absent
};
// occurrences of z (length lz) as a subsequence of s (length ls), both MSB first, ls <= 8
__device__ inline uint32_t count_small(uint32_t s, int ls, uint32_t z, int lz) {
  uint32_t dp[9] = {1, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < ls; ++i) {
    const uint32_t si = (s >> (ls - 1 - i)) & 1u;
    for (int j = (i + 1 < lz ? i + 1 : lz); j >= 1; --j)
      if (((z >> (lz - j)) & 1u) == si) dp[j] += dp[j - 1];
  }
  return dp[lz];
}
// Level l of a row set: offset and entry count.  Full layout: [2^l, 2^(l+1));
// restricted (half of every level l >= 1): [2^(l-1), 2^l), level 0 at 0.
__device__ inline int lev_off(int l, int r) { return r ? (l ? 1 << (l - 1) : 0) : (1 << l); }
__device__ inline uint32_t lev_mask_top(int l, int r) { return r ? (l ? (1u << (l - 1)) - 1u : 0u) : (1u << l) - 1u; }
// Half tables of every string h of length l <= 8: row h of (2 << l) - 1 bytes,
// T[2^d - 1 + w] = c(h, w) for |w| = d <= l (counts <= C(8,4) = 70).  Built
// once per entropy pass; every row set of the pass reads them.
__global__ void k_half_tables(int l, uint8_t *__restrict__ tab) {
  const int per = (2 << l) - 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1 << l) * per) return;
  const int h = i / per, e = i % per, d = 31 - __clz(e + 1);
  tab[i] = (uint8_t)count_small((uint32_t)h, l, (uint32_t)(e + 1 - (1 << d)), d);
}
__device__ inline const uint8_t *half_row(const uint8_t *htab, const EntropyArgs &ea, int l, uint32_t h) {
  return htab + ea.htoff[l] + (size_t)h * ((2 << l) - 1);
}
// c(s, z), |z| = d < |s|, s = (s_hi of length a, s_lo of length bl).  The split
// points that use a half whole are identity factors (a string's only
// subsequence of its own length is itself), so those two terms are compares.
__device__ inline uint32_t count_halves(const uint8_t *Th, int a, uint32_t sh, const uint8_t *Tl, int bl, uint32_t sl,
                                        uint32_t z, int d) {
  uint32_t c = 0;
  int ilo = d - bl > 0 ? d - bl : 0, ihi = d < a ? d : a;
  if (ihi == a) {  // i = a: z's first a symbols must be s_hi
    const int r = d - a;
    if ((z >> r) == sh) c += Tl[(1 << r) - 1 + (z & ((1u << r) - 1u))];
    --ihi;
  }
  if (ilo == d - bl && ilo <= ihi) {  // d - i = bl: z's last bl symbols must be s_lo
    if ((z & ((1u << bl) - 1u)) == sl) c += Th[(1 << ilo) - 1 + (z >> bl)];
    ++ilo;
  }
  for (int i = ilo; i <= ihi; ++i) {
    const int r = d - i;
    c += __umul24(Th[(1 << i) - 1 + (z >> r)], Tl[(1 << r) - 1 + (z & ((1u << r) - 1u))]);
  }
  return c;
}
// Row-set levels [lo, hi] of s into R.  top = 1: restricted to entries whose
// top bit is `bit` (A side); top = 0: restricted to entries whose bottom bit is
// `bit` (B side); r = 0: unrestricted.
__device__ inline void build_levels(const uint8_t *Th, const uint8_t *Tl, uint32_t s, int L, int lo, int hi, int r,
                                    int top, uint32_t bit, uint16_t *R) {
  if (lo > hi) return;
  const int a = (L + 1) / 2, bl = L / 2;
  const uint32_t sh = s >> bl, sl = s & ((1u << bl) - 1u);
  const int f0 = lev_off(lo, r), f1 = lev_off(hi, r) + (r ? (hi ? 1 << (hi - 1) : 1) : (1 << hi));
  for (int f = f0 + threadIdx.x; f < f1; f += blockDim.x) {
    const int l = r ? (f ? 32 - __clz(f) : 0) : 31 - __clz(f);
    const uint32_t e = (uint32_t)(f - lev_off(l, r));
    uint32_t z = e;
    if (r && l) z = top ? ((bit << (l - 1)) | e) : ((e << 1) | bit);
    R[f] = (uint16_t)count_halves(Th, a, sh, Tl, bl, sl, z, l);
  }
}

typedef unsigned short ushort2_ __attribute__((ext_vector_type(2)));
// Per-representative copies of x2's half tables in LDS, read by build_b_grouped:
// thb[2^i - 1 + w] = c(s_hi, w) in both u16 halves, tlf[2^d - 1 + w] = c(s_lo, w),
// and tlr = c(s_lo, w) in the B row set's level layout (level d at lev_off(d, r),
// restricted to w's bottom bit when r).
struct BTables {
  uint32_t thb[512];
  uint16_t tlf[512];
  uint16_t tlr[512];
};
__device__ inline void stage_b_tables(const uint8_t *htab, const EntropyArgs &ea, uint32_t x2, int r, uint32_t bit,
                                      BTables &t) {
  const int a = (ea.n2 + 1) / 2, bl = ea.n2 / 2;
  const uint8_t *Th = half_row(htab, ea, a, x2 >> bl), *Tl = half_row(htab, ea, bl, x2 & ((1u << bl) - 1u));
  const int nh = (2 << a) - 1, nl = (2 << bl) - 1, nr = r ? 1 << bl : 2 << bl;
  for (int i = threadIdx.x; i < nh; i += blockDim.x) t.thb[i] = Th[i] * 0x10001u;
  for (int i = threadIdx.x; i < nl; i += blockDim.x) t.tlf[i] = Tl[i];
  for (int f = threadIdx.x; f < nr; f += blockDim.x) {
    if (!r && !f) continue;  // the full layout starts at entry 1
    const int d = r ? (f ? 32 - __clz(f) : 0) : 31 - __clz(f);
    const uint32_t e = (uint32_t)(f - lev_off(d, r));
    const uint32_t w = (r && d) ? ((e << 1) | bit) : e;
    t.tlr[f] = Tl[(1 << d) - 1 + w];
  }
}
// B row-set levels [lo, hi] (lo >= 4 when R, else >= 3: every level holds a
// multiple of 8 entries at a 16-byte aligned offset), 8 consecutive entries per
// thread.  With d = m - i low symbols taken from s_lo, the 8 entries share their
// high part when d >= 4 (R) / 3, and their low parts are 8 consecutive tlr
// entries: one 16-byte LDS load and four packed u16 multiply-adds per split.
// Every count of a string of length <= 16 is at most C(16, 8) = 12870 < 2^16, so
// the packed sums are exact.
template <int R>
__device__ inline void build_b_grouped(const BTables &t, int L, int lo, int hi, uint32_t bit, uint16_t *B) {
  constexpr int D0 = R ? 4 : 3;
  const int a = (L + 1) / 2, bl = L / 2;
  const int f0 = lev_off(lo, R), f1 = lev_off(hi, R) + (R ? 1 << (hi - 1) : 1 << hi);
  for (int f = f0 + 8 * (int)threadIdx.x; f < f1; f += 8 * (int)blockDim.x) {
    const int m = R ? 32 - __clz(f) : 31 - __clz(f);
    const uint32_t e0 = (uint32_t)(f - lev_off(m, R));
    const uint32_t z0 = R ? ((e0 << 1) | bit) : e0;
    ushort2_ acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = ushort2_{0, 0};
    const int ilo = m - bl > 0 ? m - bl : 0, ihi = m < a ? m : a;
    for (int i = ilo; i <= ihi; ++i) {
      const int d = m - i;
      if (d >= D0) {
        const ushort2_ th = __builtin_bit_cast(ushort2_, t.thb[(1 << i) - 1 + (z0 >> d)]);
        const uint4 v = *reinterpret_cast<const uint4 *>(t.tlr + lev_off(d, R) + ((z0 & ((1u << d) - 1u)) >> R));
        acc[0] = th * __builtin_bit_cast(ushort2_, v.x) + acc[0];
        acc[1] = th * __builtin_bit_cast(ushort2_, v.y) + acc[1];
        acc[2] = th * __builtin_bit_cast(ushort2_, v.z) + acc[2];
        acc[3] = th * __builtin_bit_cast(ushort2_, v.w) + acc[3];
      } else {
        for (int j = 0; j < 8; ++j) {
          const uint32_t z = z0 + ((uint32_t)j << R);
          const uint32_t p = (t.thb[(1 << i) - 1 + (z >> d)] & 0xffffu) * t.tlf[(1 << d) - 1 + (z & ((1u << d) - 1u))];
          acc[j >> 1][j & 1] += (unsigned short)p;
        }
      }
    }
    uint4 o;
    o.x = __builtin_bit_cast(uint32_t, acc[0]);
    o.y = __builtin_bit_cast(uint32_t, acc[1]);
    o.z = __builtin_bit_cast(uint32_t, acc[2]);
    o.w = __builtin_bit_cast(uint32_t, acc[3]);
    *reinterpret_cast<uint4 *>(B + f) = o;
  }
}

__device__ inline uint32_t half_of(const uint32_t *w, int i) { return (w[i >> 1] >> (16 * (i & 1))) & 0xffffu; }

// Counts of the 16 outputs y_e = ys | (e << R), e < 16 (ys has bits R..R+3
// clear) and their sum of c ln c.  R = 1: quadrant layout (A holds the entries
// whose top bit is the quadrant's, B those whose bottom bit is).
//  - splits with m = k - j >= R + 4: the A entry is shared by the 16 outputs
//    and their B entries are 16 consecutive u16 (two 16-byte LDS loads); two
//    splits go through one v_dot2_u32_u16 per output (A pair . B pair);
//  - splits with m < R + 4 (unrolled): the outputs' A entries are consecutive
//    (or every other one for m < R), B level m has at most 8 entries;
//  - the identity factors (y_hi = x1 or y_lo = x2) are compares.
template <int R>
__device__ inline double sweep16(const EntropyArgs &ea, const uint16_t *__restrict__ A, const uint16_t *__restrict__ B,
                                 uint32_t ys, uint32_t x1, uint32_t x2, const double *lnF, const double *invF) {
  const int k = ea.k, n1 = ea.n1, n2 = ea.n2;
  const bool idA = k >= n1, idB = k >= n2;
  const int j0 = ea.jlo + (idB ? 1 : 0), j1 = ea.jhi - (idA ? 1 : 0);
  uint32_t c[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) c[e] = 0u;
  const int jb = min(j1, k - R - 4);
  int j = j0;
  for (; j + 1 <= jb; j += 2) {
    const int m = k - j;
    const uint32_t a0 = A[lev_off(j, R) + ((ys >> m) & lev_mask_top(j, R))];
    const uint32_t a1 = A[lev_off(j + 1, R) + ((ys >> (m - 1)) & lev_mask_top(j + 1, R))];
    const ushort2_ ap = __builtin_bit_cast(ushort2_, a0 | (a1 << 16));
    const uint4 *p0 = reinterpret_cast<const uint4 *>(B + lev_off(m, R) + ((ys & ((1u << m) - 1u)) >> R));
    const uint4 *p1 = reinterpret_cast<const uint4 *>(B + lev_off(m - 1, R) + ((ys & ((1u << (m - 1)) - 1u)) >> R));
    const uint4 u0 = p0[0], u1 = p0[1], v0 = p1[0], v1 = p1[1];
    const uint32_t w0[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
    const uint32_t w1[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      // low half: B_m(y_e) (split j), high half: B_{m-1}(y_e) (split j + 1)
      const uint32_t pk = __builtin_amdgcn_perm(w1[e >> 1], w0[e >> 1], (e & 1) ? 0x07060302u : 0x05040100u);
      c[e] = __builtin_amdgcn_udot2(ap, __builtin_bit_cast(ushort2_, pk), c[e], false);
    }
  }
  if (j <= jb) {
    const int m = k - j;
    const uint32_t a = A[lev_off(j, R) + ((ys >> m) & lev_mask_top(j, R))];
    const uint4 *p0 = reinterpret_cast<const uint4 *>(B + lev_off(m, R) + ((ys & ((1u << m) - 1u)) >> R));
    const uint4 u0 = p0[0], u1 = p0[1];
    const uint32_t w0[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
    for (int e = 0; e < 16; ++e) c[e] += __umul24(a, half_of(w0, e));
  }
#pragma unroll
  for (int mm = 0; mm < R + 4; ++mm) {
    const int jj = k - mm;
    if (jj < j0 || jj > j1 || jj <= jb) continue;
    const int abase = lev_off(jj, R) + (int)((ys >> mm) & lev_mask_top(jj, R));
    uint32_t aw[8], bw[4];
    // A entries of the 16 outputs: abase + ((e << R) >> mm)
    if (mm < R) {
#pragma unroll
      for (int i = 0; i < 8; ++i) aw[i] = (uint32_t)A[abase + ((2 * i) << R)] | ((uint32_t)A[abase + ((2 * i + 1) << R)] << 16);
    } else if (mm == R) {
      const uint4 *pa = reinterpret_cast<const uint4 *>(A + abase);
      const uint4 q0 = pa[0], q1 = pa[1];
      aw[0] = q0.x; aw[1] = q0.y; aw[2] = q0.z; aw[3] = q0.w; aw[4] = q1.x; aw[5] = q1.y; aw[6] = q1.z; aw[7] = q1.w;
    } else if (mm == R + 1) {
      const uint4 q0 = *reinterpret_cast<const uint4 *>(A + abase);
      aw[0] = q0.x; aw[1] = q0.y; aw[2] = q0.z; aw[3] = q0.w;
    } else if (mm == R + 2) {
      const uint2 q0 = *reinterpret_cast<const uint2 *>(A + abase);
      aw[0] = q0.x; aw[1] = q0.y;
    } else {
      aw[0] = *reinterpret_cast<const uint32_t *>(A + abase);
    }
    // B level mm (restricted): entry e mod 2^(mm - R), or its single entry
    const int boff = lev_off(mm, R);
    if (mm <= R) {
      bw[0] = B[boff];
    } else if (mm == R + 1) {
      bw[0] = *reinterpret_cast<const uint32_t *>(B + boff);
    } else if (mm == R + 2) {
      const uint2 q = *reinterpret_cast<const uint2 *>(B + boff);
      bw[0] = q.x; bw[1] = q.y;
    } else {
      const uint4 q = *reinterpret_cast<const uint4 *>(B + boff);
      bw[0] = q.x; bw[1] = q.y; bw[2] = q.z; bw[3] = q.w;
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const uint32_t av = mm < R ? half_of(aw, e) : half_of(aw, e >> (mm - R));
      const uint32_t bv = mm <= R ? bw[0] : half_of(bw, e & ((1 << (mm - R)) - 1));
      c[e] += __umul24(av, bv);
    }
  }
  double h = 0.0;
  const uint32_t m2 = (1u << n2) - 1u;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const uint32_t y = ys | ((uint32_t)e << R);
    uint32_t cy = c[e];
    if (idA && (y >> (k - n1)) == x1) {
      const int m = k - n1;
      cy += B[lev_off(m, R) + ((y & ((1u << m) - 1u)) >> R)];
    }
    if (idB && (y & m2) == x2) cy += A[lev_off(ea.jlo, R) + ((y >> n2) & lev_mask_top(ea.jlo, R))];
    if (cy > 1u) h = fma((double)cy, log_tang_int(cy, lnF, invF), h);
  }
  return h;
}

// WAVES_PER_EU 8: two 1024-thread blocks per CU (64 VGPRs) where their LDS fits,
// 4: one block per CU with up to 128 VGPRs.
template <int WAVES_PER_EU>
__global__ void __launch_bounds__(1024, WAVES_PER_EU) k_entropy_reps(EntropyArgs ea, const uint64_t *__restrict__ mask,
                                                      const uint32_t *__restrict__ wordbase,
                                                      const double *__restrict__ logtab,
                                                      const uint8_t *__restrict__ htab, double *__restrict__ Hrep) {
  extern __shared__ __attribute__((aligned(16))) uint16_t lds16[];
  __shared__ double hw[2][16];  // per-wave sums of the last two representatives
  __shared__ double lt[130];
  uint16_t *A = lds16, *B = lds16 + ea.aWords;
  const int n = ea.n, k = ea.k, n1 = ea.n1, n2 = ea.n2, rt = ea.rt, rb = ea.rb;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  for (int i = threadIdx.x; i < 130; i += blockDim.x) lt[i] = logtab[i];
  const uint32_t x1 = ea.x1base + blockIdx.x;
  const int a1 = (n1 + 1) / 2, b1 = n1 / 2, a2 = (n2 + 1) / 2, b2 = n2 / 2;
  build_levels(half_row(htab, ea, a1, x1 >> b1), half_row(htab, ea, b1, x1 & ((1u << b1) - 1u)), x1, n1, ea.aLo,
               ea.aHi, rt, 1, (uint32_t)ea.t, A);
  const double *lnF = lt, *invF = lt + 65;
  const bool idA = k >= n1, idB = k >= n2;  // identity factors: j = n1 (y_hi = x1), k - j = n2 (y_lo = x2)
  const int j0 = ea.jlo + (idB ? 1 : 0), j1 = ea.jhi - (idA ? 1 : 0);
  const int jb1 = min(j1, k - rb - 2);  // bulk splits: A entry shared by the 4 outputs, B entries consecutive
  const int je0 = max(j0, k - rb - 1);  // edge splits: m = k - j < rb + 2
  const uint32_t N2 = 1u << n2, ky = 1u << k, m2 = N2 - 1u;
  const uint32_t ybase = (rt ? ((uint32_t)ea.t << (k - 1)) : 0u) | (rb ? (uint32_t)ea.b : 0u);
  // The sum of a representative is reduced per wave into hw[parity] and
  // combined by thread 0 after the next barrier (one representative later),
  // so a representative costs two barriers.
  uint32_t prev = 0xffffffffu;
  int par = 0;
  __shared__ __attribute__((aligned(16))) BTables btab[2];  // x2's tables: this representative's and the next's
  const bool grouped = ea.gLo <= ea.bHi;
  auto next_rep = [&](uint32_t from) {  // uniform across the block
    for (uint32_t y = from; y < N2; ++y)
      if (rep_of((x1 << n2) | y, n) == ((x1 << n2) | y)) return y;
    return N2;
  };
  uint32_t x2 = next_rep(0);
  int tb = 0;
  if (grouped && x2 < N2) stage_b_tables(htab, ea, x2, rb, (uint32_t)ea.b, btab[0]);
  auto flush = [&](int p) {
    if (threadIdx.x == 0 && prev != 0xffffffffu) {
      double h = 0.0;
      for (int w = 0; w < nw; ++w) h += hw[p][w];
      const uint32_t r = rank_of_rep(prev, mask, wordbase);
      Hrep[r] = ea.accumulate ? Hrep[r] + h : h;
    }
  };
  while (x2 < N2) {
    const uint32_t x = (x1 << n2) | x2;
    __syncthreads();  // A built / previous sweep done with B; hw[par ^ 1] and btab[tb] complete
    flush(par ^ 1);
    build_levels(half_row(htab, ea, a2, x2 >> b2), half_row(htab, ea, b2, x2 & ((1u << b2) - 1u)), x2, n2, ea.bLo,
                 grouped ? ea.gLo - 1 : ea.bHi, rb, 0, (uint32_t)ea.b, B);
    if (grouped) {
      if (rb) build_b_grouped<1>(btab[tb], n2, ea.gLo, ea.bHi, (uint32_t)ea.b, B);
      else build_b_grouped<0>(btab[tb], n2, ea.gLo, ea.bHi, (uint32_t)ea.b, B);
    }
    __syncthreads();
    const uint32_t nx2 = next_rep(x2 + 1);
    if (grouped && nx2 < N2) stage_b_tables(htab, ea, nx2, rb, (uint32_t)ea.b, btab[tb ^ 1]);  // read after the next barrier
    double h = 0.0;
    if (ea.groups >= 4) {  // 16 outputs per thread (k >= 2 rb + 4)
      const int nsg = ea.groups >> 2;
      for (int sg = threadIdx.x; sg < nsg; sg += blockDim.x) {
        const uint32_t ys = ybase | ((uint32_t)sg << (4 + rb));
        h += rb ? sweep16<1>(ea, A, B, ys, x1, x2, lnF, invF) : sweep16<0>(ea, A, B, ys, x1, x2, lnF, invF);
      }
    } else
    for (int g = threadIdx.x; g < ea.groups; g += blockDim.x) {
      const uint32_t y0 = ybase | ((uint32_t)g << (2 + rb));
      uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
      for (int j = j0; j <= jb1; ++j) {
        const int m = k - j;
        const uint32_t a = A[lev_off(j, rt) + ((y0 >> m) & lev_mask_top(j, rt))];
        const uint2 bv = *reinterpret_cast<const uint2 *>(B + lev_off(m, rb) + ((y0 & ((1u << m) - 1u)) >> rb));
        c0 += __umul24(a, bv.x & 0xffffu);
        c1 += __umul24(a, bv.x >> 16);
        c2 += __umul24(a, bv.y & 0xffffu);
        c3 += __umul24(a, bv.y >> 16);
      }
      uint32_t cc[4] = {c0, c1, c2, c3};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t y = y0 | ((uint32_t)i << rb);
        if (y >= ky) continue;
        uint32_t c = cc[i];
        for (int j = je0; j <= j1; ++j) {
          const int m = k - j;
          c += __umul24(A[lev_off(j, rt) + ((y >> m) & lev_mask_top(j, rt))],
                        B[lev_off(m, rb) + ((y & ((1u << m) - 1u)) >> rb)]);
        }
        if (idA && (y >> (k - n1)) == x1) {
          const int m = k - n1;
          c += B[lev_off(m, rb) + ((y & ((1u << m) - 1u)) >> rb)];
        }
        if (idB && (y & m2) == x2) c += A[lev_off(ea.jlo, rt) + ((y >> n2) & lev_mask_top(ea.jlo, rt))];
        if (c > 1u) h = fma((double)c, log_tang_int(c, lnF, invF), h);
      }
    }
    h = wave_sum(h);
    if (lane == 0) hw[par][wave] = h;
    prev = x;
    par ^= 1;
    x2 = nx2;
    tb ^= 1;
  }
  __syncthreads();
  flush(par ^ 1);
}
int launch_entropy_reps(int n, int k, int n1, int n2, const uint64_t *mask, const uint32_t *wordbase,
                        const double *logtab, double *Hrep, hipStream_t s) {
  if (n1 > 16 || n2 > 16 || k < 1 || k >= n) return BDC_ERR_ARG;
  EntropyArgs ea{};
  ea.n = n;
  ea.k = k;
  ea.n1 = n1;
  ea.n2 = n2;
  ea.jlo = std::max(0, k - n2);
  ea.jhi = std::min(k, n1);
  ea.aLo = ea.jlo;
  ea.aHi = std::min(ea.jhi, n1 - 1);
  ea.bLo = k - ea.jhi;
  ea.bHi = std::min(k - ea.jlo, n2 - 1);
  auto words = [](int hi, int r) {  // multiples of 8 entries: 16-byte aligned B region
    const int w = r ? (hi ? 1 << hi : 1) : (2 << hi);
    return (w + 7) & ~7;
  };
  // restrict to output quadrants only when the full row sets do not fit
  const size_t kLdsCap = 136 * 1024;
  int r = 0;
  if (2 * (size_t)(words(ea.aHi, 0) + words(ea.bHi, 0)) > kLdsCap) r = 1;
  const char *fq = getenv("BDC_ENTROPY_QUADRANTS");  // tests: force the quadrant layout (k >= 4)
  if (fq && fq[0] == '1' && k >= 4) r = 1;
  if (r && k < 4) return BDC_ERR_ARG;
  ea.rt = ea.rb = r;
  // B levels with at least 8 entries are built 8 per thread from LDS copies of
  // x2's half tables (BDC_ENTROPY_GROUPED=0: every entry on its own, for tests)
  const char *fg = getenv("BDC_ENTROPY_GROUPED");
  ea.gLo = (fg && fg[0] == '0') ? ea.bHi + 1 : std::max(ea.bLo, r ? 4 : 3);
  ea.aWords = words(ea.aHi, r);
  ea.bWords = words(ea.bHi, r);
  const size_t lds = 2 * (size_t)(ea.aWords + ea.bWords);
  if (lds > kLdsCap) return BDC_ERR_ARG;
  // half tables of the lengths the row sets need (built once, read by every block)
  for (int l = 0; l < 9; ++l) ea.htoff[l] = -1;
  size_t tbytes = 0;
  for (int L : {n1, n2})
    for (int l : {(L + 1) / 2, L / 2})
      if (ea.htoff[l] < 0) {
        ea.htoff[l] = (int)tbytes;
        tbytes += ((size_t)(1u << l) * ((2u << l) - 1u) + 15) & ~(size_t)15;
      }
  uint8_t *htab = nullptr;
  if (hipMalloc(&htab, tbytes) != hipSuccess) return BDC_ERR_MEMORY;
  for (int l = 0; l < 9; ++l)
    if (ea.htoff[l] >= 0) {
      const int tot = (1 << l) * ((2 << l) - 1);
      hipLaunchKernelGGL(k_half_tables, dim3((tot + 255) / 256), dim3(256), 0, s, l, htab + ea.htoff[l]);
    }
  const int free_bits = k - 2 * r - 2;
  ea.groups = free_bits > 0 ? 1 << free_bits : 1;
  const int tasks = ea.groups >= 4 ? ea.groups / 4 : ea.groups;  // 16 or 4 outputs per thread
  int threads = 64;
  while (threads < 1024 && threads < tasks) threads *= 2;
  // representatives have top bit 0 (x <= ~x), so only x1 < N1/2 holds any
  const char *v = getenv("BDC_VERBOSE");
  const bool verbose = v && v[0] == '1';
  // one launch per quadrant (low x1 rows hold the most representatives; the block
  // scheduler balances them against the light rows), or 16 chunks of rows when
  // progress is reported
  const uint32_t X1 = 1u << (n1 - 1), chunk = (verbose && X1 >= 16) ? X1 / 16 : X1;
  const int quads = r ? 4 : 1;
  // static LDS: two BTables, the per-wave sums and the log table
  const size_t kStatic = 2 * sizeof(BTables) + 2 * 16 * sizeof(double) + 130 * sizeof(double);
  constexpr size_t kLdsPerCu = 160 * 1024;  // gfx950
  const bool two_blocks = 2 * (lds + kStatic) <= kLdsPerCu;
  for (int q = 0; q < quads; ++q) {
    ea.t = q >> 1;
    ea.b = q & 1;
    ea.accumulate = q > 0;
    for (uint32_t b = 0; b < X1; b += chunk) {
      ea.x1base = b;
      if (two_blocks)
        hipLaunchKernelGGL(k_entropy_reps<8>, dim3(chunk), dim3(threads), lds, s, ea, mask, wordbase, logtab, htab, Hrep);
      else
        hipLaunchKernelGGL(k_entropy_reps<4>, dim3(chunk), dim3(threads), lds, s, ea, mask, wordbase, logtab, htab, Hrep);
      if (hipGetLastError() != hipSuccess) {
        (void)hipFree(htab);
        return BDC_ERR_HIP;
      }
      if (verbose) {
        if (hipStreamSynchronize(s) != hipSuccess) {
          (void)hipFree(htab);
          return BDC_ERR_HIP;
        }
        fprintf(stderr, "[bdc] entropy term BDC_(%d,%d): quadrant %d/%d, x1 rows %u/%u done\n", n, k, q + 1, quads,
                b + chunk, X1);
      }
    }
  }
  const bool ok = hipStreamSynchronize(s) == hipSuccess;  // the tables are in use until the pass ends
  (void)hipFree(htab);
  return ok ? BDC_OK : BDC_ERR_HIP;
}

}  // namespace bdc
