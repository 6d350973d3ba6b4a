// Setup kernels: subsequence-count caches, stacked count operands,
// complement/reversal orbits and ranks, initial distributions, and the hoisted
// entropy term H(x)/C - ln C.  All run once per context.
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
// P[x] = val (uniform, backend/main_BAA.cu:64)
__global__ void k_fill(double *__restrict__ P, uint64_t len, double val) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) P[i] = val;
}
int launch_fill(double *P, uint64_t len, double val, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, P, len, val);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// full-space expansion of an orbit-indexed array: P[x] = scale * Q[rank(rep(x))] + add
__global__ void k_expand_orbits(int n, const double *__restrict__ Q, const uint64_t *__restrict__ mask,
                                const uint32_t *__restrict__ wordbase, double scale, double add,
                                double *__restrict__ P) {
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (1ull << n)) return;
  const uint32_t r = rep_of((uint32_t)x, n);
  P[x] = scale * Q[rank_of_rep(r, mask, wordbase)] + add;
}
int launch_expand_orbits(int n, const double *Q, const uint64_t *mask, const uint32_t *wordbase, double scale,
                         double add, double *P, hipStream_t s) {
  const uint64_t tot = 1ull << n;
  hipLaunchKernelGGL(k_expand_orbits, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, Q, mask, wordbase,
                     scale, add, P);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// Q[i] = scale * P[orbits[i]]  (scale read from a device scalar)
__global__ void k_gather_orbits(uint64_t norb, const uint32_t *__restrict__ orbits, const double *__restrict__ P,
                                const double *__restrict__ scal, double *__restrict__ Q) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < norb) Q[i] = scal[kScalPScale] * P[orbits[i]];
}
int launch_gather_orbits(uint64_t norb, const uint32_t *orbits, const double *P, const double *scal, double *Q,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_gather_orbits, dim3((unsigned)((norb + 255) / 256)), dim3(256), 0, s, norb, orbits, P, scal,
                     Q);
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
// W ln W of backend/BAA.cu:401-411, hoisted out of the iteration).
// One block per x1.  The subsequence counts of a half string s (length L) for
// every z of length d < L (level L is the identity, kept implicit) form its
// "row set" R(s)[2^d - 1 + z]; every entry is built independently from quarter
// tables of the two halves of s (lengths a = ceil(L/2), b = floor(L/2)):
//   c(s, z) = sum_i c(s_hi, first i symbols of z) c(s_lo, last d - i symbols),
// so no level-by-level pass (and no barrier per level) is needed.  The block
// keeps R(x1) and, for each representative (x1, x2), builds R(x2) and sweeps all
// y in {0,1}^k with c(x, y) = sum_j c(x1, y_hi(j)) c(x2, y_lo(k-j)), adding
// c ln c with a table-driven integer log.
struct QuarterTabs {
  int off[8];  // byte offset in LDS of the table for strings of length l (0: absent)
};
// occurrences of z (length lz) as a subsequence of s (length ls), both MSB first, ls <= 7
__device__ inline uint32_t count_small(uint32_t s, int ls, uint32_t z, int lz) {
  uint32_t dp[8] = {1, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < ls; ++i) {
    const uint32_t si = (s >> (ls - 1 - i)) & 1u;
    for (int j = (i + 1 < lz ? i + 1 : lz); j >= 1; --j)
      if (((z >> (lz - j)) & 1u) == si) dp[j] += dp[j - 1];
  }
  return dp[lz];
}
// table for length l: [s < 2^l][2^d - 1 + z, d <= l] as u8 (counts <= C(7,3) = 35)
__device__ inline int qtab_bytes(int l) { return (1 << l) * ((2 << l) - 1); }
__device__ inline uint32_t qget(const uint8_t *T, int l, uint32_t s, int d, uint32_t z) {
  return T[s * ((2 << l) - 1) + (1 << d) - 1 + z];
}
// R(s)[2^d - 1 + z] for d < L
__device__ void build_rowset(uint32_t s, int L, const uint8_t *lds8, const QuarterTabs &qt, uint16_t *R) {
  const int a = (L + 1) / 2, b = L / 2;
  const uint8_t *Ta = lds8 + qt.off[a], *Tb = lds8 + qt.off[b];
  const uint32_t sh = s >> b, sl = s & ((1u << b) - 1u);
  const int tot = (1 << L) - 1;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    const int d = 31 - __clz(e + 1);
    const uint32_t z = (uint32_t)(e + 1 - (1 << d));
    uint32_t c = 0;
    const int ilo = d - b > 0 ? d - b : 0, ihi = d < a ? d : a;
    for (int i = ilo; i <= ihi; ++i)
      c += __umul24(qget(Ta, a, sh, i, z >> (d - i)), qget(Tb, b, sl, d - i, z & ((1u << (d - i)) - 1u)));
    R[e] = (uint16_t)c;
  }
}

__global__ void __launch_bounds__(1024) k_entropy_reps(int n, int k, int n1, int n2, const uint64_t *__restrict__ mask,
                                                      const uint32_t *__restrict__ wordbase, QuarterTabs qt,
                                                      int qbytes, const double *__restrict__ logtab, uint32_t x1base,
                                                      double *__restrict__ Hrep) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds8[];
  __shared__ double red[16];
  __shared__ double lt[130];
  // layout: quarter tables (qbytes, rounded to 16), R(x1), R(x2)
  uint16_t *A = reinterpret_cast<uint16_t *>(lds8 + ((qbytes + 15) & ~15));
  uint16_t *B = A + ((1 << n1) - 1);
  for (int i = threadIdx.x; i < 130; i += blockDim.x) lt[i] = logtab[i];
  // quarter tables
  for (int l = 0; l <= 7; ++l) {
    if (qt.off[l] < 0) continue;
    uint8_t *T = lds8 + qt.off[l];
    const int per = (2 << l) - 1, tot = (1 << l) * per;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int sv = e / per, r = e % per, d = 31 - __clz(r + 1);
      T[e] = (uint8_t)count_small((uint32_t)sv, l, (uint32_t)(r + 1 - (1 << d)), d);
    }
  }
  __syncthreads();
  const uint32_t x1 = x1base + blockIdx.x;
  build_rowset(x1, n1, lds8, qt, A);
  const int jlo = k - n2 > 0 ? k - n2 : 0, jhi = k < n1 ? k : n1;
  const uint32_t N2 = 1u << n2, ny = 1u << k;
  const double *lnF = lt, *invF = lt + 65;
  for (uint32_t x2 = 0; x2 < N2; ++x2) {
    const uint32_t x = (x1 << n2) | x2;
    if (rep_of(x, n) != x) continue;  // uniform across the block
    __syncthreads();                  // R(x1) built / previous sweep done with R(x2)
    build_rowset(x2, n2, lds8, qt, B);
    __syncthreads();
    // splits with an identity factor (j = n1: y_hi must equal x1; k - j = n2:
    // y_lo must equal x2) are handled outside the main loop
    const int j0 = (k - jlo == n2) ? jlo + 1 : jlo, j1 = (jhi == n1) ? jhi - 1 : jhi;
    double h = 0.0;
    for (uint32_t y = threadIdx.x; y < ny; y += blockDim.x) {
      uint32_t c = 0;
      for (int j = j0; j <= j1; ++j) {
        const int kj = k - j;
        // counts < 2^16: 24-bit products are exact (and full rate)
        c += __umul24(A[(1 << j) - 1 + (y >> kj)], B[(1 << kj) - 1 + (y & ((1u << kj) - 1u))]);
      }
      if (jhi == n1 && (y >> (k - n1)) == x1 && k - n1 != n2) c += B[(1 << (k - n1)) - 1 + (y & ((1u << (k - n1)) - 1u))];
      if (k - jlo == n2 && (y & ((1u << n2) - 1u)) == x2 && jlo != n1) c += A[(1 << jlo) - 1 + (y >> n2)];
      if (jhi == n1 && k - n1 == n2 && y == ((x1 << n2) | x2)) c += 1u;  // k = n: both factors are identities
      if (c > 1u) h = fma((double)c, log_tang_int(c, lnF, invF), h);
    }
    h = block_sum(h, red);
    if (threadIdx.x == 0) Hrep[rank_of_rep(x, mask, wordbase)] = h;
  }
}
int launch_entropy_reps(int n, int k, int n1, int n2, const uint64_t *mask, const uint32_t *wordbase,
                        const double *logtab, double *Hrep, hipStream_t s) {
  if (n1 > 14 || n2 > 14) return BDC_ERR_ARG;
  QuarterTabs qt;
  for (int l = 0; l < 8; ++l) qt.off[l] = -1;
  int bytes = 0;
  for (int L : {n1, n2})
    for (int l : {(L + 1) / 2, L / 2})
      if (qt.off[l] < 0) {
        qt.off[l] = bytes;
        bytes += (1 << l) * ((2 << l) - 1);
      }
  const size_t lds = ((bytes + 15) & ~15) + sizeof(uint16_t) * (((1u << n1) - 1) + ((1u << n2) - 1));
  if (lds > 150 * 1024) return BDC_ERR_ARG;
  // in chunks of x1 rows, so a long precompute (large k) can report progress
  const char *v = getenv("BDC_VERBOSE");
  const bool verbose = v && v[0] == '1';
  const uint32_t N1 = 1u << n1, chunk = N1 >= 16 ? N1 / 16 : N1;
  for (uint32_t b = 0; b < N1; b += chunk) {
    hipLaunchKernelGGL(k_entropy_reps, dim3(chunk), dim3(1024), lds, s, n, k, n1, n2, mask, wordbase, qt, bytes,
                       logtab, b, Hrep);
    if (hipGetLastError() != hipSuccess) return BDC_ERR_HIP;
    if (verbose) {
      if (hipStreamSynchronize(s) != hipSuccess) return BDC_ERR_HIP;
      fprintf(stderr, "[bdc] entropy term BDC_(%d,%d): x1 rows %u/%u done\n", n, k, b + chunk, N1);
    }
  }
  return BDC_OK;
}

}  // namespace bdc
