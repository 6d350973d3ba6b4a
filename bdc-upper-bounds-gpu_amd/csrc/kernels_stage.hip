// Stage contractions with a few-deletion count factor, done as gathers.
//
// When a half-length count matrix B_m = cache_{nb,m} removes only d = nb - m
// symbols (d <= 3), its rows hold at most C(nb, d) non-zeros and the dense
// stage GEMM wastes both MFMA work and the 2^(nb+m)-entry read of the cache.
// The same numbers follow from the definition of the count:
//   c(x, z) = #{ S subset of positions, |S| = d : x with S deleted == z }
// so for any array F
//   sum_z F[z] c(x, z)            = sum_S F[del(x, S)]                 (deletion form)
//   sum_x F[x] c(x, z)            = sum_S sum_{b in {0,1}^d} F[ins(z, S, b)]  (insertion form)
// where ins(z, S, b) places the bits b at the positions S of an nb-bit word.
// Both are plain FP64 sums of the same products the GEMM forms.
#include "device_common.hpp"
#include "kernels.hpp"

namespace bdc {

// remove the bits of x at the set positions of mask (bit 0 = least significant)
__device__ inline uint32_t del_bits(uint32_t x, uint32_t mask) {
  // highest positions first so lower positions keep their index
  while (mask) {
    const int s = 31 - __clz(mask);
    x = ((x >> (s + 1)) << s) | (x & ((1u << s) - 1u));
    mask &= ~(1u << s);
  }
  return x;
}
// insert the bits of b (lowest first) at the ascending set positions of mask
__device__ inline uint32_t ins_bits(uint32_t z, uint32_t mask, uint32_t b) {
  while (mask) {
    const int s = __ffs(mask) - 1;
    z = ((z >> s) << (s + 1)) | ((b & 1u) << s) | (z & ((1u << s) - 1u));
    b >>= 1;
    mask &= mask - 1u;
  }
  return z;
}
// next d-subset of the low nb bits after mask (Gosper), 0 when exhausted
__device__ inline uint32_t next_subset(uint32_t v, int nb) {
  if (v == 0) return 0;
  const uint32_t t = v | (v - 1u);
  const uint32_t w = (t + 1u) | (((~t & -~t) - 1u) >> (__ffs(v)));
  return w < (1u << nb) ? w : 0u;
}

// The d-subsets of nb positions (C(16,3) = 560 at most here) are listed in LDS
// once per block.  A block of 8 waves covers 64 consecutive outputs: wave w
// sums the subsets with index = w mod 8 (lanes = consecutive outputs, so the
// loads stay coalesced), and the 8 partial sums are combined in a fixed order
// through LDS.  This gives the latency-bound gathers 8x the waves in flight.
constexpr int kMaxSubsets = 1024;
constexpr int kGatherWaves = 8;
__device__ inline int list_subsets(int nb, int d, uint32_t *masks) {
  __shared__ int count;
  if (threadIdx.x == 0) {
    int c = 0;
    for (uint32_t m = (1u << d) - 1u; m && c < kMaxSubsets; m = next_subset(m, nb)) masks[c++] = m;
    count = c;
  }
  __syncthreads();
  return count;
}
__device__ inline double combine_waves(double v, double *part) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  part[w * 64 + lane] = v;
  __syncthreads();
  double t = 0.0;
  if (w == 0)
#pragma unroll
    for (int q = 0; q < kGatherWaves; ++q) t += part[q * 64 + lane];
  return t;  // valid in wave 0
}

// out[r*ors + x*ocs] = sum_{|S| = d} in[r*irs + del(x, S)*ics]   for r < R, x < 2^nb
__global__ void __launch_bounds__(512) k_del_gather(const double *__restrict__ in, int64_t irs, int64_t ics,
                                                    double *__restrict__ out, int64_t ors, int64_t ocs, int R, int nb,
                                                    int d) {
  __shared__ uint32_t masks[kMaxSubsets];
  __shared__ double part[kGatherWaves * 64];
  const int ns = list_subsets(nb, d, masks);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  const int64_t nx = 1ll << nb;
  const bool live = i < (int64_t)R * nx;
  const int r = (int)(i >> nb);
  const uint32_t x = (uint32_t)(i & (nx - 1));
  double s = 0.0;
  if (live) {
    const double *row = in + r * irs;
    double s2 = 0.0;
    int t = w;
    for (; t + kGatherWaves < ns; t += 2 * kGatherWaves) {
      s += row[(int64_t)del_bits(x, masks[t]) * ics];
      s2 += row[(int64_t)del_bits(x, masks[t + kGatherWaves]) * ics];
    }
    if (t < ns) s += row[(int64_t)del_bits(x, masks[t]) * ics];
    s += s2;
  }
  s = combine_waves(s, part);
  if (live && w == 0) out[r * ors + (int64_t)x * ocs] = s;
}

// out[r*ors + z*ocs] = sum_{|S| = d} sum_b in[r*irs + ins(z, S, b)*ics]   for r < R, z < 2^(nb-d)
__global__ void __launch_bounds__(512) k_ins_gather(const double *__restrict__ in, int64_t irs, int64_t ics,
                                                    double *__restrict__ out, int64_t ors, int64_t ocs, int R, int nb,
                                                    int d) {
  __shared__ uint32_t masks[kMaxSubsets];
  __shared__ double part[kGatherWaves * 64];
  const int ns = list_subsets(nb, d, masks);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  const int nz = nb - d;
  const int64_t cnt = 1ll << nz;
  const bool live = i < (int64_t)R * cnt;
  const int r = (int)(i >> nz);
  const uint32_t z = (uint32_t)(i & (cnt - 1));
  double s = 0.0;
  if (live) {
    const double *row = in + r * irs;
    const uint32_t nbv = 1u << d;
    double s2 = 0.0;
    for (int t = w; t < ns; t += kGatherWaves) {
      const uint32_t m = masks[t];
      uint32_t b = 0;
      for (; b + 2 <= nbv; b += 2) {
        s += row[(int64_t)ins_bits(z, m, b) * ics];
        s2 += row[(int64_t)ins_bits(z, m, b + 1) * ics];
      }
      if (b < nbv) s += row[(int64_t)ins_bits(z, m, b) * ics];
    }
    s += s2;
  }
  s = combine_waves(s, part);
  if (live && w == 0) out[r * ors + (int64_t)z * ocs] = s;
}

static uint64_t n_subsets(int nb, int d) {
  uint64_t c = 1;
  for (int i = 0; i < d; ++i) c = c * (uint64_t)(nb - i) / (uint64_t)(i + 1);
  return c;
}
int launch_del_gather(const double *in, int64_t irs, int64_t ics, double *out, int64_t ors, int64_t ocs, int R, int nb,
                      int d, hipStream_t s) {
  if (d < 1 || n_subsets(nb, d) > (uint64_t)kMaxSubsets) return BDC_ERR_ARG;
  const int64_t tot = (int64_t)R << nb;
  hipLaunchKernelGGL(k_del_gather, dim3((unsigned)((tot + 63) / 64)), dim3(64 * kGatherWaves), 0, s, in, irs, ics, out,
                     ors, ocs, R, nb, d);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}
int launch_ins_gather(const double *in, int64_t irs, int64_t ics, double *out, int64_t ors, int64_t ocs, int R, int nb,
                      int d, hipStream_t s) {
  if (d < 1 || n_subsets(nb, d) > (uint64_t)kMaxSubsets) return BDC_ERR_ARG;
  const int64_t tot = (int64_t)R << (nb - d);
  hipLaunchKernelGGL(k_ins_gather, dim3((unsigned)((tot + 63) / 64)), dim3(64 * kGatherWaves), 0, s, in, irs, ics, out,
                     ors, ocs, R, nb, d);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// Wbar from Vbar by reversal (n1 = n2, L reversal-symmetric):
//   W_j[x1][z] = V_{k-j}[rev z][rev x1]
// (substitute y1 -> rev y1 in sum_y1 c(x1,y1) L(y1 z) and use c(rev u, rev v) = c(u, v),
//  L(y1 z) = L(rev z rev y1)).  One launch covers every j in JR.
__global__ void k_w_from_v(MirrorW mw, const double *__restrict__ V, int N2, int n2, double *__restrict__ W, int N1,
                           int KR, int KRp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)N1 * KR) return;
  const int x1 = (int)(i / KR), kap = (int)(i % KR);
  int s = 0;
  while (s + 1 < mw.count && kap >= mw.offR[s + 1]) ++s;
  const int z = kap - mw.offR[s], m = mw.m[s];
  // column rev x1 of Vbar, stored at x1 when Vbar's columns are bit-reversed
  const uint32_t rz = rev_bits((uint32_t)z, m), rx = mw.stored ? (uint32_t)x1 : rev_bits((uint32_t)x1, n2);
  W[(int64_t)x1 * KRp + kap] = V[(int64_t)(mw.offL[s] + rz) * N2 + rx];
}
// The same copy tiled: a block covers 64 columns kappa' of Wbar and 32 of its
// rows, taken as x1 = rev(c << sh) for 32 consecutive c (rows = 2^(n2 - sh)), so
// that rev x1 = c << sh: the Vbar reads are 32 entries of one row (stride 2^sh)
// and the Wbar writes 64-entry row segments, transposed through LDS.  With
// Vbar's columns bit-reversed the rows are x1 = c themselves and the reads
// contiguous.  (The one-thread-per-entry form reads one Vbar row per entry.)
__global__ void __launch_bounds__(256) k_w_from_v_tiled(MirrorW mw, const double *__restrict__ V, int N2, int n2,
                                                        double *__restrict__ W, int KR, int KRp, int sh) {
  __shared__ double t[64][33];
  const int k0 = blockIdx.x * 64, c0 = blockIdx.y * 32, tid = threadIdx.x;
  const int cl = tid & 31;
  for (int i = 0; i < 8; ++i) {
    const int kl = (tid >> 5) + 8 * i, kap = k0 + kl;
    if (kap < KR) {
      int s = 0;
      while (s + 1 < mw.count && kap >= mw.offR[s + 1]) ++s;
      const uint32_t rz = rev_bits((uint32_t)(kap - mw.offR[s]), mw.m[s]);
      // (Vbar with bit-reversed columns holds the entry for row x1 at column x1:
      // the block then takes the rows x1 = c0 .. c0 + 31 themselves)
      const uint32_t col = mw.stored ? (uint32_t)(c0 + cl) : (uint32_t)(c0 + cl) << sh;
      t[kl][cl] = V[(int64_t)(mw.offL[s] + rz) * N2 + col];
    }
  }
  __syncthreads();
  const int kl = tid & 63, kap = k0 + kl;
  if (kap < KR)
    for (int i = 0; i < 8; ++i) {
      const int r = (tid >> 6) + 4 * i;
      const uint32_t x1 = mw.stored ? (uint32_t)(c0 + r) : rev_bits((uint32_t)(c0 + r) << sh, n2);
      W[(int64_t)x1 * KRp + kap] = t[kl][r];
    }
}
int launch_w_from_v(const MirrorW &mw, const double *V, int N2, int n2, double *W, int N1, int KR, int KRp,
                    hipStream_t s) {
  // N1: rows of Wbar formed (all, or the half space x1 < 2^(n2 - 1))
  int b = 0;
  while (b < 31 && (1 << b) < N1) ++b;
  if ((1 << b) == N1 && N1 >= 32 && b <= n2 && N2 == (1 << n2)) {
    hipLaunchKernelGGL(k_w_from_v_tiled, dim3((unsigned)((KR + 63) / 64), (unsigned)(N1 / 32)), dim3(256), 0, s, mw, V,
                       N2, n2, W, KR, KRp, n2 - b);
  } else {
    const int64_t tot = (int64_t)N1 * KR;
    hipLaunchKernelGGL(k_w_from_v, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, mw, V, N2, n2, W, N1, KR, KRp);
  }
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// ---------------------------------------------------------------------------
// Factored stage contractions.  A count matrix between strings of lengths L and
// m = L - d is a power of the single-deletion operator:
//   c(x, z) = (D^d)(x, z) / d!,   D(u, v) = #{ i : u with bit i removed = v },
// because each set of d deleted positions is reached by exactly d! ordered
// deletion sequences (one per order), and a sequence through intermediate
// strings u_1 .. u_{d-1} is counted by the product of the D factors.  So a
// stage contraction is d sparse steps (2l or l + 1 terms per output at length
// l) on one row held in LDS, instead of streaming the 2^(L+m)-entry count
// matrix: the counts never leave the index arithmetic.
//   insertion form (long -> short): out[v] = sum_{p < l, b} in[ins(v, p, b)]
//     (inserting bit b at position p of v; each u with D(u, v) = r is hit r times)
//   deletion form (short -> long):  out[u] = sum_{i <= l} in[del(u, i)]
// All sums run in a fixed order; the 1/d! scale is applied once at the end.
__device__ inline uint32_t ins_bit(uint32_t v, int p, uint32_t b) {
  return ((v >> p) << (p + 1)) | (b << p) | (v & ((1u << p) - 1u));
}
__device__ inline uint32_t del_bit(uint32_t u, int i) { return ((u >> (i + 1)) << i) | (u & ((1u << i) - 1u)); }
__device__ inline double factorial(int d) {
  double f = 1.0;
  for (int i = 2; i <= d; ++i) f *= (double)i;
  return f;
}
__device__ inline int chain_job(const ChainJobs &jb, int b) {
  int i = 0;
  while (i + 1 < jb.count && b >= jb.row0[i + 1]) ++i;
  return i;
}

// (the two insertions at position 0 of v are the adjacent entries 2v, 2v + 1:
// one 16-byte LDS load, where two 8-byte loads at a stride of 16 bytes would
// conflict)
__device__ inline void add_pos0(double &a, const double *row, uint32_t v) {
  const double2 t = *reinterpret_cast<const double2 *>(row + 2 * v);
  a += t.x;
  a += t.y;
}
__global__ void __launch_bounds__(1024) k_chain_ins(const double *__restrict__ src, double *__restrict__ dst, int L,
                                                    ChainJobs jb) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  const int i = chain_job(jb, blockIdx.x), y = blockIdx.x - jb.row0[i], d = jb.d[i];
  const double *in = src + jb.src_off[i] + (int64_t)y * jb.src_ld[i];
  double *out = dst + jb.dst_off[i] + (int64_t)y * jb.dst_ld[i];
  const uint32_t tid = threadIdx.x, bd = blockDim.x;
  // reversed source rows: the steps run on them as they are (c(rev u, rev v) =
  // c(u, v)), and output v' lands at its natural index rev v'
  const int m = L - d, rv = jb.rev;
  auto dsti = [&](uint32_t v) { return rv ? rev_bits(v, m) : v; };
  if (d == 0) {
    for (uint32_t t = tid; t < (1u << L); t += bd) out[dsti(t)] = in[t];
    return;
  }
  const double f = factorial(d);
  // step 1 streams the long row through LDS in two halves (the whole row and
  // the step's output would not fit): an output v takes the top-position
  // insertion from each half, and its other insertions from the half of its
  // own top bit
  const uint32_t H = 1u << (L - 1);
  double *acc = lds, *half = lds + H;
  for (uint32_t h = 0; h < 2; ++h) {
    __syncthreads();  // the previous half has been read
    for (uint32_t t = tid; t < H; t += bd) half[t] = in[h * H + t];
    __syncthreads();
    for (uint32_t v = tid; v < H; v += bd) {
      double a = h ? acc[v] : 0.0;
      a += half[v];
      if (L >= 2 && (v >> (L - 2)) == h) {
        add_pos0(a, half, v - h * (H >> 1));  // ins_bit(v, 0, b) - h H = 2 (v - h H / 2) + b
        for (int p = 1; p < L - 1; ++p) {
          a += half[ins_bit(v, p, 0u) - h * H];
          a += half[ins_bit(v, p, 1u) - h * H];
        }
      }
      acc[v] = a;
    }
  }
  __syncthreads();
  if (d == 1) {
    for (uint32_t v = tid; v < H; v += bd) out[dsti(v)] = acc[v];
    return;
  }
  double *cur = acc, *nxt = half;
  int l = L - 1;  // bits of cur
  for (int s = 2; s <= d; ++s, --l) {
    const uint32_t nout = 1u << (l - 1);
    for (uint32_t v = tid; v < nout; v += bd) {
      double a = 0.0;
      add_pos0(a, cur, v);
      for (int p = 1; p < l; ++p) {
        a += cur[ins_bit(v, p, 0u)];
        a += cur[ins_bit(v, p, 1u)];
      }
      if (s == d)
        out[dsti(v)] = a / f;
      else
        nxt[v] = a;
    }
    double *t = cur;
    cur = nxt;
    nxt = t;
    __syncthreads();
  }
}

// amax (may be null): amax[block] = max |out| over the row the block writes.
// mu_part (may be null): the source is ln q uncentred; mu is the mean over its
// 2^mu_k entries from the 2^mu_k / 256 block sums in mu_part, formed exactly as
// k_q_center forms it (threads below 256 sum every 256th partial, then a block
// sum whose other waves add +0.0), and subtracted on load; block 0 records it.
__global__ void __launch_bounds__(1024) k_chain_del(const double *__restrict__ src, double *__restrict__ dst, int L,
                                                    ChainJobs jb, double *__restrict__ amax,
                                                    const double *__restrict__ mu_part, int mu_k,
                                                    double *__restrict__ scal) {
  extern __shared__ double lds[];
  __shared__ double red[16];
  const int i = chain_job(jb, blockIdx.x), y = blockIdx.x - jb.row0[i], d = jb.d[i], m = L - d;
  const double *in = src + jb.src_off[i] + (int64_t)y * jb.src_ld[i];
  double *out = dst + jb.dst_off[i] + (int64_t)y * jb.dst_ld[i];
  const uint32_t tid = threadIdx.x, bd = blockDim.x;
  double mx = 0.0, mu = 0.0;
  if (mu_part) {  // uniform
    const int np = (int)(((1u << mu_k) + 255) / 256);
    double sm = 0.0;
    if (tid < 256)
      for (int j = (int)tid; j < np; j += 256) sm += mu_part[j];
    sm = block_sum(sm, red);
    mu = sm / (double)(1u << mu_k);
    if (blockIdx.x == 0 && tid == 0) scal[kScalMu] = mu;
  }
  // reversed outputs: the steps run on the source read at reversed indices
  // (c(rev u, rev v) = c(u, v)), so output u' is the natural output rev u'
  const int rv = jb.rev;
  auto srci = [&](uint32_t t) { return rv ? rev_bits(t, m) : t; };
  if (d == 0) {
    for (uint32_t t = tid; t < (1u << L); t += bd) {
      const double v = in[srci(t)] - mu;
      out[t] = v;
      mx = fmax(mx, fabs(v));
    }
  } else {
    const double f = factorial(d);
    double *cur = lds, *nxt = lds + (1u << (L - 1));  // each region holds up to 2^(L-1) entries
    for (uint32_t t = tid; t < (1u << m); t += bd) cur[t] = in[srci(t)] - mu;
    __syncthreads();
    int l = m;  // bits of cur
    for (int s = 1; s <= d; ++s, ++l) {
      const uint32_t nout = 1u << (l + 1);
      for (uint32_t u = tid; u < nout; u += 2 * bd) {
        // two outputs per pass, u and v = u + bd (v repeats u, unstored, past
        // the end), each with its l + 1 terms in order in its own accumulator,
        // loaded four at a time ahead of the adds: the same sums as term by
        // term, with eight LDS loads in flight instead of a round trip per term
        const uint32_t v = u + bd < nout ? u + bd : u;
        double a = 0.0, b = 0.0;
        int q = 0;
        for (; q + 3 <= l; q += 4) {
          const double t0 = cur[del_bit(u, q)], t1 = cur[del_bit(u, q + 1)], t2 = cur[del_bit(u, q + 2)],
                       t3 = cur[del_bit(u, q + 3)];
          const double w0 = cur[del_bit(v, q)], w1 = cur[del_bit(v, q + 1)], w2 = cur[del_bit(v, q + 2)],
                       w3 = cur[del_bit(v, q + 3)];
          a += t0;
          a += t1;
          a += t2;
          a += t3;
          b += w0;
          b += w1;
          b += w2;
          b += w3;
        }
        for (; q <= l; ++q) {
          const double t0 = cur[del_bit(u, q)], w0 = cur[del_bit(v, q)];
          a += t0;
          b += w0;
        }
        if (s == d) {
          out[u] = a / f;
          mx = fmax(mx, fabs(a / f));
          if (v != u) {
            out[v] = b / f;
            mx = fmax(mx, fabs(b / f));
          }
        } else {
          nxt[u] = a;
          if (v != u) nxt[v] = b;
        }
      }
      double *t = cur;
      cur = nxt;
      nxt = t;
      __syncthreads();
    }
  }
  if (amax) {  // uniform: every thread of the block takes it
    mx = block_max(mx, red);
    if (tid == 0) amax[blockIdx.x] = mx;
  }
}

static int chain_check(int L, const ChainJobs &jb) {
  if (L < 1 || L > 14 || jb.count < 1 || jb.count > kMaxChainJobs || jb.row0[0] != 0) return BDC_ERR_ARG;
  for (int i = 0; i < jb.count; ++i)
    if (jb.d[i] < 0 || jb.d[i] > L || jb.row0[i + 1] <= jb.row0[i]) return BDC_ERR_ARG;
  return BDC_OK;
}
int launch_chain_ins(const double *src, double *dst, int L, const ChainJobs &jobs, hipStream_t s) {
  if (chain_check(L, jobs) != BDC_OK) return BDC_ERR_ARG;
  hipLaunchKernelGGL(k_chain_ins, dim3(jobs.row0[jobs.count]), dim3(1024), sizeof(double) << L, s, src, dst, L, jobs);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}
int launch_chain_del(const double *src, double *dst, int L, const ChainJobs &jobs, double *amax,
                     const double *mu_part, int mu_k, double *scal, hipStream_t s) {
  if (chain_check(L, jobs) != BDC_OK || (mu_part && !scal)) return BDC_ERR_ARG;
  hipLaunchKernelGGL(k_chain_del, dim3(jobs.row0[jobs.count]), dim3(1024), sizeof(double) << L, s, src, dst, L, jobs,
                     amax, mu_part, mu_k, scal);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// ---------------------------------------------------------------------------
// Stage contractions with a many-deletion count factor: a skinny FP64 product
// against a half-length count matrix (u16), which is the large operand
// (cache_{14,11} is 64 MiB).  The counts are streamed once with 16-byte loads,
// widened to FP64 in LDS and multiplied on the FP64 matrix cores
// (v_mfma_f64_16x16x4); a block holds every row of the small operand (M <= 128,
// padded to MP) against 64 columns, over one K range of the split.
//   C[z][m][n] = sum_{k in range z} A[m][k] B(k, n)
//   TRANSB = 0: B(k, n) = B[k * ldb + n]   (output-law stage: rows x2, columns z)
//   TRANSB = 1: B(k, n) = B[n * ldb + k]   (cross-operand stage: B = cache rows x2 = n)
typedef double double4_ __attribute__((ext_vector_type(4)));
template <int TRANSB, int MP>
__global__ void __launch_bounds__(256) k_count_gemm(int M, int K, const double *__restrict__ A, int64_t lda,
                                                    const uint16_t *__restrict__ B, int64_t ldb, int kc,
                                                    double *__restrict__ C, int64_t ldc, int64_t zstride) {
  constexpr int BN = 64, BK = 32, MT = MP / 16;
  constexpr int ACH = MP * BK / 2 / 256;  // 16-byte pieces (2 doubles) of the A tile per thread
  static_assert(ACH >= 1, "MP >= 16");
  __shared__ double As[BK][MP + 1];
  __shared__ double Bs[BK][BN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n0 = blockIdx.x * BN, kz = blockIdx.y;
  const int kb = kz * kc, ke = min(K, kb + kc);
  double4_ acc[MT];
#pragma unroll
  for (int i = 0; i < MT; ++i) acc[i] = (double4_){0.0, 0.0, 0.0, 0.0};
  double2 ra[ACH];
  uint4 rb;
  auto gload = [&](int k0) {
#pragma unroll
    for (int c = 0; c < ACH; ++c) {
      const int e = tid + 256 * c, m = e / (BK / 2), kp = e % (BK / 2);
      ra[c] = m < M ? *reinterpret_cast<const double2 *>(A + (int64_t)m * lda + k0 + 2 * kp) : make_double2(0.0, 0.0);
    }
    if (TRANSB == 0)
      rb = *reinterpret_cast<const uint4 *>(B + (int64_t)(k0 + (tid >> 3)) * ldb + n0 + (tid & 7) * 8);
    else
      rb = *reinterpret_cast<const uint4 *>(B + (int64_t)(n0 + (tid >> 2)) * ldb + k0 + (tid & 3) * 8);
  };
  auto sstore = [&]() {
#pragma unroll
    for (int c = 0; c < ACH; ++c) {
      const int e = tid + 256 * c, m = e / (BK / 2), kp = e % (BK / 2);
      As[2 * kp][m] = ra[c].x;
      As[2 * kp + 1][m] = ra[c].y;
    }
    const uint32_t w[4] = {rb.x, rb.y, rb.z, rb.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double v = (double)((w[q >> 1] >> (16 * (q & 1))) & 0xffffu);
      if (TRANSB == 0)
        Bs[tid >> 3][(tid & 7) * 8 + q] = v;
      else
        Bs[(tid & 3) * 8 + q][tid >> 2] = v;
    }
  };
  if (kb < ke) gload(kb);
  for (int k0 = kb; k0 < ke; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < ke) gload(k0 + BK);
#pragma unroll
    for (int k4 = 0; k4 < BK / 4; ++k4) {
      const int kr = k4 * 4 + (lane >> 4);
      const double b = Bs[kr][wave * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < MT; ++i)
        acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(As[kr][i * 16 + (lane & 15)], b, acc[i], 0, 0, 0);
    }
  }
  double *Cz = C + (int64_t)kz * zstride;
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = i * 16 + (lane >> 4) + 4 * r;
      if (row < M) Cz[(int64_t)row * ldc + n0 + wave * 16 + (lane & 15)] = acc[i][r];
    }
}

int launch_count_gemm(int M, int N, int K, const double *A, int64_t lda, const uint16_t *B, int64_t ldb, int transb,
                      double *C, int64_t ldc, int splits, int64_t zstride, hipStream_t s) {
  if (M < 1 || M > 128 || N % 64 || K % 32 || lda % 2 || ldb % 8 || splits < 1 || (K / 32) % splits)
    return BDC_ERR_ARG;
  const int kc = K / splits;
  dim3 grid(N / 64, splits);
#define CG(TB, MPP) \
  hipLaunchKernelGGL((k_count_gemm<TB, MPP>), grid, dim3(256), 0, s, M, K, A, lda, B, ldb, kc, C, ldc, zstride)
#define CG_M(TB)                     \
  if (M <= 16) CG(TB, 16);           \
  else if (M <= 32) CG(TB, 32);      \
  else if (M <= 64) CG(TB, 64);      \
  else CG(TB, 128);
  if (transb) {
    CG_M(1)
  } else {
    CG_M(0)
  }
#undef CG_M
#undef CG
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

}  // namespace bdc
