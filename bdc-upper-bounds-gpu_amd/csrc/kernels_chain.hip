// Kernels of the deletion-chain path, used when the output length k is close
// to n (k > 20: 2^k outputs are too many for the dense split contractions).
//
// With D the single-deletion operator (D(u, v) = number of positions whose
// removal turns u into v), the counts of d = n - k deletions are
// c(x, y) = (D^d / d!)(x, y): every deletion set is reached by d! orders.  The
// input x is cut into bit groups (top to bottom) and D = sum_g D_g, D_g
// deleting inside group g, so D^d / d! = sum over d_1 + ... + d_G = d of
// prod_g D_g^{d_g} / d_g!.  Level buffers R_t hold the strings after t
// deletions (2^(n-t) entries, the index being the string itself):
//   output law   R_0 = p; group by group (top first), for sources t from d-1
//                down to 0: R_{t+j} += D_g^j R_t / j!;  q = R_d / C(n,k);
//   cross term   the transpose of that program: R'_d = L (log output law);
//                group by group (bottom first), for targets t from 0 up:
//                R'_t += sum_j (D_g^T)^j R'_{t+j} / j!  (Horner form);
//                S(x) = sum_y c(x,y) L(y) = R'_0(x).
// Seen from group g, R_t is a [high][group][low] array: the low bits are the
// groups below (untouched), the high part holds the groups above after their
// t deletions.  A block owns one high row and a tile of W low columns, keeps
// the group axis in LDS and runs all deletion steps there.
//
// The entropy term H(x) = sum_y c ln c is summed over the distinct
// subsequences of x (the reference's enumeration, backend/BAA.cu:261-335), far
// fewer than 2^k when k is close to n, by a depth-first walk of their prefix
// tree with the counts in a band of d + 1 embedding counts.
#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "device_common.hpp"
#include "fast_exp.hpp"
#include "fast_log.hpp"
#include "kernels.hpp"

namespace bdc {

// u (l bits) with bit b inserted at position s (counted from the LSB)
__device__ inline uint32_t ins_bit(uint32_t v, int s, uint32_t b) {
  return ((v >> s) << (s + 1)) | (b << s) | (v & ((1u << s) - 1u));
}
// u with the bit at position s removed
__device__ inline uint32_t del_bit(uint32_t u, int s) { return ((u >> (s + 1)) << s) | (u & ((1u << s) - 1u)); }

// ------------------------------------------------------------ output law
// One source R_t viewed as [NH][2^L][2^B]; steps j = 1..jmax in LDS,
// (D^j R_t)/j! written to (or added into) R_{t+j} viewed as [NH][2^(L-j)][2^B].
// LDS: 2^L * W + 2^(L-1) * W doubles.
template <int W>
__global__ void __launch_bounds__(256) k_grp_del(GrpChain g) {
  extern __shared__ double lds[];
  const int L = g.L;
  double *bufA = lds, *bufB = lds + ((size_t)W << L);
  const uint64_t NW = 1ull << g.B;
  const uint64_t blk = blockIdx.x, h = blk / g.ntiles, col0 = (blk % g.ntiles) * W;
  const double *src = g.src + ((h << L) * NW) + col0;
  for (int idx = threadIdx.x; idx < (W << L); idx += blockDim.x) bufA[idx] = src[(uint64_t)(idx / W) * NW + idx % W];
  __syncthreads();
  double *in = bufA, *out = bufB;
  for (int j = 1; j <= g.m; ++j) {
    const int l = L - j + 1;  // input length in the group
    const int nout = W << (l - 1);
    double *dst = g.lvl[j] + ((h << (L - j)) * NW) + col0;
    const double f = g.scale[j];
    const bool acc = (g.acc >> j) & 1;
    for (int idx = threadIdx.x; idx < nout; idx += blockDim.x) {
      const uint32_t v = idx / W;
      const int c = idx % W;
      double s = 0.0;
      for (int p = 0; p < l; ++p) {
        s += in[ins_bit(v, p, 0u) * W + c];
        s += in[ins_bit(v, p, 1u) * W + c];
      }
      out[idx] = s;
      double *o = dst + (uint64_t)v * NW + c;
      *o = acc ? *o + s * f : s * f;
    }
    __syncthreads();
    double *tmp = in;
    in = out;
    out = tmp;
  }
}

// ------------------------------------------------------------ cross term
// Target R'_t viewed as [NH][2^L][2^B] gets sum_{j=1..m} (D^T)^j R'_{t+j} / j!
// by Horner: V = R'_{t+m}; V = D^T V / (j+1) + R'_{t+j} for j = m-1..1;
// R'_t (=|+=) D^T V.  A null source is zero.  LDS: 2 * 2^(L-1) * W doubles.
template <int W>
__global__ void __launch_bounds__(256) k_grp_ins(GrpChain g) {
  extern __shared__ double lds[];
  const int L = g.L, m = g.m;
  double *bufA = lds, *bufB = lds + ((size_t)W << (L > 0 ? L - 1 : 0));
  const uint64_t NW = 1ull << g.B;
  const uint64_t blk = blockIdx.x, h = blk / g.ntiles, col0 = (blk % g.ntiles) * W;
  {
    const int lm = L - m;
    const double *src = g.lvl[m] ? g.lvl[m] + ((h << lm) * NW) + col0 : nullptr;
    for (int idx = threadIdx.x; idx < (W << lm); idx += blockDim.x)
      bufA[idx] = src ? src[(uint64_t)(idx / W) * NW + idx % W] : 0.0;
  }
  __syncthreads();
  double *in = bufA, *out = bufB;
  for (int j = m - 1; j >= 1; --j) {
    const int l = L - j;  // output length
    const double *src = g.lvl[j] ? g.lvl[j] + ((h << l) * NW) + col0 : nullptr;
    const double f = g.scale[j + 1];  // 1 / (j + 1)
    for (int idx = threadIdx.x; idx < (W << l); idx += blockDim.x) {
      const uint32_t u = idx / W;
      const int c = idx % W;
      double s = 0.0;
      for (int p = 0; p < l; ++p) s += in[del_bit(u, p) * W + c];
      s *= f;
      if (src) s += src[(uint64_t)u * NW + c];
      out[idx] = s;
    }
    __syncthreads();
    double *tmp = in;
    in = out;
    out = tmp;
  }
  double *dst = g.dst + ((h << L) * NW) + col0;
  const bool acc = g.acc & 1;
  for (int idx = threadIdx.x; idx < (W << L); idx += blockDim.x) {
    const uint32_t u = idx / W;
    const int c = idx % W;
    double s = 0.0;
    for (int p = 0; p < L; ++p) s += in[del_bit(u, p) * W + c];
    double *o = dst + (uint64_t)u * NW + c;
    *o = acc ? *o + s : s;
  }
}

int launch_grp_del(const GrpChain &g, hipStream_t s) {
  if (g.m < 1 || g.m > g.L || g.m > kMaxGrpSteps) return BDC_ERR_ARG;
  const size_t lds = (((size_t)1 << g.L) + ((size_t)1 << (g.L - 1))) * g.W * sizeof(double);
  const uint64_t blocks = g.nh * g.ntiles;
  if (blocks == 0 || blocks > 0x7fffffffull || lds > 160 * 1024) return BDC_ERR_ARG;
  switch (g.W) {
    case 1: hipLaunchKernelGGL(k_grp_del<1>, dim3((unsigned)blocks), dim3(256), lds, s, g); break;
    case 2: hipLaunchKernelGGL(k_grp_del<2>, dim3((unsigned)blocks), dim3(256), lds, s, g); break;
    case 4: hipLaunchKernelGGL(k_grp_del<4>, dim3((unsigned)blocks), dim3(256), lds, s, g); break;
    case 8: hipLaunchKernelGGL(k_grp_del<8>, dim3((unsigned)blocks), dim3(256), lds, s, g); break;
    case 16: hipLaunchKernelGGL(k_grp_del<16>, dim3((unsigned)blocks), dim3(256), lds, s, g); break;
    default: return BDC_ERR_ARG;
  }
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}
int launch_grp_ins(const GrpChain &g, hipStream_t s) {
  if (g.m < 1 || g.m > g.L || g.m > kMaxGrpSteps) return BDC_ERR_ARG;
  const size_t lds = ((size_t)2 << (g.L - 1)) * g.W * sizeof(double);
  const uint64_t blocks = g.nh * g.ntiles;
  if (blocks == 0 || blocks > 0x7fffffffull || lds > 160 * 1024) return BDC_ERR_ARG;
  switch (g.W) {
    case 1: hipLaunchKernelGGL(k_grp_ins<1>, dim3((unsigned)blocks), dim3(256), lds, s, g); break;
    case 2: hipLaunchKernelGGL(k_grp_ins<2>, dim3((unsigned)blocks), dim3(256), lds, s, g); break;
    case 4: hipLaunchKernelGGL(k_grp_ins<4>, dim3((unsigned)blocks), dim3(256), lds, s, g); break;
    case 8: hipLaunchKernelGGL(k_grp_ins<8>, dim3((unsigned)blocks), dim3(256), lds, s, g); break;
    case 16: hipLaunchKernelGGL(k_grp_ins<16>, dim3((unsigned)blocks), dim3(256), lds, s, g); break;
    default: return BDC_ERR_ARG;
  }
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// q(y) = scal[slot] * inv_c * R[y];  L(y) = ln q(y) (not centred: mu = 0)
__global__ void k_chain_q(const double *__restrict__ R, uint64_t ny, double *__restrict__ scal, int slot, double inv_c,
                          double *__restrict__ q, double *__restrict__ L) {
  const double f = scal[slot] * inv_c;
  for (uint64_t y = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; y < ny; y += (uint64_t)gridDim.x * blockDim.x) {
    const double v = R[y] * f;
    q[y] = v;
    if (L) L[y] = log(v);
  }
  if (L && blockIdx.x == 0 && threadIdx.x == 0) scal[kScalMu] = 0.0;
}
int launch_chain_q(const double *R, uint64_t ny, double *scal, int slot, double inv_c, double *q, double *L,
                   hipStream_t s) {
  const unsigned blocks = (unsigned)std::min<uint64_t>(16384, (ny + 255) / 256);
  hipLaunchKernelGGL(k_chain_q, dim3(blocks), dim3(256), 0, s, R, ny, scal, slot, inv_c, q, L);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// D(x) = Hc(x) - mu - S(x)/C over all inputs.  EPI_BA: the BA update
// p~ <- p~ exp(D - dref) in place with the block partials (sum p~', max D,
// sum p~ D) of backend/BAA.cu:600-632 (finished by launch_ba_finalize);
// EPI_D: D stored, max D in the partials.  Each block owns a contiguous range.
template <int EPI>
__global__ void __launch_bounds__(256) k_chain_epi(const double *__restrict__ S, double *__restrict__ P,
                                                   const double *__restrict__ Hc, double *__restrict__ Dout,
                                                   uint64_t nx, const double *__restrict__ scal, double inv_c,
                                                   const double *__restrict__ exptab, double *__restrict__ part) {
  __shared__ double etab[32];
  __shared__ double red[48];
  if (threadIdx.x < 32) etab[threadIdx.x] = exptab[threadIdx.x];
  __syncthreads();
  const double mu = scal[kScalMu], dref = scal[kScalDRef];
  const uint64_t per = (nx + gridDim.x - 1) / gridDim.x;
  const uint64_t lo = (uint64_t)blockIdx.x * per, hi = lo + per < nx ? lo + per : nx;
  double s0 = 0.0, m1 = -INFINITY, s2 = 0.0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const double D = Hc[i] - mu - S[i] * inv_c;
    m1 = fmax(m1, D);
    if constexpr (EPI == EPI_BA) {
      const double p = P[i];
      const double w = p * exp_tang(D - dref, etab);
      P[i] = w;
      s0 += w;
      s2 += p * D;
    } else {
      Dout[i] = D;
    }
  }
  block_reduce3(s0, m1, s2, red);
  if (threadIdx.x == 0) {
    part[(int64_t)blockIdx.x * kPartStrideDev + 0] = s0;
    part[(int64_t)blockIdx.x * kPartStrideDev + 1] = m1;
    part[(int64_t)blockIdx.x * kPartStrideDev + 2] = s2;
  }
}
int launch_chain_epi(int epi, const double *S, double *P, const double *Hc, double *Dout, uint64_t nx,
                     const double *scal, double inv_c, const double *exptab, double *part, int nblocks,
                     hipStream_t s) {
  if (epi == EPI_BA)
    hipLaunchKernelGGL(k_chain_epi<EPI_BA>, dim3(nblocks), dim3(256), 0, s, S, P, Hc, Dout, nx, scal, inv_c, exptab,
                       part);
  else
    hipLaunchKernelGGL(k_chain_epi<EPI_D>, dim3(nblocks), dim3(256), 0, s, S, P, Hc, Dout, nx, scal, inv_c, exptab,
                       part);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// ------------------------------------------------------------ entropy term
// H(x) = sum_y c(x,y) ln c(x,y) over the distinct length-k subsequences y of x.
// A prefix z of t symbols carries the band G[e] = number of embeddings of z in
// x[0 .. t+e), e = 0..d (d = n - k): a prefix embeds with at most d skips, and
// appending symbol b gives G'[e] = sum_{e' <= e} [x_{t+e'} = b] G[e'].  A child
// exists iff some e with G[e] > 0 has x_{t+e} = b; the count of a full y is
// G[d] (embeddings in all of x).  Each thread walks the tree of its input
// depth first; a level where both children exist is marked pending and
// revisited later, its band recomputed from the checkpoint at the nearest
// depth multiple of 8 below it (kept in LDS).  The last two levels are summed
// in one go (up to four leaves).  One band step per loop iteration per
// thread, and a thread moves to its next input as soon as it finishes, so the
// threads of a wave stay busy whatever the tree sizes.  The c ln c terms are
// summed with Neumaier's compensation.
constexpr int kDfsBlock = 256;
constexpr int kDfsCk = 3;  // checkpoints at depths 8, 16, 24 (depth 0 is all ones)

// One instantiation per number of deletions d = DMAX (0..16): the band loops
// are unrolled to exactly d + 1 entries.
template <int DMAX>
__global__ void __launch_bounds__(kDfsBlock) k_entropy_dfs(int n, int k, uint64_t first, uint64_t norb,
                                                           const uint32_t *__restrict__ orbits,
                                                           const double *__restrict__ logtab,
                                                           double *__restrict__ Hrep) {
  __shared__ double lt[130];
  __shared__ uint32_t ck[kDfsCk * (DMAX + 1) * kDfsBlock];
  for (int i = threadIdx.x; i < 130; i += blockDim.x) lt[i] = logtab[i];
  __syncthreads();
  const double *lnF = lt, *invF = lt + 65;
  constexpr int d = DMAX;  // n - k
  const uint32_t wmask = (2u << d) - 1u;  // d + 1 bits
  const int tb = k >= 2 ? k - 2 : 0;      // depth of the final two-level (or one-level) sum
  const int tid = threadIdx.x;
  const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = first + (uint64_t)blockIdx.x * blockDim.x + tid;  // orbits [first, norb)
  if (i >= norb) return;
  uint32_t x = orbits[i];
  uint32_t G[DMAX + 1];
#pragma unroll
  for (int e = 0; e <= DMAX; ++e) G[e] = 1u;
  int t = 0, mode = 0, tr = 0;
  uint32_t yp = 0, pending = 0;
  double hs = 0.0, hc = 0.0;
  // bits x_{tt}..x_{tt+d} (MSB-first positions), x_{tt} the highest; tt <= k - 1
  auto window = [&](int tt) -> uint32_t { return (x >> (n - 1 - tt - d)) & wmask; };
  auto add = [&](uint32_t c) {
    if (c > 1u) {
      const double v = (double)c * log_tang_int(c, lnF, invF);
      const double s = hs + v;
      hc += fabs(hs) >= fabs(v) ? (hs - s) + v : (v - s) + hs;
      hs = s;
    }
  };
  // sum of G[e] over the e whose bit in w (aligned as a window) is set
  auto masked = [&](const uint32_t *Gv, uint32_t w) -> uint32_t {
    uint32_t acc = 0u;
#pragma unroll
    for (int e = 0; e <= DMAX; ++e)
      if (e <= d) acc += ((w >> (d - e)) & 1u) ? Gv[e] : 0u;
    return acc;
  };
  auto step = [&](uint32_t *Gv, uint32_t w) {
    uint32_t acc = 0u;
#pragma unroll
    for (int e = 0; e <= DMAX; ++e)
      if (e <= d) {
        acc += ((w >> (d - e)) & 1u) ? Gv[e] : 0u;
        Gv[e] = acc;
      }
  };
  auto nonzero = [&](const uint32_t *Gv) -> uint32_t {
    uint32_t m = 0u;
#pragma unroll
    for (int e = 0; e <= DMAX; ++e)
      if (e <= d) m |= (Gv[e] != 0u ? 1u : 0u) << (d - e);
    return m;
  };
  while (true) {
    if (mode == 0 && t == tb) {
      const uint32_t w = window(t);
      if (k - tb == 1) {
        add(masked(G, ~w & wmask));
        add(masked(G, w));
      } else {
        const uint32_t nz = nonzero(G), w2 = window(t + 1);
        for (uint32_t a = 0; a < 2; ++a) {
          const uint32_t wa = a ? w : (~w & wmask);
          if (!(wa & nz)) continue;
          uint32_t Ga[DMAX + 1];
#pragma unroll
          for (int e = 0; e <= DMAX; ++e) Ga[e] = G[e];
          step(Ga, wa);
          add(masked(Ga, ~w2 & wmask));
          add(masked(Ga, w2));
        }
      }
      if (pending == 0u) {
        Hrep[i] = hs + hc;
        i += T;
        if (i >= norb) break;
        x = orbits[i];
#pragma unroll
        for (int e = 0; e <= DMAX; ++e) G[e] = 1u;
        t = 0;
        yp = 0u;
        hs = hc = 0.0;
        continue;
      }
      tr = 31 - __clz(pending);
      pending &= ~(1u << tr);
      yp >>= (t - tr);
      const int c0 = tr & ~7;
      if (c0 == 0) {
#pragma unroll
        for (int e = 0; e <= DMAX; ++e) G[e] = 1u;
      } else {
#pragma unroll
        for (int e = 0; e <= DMAX; ++e) G[e] = ck[(((c0 >> 3) - 1) * (DMAX + 1) + e) * kDfsBlock + tid];
      }
      t = c0;
      mode = 1;
    }
    // one band step from depth t
    const uint32_t w = window(t);
    uint32_t b;
    if (mode == 1) {
      if (t < tr) {
        b = (yp >> (tr - 1 - t)) & 1u;
      } else {
        b = 1u;
        mode = 0;
        yp = (yp << 1) | 1u;
      }
    } else {
      const uint32_t nz = nonzero(G);
      const bool v0 = (~w & wmask & nz) != 0u, v1 = (w & nz) != 0u;
      b = v0 ? 0u : 1u;
      if (v0 && v1) pending |= 1u << t;
      yp = (yp << 1) | b;
    }
    step(G, b ? w : (~w & wmask));
    ++t;
    if ((t & 7) == 0 && t < tb && (t >> 3) <= kDfsCk) {
#pragma unroll
      for (int e = 0; e <= DMAX; ++e) ck[(((t >> 3) - 1) * (DMAX + 1) + e) * kDfsBlock + tid] = G[e];
    }
  }
}

int launch_entropy_dfs(int n, int k, uint64_t norb, const uint32_t *orbits, const double *logtab, double *Hrep,
                       hipStream_t s) {
  const int d = n - k;
  if (k < 1 || d < 0 || k > 31 || n > 32 || d > 16) return BDC_ERR_ARG;
  // 16 chunks with a progress line each when BDC_VERBOSE=1 (minutes at n = 31)
  const char *v = getenv("BDC_VERBOSE");
  const bool verbose = v && v[0] == '1';
  const uint64_t chunk = verbose && norb >= 16 * kDfsBlock ? (norb + 15) / 16 : norb;
  for (uint64_t first = 0; first < norb; first += chunk) {
    const uint64_t last = std::min(norb, first + chunk);
    const uint64_t want = (last - first + kDfsBlock - 1) / kDfsBlock;
    const unsigned blocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(want, 2048));
    const auto launch = [&](auto dd) {
      hipLaunchKernelGGL(k_entropy_dfs<decltype(dd)::value>, dim3(blocks), dim3(kDfsBlock), 0, s, n, k, first, last,
                         orbits, logtab, Hrep);
    };
    switch (d) {
      case 0: launch(std::integral_constant<int, 0>{}); break;
      case 1: launch(std::integral_constant<int, 1>{}); break;
      case 2: launch(std::integral_constant<int, 2>{}); break;
      case 3: launch(std::integral_constant<int, 3>{}); break;
      case 4: launch(std::integral_constant<int, 4>{}); break;
      case 5: launch(std::integral_constant<int, 5>{}); break;
      case 6: launch(std::integral_constant<int, 6>{}); break;
      case 7: launch(std::integral_constant<int, 7>{}); break;
      case 8: launch(std::integral_constant<int, 8>{}); break;
      case 9: launch(std::integral_constant<int, 9>{}); break;
      case 10: launch(std::integral_constant<int, 10>{}); break;
      case 11: launch(std::integral_constant<int, 11>{}); break;
      case 12: launch(std::integral_constant<int, 12>{}); break;
      case 13: launch(std::integral_constant<int, 13>{}); break;
      case 14: launch(std::integral_constant<int, 14>{}); break;
      case 15: launch(std::integral_constant<int, 15>{}); break;
      default: launch(std::integral_constant<int, 16>{}); break;
    }
    if (hipGetLastError() != hipSuccess) return BDC_ERR_HIP;
    if (verbose) {
      if (hipStreamSynchronize(s) != hipSuccess) return BDC_ERR_HIP;
      fprintf(stderr, "[bdc] entropy term BDC_(%d,%d) by distinct subsequences: %llu / %llu representatives done\n", n,
              k, (unsigned long long)last, (unsigned long long)norb);
    }
  }
  return BDC_OK;
}

}  // namespace bdc
