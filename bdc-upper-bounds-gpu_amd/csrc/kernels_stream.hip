// Streaming and reduction kernels of a BA step: the output-law combine with
// ln q, the finalize of the fused update, the dual-bound weights and Adam.
#include <algorithm>

#include "device_common.hpp"
#include "kernels.hpp"

namespace bdc {

// Reduce the per-block partials of the fused update and advance the device
// scalars.  With p = pscale * P~ and P~' = P~ exp(D - dref):
//   Z_t   = sum_x p_t e^{D_t} = pscale e^{dref} sum P~'        (BAA.cu:618-626)
//   I(p_t)= sum_x p_t D_t                                      (BAA.cu:820-913)
//   mx_t  = max_x log2(p_{t+1}/p_t) = (max D_t - ln Z_t)/ln 2   (BAA.cu:916-946)
//   pscale' = 1 / sum P~',  dref' = ln Z_t.
// Partials are combined in a fixed order, so the result is deterministic.
// mult = 2 when the update ran on the half space x1 < N1/2: the sums over the
// other half (complement images, p(~x) = p(x), D(~x) = D(x)) are equal.
__global__ void __launch_bounds__(256) k_ba_finalize(const double *__restrict__ part, int nblocks, double mult,
                                                     double *__restrict__ scal, double *__restrict__ hist,
                                                     int hist_idx) {
  __shared__ double red[48];
  double s0 = 0.0, m1 = -INFINITY, s2 = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    s0 += part[(int64_t)b * kPartStrideDev + 0];
    m1 = fmax(m1, part[(int64_t)b * kPartStrideDev + 1]);
    s2 += part[(int64_t)b * kPartStrideDev + 2];
  }
  block_reduce3(s0, m1, s2, red);
  s0 *= mult;
  s2 *= mult;
  if (threadIdx.x == 0) {
    const double pscale = scal[kScalPScale], dref = scal[kScalDRef];
    const double lnz = log(pscale) + dref + log(s0);
    const double info = pscale * s2;
    const double mx = (m1 - lnz) / 0.69314718055994530942;
    scal[kScalPScale] = 1.0 / s0;
    scal[kScalDRef] = lnz;
    scal[kScalLastMx] = mx;
    scal[kScalIter] += 1.0;
    scal[kScalDMax] = m1;
    if (hist) {
      double *h = hist + 4 * (int64_t)hist_idx;
      h[0] = mx;
      h[1] = m1;
      h[2] = lnz;
      h[3] = info;
    }
  }
}
// First level of the finalize reduction: block g combines the partials of a
// contiguous chunk (fixed order, so the total is deterministic).
__global__ void __launch_bounds__(256) k_part_reduce(const double *__restrict__ part, int nblocks, int chunk,
                                                     double *__restrict__ out) {
  __shared__ double red[48];
  const int b0 = blockIdx.x * chunk, b1 = min(nblocks, b0 + chunk);
  double s0 = 0.0, m1 = -INFINITY, s2 = 0.0;
  for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
    s0 += part[(int64_t)b * kPartStrideDev + 0];
    m1 = fmax(m1, part[(int64_t)b * kPartStrideDev + 1]);
    s2 += part[(int64_t)b * kPartStrideDev + 2];
  }
  block_reduce3(s0, m1, s2, red);
  if (threadIdx.x == 0) {
    out[blockIdx.x * kPartStrideDev + 0] = s0;
    out[blockIdx.x * kPartStrideDev + 1] = m1;
    out[blockIdx.x * kPartStrideDev + 2] = s2;
  }
}
int launch_ba_finalize(const double *part, int nblocks, double mult, double *scal, double *hist, int hist_idx,
                       double *scratch, hipStream_t s) {
  const int groups = std::min(256, (nblocks + 255) / 256);
  const int chunk = (nblocks + groups - 1) / groups;
  hipLaunchKernelGGL(k_part_reduce, dim3(groups), dim3(256), 0, s, part, nblocks, chunk, scratch);
  hipLaunchKernelGGL(k_ba_finalize, dim3(1), dim3(256), 0, s, scratch, groups, mult, scal, hist, hist_idx);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// q(y) = pscale/C * (sum_{j in JL} q_j(y) + sum_{j in JR} q_j(y)).  For a
// reversal-symmetric p the JR slices equal JL slices read at rev(y)
// (q_j(y) = q_{k-j}(rev y)), which the mirror mode uses.  In complement mode
// the slices hold R_j, the contraction over the half space x1 < N1/2 only,
// and q_j(y) = R_j(y) + R_j(~y) (the other half is {~x}, c(~x, y) = c(x, ~y)).
// Then
// L(y) = ln q(y) - mu with mu the mean of ln q (any constant is exact since
// sum_y W(y|x) = 1; centring keeps |L| small for the split contraction).
// Grid pass: q(y), ln q(y) into Lout, per-block sums of ln q.
__global__ void __launch_bounds__(256) k_q_combine(QCombine qc, double inv_c, const double *__restrict__ scal,
                                                   double *__restrict__ qout, double *__restrict__ Lout,
                                                   double *__restrict__ part) {
  __shared__ double red[16];
  const uint32_t ny = 1u << qc.k;
  const double f = scal[qc.scale_slot] * inv_c;
  const uint32_t cm = ny - 1u;
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  double lq = 0.0;
  if (y < ny) {
    // Slice sums in a fixed order.  The loops have no branch in their bodies
    // and are unrolled, so the slice loads are issued ahead of the adds (one
    // grid of 2^k threads has too few waves to hide a load per add); in the
    // mirror loop an unselected slice adds +0.0, which leaves the sum of
    // nonnegative terms (from +0.0) bitwise unchanged.
    double s = 0.0;
    auto slices = [&](const double *__restrict__ base, int n, uint32_t yy, uint64_t sel) {
      if (qc.comp) {
#pragma unroll 4
        for (int i = 0; i < n; ++i) {
          const bool on = (sel >> i) & 1ull;
          const double a = base[(int64_t)i * ny + yy], b = base[(int64_t)i * ny + (yy ^ cm)];
          s += on ? a : 0.0;
          s += on ? b : 0.0;
        }
      } else {
#pragma unroll 4
        for (int i = 0; i < n; ++i) {
          const bool on = (sel >> i) & 1ull;
          const double a = base[(int64_t)i * ny + yy];
          s += on ? a : 0.0;
        }
      }
    };
    slices(qc.qL, qc.nL, y, ~0ull);
    if (qc.mirror)
      slices(qc.qL, qc.nL, rev_bits(y, qc.k), qc.mirrorSel);
    else
      slices(qc.qR, qc.nR, y, ~0ull);
    const double q = s * f;
    qout[y] = q;
    lq = log(q);
    Lout[y] = lq;
  }
  lq = block_sum(lq, red);
  if (threadIdx.x == 0) part[blockIdx.x] = lq;
}
// Centre: every block sums the same partials in the same order (so all agree
// on mu), then L = ln q - mu.
__global__ void __launch_bounds__(256) k_q_center(int k, const double *__restrict__ part, int np,
                                                  double *__restrict__ scal, double *__restrict__ Lout) {
  __shared__ double red[16];
  const uint32_t ny = 1u << k;
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  s = block_sum(s, red);
  const double mu = s / (double)ny;
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y < ny) Lout[y] -= mu;
  if (blockIdx.x == 0 && threadIdx.x == 0) scal[kScalMu] = mu;
}
int launch_q_combine(const QCombine &qc, double inv_c, double *scal, double *qout, double *Lout, double *part,
                     int center, hipStream_t s) {
  const uint32_t ny = 1u << qc.k;
  const int nb = (int)((ny + 255) / 256);
  hipLaunchKernelGGL(k_q_combine, dim3(nb), dim3(256), 0, s, qc, inv_c, scal, qout, Lout, part);
  if (center) hipLaunchKernelGGL(k_q_center, dim3(nb), dim3(256), 0, s, qc.k, part, nb, scal, Lout);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}
int launch_q_center(int k, const double *part, double *scal, double *Lout, hipStream_t s) {
  const int nb = (int)(((1u << k) + 255) / 256);
  hipLaunchKernelGGL(k_q_center, dim3(nb), dim3(256), 0, s, k, part, nb, scal, Lout);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// q = softmax(theta), L = ln q - mu
__global__ void __launch_bounds__(1024) k_softmax_log(const double *__restrict__ theta, int k, double *__restrict__ q,
                                                      double *__restrict__ L, double *__restrict__ scal) {
  __shared__ double red[16];
  const uint32_t ny = 1u << k;
  double m = -INFINITY;
  for (uint32_t y = threadIdx.x; y < ny; y += blockDim.x) m = fmax(m, theta[y]);
  m = block_max(m, red);
  double s = 0.0, st = 0.0;
  for (uint32_t y = threadIdx.x; y < ny; y += blockDim.x) {
    s += exp(theta[y] - m);
    st += theta[y];
  }
  s = block_sum(s, red);
  st = block_sum(st, red);
  const double lse = m + log(s);
  const double mu = st / (double)ny - lse;
  for (uint32_t y = threadIdx.x; y < ny; y += blockDim.x) {
    const double lq = theta[y] - lse;
    q[y] = exp(lq);
    L[y] = lq - mu;
  }
  if (threadIdx.x == 0) scal[kScalMu] = mu;
}
int launch_softmax_log(const double *theta, int k, double *q, double *L, double *scal, hipStream_t s) {
  hipLaunchKernelGGL(k_softmax_log, dim3(1), dim3(1024), 0, s, theta, k, q, L, scal);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// Dual bound weights.  Stage 1 (one block): Dmax from the D-pass partials.
// Stage 2 (grid): e(x) = exp(beta (D(x) - Dmax)) in place, block sums.
// Stage 3 (one block): U = Dmax + ln(sum e)/beta; the dual-scale slot gets
// 1/sum e so that pi = scale * e feeds the output law like a distribution.
__global__ void __launch_bounds__(256) k_dual_max(const double *__restrict__ part, int nblocks,
                                                  double *__restrict__ scal) {
  __shared__ double red[16];
  double m = -INFINITY;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) m = fmax(m, part[(int64_t)b * kPartStrideDev + 1]);
  m = block_max(m, red);
  if (threadIdx.x == 0) scal[kScalDMax] = m;
}
__global__ void __launch_bounds__(256) k_dual_exp(double *__restrict__ D, uint64_t nx, double beta,
                                                  const double *__restrict__ scal, double *__restrict__ part) {
  __shared__ double red[16];
  const double dm = scal[kScalDMax];
  double s = 0.0;
  const uint64_t per = (nx + gridDim.x - 1) / gridDim.x;
  const uint64_t b = (uint64_t)blockIdx.x * per, e = b + per < nx ? b + per : nx;
  for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    const double v = exp(beta * (D[i] - dm));
    D[i] = v;
    s += v;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void __launch_bounds__(256) k_dual_final(const double *__restrict__ part, int np, double beta, double mult,
                                                    double *__restrict__ scal, double *__restrict__ u_out) {
  __shared__ double red[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  s = mult * block_sum(s, red);
  if (threadIdx.x == 0) {
    scal[kScalSumW] = s;
    scal[kScalDualScale] = 1.0 / s;
    *u_out = scal[kScalDMax] + log(s) / beta;
  }
}
int launch_dual_weights(double *D, uint64_t nx, double beta, double mult, double *part, int nblocks, double *scal,
                        double *u_out, hipStream_t s) {
  hipLaunchKernelGGL(k_dual_max, dim3(1), dim3(256), 0, s, part, nblocks, scal);
  const int np = 1024;
  hipLaunchKernelGGL(k_dual_exp, dim3(np), dim3(256), 0, s, D, nx, beta, scal, part);
  hipLaunchKernelGGL(k_dual_final, dim3(1), dim3(256), 0, s, part, np, beta, mult, scal, u_out);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// dst[y] = mean of src over the complement/reversal orbit of y (the members
// summed in ascending index order, so every member gets the same value)
__global__ void k_orbit_mean(const double *__restrict__ src, int k, double *__restrict__ dst) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= (1u << k)) return;
  uint32_t m[4] = {y, comp_bits(y, k), rev_bits(y, k), rev_bits(comp_bits(y, k), k)};
  for (int i = 1; i < 4; ++i)  // sort the four indices
    for (int j = i; j > 0 && m[j] < m[j - 1]; --j) {
      const uint32_t t = m[j];
      m[j] = m[j - 1];
      m[j - 1] = t;
    }
  dst[y] = 0.25 * (((src[m[0]] + src[m[1]]) + src[m[2]]) + src[m[3]]);
}
int launch_orbit_mean(const double *src, int k, double *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_orbit_mean, dim3(((1u << k) + 255) / 256), dim3(256), 0, s, src, k, dst);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// Adam (Kingma & Ba) with bias correction, descending on U
__global__ void k_adam(double *__restrict__ th, const double *__restrict__ g, double *__restrict__ m,
                       double *__restrict__ v, uint64_t len, double lr, double b1, double b2, double eps, double c1,
                       double c2) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const double gi = g[i];
  const double mi = b1 * m[i] + (1.0 - b1) * gi;
  const double vi = b2 * v[i] + (1.0 - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr * (mi / c1) / (sqrt(vi / c2) + eps);
}
int launch_adam(double *theta, const double *grad, double *m, double *v, uint64_t len, double lr, double b1, double b2,
                double eps, int step, hipStream_t s) {
  const double c1 = 1.0 - pow(b1, step), c2 = 1.0 - pow(b2, step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, theta, grad, m, v, len, lr, b1, b2,
                     eps, c1, c2);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// out = a - scale_b * b
__global__ void k_vec_sub(const double *__restrict__ a, const double *__restrict__ b, double sb,
                          double *__restrict__ out, uint64_t len) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) out[i] = a[i] - sb * b[i];
}
int launch_vec_sub(const double *a, const double *b, double scale_b, double *out, uint64_t len, hipStream_t s) {
  hipLaunchKernelGGL(k_vec_sub, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, a, b, scale_b, out, len);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// dst[r*drs + c*dcs] = src[r*srs + c*scs] for r < R, c < Cc
__global__ void k_copy2d(const double *__restrict__ src, int64_t srs, int64_t scs, double *__restrict__ dst,
                         int64_t drs, int64_t dcs, int64_t R, int64_t Cc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R * Cc) return;
  const int64_t r = i / Cc, c = i % Cc;
  dst[r * drs + c * dcs] = src[r * srs + c * scs];
}
int launch_copy2d(const double *src, int64_t srs, int64_t scs, double *dst, int64_t drs, int64_t dcs, int64_t R,
                  int64_t Cc, hipStream_t s) {
  const int64_t tot = R * Cc;
  hipLaunchKernelGGL(k_copy2d, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, src, srs, scs, dst, drs, dcs, R, Cc);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// sum_i a_i * b_i (b may be null: plain sum) -> out[0]; max_i b_i -> out[1] when b given
__global__ void __launch_bounds__(256) k_dot_partial(const double *__restrict__ a, const double *__restrict__ b,
                                                     uint64_t len, double *__restrict__ part) {
  __shared__ double red[48];
  double s0 = 0.0, m1 = -INFINITY, s2 = 0.0;
  const uint64_t per = (len + gridDim.x - 1) / gridDim.x;
  const uint64_t lo = (uint64_t)blockIdx.x * per, hi = lo + per < len ? lo + per : len;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const double av = a[i];
    s2 += av;
    if (b) {
      s0 += av * b[i];
      m1 = fmax(m1, b[i]);
    }
  }
  block_reduce3(s0, m1, s2, red);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = s0;
    part[3 * blockIdx.x + 1] = m1;
    part[3 * blockIdx.x + 2] = s2;
  }
}
__global__ void __launch_bounds__(256) k_dot_final(const double *__restrict__ part, int np, double *__restrict__ out) {
  __shared__ double red[48];
  double s0 = 0.0, m1 = -INFINITY, s2 = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    s0 += part[3 * i];
    m1 = fmax(m1, part[3 * i + 1]);
    s2 += part[3 * i + 2];
  }
  block_reduce3(s0, m1, s2, red);
  if (threadIdx.x == 0) {
    out[0] = s0;
    out[1] = m1;
    out[2] = s2;
  }
}
int launch_dot(const double *a, const double *b, uint64_t len, double *part, double *out, hipStream_t s) {
  const int np = 1024;
  hipLaunchKernelGGL(k_dot_partial, dim3(np), dim3(256), 0, s, a, b, len, part);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, part, np, out);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// Complement image of the half space: P[~x] = P[x] for x < 2^n / 2
// (~x = 2^n - 1 - x), restoring the rows x1 >= N1/2 after half-space steps.
__global__ void k_reflect_half(double *__restrict__ P, uint64_t nx) {
  const uint64_t half = nx >> 1;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half; i += (uint64_t)gridDim.x * blockDim.x)
    P[nx - 1 - i] = P[i];
}
int launch_reflect_half(double *P, uint64_t nx, hipStream_t s) {
  const uint64_t half = nx >> 1;
  const unsigned blocks = (unsigned)std::min<uint64_t>(8192, (half + 255) / 256);
  hipLaunchKernelGGL(k_reflect_half, dim3(blocks), dim3(256), 0, s, P, nx);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

}  // namespace bdc
