// ln(c) in FP64 for integers c >= 1, table-driven after Tang, "Table-driven
// implementation of the logarithm function in IEEE floating-point
// arithmetic", ACM TOMS 16(4), 1990:
//   c = 2^e m, m in [1, 2);  F = 1 + j/64 nearest to m;  f = m - F (exact)
//   ln c = e ln2 + ln F + log1p(f / F),  |f / F| <= 2^-7,
// with log1p by its degree-7 Taylor polynomial (truncation below 2^-59),
// ln F and 1/F from a 65-entry table computed on the host in long double,
// and ln2 split into a 32-bit head (exact products with e) and a tail.
#pragma once
#include <hip/hip_runtime.h>

namespace bdc {

struct LogTable {
  double lnF[65];   // ln(1 + j/64)
  double invF[65];  // 1 / (1 + j/64)
};

__device__ inline double log_tang_int(uint32_t c, const double *__restrict__ lnF, const double *__restrict__ invF) {
  constexpr double kLn2Hi = 0.6931471806019545;     // ln 2 rounded to 32 bits
  constexpr double kLn2Lo = -4.2009150726810846e-11;  // ln 2 - kLn2Hi
  int e;
  const double m = 2.0 * frexp((double)c, &e);  // (double)c = m 2^(e-1), m in [1, 2)
  e -= 1;
  const int j = (int)rint((m - 1.0) * 64.0);
  const double f = m - (1.0 + j * (1.0 / 64.0));
  const double u = f * invF[j];
  double p = 1.0 / 7.0;
  p = fma(p, u, -1.0 / 6.0);
  p = fma(p, u, 1.0 / 5.0);
  p = fma(p, u, -1.0 / 4.0);
  p = fma(p, u, 1.0 / 3.0);
  p = fma(p, u, -0.5);
  p = fma(p, u, 1.0);
  p *= u;
  return e * kLn2Hi + (lnF[j] + fma((double)e, kLn2Lo, p));
}

}  // namespace bdc
