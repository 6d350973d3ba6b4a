/* Host-only exercise of the CPU oracle (oracle/bdc_oracle.c) built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_host_sanitizers.py: every entry point at small sizes, including
 * the edge cases k = 1, k = n - 1, k = n and n = 32 for the single-input sums. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/bdc_oracle.h"

int main(void) {
  int bad = 0;
  uint32_t *c = malloc(sizeof(uint32_t) << 16); /* n = k = 8: 2^16 counts */
  for (int n = 1; n <= 8; ++n)
    for (int k = 0; k <= n; ++k) bad |= or_fill_cache(n, k, c) != 0;
  free(c);
  bad |= or_orbit_list(12, NULL) != (4096 + 64 + 64) / 4;
  int cases[][2] = {{10, 1}, {10, 5}, {11, 10}, {9, 9}, {12, 7}};
  for (unsigned t = 0; t < sizeof cases / sizeof cases[0]; ++t) {
    int n = cases[t][0], k = cases[t][1];
    or_ctx *ctx = or_ctx_create(n, k);
    if (!ctx) return 2;
    double mx = 0;
    for (int i = 0; i < 3; ++i) mx = or_ctx_iterate(ctx);
    uint64_t no = or_ctx_num_orbits(ctx);
    double *kl = malloc(no * sizeof(double)), *ld = malloc((sizeof(double) << (k - 1)) + 8);
    or_ctx_kl(ctx, kl);
    or_ctx_log_den(ctx, ld);
    double r = or_ctx_rate(ctx), m = or_ctx_max_kl(ctx), s = or_ctx_sample_pass(ctx, no / 3, no / 2);
    bad |= !(isfinite(mx) && isfinite(r) && isfinite(m) && isfinite(s) && r <= m + 1e-12);
    free(kl);
    free(ld);
    or_ctx_destroy(ctx);
  }
  double w[256], q[16], D[256], th[16], g[16];
  for (int i = 0; i < 256; ++i) w[i] = 1.0 / 256;
  or_full_output_law(8, 4, w, q);
  for (int y = 0; y < 16; ++y) th[y] = log(q[y]);
  or_full_kl(8, 4, th, D);
  double U = or_full_dual(8, 4, th, 64.0, g);
  bad |= !isfinite(U);
  double lq[1 << 10];
  for (int y = 0; y < (1 << 10); ++y) lq[y] = -10 * log(2.0);
  bad |= !isfinite(or_entropy_x(32, 10, 0xDEADBEEFu)) || !isfinite(or_kl_x(32, 10, 0x12345678u, lq));
  bad |= !isfinite(or_entropy_x(20, 19, 0xABCDEu)) || !isfinite(or_kl_x(11, 10, 0x5A5u, lq));
  printf("oracle checks: %s\n", bad ? "FAILED" : "ok");
  return bad;
}
