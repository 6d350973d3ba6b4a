/*
 * bdc_oracle.c — CPU restatement (FP64, OpenMP) of the reference Blahut–Arimoto
 * method for the binary deletion channel BDC_{n,k}.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker.  Only tests/, the smoke entry and
 * the bench's cpu_baseline leg load it.  The product library never links it and
 * has no CPU fallback that routes here.
 *
 * Each function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).  How each part is pinned:
 *   - or_count / or_fill_cache: byte-compared against files written by the
 *     reference generator built from source (oracle/Makefile -> oracle/_ref/),
 *     digests committed under tests/golden/, plus brute force over position
 *     subsets and the row/column-sum identities (tests/test_oracle_counts.py).
 *   - or_rep / or_orbit_size / or_orbit_list: Burnside orbit counts and a
 *     brute-force orbit enumeration (tests/test_oracle_symmetry.py).
 *   - or_ctx_* (reference BA iteration, bounds, stopping): a dense textbook
 *     Blahut–Arimoto in numpy on the full 2^n x 2^k channel matrix for n <= 12,
 *     the closed forms C_{n,1} = 1 bit and C_{n,n} = n bits, monotone I(p_t) and
 *     the sandwich log Z_t <= I(p_{t+1}) <= max_x D(x||q_t)
 *     (tests/test_oracle_ba.py).
 *   - or_full_dual: central finite differences and the closed form
 *     dU/dtheta = q - W^T pi (tests/test_oracle_dual.py).
 *   - or_entropy_x / or_kl_x: the dense channel, a brute-force sweep over every
 *     output (k close to n and k = n), the closed forms for x = 0^a 1^b and for
 *     one deletion (sum over runs of l ln l, at n = 29), and D(x||uniform)
 *     (tests/test_oracle_counts.py).
 *   - or_sparse_output_law: or_full_output_law on the same weights.
 *   - or_ctx_sample_pass: its checksum against a dense textbook BA over all
 *     orbits and over a slice (tests/test_oracle_ba.py).
 */
#include "bdc_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX_N 33 /* backend/config.hpp:24 */

/* ------------------------------------------------------------------ basics */

/* backend/BAA.cu:15-24 */
uint32_t or_nck(uint32_t n, uint32_t k) {
  if (k > n) return 0;
  if (k > n - k) k = n - k;
  uint64_t r = 1;
  for (uint32_t i = 1; i <= k; ++i) r = r * (n - i + 1) / i;
  return (uint32_t)r;
}

/* backend/generate_bit_transition_cache.cpp:23-44: dp[i][j] = number of ways to
 * form the first j symbols of y from the first i symbols of x.  Kept as one
 * rolling row (descending j) — same recurrence, same uint32 wrap-around. */
uint32_t or_count(uint32_t x, int n, uint32_t y, int k) {
  uint32_t dp[OR_MAX_N + 1];
  dp[0] = 1;
  for (int j = 1; j <= k; ++j) dp[j] = 0;
  for (int i = 1; i <= n; ++i) {
    uint32_t xi = (x >> (n - i)) & 1u;
    for (int j = (i < k ? i : k); j >= 1; --j) {
      uint32_t yj = (y >> (k - j)) & 1u;
      if (xi == yj) dp[j] += dp[j - 1];
    }
  }
  return dp[k];
}

/* generate_bit_transition_cache.cpp:56-72 with cache_io.h:17-26 layout:
 * row x (n bits) major, column y (k bits), index (x << k) | y. */
int or_fill_cache(int n, int k, uint32_t *out) {
  if (n < 0 || k < 0 || n > 20 || k > n) return -1;
  const uint64_t nx = 1ull << n, ny = 1ull << k;
#pragma omp parallel for schedule(dynamic, 64)
  for (uint64_t x = 0; x < nx; ++x)
    for (uint64_t y = 0; y < ny; ++y) out[(x << k) | y] = or_count((uint32_t)x, n, (uint32_t)y, k);
  return 0;
}

/* ---------------------------------------------------------------- symmetry */

/* backend/BAA.cu:195-202 */
static uint32_t rev32(uint32_t x) {
  x = ((x & 0x55555555u) << 1) | ((x & 0xAAAAAAAAu) >> 1);
  x = ((x & 0x33333333u) << 2) | ((x & 0xCCCCCCCCu) >> 2);
  x = ((x & 0x0F0F0F0Fu) << 4) | ((x & 0xF0F0F0F0u) >> 4);
  x = ((x & 0x00FF00FFu) << 8) | ((x & 0xFF00FF00u) >> 8);
  return (x << 16) | (x >> 16);
}
/* backend/BAA.cu:205-213 (k = 0 handled: the empty string is its own image) */
static uint32_t comp_k(uint32_t y, int k) { return k == 32 ? ~y : (y ^ ((1u << k) - 1u)); }
static uint32_t rev_k(uint32_t y, int k) { return k == 0 ? 0u : rev32(y << (32 - k)); }

/* backend/BAA.cu:216-227 */
uint32_t or_rep(uint32_t x, int n) {
  uint32_t r = rev_k(x, n), c = comp_k(x, n), rc = rev_k(c, n);
  uint32_t m1 = r < x ? r : x, m2 = c < rc ? c : rc;
  return m2 < m1 ? m2 : m1;
}
/* backend/BAA.cu:230-241 */
int or_orbit_size(uint32_t x, int n) {
  uint32_t r = rev_k(x, n), rc = rev_k(comp_k(x, n), n);
  return (x == r || x == rc) ? 2 : 4;
}
/* backend/main_BAA.cu:46-50 (uint64 loop bound so n = 32 is defined) */
uint64_t or_orbit_list(int n, uint32_t *out) {
  uint64_t cnt = 0, nx = 1ull << n;
  for (uint64_t x = 0; x < nx; ++x)
    if (or_rep((uint32_t)x, n) == (uint32_t)x) {
      if (out) out[cnt] = (uint32_t)x;
      ++cnt;
    }
  return cnt;
}

/* ------------------------------------------------- per-input DP + unranking */

typedef struct {
  uint8_t nextpos[OR_MAX_N + 1][2];
  uint32_t cnt[OR_MAX_N + 2][OR_MAX_N + 1];
} dp_tab;

/* backend/BAA.cu:261-303: nextpos[i][b] = first position >= i holding b (n if
 * none); cnt[i][r] = number of distinct length-r subsequences of x[i..n). */
static void fill_dp(uint32_t x, int n, int k, dp_tab *t) {
  for (int b = 0; b < 2; ++b) {
    t->nextpos[n][b] = (uint8_t)n;
    int last = n;
    for (int i = n - 1; i >= 0; --i) {
      if ((int)((x >> (n - i - 1)) & 1u) == b) last = i;
      t->nextpos[i][b] = (uint8_t)last;
    }
  }
  for (int i = 0; i <= n + 1; ++i) {
    t->cnt[i][0] = 1;
    for (int r = 1; r <= k; ++r) t->cnt[i][r] = 0;
  }
  for (int i = n - 1; i >= 0; --i)
    for (int r = 1; r <= k; ++r) {
      uint32_t s = 0;
      int j0 = t->nextpos[i][0], j1 = t->nextpos[i][1];
      if (j0 < n) s += t->cnt[j0 + 1][r - 1];
      if (j1 < n) s += t->cnt[j1 + 1][r - 1];
      t->cnt[i][r] = s;
    }
}

/* backend/BAA.cu:304-335: the rank-th distinct length-k subsequence of x in
 * lexicographic order (0 < 1), MSB first. */
static uint32_t unrank(int n, int k, uint32_t rank, const dp_tab *t) {
  int i = 0, rem = k;
  uint32_t y = 0;
  while (rem > 0) {
    int j0 = t->nextpos[i][0];
    uint32_t c0 = j0 < n ? t->cnt[j0 + 1][rem - 1] : 0;
    if (rank < c0) {
      i = j0 + 1;
      --rem;
      continue;
    }
    rank -= c0;
    int j1 = t->nextpos[i][1];
    y |= 1u << (rem - 1);
    i = j1 + 1;
    --rem;
  }
  return y;
}

/* ------------------------------------------------------ reference BA context */

struct or_ctx {
  int n, k, n1, n2;
  uint64_t norb;
  uint32_t *orbits; /* ascending representatives */
  double *Q;        /* per-string probability of each orbit (main_BAA.cu:53-65) */
  uint32_t *cache1[OR_MAX_N]; /* cache_{n1,j}, j < n1 */
  uint32_t *cache2[OR_MAX_N]; /* cache_{n2,m}, m < n2 */
  double inv_c, log_c;        /* 1/C(n,k), ln C(n,k) (BAA.cu:96-111) */
};

/* backend/BAA.cu:340-395: c(x,y) by the split sum over j = symbols of y taken
 * from the high n1 bits of x; identity factor when a half is consumed whole. */
static uint32_t split_count(const or_ctx *c, uint32_t x, uint32_t y) {
  const int n1 = c->n1, n2 = c->n2, k = c->k;
  uint32_t x1 = x >> n2, x2 = x & ((1u << n2) - 1u);
  uint32_t total = 0;
  for (int j = 0; j <= k; ++j) {
    if (j > n1 || k - j > n2) continue;
    uint32_t y1 = (k - j) >= 32 ? 0 : (y >> (k - j));
    uint32_t y2 = y & ((k - j) >= 32 ? 0xFFFFFFFFu : ((1u << (k - j)) - 1u));
    uint32_t a = (j == n1) ? (uint32_t)(x1 == y1) : c->cache1[j][((uint64_t)x1 << j) | y1];
    uint32_t b = (k - j == n2) ? (uint32_t)(x2 == y2) : c->cache2[k - j][((uint64_t)x2 << (k - j)) | y2];
    total += a * b;
  }
  return total;
}

or_ctx *or_ctx_create(int n, int k) {
  if (n < 1 || n > 30 || k < 1 || k > n) return NULL;
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->n = n;
  c->k = k;
  c->n1 = (n + 1) / 2; /* BAA.cu:341-342 */
  c->n2 = n - c->n1;
  c->norb = or_orbit_list(n, NULL);
  c->orbits = (uint32_t *)malloc(c->norb * sizeof(uint32_t));
  or_orbit_list(n, c->orbits);
  c->Q = (double *)malloc(c->norb * sizeof(double));
  for (uint64_t i = 0; i < c->norb; ++i) c->Q[i] = 1.0 / (double)(1ull << n); /* main_BAA.cu:64 */
  for (int j = 0; j < c->n1; ++j) {
    c->cache1[j] = (uint32_t *)malloc(sizeof(uint32_t) << (c->n1 + j));
    or_fill_cache(c->n1, j, c->cache1[j]);
  }
  for (int m = 0; m < c->n2; ++m) {
    if (c->n2 == c->n1) {
      c->cache2[m] = c->cache1[m];
    } else {
      c->cache2[m] = (uint32_t *)malloc(sizeof(uint32_t) << (c->n2 + m));
      or_fill_cache(c->n2, m, c->cache2[m]);
    }
  }
  double nck = (double)or_nck((uint32_t)n, (uint32_t)k);
  c->inv_c = 1.0 / nck;
  c->log_c = log(nck);
  return c;
}

void or_ctx_destroy(or_ctx *c) {
  if (!c) return;
  for (int j = 0; j < c->n1; ++j) free(c->cache1[j]);
  if (c->n2 != c->n1)
    for (int m = 0; m < c->n2; ++m) free(c->cache2[m]);
  free(c->orbits);
  free(c->Q);
  free(c);
}

uint64_t or_ctx_num_orbits(const or_ctx *c) { return c->norb; }
void or_ctx_orbits(const or_ctx *c, uint32_t *out) { memcpy(out, c->orbits, c->norb * sizeof(uint32_t)); }
void or_ctx_set_q(or_ctx *c, const double *Q) { memcpy(c->Q, Q, c->norb * sizeof(double)); }
void or_ctx_get_q(const or_ctx *c, double *Q) { memcpy(Q, c->Q, c->norb * sizeof(double)); }

/* backend/BAA.cu:429-485 (compute_den_kernel) + 642-724 (compute_all_log_D) +
 * 573-580 (apply_log_kernel: zeros stay zero).  Thread-private accumulators
 * replace the FP64 atomics; representatives of y are < 2^(k-1) (BAA.cu:651-652). */
static void log_den_of(const or_ctx *c, const double *Q, double *log_den) {
  const int n = c->n, k = c->k;
  const uint64_t nr = 1ull << (k - 1);
  memset(log_den, 0, nr * sizeof(double));
#pragma omp parallel
  {
    double *acc = (double *)calloc(nr, sizeof(double));
    dp_tab t;
#pragma omp for schedule(dynamic, 16)
    for (uint64_t i = 0; i < c->norb; ++i) {
      uint32_t x = c->orbits[i];
      fill_dp(x, n, k, &t);
      double ox = (double)or_orbit_size(x, n);
      uint32_t nd = t.cnt[0][k];
      for (uint32_t r = 0; r < nd; ++r) {
        uint32_t y = unrank(n, k, r, &t);
        uint32_t ry = or_rep(y, k);
        double factor = ox / (double)or_orbit_size(y, k);
        acc[ry] += factor * (Q[i] * (double)split_count(c, x, y) * c->inv_c);
      }
    }
#pragma omp critical
    for (uint64_t y = 0; y < nr; ++y) log_den[y] += acc[y];
    free(acc);
  }
  for (uint64_t y = 0; y < nr; ++y)
    if (log_den[y] != 0.0) log_den[y] = log(log_den[y]);
}

void or_ctx_log_den(or_ctx *c, double *log_den) { log_den_of(c, c->Q, log_den); }

/* backend/BAA.cu:401-411 + 487-530: per orbit, sum over distinct y of
 * W (ln W + ln Q_x - ln q(rep y)) = ln Q_x + D(x||q).  With with_logq = 0 the
 * ln Q_x term is dropped, giving D(x||q) as in KL_term (BAA.cu:413-425, 533-571). */
static void per_x_terms(const or_ctx *c, const double *log_den, int with_logq, double *out) {
  const int n = c->n, k = c->k;
#pragma omp parallel
  {
    dp_tab t;
#pragma omp for schedule(dynamic, 16)
    for (uint64_t i = 0; i < c->norb; ++i) {
      uint32_t x = c->orbits[i];
      fill_dp(x, n, k, &t);
      double lq = 0.0;
      if (with_logq) lq = c->Q[i] != 0.0 ? log(c->Q[i]) : 0.0; /* apply_log_kernel */
      uint32_t nd = t.cnt[0][k];
      double s = 0.0;
      for (uint32_t r = 0; r < nd; ++r) {
        uint32_t y = unrank(n, k, r, &t);
        uint32_t cnt = split_count(c, x, y);
        double w = (double)cnt * c->inv_c;
        double lw = (cnt ? log((double)cnt) : 0.0) - c->log_c; /* BAA.cu:151-157 ln 0 := 0 */
        s += w * (lw + lq - log_den[or_rep(y, k)]);
      }
      out[i] = s;
    }
  }
}

/* backend/BAA.cu:590-639 then backend/BAA.cu:916-946 */
double or_ctx_iterate(or_ctx *c) {
  const uint64_t N = c->norb;
  double *log_den = (double *)malloc(sizeof(double) << (c->k - 1));
  double *logw = (double *)malloc(N * sizeof(double));
  log_den_of(c, c->Q, log_den);
  per_x_terms(c, log_den, 1, logw);
  double mx = -INFINITY;
  for (uint64_t i = 0; i < N; ++i)
    if (logw[i] > mx) mx = logw[i];
  double wsum = 0.0;
  for (uint64_t i = 0; i < N; ++i) {
    logw[i] = exp(logw[i] - mx);
    wsum += (double)or_orbit_size(c->orbits[i], c->n) * logw[i]; /* BAA.cuh:66-79 */
  }
  double dev = -INFINITY;
  for (uint64_t i = 0; i < N; ++i) {
    double qn = logw[i] / wsum;
    double d = log2(qn) - log2(c->Q[i]);
    if (d > dev) dev = d;
    c->Q[i] = qn;
  }
  free(log_den);
  free(logw);
  return dev;
}

void or_ctx_kl(or_ctx *c, double *kl) {
  double *log_den = (double *)malloc(sizeof(double) << (c->k - 1));
  log_den_of(c, c->Q, log_den);
  per_x_terms(c, log_den, 0, kl);
  free(log_den);
}

/* backend/BAA.cu:820-913 (compute_rate_terms + reduce), main_BAA.cu:123-124 */
double or_ctx_rate(or_ctx *c) {
  double *kl = (double *)malloc(c->norb * sizeof(double));
  or_ctx_kl(c, kl);
  double r = 0.0;
  for (uint64_t i = 0; i < c->norb; ++i) r += (double)or_orbit_size(c->orbits[i], c->n) * c->Q[i] * kl[i];
  free(kl);
  return r;
}

/* certified upper bound max_x D(x||q): no reference counterpart (the reference
 * reports I(p) + mx, README.md:176-183); computed from the same per-x terms */
double or_ctx_max_kl(or_ctx *c) {
  double *kl = (double *)malloc(c->norb * sizeof(double));
  or_ctx_kl(c, kl);
  double m = -INFINITY;
  for (uint64_t i = 0; i < c->norb; ++i)
    if (kl[i] > m) m = kl[i];
  free(kl);
  return m;
}

/* both enumeration passes of one reference iteration, restricted to a slice of
 * orbits (CPU-baseline sampling); returns a checksum so the work is not elided */
double or_ctx_sample_pass(or_ctx *c, uint64_t first, uint64_t count) {
  const int n = c->n, k = c->k;
  if (first >= c->norb) return 0.0;
  if (first + count > c->norb) count = c->norb - first;
  const uint64_t nr = 1ull << (k - 1);
  double *den = (double *)calloc(nr, sizeof(double));
  double check = 0.0;
#pragma omp parallel
  {
    double *acc = (double *)calloc(nr, sizeof(double));
    dp_tab t;
#pragma omp for schedule(dynamic, 16)
    for (uint64_t i = first; i < first + count; ++i) {
      uint32_t x = c->orbits[i];
      fill_dp(x, n, k, &t);
      double ox = (double)or_orbit_size(x, n);
      uint32_t nd = t.cnt[0][k];
      for (uint32_t r = 0; r < nd; ++r) {
        uint32_t y = unrank(n, k, r, &t);
        acc[or_rep(y, k)] += ox / (double)or_orbit_size(y, k) * (c->Q[i] * (double)split_count(c, x, y) * c->inv_c);
      }
    }
#pragma omp critical
    for (uint64_t y = 0; y < nr; ++y) den[y] += acc[y];
    free(acc);
  }
  for (uint64_t y = 0; y < nr; ++y) den[y] = den[y] != 0.0 ? log(den[y]) : 0.0;
#pragma omp parallel reduction(+ : check)
  {
    dp_tab t;
#pragma omp for schedule(dynamic, 16)
    for (uint64_t i = first; i < first + count; ++i) {
      uint32_t x = c->orbits[i];
      fill_dp(x, n, k, &t);
      double lq = log(c->Q[i]);
      uint32_t nd = t.cnt[0][k];
      double s = 0.0;
      for (uint32_t r = 0; r < nd; ++r) {
        uint32_t y = unrank(n, k, r, &t);
        uint32_t cnt = split_count(c, x, y);
        double w = (double)cnt * c->inv_c;
        s += w * (log((double)cnt) - c->log_c + lq - den[or_rep(y, k)]);
      }
      check += s;
    }
  }
  free(den);
  return check;
}

/* ------------------------------------------------------ full-space helpers */

void or_full_output_law(int n, int k, const double *w, double *q) {
  const uint64_t nx = 1ull << n, ny = 1ull << k;
  const double inv_c = 1.0 / (double)or_nck((uint32_t)n, (uint32_t)k);
  memset(q, 0, ny * sizeof(double));
#pragma omp parallel
  {
    double *acc = (double *)calloc(ny, sizeof(double));
#pragma omp for schedule(static)
    for (uint64_t x = 0; x < nx; ++x) {
      if (w[x] == 0.0) continue;
      for (uint64_t y = 0; y < ny; ++y) {
        uint32_t c = or_count((uint32_t)x, n, (uint32_t)y, k);
        if (c) acc[y] += w[x] * (double)c * inv_c;
      }
    }
#pragma omp critical
    for (uint64_t y = 0; y < ny; ++y) q[y] += acc[y];
    free(acc);
  }
}

void or_full_kl(int n, int k, const double *logq, double *D) {
  const uint64_t nx = 1ull << n, ny = 1ull << k;
  const double nck = (double)or_nck((uint32_t)n, (uint32_t)k);
  const double inv_c = 1.0 / nck, log_c = log(nck);
#pragma omp parallel for schedule(static)
  for (uint64_t x = 0; x < nx; ++x) {
    double s = 0.0;
    for (uint64_t y = 0; y < ny; ++y) {
      uint32_t c = or_count((uint32_t)x, n, (uint32_t)y, k);
      if (c) s += (double)c * inv_c * (log((double)c) - log_c - logq[y]);
    }
    D[x] = s;
  }
}

/* U_beta and its gradient in closed form: with q = softmax(theta) and
 * pi = softmax_x(beta D(x||q)), dU/dtheta_y = q(y) - sum_x pi(x) W(y|x).
 * No reference counterpart. */
double or_full_dual(int n, int k, const double *theta, double beta, double *grad) {
  const uint64_t nx = 1ull << n, ny = 1ull << k;
  double tmax = -INFINITY;
  for (uint64_t y = 0; y < ny; ++y)
    if (theta[y] > tmax) tmax = theta[y];
  double zs = 0.0;
  for (uint64_t y = 0; y < ny; ++y) zs += exp(theta[y] - tmax);
  const double lse = tmax + log(zs);
  double *logq = (double *)malloc(ny * sizeof(double));
  for (uint64_t y = 0; y < ny; ++y) logq[y] = theta[y] - lse;
  double *D = (double *)malloc(nx * sizeof(double));
  or_full_kl(n, k, logq, D);
  double dmax = -INFINITY;
  for (uint64_t x = 0; x < nx; ++x)
    if (D[x] > dmax) dmax = D[x];
  double s = 0.0;
  for (uint64_t x = 0; x < nx; ++x) s += exp(beta * (D[x] - dmax));
  const double U = dmax + log(s) / beta;
  if (grad) {
    for (uint64_t x = 0; x < nx; ++x) D[x] = exp(beta * (D[x] - dmax)) / s; /* pi */
    or_full_output_law(n, k, D, grad);
    for (uint64_t y = 0; y < ny; ++y) grad[y] = exp(logq[y]) - grad[y];
  }
  free(D);
  free(logq);
  return U;
}

int or_num_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------- per-input sums */

/* Single-input sums over the outputs y with c(x,y) > 0, visited in ascending
 * order through the reference's own enumeration of distinct subsequences
 * (fill_DP_tables + unrank_k, backend/BAA.cu:261-335): the same nonzero terms in
 * the same order as a sweep over all 2^k outputs, at a cost of (number of
 * distinct subsequences) x O(nk) instead of 2^k x O(nk), so k close to n is
 * practical (a few thousand outputs per input at n = 29, k = 24). */

/* sum_y c(x,y) ln c(x,y) in nats: the W ln W part of log_W_term
 * (backend/BAA.cu:401-411) times C(n,k), for a single input x (n <= 32).
 * Terms with c <= 1 contribute 0. */
double or_entropy_x(int n, int k, uint32_t x) {
  /* extended-precision accumulation: at n = 32 the sum has up to 2^16 terms and
   * reaches ~6e9, where a plain FP64 running sum loses ~1e-12 relative */
  dp_tab t;
  fill_dp(x, n, k, &t);
  const uint32_t nd = t.cnt[0][k];
  long double h = 0.0L;
  for (uint32_t r = 0; r < nd; ++r) {
    const uint32_t c = or_count(x, n, unrank(n, k, r, &t), k);
    if (c > 1) h += (long double)c * logl((long double)c);
  }
  return (double)h;
}

/* D(x||q) = sum_y W(y|x) (ln W(y|x) - ln q(y)) in nats for a single input x,
 * W = c / C(n,k) (compute_KL_kernel / KL_term, backend/BAA.cu:419-425, 533-571);
 * logq holds ln q over all 2^k outputs */
double or_kl_x(int n, int k, uint32_t x, const double *logq) {
  dp_tab t;
  fill_dp(x, n, k, &t);
  const uint32_t nd = t.cnt[0][k];
  const long double C = (long double)or_nck((uint32_t)n, (uint32_t)k), lnC = logl(C);
  long double d = 0.0L;
  for (uint32_t r = 0; r < nd; ++r) {
    const uint32_t y = unrank(n, k, r, &t);
    const uint32_t c = or_count(x, n, y, k);
    d += (c / C) * (logl((long double)c) - lnC - (long double)logq[y]);
  }
  return (double)d;
}

/* q(y) = sum_i w[i] c(xs[i], y) / C(n,k) over all 2^k outputs for a weight
 * vector supported on the inputs xs[0..count) (compute_den_kernel,
 * backend/BAA.cu:429-485, restricted to a support).  The counts of one input
 * for all outputs come from one pass over its symbols: cnt_t[z] = embeddings
 * of z (t symbols) in the prefix read so far, and reading symbol b adds
 * cnt_{t-1}[z'] to cnt_t[z' b] (longest t first), the recurrence of
 * generate_bit_transition_cache.cpp:23-44 run for every output at once. */
void or_sparse_output_law(int n, int k, uint64_t count, const uint32_t *xs, const double *w, double *q) {
  const uint64_t ny = 1ull << k;
  const double inv_c = 1.0 / (double)or_nck((uint32_t)n, (uint32_t)k);
  memset(q, 0, ny * sizeof(double));
#pragma omp parallel
  {
    double *acc = (double *)calloc(ny, sizeof(double));
    uint32_t *cnt = (uint32_t *)malloc(sizeof(uint32_t) << (k + 1)); /* level t at offset 2^t */
#pragma omp for schedule(dynamic, 4)
    for (uint64_t i = 0; i < count; ++i) {
      memset(cnt, 0, sizeof(uint32_t) << (k + 1));
      cnt[1] = 1; /* the empty string */
      for (int pos = 0; pos < n; ++pos) {
        const uint32_t b = (xs[i] >> (n - 1 - pos)) & 1u;
        for (int tl = (pos + 1 < k ? pos + 1 : k); tl >= 1; --tl) {
          uint32_t *dst = cnt + (1ull << tl);
          const uint32_t *src = cnt + (1ull << (tl - 1));
          for (uint64_t z = 0; z < (1ull << (tl - 1)); ++z) dst[(z << 1) | b] += src[z];
        }
      }
      const uint32_t *ck = cnt + ny;
      for (uint64_t y = 0; y < ny; ++y)
        if (ck[y]) acc[y] += w[i] * (double)ck[y] * inv_c;
    }
#pragma omp critical
    for (uint64_t y = 0; y < ny; ++y) q[y] += acc[y];
    free(acc);
    free(cnt);
  }
}
