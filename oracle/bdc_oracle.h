/*
 * bdc_oracle.h — CPU oracle for the BDC_{n,k} Blahut–Arimoto hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library never does.
 *
 * All citations are to the reference repository Pintz45/BDC-Upper-Bounds-GPU.
 * Bit order is MSB-first everywhere: bit (n-1-i) of x is the i-th symbol x_i
 * (backend/generate_bit_transition_cache.cpp:15, backend/BAA.cu:270-271).
 */
#ifndef BDC_ORACLE_H
#define BDC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* binomial coefficient as in backend/BAA.cu:15-24 (uint64 intermediate, uint32 result) */
uint32_t or_nck(uint32_t n, uint32_t k);
/* c(x,y): occurrences of y (k bits) as a subsequence of x (n bits);
 * the O(nk) prefix DP of backend/generate_bit_transition_cache.cpp:23-44 */
uint32_t or_count(uint32_t x, int n, uint32_t y, int k);
/* cache_{n,k}: 2^(n+k) uint32 counts, index (x<<k)|y (generate_bit_transition_cache.cpp:56-72) */
int or_fill_cache(int n, int k, uint32_t *out);

/* symmetry group {id, reverse, complement, reverse∘complement} (backend/BAA.cu:195-258) */
uint32_t or_rep(uint32_t x, int n);
int or_orbit_size(uint32_t x, int n);
/* ascending list of representatives (backend/main_BAA.cu:46-50); returns the count.
 * out may be NULL to only count. */
uint64_t or_orbit_list(int n, uint32_t *out);

/* ---- reference-method Blahut–Arimoto context (per-orbit enumeration) ---- */
typedef struct or_ctx or_ctx;
or_ctx *or_ctx_create(int n, int k);
void or_ctx_destroy(or_ctx *c);
uint64_t or_ctx_num_orbits(const or_ctx *c);
void or_ctx_orbits(const or_ctx *c, uint32_t *out);
void or_ctx_set_q(or_ctx *c, const double *Q);
void or_ctx_get_q(const or_ctx *c, double *Q);
/* one reference BA iteration (backend/BAA.cu:590-639) followed by the stopping
 * quantity mx = max log2(Q'/Q) (backend/BAA.cu:916-946). Returns mx. */
double or_ctx_iterate(or_ctx *c);
/* log q(y) at representatives (array of 2^(k-1)), as compute_all_log_D
 * (backend/BAA.cu:642-724, zeros stay zero per apply_log_kernel :573-580) */
void or_ctx_log_den(or_ctx *c, double *log_den);
/* per-orbit D(x||q) in nats, q from the current Q (compute_KL_kernel, BAA.cu:533-571) */
void or_ctx_kl(or_ctx *c, double *kl);
/* I(p) in nats = sum |O_x| Q_x D(x||q)  (compute_rate, BAA.cu:820-913) */
double or_ctx_rate(or_ctx *c);
/* max_x D(x||q) in nats for q from the current Q (certified upper bound) */
double or_ctx_max_kl(or_ctx *c);
/* Reference-method work on a subset of orbits, for CPU-baseline timing:
 * runs both enumeration passes (denominator and log W) over orbits
 * [first, first+count) and returns a checksum. */
double or_ctx_sample_pass(or_ctx *c, uint64_t first, uint64_t count);

/* ---- full-space helpers (no symmetry assumption) ---- */
/* q(y) = sum_x w(x) c(x,y)/C(n,k) over all 2^n inputs; w has 2^n entries, q has 2^k */
void or_full_output_law(int n, int k, const double *w, double *q);
/* D(x||q) for all 2^n x given log q over all 2^k y */
void or_full_kl(int n, int k, const double *logq, double *D);
/* smoothed dual bound U_beta(theta) = beta^-1 log sum_x exp(beta D(x||softmax theta));
 * writes dU/dtheta (2^k) when grad != NULL.  Returns U. */
double or_full_dual(int n, int k, const double *theta, double beta, double *grad);

/* single-input sums (any n <= 32, k <= n): sum_y c ln c, and D(x||q) given ln q
 * (2^k), over the distinct subsequences of x in ascending order */
double or_entropy_x(int n, int k, uint32_t x);
double or_kl_x(int n, int k, uint32_t x, const double *logq);
/* output law of a weight vector supported on count inputs xs (k <= 24) */
void or_sparse_output_law(int n, int k, uint64_t count, const uint32_t *xs, const double *w, double *q);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
