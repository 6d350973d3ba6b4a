/*
 * bdc.h — C ABI of the MI355X Blahut–Arimoto engine for the binary deletion
 * channel BDC_{n,k} (libbdc.so, built from bdc-upper-bounds-gpu_amd/csrc/).
 *
 * The reference (Pintz45/BDC-Upper-Bounds-GPU) has no plugin or FFI layer; its
 * seam is the host function set declared in backend/BAA.cuh:16-29 and
 * backend/BAA.cuh:295-362, called only by compute_C_n_k (backend/main_BAA.cu).
 * Each entry point below names the reference function(s) it replaces.
 *
 * Conventions (all entry points):
 *   - plain C types and pointers only; host buffers are borrowed and copied;
 *   - return BDC_OK (0) or a negative status; the message of the last failure
 *     on the calling thread is available from bdc_last_error();
 *   - nothing calls exit() (the reference exits on every CUDA error,
 *     backend/BAA.cuh:5-13, and on a missing file, backend/utils.cc:46-54);
 *   - one context per thread; calls on one context are serialised by the
 *     caller; each context owns one HIP stream and all its device memory.
 * Bit order is MSB-first: bit (n-1-i) of an input x is its i-th symbol.
 * Probabilities are per string: p(x) for every x of an orbit is the orbit's
 * value, and sum_orbits |O| * value = 1 (backend/main_BAA.cu:64, BAA.cuh:66-79).
 */
#ifndef BDC_H
#define BDC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BDC_OK = 0,
  BDC_ERR_ARG = -1,      /* invalid argument (n, k, length, precision ...) */
  BDC_ERR_HIP = -2,      /* HIP runtime error */
  BDC_ERR_MEMORY = -3,   /* device allocation failed / insufficient HBM */
  BDC_ERR_IO = -4,       /* file missing, short or malformed */
  BDC_ERR_STATE = -5,    /* call out of order (e.g. iterate before p is set) */
  BDC_ERR_NUMERIC = -6,  /* a precondition of exact split arithmetic failed */
};

enum {
  BDC_PREC_FP64 = 0,  /* plain FP64 MFMA contractions */
  BDC_PREC_SPLIT16 = 1, /* split-FP16 MFMA (exact integer limbs, FP32 accumulate) */
};

typedef struct bdc_ctx bdc_ctx;

/* Per-iteration record.  All values in nats except mx_bits. */
typedef struct {
  double mx_bits; /* max_x log2(p_{t+1}(x)/p_t(x)) — the reference stopping quantity (BAA.cu:916-946) */
  double max_d;   /* max_x D(x||q_t) */
  double log_z;   /* ln sum_x p_t(x) exp(D(x||q_t)) */
  double info;    /* I(p_t) = sum_x p_t(x) D(x||q_t) */
} bdc_iter_stats;

/* Bounds of the current p, evaluated entirely in FP64.  Nats. */
typedef struct {
  double lower;      /* I(p) — the reference's reported rate (main_BAA.cu:123-125) */
  double upper_cert; /* max_x D(x||q(p)) — certified upper bound on C_{n,k} */
  double sum_p;      /* sum_x p(x) (should be 1) */
  double upper_ref;  /* I(p) + mx of the step that produced p (the reference's printed UB,
                        main_BAA.cu:123-130); NaN when p was set rather than iterated */
} bdc_certificate;

const char *bdc_last_error(void);
int bdc_version(void);

/* ---- caches: generate_bit_transition_cache.cpp:23-77, cache_io.h:7-44 ---- */
/* c(x,y) for all x in {0,1}^n, y in {0,1}^k into out (2^(n+k) uint32, index
 * (x<<k)|y), computed by the subsequence-count DP on the GPU.  0 <= k <= n <= 20. */
int bdc_generate_cache(int n, int k, uint32_t *out);
/* Writes <dir>/cache_{n'}_{k'} for 0 <= k' < n' <= n_max, byte-identical to the
 * reference generator's files (raw native-endian uint32, no header). */
int bdc_write_caches(int n_max, const char *dir);

/* ---- context: load_cache (BAA.cu:183-188) + orbit list (main_BAA.cu:46-50) ---- */
/* Creates an engine for BDC_{n,k}: 2 <= n <= 32, 1 <= k <= n, k <= 30 (see
 * bdc_supported()).  cache_dir: directory holding cache_{n'}_{k'} files to load
 * (existence, exact size and counts <= 65535 checked before any device
 * allocation: BDC_ERR_IO otherwise); NULL or "" generates the needed
 * half-length caches on the GPU.  k > 20 and k = n read no caches (the
 * deletion-chain path). */
int bdc_create(int n, int k, const char *cache_dir, bdc_ctx **out);
int bdc_destroy(bdc_ctx *ctx);
/* 1 if (n,k) is supported by this build, else 0 (message in bdc_last_error) */
int bdc_supported(int n, int k);
/* number of complement/reversal orbits = length of a dist file */
int bdc_num_orbits(bdc_ctx *ctx, uint64_t *out);
/* ascending representatives (dist-file line order) */
int bdc_get_orbits(bdc_ctx *ctx, uint32_t *out, uint64_t len);

/* ---- distribution: main_BAA.cu:53-65 (init) and :3-18 / :83-96 (file) ---- */
int bdc_set_p_uniform(bdc_ctx *ctx);                              /* 1/2^n per string (main_BAA.cu:64) */
int bdc_set_p_orbits(bdc_ctx *ctx, const double *q, uint64_t len); /* dist-file order */
int bdc_get_p_orbits(bdc_ctx *ctx, double *q, uint64_t len);
/* Dirichlet(alpha) over orbit masses with seed (no reference counterpart):
 * orbit mass m_i ~ Gamma(alpha) (Marsaglia–Tsang on a counter-based PRNG),
 * per-string value m_i / (|O_i| * sum m).  Deterministic in (n, alpha, seed). */
int bdc_set_p_dirichlet(bdc_ctx *ctx, double alpha, uint64_t seed);

/* ---- iteration: baa_iteration (BAA.cu:590-639) + compute_max_deviation (:916-946) ---- */
/* Runs `iters` BA steps at the given precision; stats (may be NULL) receives one
 * record per step.  If tol > 0 it stops early after the first step with
 * mx_bits < tol (main_BAA.cu:73) and *done receives the steps run.  With
 * tol > 0 on small problems the steps run in batches from a device snapshot
 * of the step state (up to 256 MiB, kept by the context) and the batch that
 * crosses the stopping point is rerun up to it: the result is bitwise that
 * of one step per call. */
int bdc_iterate(bdc_ctx *ctx, int precision, int iters, double tol, bdc_iter_stats *stats, int *done);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): `iters` BA steps exactly as bdc_iterate with tol = 0, but
 * the per-step statistics live in a file-scope device array of 4096 steps
 * (128 KiB per GPU, held until the process ends) instead of a buffer the
 * context grows, and no call on its path allocates (a step that would, fails
 * with BDC_ERR_STATE).  iters must be in [0, 4096].  split16 needs its operands
 * prepared by one earlier bdc_iterate or bdc_split_config on the context.
 * The array is shared by every call on one GPU: calls that overlap on one GPU
 * race.  stats (may be NULL) receives one record per step. */
int bdc_iterate_noalloc(bdc_ctx *ctx, int precision, int iters, bdc_iter_stats *stats);

/* ---- bounds: compute_all_log_D + compute_rate (BAA.cu:642-724, 836-913) ---- */
int bdc_certify(bdc_ctx *ctx, bdc_certificate *out);

/* ---- building blocks (compute_den_kernel BAA.cu:429-485, compute_KL_kernel :533-571) ---- */
/* q(y) = sum_x w(x) c(x,y) / C(n,k) for arbitrary full-space weights w (2^n) -> q (2^k) */
int bdc_output_law(bdc_ctx *ctx, int precision, const double *w, double *q);
/* D(x||q) for all x (2^n) from ln q over all y (2^k) */
int bdc_kl(bdc_ctx *ctx, int precision, const double *log_q, double *d);
/* H(x)/C(n,k) - ln C(n,k) = sum_y W ln W for all x (2^n) — the hoisted entropy term
 * (the per-pair W ln W of log_W_term, BAA.cu:401-411) */
int bdc_entropy_term(bdc_ctx *ctx, double *h);
/* Values at selected inputs x[0..len) (len <= 2^24), for checks where the
 * full 2^n arrays are too large to copy (n up to 32): what = BDC_SAMPLE_ENTROPY
 * gives H(x)/C(n,k) - ln C(n,k) (as bdc_entropy_term), BDC_SAMPLE_P gives p(x)
 * (per string), BDC_SAMPLE_KL gives D(x||q) from the last bdc_certify or
 * bdc_certify_q.  No
 * reference counterpart (the reference only ever copies whole arrays,
 * BAA.cu:642-724). */
enum { BDC_SAMPLE_ENTROPY = 0, BDC_SAMPLE_P = 1, BDC_SAMPLE_KL = 2 };
int bdc_sample_inputs(bdc_ctx *ctx, int what, const uint32_t *x, uint64_t len, double *out);

/* ---- smoothed dual bound and its adjoint (no reference counterpart) ---- */
/* U_beta(theta) = beta^-1 log sum_x exp(beta D(x||softmax theta)) and, when grad
 * is not NULL, dU/dtheta = softmax(theta) - W^T pi, pi = softmax_x(beta D).
 * theta has 2^k entries. */
int bdc_dual(bdc_ctx *ctx, int precision, const double *theta, double beta, double *u, double *grad);
/* Adam on theta = log q (device resident): `steps` steps at learning rate lr with
 * the default Adam constants (0.9, 0.999, 1e-8).  theta is read and written
 * (2^k).  u_hist (may be NULL) receives U_beta before each step.  When n is even
 * and theta is invariant under y -> ~y and y -> rev y to 1e-9 relative (ln q of a
 * BA iterate is), theta is first averaged over each such output orbit and the
 * steps run on the half input space with the BA step's mirror rule, the
 * gradient averaged over each orbit so theta stays symmetric.  U_beta exceeds
 * max_x D by at most n ln 2 / beta nats, so beta must make that small against the
 * improvement sought: from ln q of 200 BA steps at n = 28, k = 18, beta = 16384
 * and lr 1e-3 lower the certified bound faster per pass than BA; beta <= 1024
 * does not. */
int bdc_refine_dual(bdc_ctx *ctx, int precision, double *theta, double beta, int steps, double lr, double *u_hist);
/* Certified upper bound max_x D(x||softmax theta) in FP64, nats. */
int bdc_certify_q(bdc_ctx *ctx, const double *theta, double *upper);
/* theta = ln q(p) for the current p (FP64), 2^k entries */
int bdc_log_output(bdc_ctx *ctx, double *theta);

/* ---- timing / introspection ---- */
/* synchronises the context stream */
int bdc_sync(bdc_ctx *ctx);
/* average device duration (ms) of the named kernel over its last `n` launches
 * measured with HIP events on the context stream; names: "g1", "g2" */
int bdc_kernel_time(bdc_ctx *ctx, const char *name, double *avg_ms, int *count);
int bdc_reset_kernel_timers(bdc_ctx *ctx, int enable);
/* algorithmic work of one BA step at the given precision: flops of the two big
 * contractions (counted as FP64-equivalent multiply-adds x 2, whatever the limb
 * count), and the HBM bytes of one launch of the cross-term kernel (G2): p read
 * and written and the entropy term read (24 B per input of the step; in the
 * quarter mode of the split step, 16 B per input of the contracted tiles plus
 * 8 B written per input of the half space) */
int bdc_work_model(bdc_ctx *ctx, int precision, double *g1_flops, double *g2_flops, double *g2_bytes);
/* split16 configuration (fixed on first split16 use, prepared here if needed):
 * limbs and bits per limb of p in the output-law contraction (g1) and of the
 * cross-term operands (g2), and the counts fp16 cannot hold (corrected from
 * sparse residual lists).  No counterpart in the reference (it has no split path). */
int bdc_split_config(bdc_ctx *ctx, int *g1_limbs, int *g1_bits, int *g2_limbs, int *g2_bits, int *residuals);

#ifdef __cplusplus
}
#endif
#endif
