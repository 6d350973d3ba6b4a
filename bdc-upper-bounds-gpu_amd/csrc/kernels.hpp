// Host-side launch wrappers of the device kernels (one stream per call).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/bdc.h"

namespace bdc {

// ---- setup (kernels_setup.hip) ----
int launch_count_cache_u32(int n, int k, uint32_t *out, hipStream_t s);
int launch_count_cache_u16(int n, int k, uint16_t *out, hipStream_t s);
int launch_stack_cols(const uint16_t *src, int m, uint64_t R, uint16_t *dst, int ldd, int off, hipStream_t s);
int launch_counts_to_half(const uint16_t *src, uint64_t R, uint64_t C, int transpose, _Float16 *dst, hipStream_t s);
int launch_repmask(int n, uint64_t nwords, uint64_t *mask, uint32_t *pop, hipStream_t s);
int launch_exclusive_scan(const uint32_t *in, uint64_t len, uint32_t *out, uint32_t *tmp_bsum, uint32_t *total,
                          hipStream_t s);
int launch_orbit_list(uint64_t nwords, const uint64_t *mask, const uint32_t *wordbase, uint32_t *orbits, hipStream_t s);
int launch_fill(double *P, uint64_t len, double val, hipStream_t s);
// cr: bits of the column index stored reversed (0: natural order), see store_index
int launch_expand_orbits(int n, const double *Q, const uint64_t *mask, const uint32_t *wordbase, double scale,
                         double add, int cr, double *P, hipStream_t s);
int launch_gather_orbits(uint64_t norb, const uint32_t *orbits, const double *P, const double *scal, int cr, double *Q,
                         hipStream_t s);
// out[r][c] = in[r][rev(c)] for rows of 2^h doubles (h <= 14)
int launch_permute_cols(const double *in, double *out, uint64_t rows, int h, hipStream_t s);
// out[r][:] = in[rev(r)][:] for 2^h rows of width u16
int launch_permute_rows_u16(const uint16_t *in, uint16_t *out, int h, int width, hipStream_t s);
int launch_dirichlet_orbits(int n, uint64_t norb, const uint32_t *orbits, double alpha, uint64_t seed, double *Q,
                            double *mass, hipStream_t s);
int launch_sum(const double *v, uint64_t len, double *part, int np, double *out, hipStream_t s);
// out[i] = scal[slot] * src[idx[i]] (scale 1 when slot < 0), len <= 2^28
int launch_gather_at(const double *src, const uint32_t *idx, uint64_t len, const double *scal, int slot, int cr,
                     double *out, hipStream_t s);
// logtab: ln(1 + j/64) then 1/(1 + j/64), j = 0..64 (130 doubles)
int launch_entropy_reps(int n, int k, int n1, int n2, const uint64_t *mask, const uint32_t *wordbase,
                        const double *logtab, double *Hrep, hipStream_t s);

// ---- FP64 contractions (kernels_gemm64.hip) ----
// Operand element (r, c) at ptr[r*rs + c*cs]; type 0 = f64, 1 = u16 counts.
struct Opnd {
  const void *p;
  int64_t rs, cs;
  int type;
};
enum { EPI_STORE = 0, EPI_BA = 1, EPI_D = 2 };
struct EpiArgs {
  double *P;          // p-tilde, updated in place (EPI_BA)
  const double *Hc;   // H(x)/C - ln C
  double *D;          // D output (EPI_D)
  double *part;       // per-block partials
  const double *scal; // device scalars
  double inv_c;
  int64_t ldx;        // row stride of the x-indexed arrays (N2)
  const double *exptab;  // 2^(j/32), j < 32
};
// C[M x N] (=) sum_phase A_ph[M x K_ph] B_ph[K_ph x N].  Phase 2 is skipped when K2 == 0.
// splits > 1: C + z*split_stride receives the partial over the z-th K range of phase 1
// (phase 2 must be empty).
int gemm64(int M, int N, const Opnd &A1, const Opnd &B1, int K1, const Opnd &A2, const Opnd &B2, int K2, double *C,
           int64_t crs, int64_t ccs, int splits, int64_t split_stride, int epi, const EpiArgs *ea, hipStream_t s,
           int *nblocks_out);
// The two large products of a BA step on tile-aligned operands (128 x 128
// tiles, K stages of 16): the output law C[z] = A^T X over the z-th K range
// (A u16 [K][lda], X f64 [K][ldb]) and the cross term with the BA epilogue
// (A1 u16 [M][lda1] x B1 f64 [K1][ldb1] + A2 f64 [M][lda2] x B2^T, B2 u16 [N][ldb2]).
bool g64_g1_fits(int Mpad, int N, int K, int splits);
int launch_g64_g1(int Mpad, int N, int K, const uint16_t *A, int64_t lda, const double *X, int64_t ldb, double *C,
                  int64_t ldc, int splits, int64_t split_stride, hipStream_t s);
bool g64_g2_fits(int M, int N, int K1, int K2);
int launch_g64_g2(int M, int N, const uint16_t *A1, int64_t lda1, const double *B1, int64_t ldb1, int K1,
                  const double *A2, int64_t lda2, const uint16_t *B2, int64_t ldb2, int K2, int epi,
                  const EpiArgs &ea, hipStream_t s, int *nb);
// C[i] = sum_z part[z * stride + i] for i < len
int launch_reduce_splits(const double *part, int splits, int64_t stride, uint64_t len, double *C, hipStream_t s);

// ---- streaming (kernels_stream.hip) ----
// finalize the fused BA epilogue: reduce block partials (times mult), update scalars, write history
// (scratch: 256 * 4 doubles for the first reduction level)
int launch_ba_finalize(const double *part, int nblocks, double mult, double *scal, double *hist, int hist_idx,
                       double *scratch, hipStream_t s);
// P[2^n - 1 - x] = P[x] for x < 2^n / 2
int launch_reflect_half(double *P, uint64_t nx, hipStream_t s);
// q(y) = pscale * inv_c * (sum_{j in JL} qpart[jl][y] + sum_{j in JR} ...) ; L = ln q - mu
struct QCombine {
  int k;
  int nL, nR;                  // number of JL / JR slices
  const double *qL;            // nL slices of 2^k (indexed by y)
  const double *qR;            // nR slices of 2^k (general path), or null for the mirror rule
  int mirror;                  // 1: JR part = sum over qL slices selected by mirrorSel at rev(y)
  unsigned long long mirrorSel; // bit i set: slice i of qL contributes to the mirrored part
  int scale_slot;              // device scalar holding the scale of the weights (p = scale * P)
  int comp;                    // 1: slices are half-space sums R_j; q_j(y) = R_j(y) + R_j(~y)
};
// part: 2^k / 256 doubles of scratch (the block sums of ln q).  center = 0
// leaves L = ln q uncentred: launch_q_center, or a deletion chain given the
// block sums (launch_chain_del), centres it later
int launch_q_combine(const QCombine &qc, double inv_c, double *scal, double *qout, double *Lout, double *part,
                     int center, hipStream_t s);
int launch_q_center(int k, const double *part, double *scal, double *Lout, hipStream_t s);
// theta -> q = softmax(theta), L = ln q - mu (dual bound)
int launch_softmax_log(const double *theta, int k, double *q, double *L, double *scal, hipStream_t s);
// D-pass reductions for the dual: max D, then pi = softmax(beta D) and U over
// nx inputs, each standing for `mult` inputs (2: the half space of a symmetric q)
int launch_dual_weights(double *D, uint64_t nx, double beta, double mult, double *part, int nblocks, double *scal,
                        double *u_out, hipStream_t s);
// dst[y] = mean of src over the complement/reversal orbit of y (2^k entries)
int launch_orbit_mean(const double *src, int k, double *dst, hipStream_t s);
int launch_adam(double *theta, const double *grad, double *m, double *v, uint64_t len, double lr, double b1, double b2,
                double eps, int step, hipStream_t s);
int launch_copy2d(const double *src, int64_t srs, int64_t scs, double *dst, int64_t drs, int64_t dcs, int64_t R,
                  int64_t Cc, hipStream_t s);
// part needs 3*1024 doubles; out[0] = sum a*b, out[1] = max b, out[2] = sum a
int launch_dot(const double *a, const double *b, uint64_t len, double *part, double *out, hipStream_t s);
int launch_vec_sub(const double *a, const double *b, double scale_b, double *out, uint64_t len, hipStream_t s);

// ---- few-deletion stage contractions (kernels_stage.hip) ----
int launch_del_gather(const double *in, int64_t irs, int64_t ics, double *out, int64_t ors, int64_t ocs, int R, int nb,
                      int d, hipStream_t s);
int launch_ins_gather(const double *in, int64_t irs, int64_t ics, double *out, int64_t ors, int64_t ocs, int R, int nb,
                      int d, hipStream_t s);
constexpr int kMaxGatherDeletions = 2;   // deletion form: C(n,d) terms per output (d = 3 streams the counts faster)
constexpr int kMaxGatherInsertions = 2;  // insertion form: 2^d C(n,d) terms per output
// C + z * zstride (row stride ldc) = A[M x K-range z] x B over count matrices (u16):
// transb = 0: B(k, n) = B[k * ldb + n]; transb = 1: B(k, n) = B[n * ldb + k].
// Needs M <= 128, N % 64 == 0, K % 32 == 0 and (K / 32) % splits == 0; else BDC_ERR_ARG.
int launch_count_gemm(int M, int N, int K, const double *A, int64_t lda, const uint16_t *B, int64_t ldb, int transb,
                      double *C, int64_t ldc, int splits, int64_t zstride, hipStream_t s);
// ---- factored stage contractions (kernels_stage.hip) ----
// One block per row of every job, one launch per stage.  Row y of job i reads
// src + src_off[i] + y * src_ld[i] and writes dst + dst_off[i] + y * dst_ld[i].
//   chain_ins: long rows (2^L entries) to short rows (2^m, m = L - d):
//     out[z] = sum_x in[x] c(x, z) over the half-length counts c = cache_{L,m}
//   chain_del: short rows (2^m) to long rows (2^L):
//     out[x] = sum_z c(x, z) in[z]
// Needs 2^L doubles of LDS (L <= 14) and at most kMaxChainJobs jobs; else BDC_ERR_ARG.
constexpr int kMaxChainJobs = 24;
struct ChainJobs {
  int count;
  int row0[kMaxChainJobs + 1];  // first block of each job (prefix sum of the row counts)
  int d[kMaxChainJobs];         // deletions L - m
  int64_t src_off[kMaxChainJobs], src_ld[kMaxChainJobs], dst_off[kMaxChainJobs], dst_ld[kMaxChainJobs];
  // 1: the long strings (L bits) are in bit-reversed order, the short ones in
  // natural order (insertion form: source rows reversed, outputs natural;
  // deletion form: sources natural, outputs reversed)
  int rev;
};
int launch_chain_ins(const double *src, double *dst, int L, const ChainJobs &jobs, hipStream_t s);
// amax (may be null): one double per block (row), the largest |entry| written
// mu_part (may be null): src is ln q uncentred and mu_part its 2^k/256 block
// sums; the chain subtracts their mean mu on load and records it in scal
int launch_chain_del(const double *src, double *dst, int L, const ChainJobs &jobs, double *amax,
                     const double *mu_part, int mu_k, double *scal, hipStream_t s);
struct MirrorW {  // for each j in JR: column offset in Wbar, m = k - j, row offset of V_{k-j} in Vbar
  int count;
  int offR[18], m[18], offL[18];
  int stored;  // 1: the columns of Vbar are in bit-reversed order
};
int launch_w_from_v(const MirrorW &mw, const double *V, int N2, int n2, double *W, int N1, int KR, int KRp,
                    hipStream_t s);

// ---- deletion-chain path for k close to n (kernels_chain.hip) ----
// One group of bits of a level buffer viewed as [nh][2^L][2^B], processed by
// blocks of one high row and W low columns (W a power of two <= 16, W <= 2^B).
constexpr int kMaxGrpSteps = 13;  // group length <= 13 (2^13 doubles of LDS per column)
struct GrpChain {
  const double *src;                   // del: R_t
  double *dst;                         // ins: target R'_t
  double *lvl[kMaxGrpSteps + 1];       // del: R_{t+j} (targets); ins: R'_{t+j} (sources, null = zero)
  double scale[kMaxGrpSteps + 2];      // del: 1/j!; ins: 1/j
  uint32_t acc;                        // del: bit j set = add into R_{t+j}; ins: bit 0 = add into R'_t
  int L, B, m, W;                      // group bits, low bits, steps, tile width
  uint64_t nh, ntiles;                 // high rows, 2^B / W
};
// output law: (D_g^j R_t) / j! into R_{t+j}, j = 1..m
int launch_grp_del(const GrpChain &g, hipStream_t s);
// cross term: R'_t (=|+=) sum_{j=1..m} (D_g^T)^j R'_{t+j} / j!
int launch_grp_ins(const GrpChain &g, hipStream_t s);
// q = scal[slot] * inv_c * R, L = ln q (L may be null); sets the centring mu to 0
int launch_chain_q(const double *R, uint64_t ny, double *scal, int slot, double inv_c, double *q, double *L,
                   hipStream_t s);
// D = Hc - mu - S inv_c; EPI_BA: fused update of P with block partials; EPI_D: D stored
int launch_chain_epi(int epi, const double *S, double *P, const double *Hc, double *Dout, uint64_t nx,
                     const double *scal, double inv_c, const double *exptab, double *part, int nblocks,
                     hipStream_t s);
// H(x) = sum_y c ln c at every representative (ascending orbit list) by a
// walk of the distinct subsequences; n - k <= 16
int launch_entropy_dfs(int n, int k, uint64_t norb, const uint32_t *orbits, const double *logtab, double *Hrep,
                       hipStream_t s);

// ---- split-FP16 contractions (kernels_split.hip) ----
// exponents of the 128 x 32 tiles of a P-like array (rows x1 < M, N2 columns), row-major by tile
int launch_tile_exp(const double *P, int M, int N2, int bits, int32_t *texp, hipStream_t s);
// np_pre > 0: part[0 .. np_pre) already holds block maxima of |V| and |W| (W a
// permutation of entries of V), so the max pass over V and W is skipped.
// fill (may be null): fill_n int32 set to INT_MIN on the way (the P tile
// exponents before a quarter step)
int launch_split_vw(const double *V, int KL, int KLp, int N2, const double *W, int M, int KR, int KRp, int limbs,
                    int bits, _Float16 *vT, _Float16 *wl, double *part, int np_pre, double *scal, int32_t *fill,
                    int fill_n, hipStream_t s);
// launch_w_from_v (the M rows of Wbar under the mirror rule, from Vm) and
// launch_split_vw in one launch, when the tiled form of the former applies and
// the np block maxima of |Vbar| are given (else the two launches)
int launch_vw_limbs(const double *V, int KL, int KLp, int N2, const MirrorW &mw, const double *Vm, int n2, double *W,
                    int M, int KR, int KRp, int limbs, int bits, _Float16 *vT, _Float16 *wl, double *part, int np,
                    double *scal, int32_t *fill, int fill_n, hipStream_t s);
// ks partial slabs (ks * KLp * ncols FP64) of Tbar = aT X over the first Kdim rows of the FP64 array X
// (row stride ncols; split into limbs on the fly with the tile exponents texp).  aT: Abar^T in MFMA
// operand order over K = lda rows of Abar (abar_operand_order), 16-byte aligned
// tri: X = P of the quarter-layout BA step (Kdim = ncols / 2), read from the
// triangle of canonical elements (the others may be stale).  merge (limbs >= 4):
// the last two limbs share one FP32 accumulator; fields: five limbs of 7 to 9
// bits (with merge) are taken as bit fields of the rounded value; uniform_loops
// (tri): the stages wholly inside the triangle and wholly outside it run
// copies of the loop without the per-element source select
int launch_split_g1(int KLp, int Kdim, int lda, int ncols, const _Float16 *aT, const double *X, const int32_t *texp,
                    int limbs, int bits, int ks, double *Tpart, int tri, int merge, int fields, int uniform_loops,
                    hipStream_t s);
// residual corrections of G1 from the rows x1 < rows (tri: X read from the triangle, as launch_split_g1)
int launch_resid_g1(const int32_t *res, int nres, const double *X, int rows, int N2, int tri, double *T,
                    hipStream_t s);
struct SplitG2Args {
  int N1, N2, M, KL, KLp, KR, KRp;  // M: rows x1 computed (< N1 in the half-space step)
  const _Float16 *A16, *vT, *wl, *B16;  // A16 and wl in MFMA operand order (32-row x 16-K chunks)
  int limbs, bits, epi;
  const uint8_t *blkA, *blkB;
  const int32_t *rowptrA, *entA, *rowptrB, *entB;
  const double *V64, *W64;
  int32_t *texp_next;  // BA: exponents of the new P tiles
  int bits_p;
  const double *exptab;
  const int32_t *res_tiles;  // 128 x 32 tiles (ty * N2/32 + tx) holding a residual, run by a second launch
  int nres_tiles;
  // quarter mode (BA step, columns stored bit-reversed, M = N1/2 = N2/2): the
  // tiles holding representatives, without and with residuals
  int quarter;
  const int32_t *q_tiles, *qres_tiles;
  int nq_tiles, nqres_tiles;
  const int32_t *qall_tiles;  // the residual tiles, then the others
  int mirror;  // quarter mode: also write the mirror of every canonical element
  int texp_reset;  // quarter mode: texp_next already reset to INT_MIN (by launch_split_vw)
};
int launch_split_g2(const SplitG2Args &a, const EpiArgs &ea, hipStream_t s, int *nb);
// the non-canonical elements of a quarter-layout half space (N/2 x N) from their partners
int launch_fill_mirror(double *P, int N, hipStream_t s);

}  // namespace bdc
