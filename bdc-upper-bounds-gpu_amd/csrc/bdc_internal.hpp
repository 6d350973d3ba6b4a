// Internal declarations of the MI355X Blahut–Arimoto engine (gfx950 only).
//
// Problem:   BDC_{n,k}, inputs x in {0,1}^n split as x = (x1, x2) with x1 the
//            high n1 = ceil(n/2) bits and x2 the low n2 = n - n1 bits
//            (backend/BAA.cu:340-349).  P is p laid out as an N1 x N2 row-major
//            FP64 matrix (N1 = 2^n1 rows x1, N2 = 2^n2 columns x2).
// Counts:    A_j = cache_{n1,j} (N1 x 2^j), B_m = cache_{n2,m} (N2 x 2^m), and
//            c(x,y) = sum_j A_j[x1, y>>(k-j)] * B_{k-j}[x2, y mod 2^(k-j)]
//            (backend/BAA.cu:355-392).
// Splits j are partitioned into JL = {j : 2j <= k} (contract x1 first) and
// JR = {j : 2j > k} (contract x2 first); stacked count operands
//   Abar = [A_j, j in JL]          (N1 x KL)
//   Bbar = [B_{k-j}, j in JR]      (N2 x KR)
// make the two large contractions of one BA step
//   G1:  Tbar = Abar^T P  (KL x N2) and Ubar = Bbar^T P^T (KR x N1)
//   G2:  S    = [Abar | Wbar] [Vbar ; Bbar^T]   (N1 x N2, K = KL + KR)
// with the small stage GEMMs q_j = T_j B_{k-j} (j in JL), q_j = A_j^T U_{k-j}
// (j in JR), V_j = L_j B_{k-j}^T (j in JL), W_j = A_j L_j (j in JR), where
// L = ln q - mu is the centred log output law.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bdc.h"

namespace bdc {

constexpr int kMaxN = 32;       // backend/config.hpp:24 (MAX_N = 33 -> n <= 32)
constexpr int kMaxHalf = 16;    // half-length caches up to n' = 16
constexpr int kPartStride = 4;  // doubles per block partial in the fused epilogue
constexpr int kDenseMaxK = 20;  // largest k of the dense split contractions (2^k outputs per split)
constexpr int kMaxGroups = 4;   // bit groups of the deletion-chain path (n <= 32: at most 3)

void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define BDC_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::bdc::fail(BDC_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,        \
                         hipGetErrorString(e_));                                        \
  } while (0)

#define BDC_TRY(expr)               \
  do {                              \
    int r_ = (expr);                \
    if (r_ != BDC_OK) return r_;    \
  } while (0)

// Device buffer owned by a context (hipMalloc / hipFree).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int alloc(size_t b);
  void release();
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// One small stage GEMM C[M x N] = sum_K A[m,k] B[k,n]; element (r,c) of an
// operand lives at ptr[r*rs + c*cs].  Type codes: 0 = f64, 1 = u16 counts.
struct StageGemm {
  const void *a;
  int64_t ars, acs;
  int atype;
  const void *b;
  int64_t brs, bcs;
  int btype;
  double *c;
  int64_t crs, ccs;
  int M, N, K;
};

struct Timer {
  const char *name;
  std::vector<hipEvent_t> start, stop;
  int used = 0;
};

struct Split {  // split-FP16 configuration, fixed on first use
  int g1_limbs = 4, g1_bits = 9, g1_chunk = 128;
  int g2_limbs = 4, g2_bits = 11;
  bool ok = false;
};

}  // namespace bdc

struct bdc_ctx {
  int n, k, n1, n2;
  int N1, N2;             // 2^n1, 2^n2
  uint64_t nx;            // 2^n
  uint64_t norb;          // number of orbits
  double nck, inv_c, log_c;
  hipStream_t stream = nullptr;
  int device = 0;

  // split sets
  std::vector<int> JL, JR;      // j values (ascending)
  std::vector<int> offL, offR;  // column offsets of A_j (j in JL) in Abar, of B_{k-j} (j in JR) in Bbar
  int KL = 0, KR = 0, KLp = 0, KRp = 0, KT = 0;

  // counts (uint16) per half length: cnt1[j] = cache_{n1,j} (j < n1), cnt2[m] = cache_{n2,m}
  bdc::DevBuf cnt1[bdc::kMaxHalf + 1], cnt2[bdc::kMaxHalf + 1];
  bdc::DevBuf abar16;   // Abar as u16, N1 x KLp (zero padded)
  bdc::DevBuf bbar16;   // Bbar as u16, N2 x KRp
  bdc::DevBuf abarh;    // Abar as fp16 in MFMA operand order (the cross term; counts rounded to fp16,
                        // residual lists below), N1 x KLp
  bdc::DevBuf abarTh;   // Abar^T as fp16 in MFMA operand order (the output law), KLp up to 256s x N1
  bdc::DevBuf bbarh;    // Bbar as fp16, N2 x KRp

  // symmetry / orbits
  bdc::DevBuf repmask;   // 2^n bits (uint64 words), bit x set iff rep(x) == x
  bdc::DevBuf wordbase;  // exclusive prefix popcount per 64-bit word (uint32 for n <= 30, uint64 beyond)
  bdc::DevBuf orbits;    // ascending representatives (uint32)

  // per-input state (full space, N1 x N2 row-major)
  bdc::DevBuf P;   // p-tilde (unnormalised), FP64; p = pscale * P
  bdc::DevBuf Hc;  // H(x)/C - ln C, FP64
  bool have_p = false;
  bool p_mirror_stale = false;
  bool texp_reset = false;  // the last split of V reset ptexp for a quarter cross term  // quarter steps left the non-canonical half-space elements stale
  double last_mx_bits = NAN;  // mx of the step that produced the current p (NaN if p was set)
  bool have_h = false;
  bool have_d = false;  // dvec holds D(x||q) of the last bdc_certify
  bool stage_factored = true;  // stage contractions as factored single-deletion steps (BDC_STAGE_FACTORED=0: count matrices)

  // contraction buffers
  bdc::DevBuf T;      // Tbar (KLp x N2) FP64
  bdc::DevBuf U;      // Ubar (KRp x N1) FP64 (general path)
  bdc::DevBuf qpart;  // per-j output-law slices (|J| x 2^k) FP64
  bdc::DevBuf qv;     // q (2^k) FP64, then L = ln q - mu
  bdc::DevBuf qsum;   // block sums of ln q (2^k / 256) from the output-law combine
  bool l_uncentred = false;  // L holds ln q; its centring is left to the stage contraction
  bdc::DevBuf g2a;    // Wbar part: N1 x KRp FP64
  bdc::DevBuf g2b;    // Vbar part: KLp x N2 FP64
  bdc::DevBuf splitk; // split-K partials for stage GEMMs
  bdc::DevBuf g64part; // split-K slabs of the aligned FP64 output law
  // Column order of the x-indexed arrays (P, Hc, D): with cr > 0 the low cr
  // bits of x (the column x2) are stored bit-reversed (n1 = n2 = cr), so the
  // reversal x -> rev x is the transpose of the N1 x N2 matrix and the
  // complement a point reflection.  The factored stage contractions read Tbar
  // and write Vbar in that order themselves (g2b_stored); the other stage forms
  // go through T2 / V2, Tbar with natural columns and Vbar with stored columns.
  int cr = 0;
  bool g2b_stored = false;
  bdc::DevBuf T2, V2;
  bdc::DevBuf qtiles[3];     // quarter-mode tile lists of the BA step: without / with residuals / both (launch order)
  int nqtiles[2] = {0, 0};
  size_t splitk_elems = 0;
  bdc::DevBuf part;   // per-block partials of the fused epilogue
  size_t part_blocks = 0;
  bdc::DevBuf scal;   // device scalars (see kScal* in kernels)
  bdc::DevBuf exptab; // 2^(j/32), j < 32, for the update's exp
  bdc::DevBuf hist;   // per-iteration stats (device)
  int hist_cap = 0;
  // snapshot of what a BA step changes (the rows of P it updates, the device
  // scalars, the P tile exponents), taken before a batch of steps that a
  // stopping rule may cut short
  bdc::DevBuf ck_p, ck_s, ck_t;

  // split-FP16 buffers
  bdc::Split split;
  bdc::DevBuf ptexp;   // per-tile exponents of P (int32), one per 128 x 64 tile
  bdc::DevBuf vlimb;   // Vbar limbs transposed, fp16 [L][N2][KLp]
  bdc::DevBuf wlimb;   // Wbar limbs fp16, L x (rows x KRp) in MFMA operand order (wl_index)
  bdc::DevBuf residA;  // (x1, kappa, r) triples of Abar counts fp16 cannot hold
  bdc::DevBuf rowptrA, entA, rowptrB, entB;  // the same residuals as CSR over x1 (Abar) and x2 (Bbar)
  bdc::DevBuf blkA, blkB;  // flags per 128 rows of Abar / 64 rows of Bbar: any residual there
  bdc::DevBuf restiles[2];  // G2 tiles (128 x 32) touched by a residual: over N1 rows, over N1/2 rows
  int nrestiles[2] = {0, 0};
  bdc::DevBuf g1part;      // split-K partial slabs of Tbar
  bool plimb_valid = false;  // ptexp holds the tile exponents of the current P (rows x1 < plimb_rows)
  int plimb_rows = 0;
  int n_resid_a = 0, n_resid_b = 0;
  double g1_chunk_sum = 0, g2_row_sum = 0;  // exactness bounds (max count mass per FP32 accumulation)

  // deletion-chain path (k > 20 or k = n): bit groups top to bottom and the
  // level buffer (slot t: strings after t deletions, 2^(n-t) doubles at lvl_off[t])
  bool chain = false;
  int ngrp = 0, grpL[bdc::kMaxGroups] = {}, grpB[bdc::kMaxGroups] = {};
  bdc::DevBuf lvl;
  std::vector<uint64_t> lvl_off;

  // dual refinement state
  bdc::DevBuf theta, adam_m, adam_v, grad, dvec;

  const uint16_t *c1(int j) const { return cnt1[j].as<uint16_t>(); }
  const uint16_t *c2(int m) const { return (n2 == n1 ? cnt1[m] : cnt2[m]).as<uint16_t>(); }

  // timers
  bool timing = false;
  bdc::Timer t_g1{"g1"}, t_g2{"g2"};
};

namespace bdc {
struct EpiArgs;
// split-FP16 path (split.hip)
int split_prepare(bdc_ctx *c);
int split_output_law_g1(bdc_ctx *c, const double *X, int rows);
// reset_texp: the quarter BA step's cross term follows (its P tile exponents are reset on the way)
int split_cross_operands(bdc_ctx *c, const double *L, bool mirror, int rows, bool reset_texp = false);
int split_cross_term(bdc_ctx *c, int epi, const EpiArgs &ea, int *nblocks, int rows);
}  // namespace bdc
