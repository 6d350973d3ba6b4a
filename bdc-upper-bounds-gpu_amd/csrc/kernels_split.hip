// Split-FP16 contractions on CDNA4 FP16 MFMA (v_mfma_f32_32x32x16_f16) with
// exact integer arithmetic in the FP32 accumulator.
//
// Operands are split Ozaki-style into integer-valued FP16 limbs with a power
// of two scale:  v = 2^e (l_0 + l_1 2^-b + l_2 2^-2b + ...), |l_i| <= 2^(b-1),
// l_0 rounded to nearest so every later limb carries the exact remainder.
// Counts are exact FP16 integers (values FP16 cannot hold are rounded and their
// difference applied from a sparse residual list).  Each limb product
// count x limb is an integer below 2^22, and the limb width b and the K extent
// accumulated in FP32 before promotion to FP64 are chosen so that every
// partial sum stays below 2^24 in magnitude — the FP32 accumulator then adds
// integers exactly, whatever order the matrix core uses.  The error left is the
// truncation of the FP64 operand to L*b bits, and in the output law with four
// or more limbs the FP32 rounding of the accumulator the last two limbs share
// (below that truncation, see k_split_g1).
//
// P tiles are 128 (x1) x 32 (x2) with one exponent each.  The output-law
// contraction reads P limbs transposed ([l][x2][x1]) so that its K index x1 is
// contiguous.  The BA step keeps only the rows x1 < N1/2 of P (the other half
// is its complement image), so K runs over those rows.
#include <climits>
#include <type_traits>

#include "device_common.hpp"
#include "fast_exp.hpp"
#include "kernels.hpp"

namespace bdc {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef _Float16 half2_ __attribute__((ext_vector_type(2)));

// Integer-valued doubles of magnitude <= 2^11 to FP16, exactly: their
// conversion to float is exact, and v_cvt_pkrtz_f16_f32 packs two floats
// (round toward zero, exact for such integers).  A plain double -> half cast
// has no instruction and expands into ~150 integer operations.
__device__ inline half2_ int_to_half2(double a, double b) {
  return __builtin_bit_cast(half2_, __builtin_amdgcn_cvt_pkrtz((float)a, (float)b));
}
__device__ inline _Float16 int_to_half(double a) { return int_to_half2(a, 0.0)[0]; }
typedef float float16_ __attribute__((ext_vector_type(16)));

// ---- limbs by bit-field extraction (L limbs of B <= 9 bits, L B + 1 <= 51)
// Adding the constant ix_magic<L, B>() to a scaled value v (|v| <= 2^(B-1))
// rounds v to F = (L - 1) B fractional bits (to nearest, ties to even: the
// limbs of the rint chain represent the same value) and leaves J + Bias in the
// low mantissa bits, J = v 2^F as an integer and Bias = 2^(B-1) in every B-bit
// digit position.  Each biased digit is then a bit field: B bits (B + 1 for the
// leading one, whose biased value reaches 2^B), and limb l is field (L-1-l)
// minus 2^(B-1), in [-2^(B-1), 2^(B-1)].  The field becomes an FP16 by writing
// it under the exponent of 1024 (the mantissa holds 10 bits) and subtracting
// 1024 + 2^(B-1) exactly: four integer operations and one packed FP16 one per
// pair of values and limb, where the rint chain takes nine FP64 operations and
// conversions.
template <int L, int B>
__device__ constexpr double ix_magic() {
  uint64_t bias = 0;
  for (int j = 0; j < L; ++j) bias += (1ull << (B - 1)) << (j * B);
  return (double)((3ull << 51) + bias) / (double)(1ull << ((L - 1) * B));
}
__device__ inline uint32_t rotr32(uint32_t x, int s) { return s ? __builtin_amdgcn_alignbit(x, x, s) : x; }
// limb l of two values that carry ix_magic (a in the low half, b in the high
// half); with S the limb is scaled by 2^-B (exact: a multiple of 2^-B of
// magnitude at most 1/2)
template <int L, int B, bool S>
__device__ inline half2_ ix_limb(int l, double a, double b) {
  static_assert(B >= 2 && B <= 9 && L * B + 1 <= 51, "fields must fit the mantissa and an FP16 mantissa");
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  const uint32_t la = (uint32_t)ua, ha = (uint32_t)(ua >> 32), lb = (uint32_t)ub, hb = (uint32_t)(ub >> 32);
  const int o = (L - 1 - l) * B;  // bit offset of the field
  const uint32_t m = (l == 0 ? (2u << B) : (1u << B)) - 1u;
  uint32_t fa, fb;  // the field of a at bit 0, that of b at bit 16
  if (o >= 32) {
    fa = rotr32(ha, o - 32);
    fb = rotr32(hb, (o - 48) & 31);
  } else if (o >= 16) {
    fa = __builtin_amdgcn_alignbit(ha, la, o);
    fb = o == 16 ? lb : __builtin_amdgcn_alignbit(hb, lb, o - 16);
  } else {
    fa = o ? __builtin_amdgcn_alignbit(ha, la, o) : la;
    fb = rotr32(lb, (o - 16) & 31);
  }
  const uint32_t w = (fb & (m << 16)) | ((fa & m) | 0x64006400u);
  const half2_ h = __builtin_bit_cast(half2_, w);
  constexpr _Float16 off = (_Float16)(1024 + (1 << (B - 1)));
  if constexpr (S) {
    constexpr _Float16 r = (_Float16)(1.0f / (float)(1 << B)), c = (_Float16)(-(float)(1024 + (1 << (B - 1))) / (float)(1 << B));
    return __builtin_elementwise_fma(h, (half2_){r, r}, (half2_){c, c});
  }
  return h - (half2_){off, off};
}

// sources of a stage of p in the output law's triangle reads
enum : int { kGeneral = 0, kInPlace = 1, kTransposed = 2 };

// row of the 32x32 MFMA result held by `lane` in register r (verified layout)
__device__ inline int mfma32_row(int lane, int r) { return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); }

__device__ inline int tile_exponent(double mx, int bits) {
  // smallest e with mx <= 2^(e + bits - 1)
  if (!(mx > 0.0)) return 0;
  return ilogb(mx) + 1 - (bits - 1);
}

// ----------------------------------------------------------------- P tile exponents
// Exponent of every 128 x 32 tile of a P-like array (rows x1 < M, N2 columns):
// the smallest e with max |v| <= 2^(e + bits - 1).  One 128 x 64 block (two
// tiles) per workgroup; thread tid always holds column tid % 64, so lanes 0-31
// of a wave cover the first tile and 32-63 the second.
__global__ void __launch_bounds__(256) k_tile_exp(const double *__restrict__ P, int N2, int bits, int tcols,
                                                  int32_t *__restrict__ texp) {
  __shared__ double red[8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, half = lane >> 5;
  const int m0 = blockIdx.y * 128, n0 = blockIdx.x * 64;
  double mx = 0.0;
#pragma unroll 8
  for (int r = 0; r < 32; ++r) {
    const int e = tid + 256 * r, row = e >> 6, col = e & 63;
    mx = fmax(mx, fabs(P[(int64_t)(m0 + row) * N2 + n0 + col]));
  }
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));  // within each 32-lane half
  if ((lane & 31) == 0) red[wave * 2 + half] = mx;
  __syncthreads();
  mx = fmax(fmax(red[half], red[2 + half]), fmax(red[4 + half], red[6 + half]));
  if (wave == 0 && (lane & 31) == 0) texp[(m0 / 128) * tcols + n0 / 32 + half] = tile_exponent(mx, bits);
}

// ----------------------------------------------------------------- V/W limbs
// max |V| over the KL x N2 block and |W| over the M x KR block
__global__ void __launch_bounds__(256) k_absmax_vw(const double *__restrict__ V, int64_t vlen, const double *__restrict__ W,
                                                   int M, int KR, int KRp, double *__restrict__ part) {
  __shared__ double red[16];
  double m = 0.0;
  const int64_t wlen = (int64_t)M * KR;
  const int64_t tot = vlen + wlen;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
    double a;
    if (i < vlen) {
      a = V[i];
    } else {
      const int64_t j = i - vlen;
      a = W[(j / KR) * KRp + (j % KR)];
    }
    m = fmax(m, fabs(a));
  }
  m = block_max(m, red);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}
// V (KLp x N2, row kappa) -> vT[l][x2][kappa] (N2 x KLp), 64x64 tiles through
// LDS.  The limb scale is the exponent of the maximum over the np block maxima
// of |V| and |W| in part (every block reduces them alike; block 0 records it in
// the scalars for the W split and the cross term).  fill: fill_n entries of
// fill set to INT_MIN by the grid (the P tile exponents reset before a
// quarter step, which accumulates them with atomic maxima).
template <int L>
__global__ void __launch_bounds__(256) k_split_v(const double *__restrict__ V, int KL, int KLp, int N2, int bits,
                                                 const double *__restrict__ part, int np, double *__restrict__ scal,
                                                 _Float16 *__restrict__ vT, int32_t *__restrict__ fill, int fill_n) {
  __shared__ double t[64][65];
  __shared__ double red[16];
  const int k0 = blockIdx.y * 64, x0 = blockIdx.x * 64;
  {
    const int nthreads = gridDim.x * gridDim.y * blockDim.x;
    for (int i = (blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x; i < fill_n; i += nthreads)
      fill[i] = INT_MIN;
  }
  double m = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) m = fmax(m, part[i]);
  m = block_max(m, red);
  const int e = tile_exponent(m, bits);
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) scal[kScalVScale] = (double)e;
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int kk = i >> 6, xx = i & 63;
    t[kk][xx] = (k0 + kk < KL) ? ldexp(V[(int64_t)(k0 + kk) * N2 + x0 + xx], -e) : 0.0;
  }
  __syncthreads();
  const double radix = ldexp(1.0, bits);
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int xx = i >> 6, kk = i & 63;
    double v = t[kk][xx];
    for (int l = 0; l < L; ++l) {
      const double li = rint(v);
      vT[(int64_t)l * N2 * KLp + (int64_t)(x0 + xx) * KLp + k0 + kk] = int_to_half(li);
      v = (v - li) * radix;
    }
  }
}
// The W limbs as the cross term reads them, in MFMA operand order: limb l
// (len = M KRp values each), row x1, column c in 1 KiB chunks of 32 rows x 16
// columns, chunk (x1 / 32, c / 16) at ((x1 / 32) (KRp / 16) + c / 16) 512,
// lane (x1 % 32) + 32 ((c % 16) / 8), value c % 8
__device__ inline int64_t wl_index(int l, int64_t len, int x1, int c, int KRp) {
  return l * len + ((int64_t)(x1 >> 5) * (KRp >> 4) + (c >> 4)) * 512 + ((x1 & 31) + 32 * ((c & 15) >> 3)) * 8 + (c & 7);
}
// W (M x KRp) -> the limbs wl (wl_index)
template <int L>
__global__ void k_split_w(const double *__restrict__ W, int64_t len, int KR, int KRp, int bits,
                          const double *__restrict__ scal, _Float16 *__restrict__ wl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const int e = (int)scal[kScalVScale];
  const int c = (int)(i % KRp), x1 = (int)(i / KRp);
  double v = c < KR ? ldexp(W[i], -e) : 0.0;
  const double radix = ldexp(1.0, bits);
  for (int l = 0; l < L; ++l) {
    const double li = rint(v);
    wl[wl_index(l, len, x1, c, KRp)] = int_to_half(li);
    v = (v - li) * radix;
  }
}

// The V limbs and, under the mirror rule (W a permutation of the entries of
// V), W and its limbs in one launch: blocks [0, nbv) split 64 x 64 tiles of V
// as k_split_v; the others form 64 (kappa') x 32 (x1) tiles of W from Vm
// through LDS (as k_w_from_v_tiled: V rows offL + rev(kappa' - offR), x1 the
// columns c0 .. c0 + 31, or their bit reversals for Vm in natural order) and
// split them as k_split_w, writing zero limbs in the padding columns.  The same
// arithmetic on the same values: bitwise the limbs of the separate kernels.
template <int L>
__global__ void __launch_bounds__(256) k_vw_limbs(const double *__restrict__ V, int KL, int KLp, int N2, int bits,
                                                  const double *__restrict__ part, int np, double *__restrict__ scal,
                                                  _Float16 *__restrict__ vT, int32_t *__restrict__ fill, int fill_n,
                                                  MirrorW mw, const double *__restrict__ Vm, int n2, int sh,
                                                  double *__restrict__ W, _Float16 *__restrict__ wl, int M, int KR,
                                                  int KRp, int nbv) {
  __shared__ double t[64][65];
  __shared__ double red[16];
  const int tid = threadIdx.x;
  for (int i = blockIdx.x * blockDim.x + tid; i < fill_n; i += gridDim.x * blockDim.x) fill[i] = INT_MIN;
  double m = 0.0;
  for (int i = tid; i < np; i += blockDim.x) m = fmax(m, part[i]);
  m = block_max(m, red);
  const int e = tile_exponent(m, bits);
  if (blockIdx.x == 0 && tid == 0) scal[kScalVScale] = (double)e;
  const double radix = ldexp(1.0, bits);
  if ((int)blockIdx.x < nbv) {
    const int ntx = N2 / 64, x0 = ((int)blockIdx.x % ntx) * 64, k0 = ((int)blockIdx.x / ntx) * 64;
    for (int i = tid; i < 64 * 64; i += 256) {
      const int kk = i >> 6, xx = i & 63;
      t[kk][xx] = (k0 + kk < KL) ? ldexp(V[(int64_t)(k0 + kk) * N2 + x0 + xx], -e) : 0.0;
    }
    __syncthreads();
    for (int i = tid; i < 64 * 64; i += 256) {
      const int xx = i >> 6, kk = i & 63;
      double v = t[kk][xx];
      for (int l = 0; l < L; ++l) {
        const double li = rint(v);
        vT[(int64_t)l * N2 * KLp + (int64_t)(x0 + xx) * KLp + k0 + kk] = int_to_half(li);
        v = (v - li) * radix;
      }
    }
    return;
  }
  const int b = (int)blockIdx.x - nbv, nkw = (KR + 63) / 64;
  const int k0 = (b % nkw) * 64, c0 = (b / nkw) * 32, cl = tid & 31;
  for (int i = 0; i < 8; ++i) {
    const int kl = (tid >> 5) + 8 * i, kap = k0 + kl;
    if (kap < KR) {
      int s = 0;
      while (s + 1 < mw.count && kap >= mw.offR[s + 1]) ++s;
      const uint32_t rz = rev_bits((uint32_t)(kap - mw.offR[s]), mw.m[s]);
      const uint32_t col = mw.stored ? (uint32_t)(c0 + cl) : (uint32_t)(c0 + cl) << sh;
      t[kl][cl] = Vm[(int64_t)(mw.offL[s] + rz) * N2 + col];
    }
  }
  __syncthreads();
  const int kl = tid & 63, kap = k0 + kl;
  if (kap >= KRp) return;
  const int64_t len = (int64_t)M * KRp;
  for (int i = 0; i < 8; ++i) {
    const int r = (tid >> 6) + 4 * i;
    const uint32_t x1 = mw.stored ? (uint32_t)(c0 + r) : rev_bits((uint32_t)(c0 + r) << sh, n2);
    const int64_t o = (int64_t)x1 * KRp + kap;
    double v = 0.0;
    if (kap < KR) {
      W[o] = t[kl][r];
      v = ldexp(t[kl][r], -e);
    }
    for (int l = 0; l < L; ++l) {
      const double li = rint(v);
      wl[wl_index(l, len, (int)x1, kap, KRp)] = int_to_half(li);
      v = (v - li) * radix;
    }
  }
}

// ----------------------------------------------------------------- G1
// Tbar[kappa][c] = sum_k aT[kappa][k] X[k][c] over the first Kdim rows of the
// FP64 array X (row stride ncols); the counts aT come in MFMA operand order
// (split_prepare), lda = the K extent of that array (>= Kdim: the half-space
// step reads the leading rows of Abar).
// Block: 256 kappa x 32 columns (all of Tbar's rows for KLp <= 256, so X is
// read once), 8 waves of 32 x 32 with all L limbs.  A K stage covers 64 rows
// of X: each wave loads its counts straight into its operand registers, four
// contiguous 1 KiB chunks; the 64 x 32 block of X is read as 16-byte pairs,
// split (one stage ahead) into L integer limbs with the exponent of its
// 128 x 32 tile, and written to LDS transposed (column-major); rows of 144
// bytes keep the MFMA operand reads conflict-free.  The FP32 sums (exact integers) are
// promoted to FP64 every 128 K rows, the height of an exponent tile.  ks > 1
// splits K (partial slabs, reduced afterwards).
//
// TRI (the BA step in the quarter layout: X = P with its columns stored
// bit-reversed, Kdim = N/2 rows of N = ncols columns, and only the canonical
// elements current, r <= c left of the middle column and r + c <= N - 1 right
// of it): column c of P is read from the triangle.  Rows above the diagonal
// band of the block's 32 columns are read in place; rows below it from the
// partner element, P[r][c] = P[c][r] (left) or P[N-1-c][N-1-r] (right), i.e.
// transposed, 8 bytes per element; the one stage holding the band picks each
// element's source.  texp[chunk][strip] is the exponent of the 32 x 128 values
// read, which the quarter cross term accumulates from the tiles they come from
// (and which equals the tile exponent of the complete P).
//
// MG (L >= 4): the last two limbs share one FP32 accumulator, the last one
// entering as l 2^-b (an exact fp16).  That accumulator's sums are no longer
// exact integers: over a 128-row chunk its FP32 roundings add at most about
// 32 units of limb L - 2, 2^5 2^(-b (L - 2)) in the units of p's tile maximum,
// below the truncation of p to L limbs (the chunk's count mass times
// 2^(-b (L - 1)), 2^-21 against 2^-22 at (28,14)), and the promotion converts
// one accumulator fewer per element.
//
// IB > 0 (= bits, with MG): the limbs come from bit fields of the value
// rounded to its last limb (ix_limb), not from the rint chain.
// The body, with the side of the block's columns (SIDE 1: left of the middle,
// 2: right; 0: either, only without TRI) a compile-time constant.
template <int L, bool TRI, bool MG, int IB, int SIDE, int ROWS, int LD>
__device__ __forceinline__ void split_g1_body(_Float16 (&sm)[2][ROWS * LD], int KLp, int Kdim, int lda, int ncols,
                                              const half8 *__restrict__ aO, const double *__restrict__ X,
                                              const int32_t *__restrict__ texp, int bits, int ks,
                                              double *__restrict__ Tpart, int uniform_loops) {
  constexpr int BM = 256, BN = 32, BK = 64;
  static_assert(ROWS == L * BN && LD == 72, "stage buffer shape");
  static_assert(SIDE != 0 || !TRI, "the triangle reads depend on the side");
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int kz = blockIdx.x % ks;
  const int tile = blockIdx.x / ks;
  const int ntn = ncols / BN;
  const int n0 = (tile % ntn) * BN, m0 = (tile / ntn) * BM;
  const int tcols = ncols / 32;
  const int krows = Kdim / ks, kbase = kz * krows;
  const int lr = lane & 31, lk = 8 * (lane >> 5);

  // counts: aO holds Abar^T in MFMA operand order, 1 KiB per 32 kappa x 16 K
  // chunk (lane l: row l % 32, K columns 8 (l / 32) .. + 7), chunks of a kappa
  // tile consecutive along K (lda / 16 of them), so a wave's operand of a
  // stage is four contiguous 1 KiB reads.  X: thread takes rows
  // 2 rp, 2 rp + 1 and columns 2 cp, 2 cp + 1 of the stage; a wave covers 8
  // values of each (rp = 8 (wave / 2) + lane % 8, cp = 8 (wave % 2) + lane / 8),
  // so the transposed 4-byte limb stores (LDS word 36 (2 cp + q) + rp) fall in
  // 8 (lane / 8) + lane % 8 + const: a different bank for every lane, whether
  // the LDS has 32 or 64 banks.  Wave-uniform bases plus 32-bit per-lane byte
  // offsets that carry the stage offset (a loop-invariant sum would be hoisted
  // as VGPR pointers).
  const int rp = 8 * (wave >> 1) + (lane & 7), cp = 8 * (wave & 1) + (lane >> 3);
  const char *cA = reinterpret_cast<const char *>(aO + ((int64_t)(m0 / 32 + wave) * (lda / 16) + kbase / 16) * 64);
  const uint32_t offA = (uint32_t)lane * 16u;
  const char *cX = reinterpret_cast<const char *>(X + (int64_t)kbase * ncols + n0);
  const uint32_t offX = (uint32_t)(2 * rp * ncols + 2 * cp) * 8u;
  // triangle reads.  Byte offsets from X (< 2^30: quarter layouts have n <= 28): element (r, c) in
  // place at ((r0 + i) ncols + n0 + q) 8 + offX, transposed at ut(r0, i, q) +
  // offT; it is canonical iff dq[q] <= -(r0 + i).  Right of the middle offT
  // can be negative (wrapped): add it to ut in 32 bits, never to a pointer.
  constexpr bool left = SIDE == 1;
  const uint32_t offT = left ? (uint32_t)(2 * cp * ncols + 2 * rp) * 8u
                             : (uint32_t)((30 - 2 * cp) * ncols + 30 - 2 * rp) * 8u;
  auto ut = [&](int r0, int i, int q) -> uint32_t {
    return (uint32_t)(left ? ((int64_t)(n0 + q) * ncols + r0 + i) * 8
                           : ((int64_t)(ncols - 31 - n0 - q) * ncols + ncols - 31 - r0 - i) * 8);
  };
  int dq[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) dq[q] = 2 * rp - (left ? n0 + 2 * cp + q : ncols - 1 - n0 - 2 * cp - q);
  // Prefetch: the counts (L2 hits: every block reads the same rows of aO) are
  // loaded straight into the MFMA operand registers, each 16-K chunk of stage
  // s + 1 as soon as the multiplications of stage s have read that chunk (they
  // never pass through LDS); p (an HBM stream, too few
  // bytes in flight per CU to cover its latency with one stage) comes through
  // a ring of three slots loaded three stages ahead, and is split into limbs
  // one stage before it is stored, while the previous stage is multiplied (the
  // FP64 split then overlaps the matrix cores instead of sitting between the
  // two barriers of a stage).  Stage s's p is in slot s % 3; the loop is
  // unrolled so slots are compile-time, and counts and p are issued in stage
  // order, so waiting for the counts of stage s leaves the later p loads
  // outstanding.
  constexpr int D = 3;
  half8 ra[4];
  double2 rx[D][2];
  int et[D];
  // (aO has a zero row for every kappa up to the next multiple of 256)
  auto aload = [&](int s, int kk) {
    ra[kk] = *reinterpret_cast<const half8 *>(cA + (offA + (uint32_t)(s * (BK / 16) + kk) * 1024u));
  };
  // where: kGeneral picks each element's source; kInPlace (every element of
  // the stage canonical) reads as without TRI; kTransposed (none canonical)
  // reads from the partners as 16-byte pairs
  auto gload_x = [&](int s, auto slot, auto where) {
    constexpr int u = decltype(slot)::value;
    constexpr int w = decltype(where)::value;
    const char *xs = cX + (int64_t)s * BK * ncols * 8;  // stage base (64-bit, uniform)
    const int r0 = kbase + s * BK;
    if constexpr (!TRI || w == kInPlace) {
#pragma unroll
      for (int i = 0; i < 2; ++i) rx[u][i] = *reinterpret_cast<const double2 *>(xs + (int64_t)i * ncols * 8 + offX);
    } else if constexpr (w == kTransposed) {
      // rows r0 + 2 rp + i of column c are adjacent in the partner row (left:
      // row c from column r0 + 2 rp; right: row N - 1 - c from column
      // N - 2 - r0 - 2 rp, descending): one 16-byte load per column, kept in
      // the order loaded (a reordering move here would wait for the load)
      const char *Xc = reinterpret_cast<const char *>(X);
#pragma unroll
      for (int q = 0; q < 2; ++q) rx[u][q] = *reinterpret_cast<const double2 *>(Xc + (ut(r0, left ? 0 : 1, q) + offT));
    } else {  // each element in place or transposed (a branch on the stage's
              // position here makes the unrolled loop spill)
      const char *Xc = reinterpret_cast<const char *>(X);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        double d[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const uint32_t on = (uint32_t)(((int64_t)(r0 + i) * ncols + n0 + q) * 8) + offX;
          const uint32_t ot = ut(r0, i, q) + offT;
          d[q] = *reinterpret_cast<const double *>(Xc + (dq[q] <= -(r0 + i) ? on : ot));
        }
        rx[u][i].x = d[0];
        rx[u][i].y = d[1];
      }
    }
    // (a chunk of zeros may have no exponent, INT_MIN: any value below the
    // FP64 range keeps its limbs and scales zero)
    et[u] = max(texp[(r0 / 128) * tcols + n0 / 32], -1100);
  };
  const double radix = ldexp(1.0, bits), rradix = ldexp(1.0, -bits);
  // limbs of rows (2 rp, 2 rp + 1) of column 2 cp + q, one packed half2 per
  // limb: split_begin scales the p values of a slot by its tile exponent,
  // split_limb(l) takes limb l off them (called for l = 0, 1, ... in order)
  half2_ hl[L][2];
  // (tp: the slot holds each column's two rows as loaded in the partner row,
  // rx[u][q] = lower address first: rows (0, 1) left of the middle, (1, 0)
  // right of it; else rx[u][i] = the row's two columns)
  auto split_begin = [&](auto slot, double(&v)[2][2], auto tp) {
    constexpr int u = decltype(slot)::value;
    const int e = et[u];
    if constexpr (decltype(tp)::value) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        v[q][0] = ldexp(left ? rx[u][q].x : rx[u][q].y, -e);
        v[q][1] = ldexp(left ? rx[u][q].y : rx[u][q].x, -e);
      }
    } else {
      v[0][0] = ldexp(rx[u][0].x, -e);
      v[0][1] = ldexp(rx[u][1].x, -e);
      v[1][0] = ldexp(rx[u][0].y, -e);
      v[1][1] = ldexp(rx[u][1].y, -e);
    }
    if constexpr (IB > 0) {
      constexpr double mg = ix_magic<L, IB ? IB : 2>();
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i) v[q][i] += mg;
    }
  };
  auto split_limb = [&](int l, double(&v)[2][2]) {
    if constexpr (IB > 0) {
      static_assert(MG, "the bit-field limbs come with the merged accumulator");
#pragma unroll
      for (int q = 0; q < 2; ++q)
        hl[l][q] = l == L - 1 ? ix_limb<L, IB ? IB : 2, true>(l, v[q][0], v[q][1])
                              : ix_limb<L, IB ? IB : 2, false>(l, v[q][0], v[q][1]);
      return;
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const double l0 = rint(v[q][0]), l1 = rint(v[q][1]);
      if (MG && l == L - 1)
        hl[l][q] = int_to_half2(l0 * rradix, l1 * rradix);
      else
        hl[l][q] = int_to_half2(l0, l1);
      v[q][0] = (v[q][0] - l0) * radix;
      v[q][1] = (v[q][1] - l1) * radix;
    }
  };
  auto sstore = [&](auto buf) {
    _Float16 *b = sm[decltype(buf)::value];
#pragma unroll
    for (int l = 0; l < L; ++l)
#pragma unroll
      for (int q = 0; q < 2; ++q) *reinterpret_cast<half2_ *>(&b[(l * BN + 2 * cp + q) * LD + 2 * rp]) = hl[l][q];
  };
  auto opnd = [&](auto buf, int row, int kk) -> half8 {
    return *reinterpret_cast<const half8 *>(&sm[decltype(buf)::value][row * LD + kk * 16 + lk]);
  };

  constexpr int NA = MG ? L - 1 : L;  // accumulators
  float16_ acc[NA];
  double acc64[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc64[r] = 0.0;
#pragma unroll
  for (int l = 0; l < NA; ++l) acc[l] = (float16_){0};
  const int nst = krows / BK;
  // Stage s is multiplied from LDS buffer s % 2 while the limbs of stage s + 1
  // (already in registers) are stored into the other buffer, so one barrier
  // per stage suffices: the barrier at the end of stage s - 1 both frees
  // buffer (s + 1) % 2 and publishes stage s.  Stage s: store the limbs of
  // stage s + 1; load p of stage s + 4 into slot `ld`; multiply stage s,
  // loading each count chunk of stage s + 1 once its multiplications are
  // issued and splitting p of stage s + 2 (slot `sp`) between the groups of
  // multiplications; promote every 128 rows (odd s) with the
  // exponent recorded when stage s was split (eh, by stage parity).  Past the
  // end the last stage is loaded again, so the body has no branch (`odd` is a
  // compile-time parity: a stage is one basic block, and the split can be
  // scheduled beside the multiplications).
  int eh[2];
  auto step = [&](int s, auto ld, auto sp, auto odd, auto where) {
    constexpr int b = decltype(odd)::value ? 1 : 0;
    using Cur = std::integral_constant<int, b>;
    using Nxt = std::integral_constant<int, 1 - b>;
    const int e = eh[b];
    sstore(Nxt{});
    gload_x(min(s + 4, nst - 1), ld, where);
    double v[2][2];
    split_begin(sp, v, std::integral_constant<bool, decltype(where)::value == kTransposed>{});
    eh[b] = et[decltype(sp)::value];
    // the limb operands of each group are read from LDS one group ahead
    half8 bop[2][L];
#pragma unroll
    for (int l = 0; l < L; ++l) bop[0][l] = opnd(Cur{}, l * BN + lr, 0);
#pragma unroll
    for (int kk = 0; kk < BK / 16; ++kk) {
      const half8 a = ra[kk];
      if (kk + 1 < BK / 16) {
#pragma unroll
        for (int l = 0; l < L; ++l) bop[(kk + 1) & 1][l] = opnd(Cur{}, l * BN + lr, kk + 1);
      }
#pragma unroll
      for (int l = 0; l < L; ++l) {
        const int la = l < NA ? l : NA - 1;
        acc[la] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, bop[kk & 1][l], acc[la], 0, 0, 0);
      }
      aload(min(s + 1, nst - 1), kk);
      // a share of the next stage's limbs per group of multiplications; the
      // scheduling barrier keeps it there (left alone, the scheduler sinks the
      // split past the barrier, to the LDS stores that consume it)
#pragma unroll
      for (int l = (kk * L) / (BK / 16); l < ((kk + 1) * L) / (BK / 16); ++l) split_limb(l, v);
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (decltype(odd)::value) {  // 128 K rows done: promote the exact integers to FP64
      // (Horner in units of the last accumulator: exact while the sums span at
      // most 53 bits, then one rounding into acc64)
      const double sc = ldexp(1.0, e - bits * (NA - 1));
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        double t = (double)acc[0][r];
#pragma unroll
        for (int l = 1; l < NA; ++l) t = fma(t, radix, (double)acc[l][r]);
        acc64[r] = fma(t, sc, acc64[r]);
      }
#pragma unroll
      for (int l = 0; l < NA; ++l) acc[l] = (float16_){0};
    }
    __syncthreads();
  };
  using U0 = std::integral_constant<int, 0>;
  using U1 = std::integral_constant<int, 1>;
  using U2 = std::integral_constant<int, 2>;
  static_assert(D == 3, "the unrolled ring below has three slots");
  using E0 = std::false_type;
  using E1 = std::true_type;
  // prologue: the limbs of stage 0 into buffer 0; the counts of stage 0 and
  // the limbs of stage 1 in registers; p of stages 2 and 3 in slots 2 and 0
  using GN0 = std::integral_constant<int, kGeneral>;
  gload_x(0, U0{}, GN0{});
  gload_x(min(1, nst - 1), U1{}, GN0{});
  gload_x(min(2, nst - 1), U2{}, GN0{});
  {
    double v[2][2];
    split_begin(U0{}, v, std::false_type{});
    eh[0] = et[0];
#pragma unroll
    for (int l = 0; l < L; ++l) split_limb(l, v);
  }
  sstore(U0{});
#pragma unroll
  for (int kk = 0; kk < BK / 16; ++kk) aload(0, kk);
  {
    double v[2][2];
    split_begin(U1{}, v, std::false_type{});
    eh[1] = et[1];
#pragma unroll
    for (int l = 0; l < L; ++l) split_limb(l, v);
  }
  gload_x(min(3, nst - 1), U0{}, GN0{});
  __syncthreads();
  // six stages per iteration (three ring slots, promotion every second stage):
  // stage s0 + i loads into slot (i + 1) % 3, splits slot (i + 2) % 3 and has
  // parity i % 2, since s0 is a multiple of 6
  int s0 = 0;
  auto six = [&](int s, auto where) {
    step(s, U1{}, U2{}, E0{}, where);
    step(s + 1, U2{}, U0{}, E1{}, where);
    step(s + 2, U0{}, U1{}, E0{}, where);
    step(s + 3, U1{}, U2{}, E1{}, where);
    step(s + 4, U2{}, U0{}, E0{}, where);
    step(s + 5, U0{}, U1{}, E1{}, where);
  };
  if constexpr (TRI) {
    // A strip's rows above the diagonal band of its columns lie in the
    // triangle, those below it are all read from their partners (rows
    // r <= cmin of every column are canonical, rows r > cmax of none): while
    // the stages loaded (four ahead) are all of one kind, a copy of the loop
    // reads them without the per-element source select.  The two slots loaded
    // and not yet split when the transposed copy starts and ends (ring slots 2
    // and 0) are reordered between the two register orders there.
    const int cmin = left ? n0 : ncols - 32 - n0, cmax = left ? n0 + 31 : ncols - 1 - n0;
    const int sin = uniform_loops && cmin - (BK - 1) - kbase >= 0 ? (cmin - (BK - 1) - kbase) / BK + 1 : 0;
    const int str = !uniform_loops ? nst : cmax - kbase >= 0 ? (cmax - kbase) / BK + 1 : 0;
#pragma unroll 1
    for (; s0 + 6 <= nst && s0 + 10 <= sin; s0 += 6) six(s0, std::integral_constant<int, kInPlace>{});
#pragma unroll 1
    for (; s0 + 6 <= nst && s0 + 4 < str; s0 += 6) six(s0, std::integral_constant<int, kGeneral>{});
    if (s0 + 6 <= nst) {
      auto to_tp = [&](auto slot) {  // rx[u][i] (columns) -> rx[u][q] (rows, as loaded)
        constexpr int u = decltype(slot)::value;
        const double2 a = rx[u][0], b = rx[u][1];
        rx[u][0] = left ? make_double2(a.x, b.x) : make_double2(b.x, a.x);
        rx[u][1] = left ? make_double2(a.y, b.y) : make_double2(b.y, a.y);
      };
      auto from_tp = [&](auto slot) {
        constexpr int u = decltype(slot)::value;
        const double2 a = rx[u][0], b = rx[u][1];
        rx[u][0] = left ? make_double2(a.x, b.x) : make_double2(a.y, b.y);
        rx[u][1] = left ? make_double2(a.y, b.y) : make_double2(a.x, b.x);
      };
      to_tp(U2{});
      to_tp(U0{});
#pragma unroll 1
      for (; s0 + 6 <= nst; s0 += 6) six(s0, std::integral_constant<int, kTransposed>{});
      from_tp(U2{});
      from_tp(U0{});
    }
  } else {
#pragma unroll 1
    for (; s0 + 6 <= nst; s0 += 6) six(s0, std::integral_constant<int, kInPlace>{});
  }
  // the last nst mod 6 stages (nst is even: K rows come in multiples of 128)
  using GN = std::integral_constant<int, kGeneral>;
  if (s0 < nst) {
    step(s0, U1{}, U2{}, E0{}, GN{});
    step(s0 + 1, U2{}, U0{}, E1{}, GN{});
  }
  if (s0 + 2 < nst) {
    step(s0 + 2, U0{}, U1{}, E0{}, GN{});
    step(s0 + 3, U1{}, U2{}, E1{}, GN{});
  }
  double *T = Tpart + (int64_t)kz * KLp * ncols;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + wave * 32 + mfma32_row(lane, r);
    if (row < KLp) T[(int64_t)row * ncols + n0 + lr] = acc64[r];
  }
}

template <int L, bool TRI, bool MG, int IB>
__global__ void __launch_bounds__(512) k_split_g1(int KLp, int Kdim, int lda, int ncols, const half8 *__restrict__ aT,
                                                  const double *__restrict__ X, const int32_t *__restrict__ texp,
                                                  int bits, int ks, double *__restrict__ Tpart, int uniform_loops) {
  constexpr int LD = 72, ROWS = L * 32;  // L limb slabs of 32 rows (x2)
  __shared__ __attribute__((aligned(16))) _Float16 sm[2][ROWS * LD];  // two stage buffers
  if constexpr (TRI) {
    const int n0 = ((int)(blockIdx.x / ks) % (ncols / 32)) * 32;  // (as in the body)
    if (n0 < ncols / 2)
      split_g1_body<L, TRI, MG, IB, 1, ROWS, LD>(sm, KLp, Kdim, lda, ncols, aT, X, texp, bits, ks, Tpart, uniform_loops);
    else
      split_g1_body<L, TRI, MG, IB, 2, ROWS, LD>(sm, KLp, Kdim, lda, ncols, aT, X, texp, bits, ks, Tpart, uniform_loops);
  } else {
    split_g1_body<L, TRI, MG, IB, 0, ROWS, LD>(sm, KLp, Kdim, lda, ncols, aT, X, texp, bits, ks, Tpart, uniform_loops);
  }
}

// T[kappa][c] += r * X(x1, c) for the residuals of the rounded counts in the
// rows x1 < rows.  One thread owns a column and walks every residual, so
// updates never race.  tri: X is a quarter-layout P whose non-canonical
// elements may be stale; X(x1, c) is read from the canonical partner.
__global__ void k_resid_g1(const int32_t *__restrict__ res, int nres, const double *__restrict__ X, int rows, int N2,
                           int tri, double *__restrict__ T) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N2) return;
  const bool left = c < N2 / 2;
  for (int e = 0; e < nres; ++e) {
    const int x1 = res[3 * e], kap = res[3 * e + 1], r = res[3 * e + 2];
    if (x1 >= rows) continue;
    int64_t at = (int64_t)x1 * N2 + c;
    if (tri && !(left ? x1 <= c : x1 + c <= N2 - 1))
      at = left ? (int64_t)c * N2 + x1 : (int64_t)(N2 - 1 - c) * N2 + (N2 - 1 - x1);
    T[(int64_t)kap * N2 + c] += (double)r * X[at];
  }
}

// ----------------------------------------------------------------- G2
struct G2Resid {
  const uint8_t *blkA;     // N1/128: block of x1 rows has a residual
  const uint8_t *blkB;     // N2/64: block of x2 rows (of Bbar) has a residual
  const int32_t *rowptrA;  // N1 + 1
  const int32_t *entA;     // (kappa, r) pairs
  const int32_t *rowptrB;  // N2 + 1
  const int32_t *entB;     // (kappa', r) pairs
  const double *V64;       // KLp x N2
  const double *W64;       // M x KRp
  int KRp;
};
struct G2Out {
  int32_t *texp;  // exponents of the new P tiles (BA epilogue)
  int bits_p;
  int tcols;      // exponent tiles per row block (N2 / 32)
  int patches;  // 1: XCD patch mapping of tiles
  int M;        // rows x1 computed (N1/2 in the half-space BA step, else N1)
  const double *exptab;
  int part_off;  // quarter mode: partials at part[part_off + blockIdx.x]
  int mirror;    // quarter mode: also write each canonical element's mirror
  int nres_first;  // RES == 2: the first nres_first blocks' tiles have residuals
};

// S[x1][x2] = sum_kappa A16[x1][kappa] V[kappa][x2] + sum_kappa' W[x1][kappa'] B16[x2][kappa']
// with V, W in L limbs of one global scale.  Block 128 x 32 (one P exponent
// tile), 4 waves of 32 x 32: 16 L accumulators per lane leave room for two
// blocks per CU, so one block's epilogue (FP64 VALU and HBM traffic) overlaps
// the other's contraction.
// The operands each wave owns (its 32 rows of A16 and of the W limbs) are
// kept in MFMA operand order (1 KiB per 32 rows x 16 K chunk) and loaded
// straight into the operand registers, each chunk of stage s + 1 once the
// multiplications of stage s have read it.  The operands all four waves share
// (the V limbs, B16) are K-contiguous: a K stage of 64 is one 128-byte line per
// row, fetched by 8 lanes with 16-byte loads and staged through registers into
// one LDS buffer whose rows are padded to 144 bytes (conflict-free for the
// 16-byte writes and for the MFMA operand reads); the next stage's loads are in
// flight while the current one is multiplied.
//   phase 1 (K = kappa): A16 x the V limbs (L x 32 LDS rows)
//   phase 2 (K = kappa'): the W limbs x B16 (32 LDS rows)
// The H and P loads of the epilogue are issued before the last stage's
// multiplications, which hide part of their HBM latency.  Tiles are dealt XCD
// by XCD in 8 x 8 patches so the operand rows of a patch stay in that XCD's L2.
// Epilogue: D, and for the BA step the update p~ exp(D - dref), the block
// partials and the exponent of the new P tile.
template <int L>
struct G2Rows {
  static constexpr int P1 = 32 * L;  // phase-1 LDS rows: the V limbs
  static constexpr int P2 = 32;      // phase-2 LDS rows: Bbar
  static constexpr int LD = 72;      // halves per LDS row (64 + 8 pad)
  // (the quarter epilogue stages 128 rows of 33 doubles in the same array)
  static constexpr int HALVES = P1 * LD > 128 * 33 * 4 ? P1 * LD : 128 * 33 * 4;
};

// Quarter mode (QTR, the BA step with the columns of P stored bit-reversed, so
// that P[r][c] = P[c][r] and P[r][c] = P[~r][~c]): only the tiles listed hold
// representatives.  An element is canonical when r <= c (left half of the
// columns) or r + c <= N - 1 (right half); a canonical element is written to
// its place and to its mirror ((c, r) left, (~c, ~r) right, staged through LDS
// so the mirror rows are written contiguously) and counts twice in the sums,
// once when it is its own mirror; other elements of a listed tile are written
// by their canonical partner.  Tile exponents of the new P are combined with
// atomic maxima (the array is reset to INT_MIN before the launch).  The
// quarter variant runs at two waves per SIMD: its epilogue needs ~180 VGPRs,
// and at three waves the four-limb kernel spills in the main loop.
// RES: 0 no residual tiles, 1 every tile has residuals, 2 the first
// go.nres_first blocks' tiles do.
template <int L, int EPI, int RES, bool QTR>
__global__ void __launch_bounds__(256, QTR ? 2 : 3) k_split_g2(int N2, int KLp, int KRp, const _Float16 *__restrict__ A16,
                                                    const _Float16 *__restrict__ vT, const _Float16 *__restrict__ wl,
                                                    const _Float16 *__restrict__ B16, int bits, EpiArgs ea,
                                                    G2Resid rs, G2Out go, const int32_t *__restrict__ tiles) {
  using G = G2Rows<L>;
  constexpr int BM = 128, BN = 32, BK = 64, LD = G::LD;
  __shared__ __attribute__((aligned(16))) _Float16 sm[G::HALVES];
  __shared__ double red[32];
  // the exp table lives in LDS: a global table could alias the P stores, which
  // would order every lookup behind the previous element's store
  __shared__ double etab[32];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int TX = N2 / BN;
  int ty, tx;
  if constexpr (RES != 0 || QTR) {  // tiles from a list (residual tiles, or the quarter's tiles)
    ty = tiles[blockIdx.x] / TX;
    tx = tiles[blockIdx.x] % TX;
  } else if (go.patches) {
    const int b = blockIdx.x, xcd = b & 7, sq = b >> 3;
    const int patch = (sq >> 6) * 8 + xcd, w = sq & 63;
    ty = (patch / (TX >> 3)) * 8 + (w >> 3);
    tx = (patch % (TX >> 3)) * 8 + (w & 7);
  } else {
    ty = blockIdx.x / TX;
    tx = blockIdx.x % TX;
  }
  const int m0 = ty * BM, n0 = tx * BN;
  // tiles touched by a residual of the rounded counts run in the second launch
  if (!RES && !QTR && (rs.blkA[m0 / 128] || rs.blkB[n0 / 64])) return;
  if (EPI == EPI_BA && tid < 32) etab[tid] = go.exptab[tid];
  const int nst1 = KLp / BK, nst2 = KRp / BK, S = nst1 + nst2;  // K stages: A16 V, then W B16^T
  const int q8 = (tid & 7) * 8, rbase = tid >> 3;  // 16-byte piece of a row, first row of this thread
  const int lr = lane & 31, lk = 8 * (lane >> 5);

  // LDS side: thread rows rbase + 32 i (phase 1: V limb i; phase 2: B16, i = 0),
  // a wave-uniform base per operand plus one 32-bit per-lane byte offset (the
  // stage offset goes into it: a loop-invariant base + lane offset would be
  // hoisted as one 64-bit VGPR pointer per row block)
  half8 rg[L];
  const uint32_t off1 = (uint32_t)(rbase * KLp + q8) * 2u, off2 = (uint32_t)(rbase * KRp + q8) * 2u;
  const char *cV = reinterpret_cast<const char *>(vT + (int64_t)n0 * KLp);
  const char *cB = reinterpret_cast<const char *>(B16 + (int64_t)n0 * KRp);
  auto gload1 = [&](int s) {
    const uint32_t v = off1 + (uint32_t)(s * BK * 2);
#pragma unroll
    for (int i = 0; i < L; ++i) rg[i] = *reinterpret_cast<const half8 *>(cV + (int64_t)i * N2 * KLp * 2 + v);
  };
  auto gload2 = [&](int kst) { rg[0] = *reinterpret_cast<const half8 *>(cB + (off2 + (uint32_t)(kst * BK * 2))); };
  auto sstore = [&](int s) {
    const int nr = s < nst1 ? L : 1;
#pragma unroll
    for (int i = 0; i < L; ++i)
      if (i < nr) *reinterpret_cast<half8 *>(&sm[(rbase + 32 * i) * LD + q8]) = rg[i];
  };
  auto opnd = [&](int row, int ks) -> half8 { return *reinterpret_cast<const half8 *>(&sm[row * LD + ks * 16 + lk]); };
  // register side: this wave's 32 rows (phase 1: ra[0][ks] = A16 chunk ks of
  // the stage; phase 2: ra[l][ks] = W limb l chunk ks)
  half8 ra[L][4];
  const char *cAo = reinterpret_cast<const char *>(A16 + (int64_t)(m0 / 32 + wave) * (KLp / 16) * 512);
  const char *cWo = reinterpret_cast<const char *>(wl + (int64_t)(m0 / 32 + wave) * (KRp / 16) * 512);
  const int64_t wlimb = (int64_t)go.M * KRp * 2;  // bytes per W limb
  const uint32_t offo = (uint32_t)lane * 16u;
  auto aload1 = [&](int s, int ks) {
    ra[0][ks] = *reinterpret_cast<const half8 *>(cAo + (offo + (uint32_t)(s * (BK / 16) + ks) * 1024u));
  };
  auto aload2 = [&](int kst, int ks) {
#pragma unroll
    for (int l = 0; l < L; ++l)
      ra[l][ks] = *reinterpret_cast<const half8 *>(cWo + l * wlimb + (offo + (uint32_t)(kst * (BK / 16) + ks) * 1024u));
  };

  float16_ acc[L];
#pragma unroll
  for (int l = 0; l < L; ++l) acc[l] = (float16_){0};

  // lane element r of the tile: row m0 + 32 wave + mfma32_row(lane, r), column n0 + lane % 32,
  // i.e. a wave-uniform base plus the lane offset lo plus a uniform offset per element
  const int64_t ldx = ea.ldx;
  const int64_t ubase = (int64_t)(m0 + wave * 32) * ldx + n0;
  const int lo = (lane >> 5) * 4 * (int)ldx + lr;
  const double *Hu = ea.Hc + ubase;
  double *Pu = EPI == EPI_BA ? ea.P + ubase : ea.D + ubase;
  double hv[16], pv[16];

  // nxt: which stage's register operands to load behind the multiplications
  // (0: none, 1: phase-1 stage s + 1, 2: phase-2 stage 0 / stage kst + 1)
  auto mult1 = [&](int s, auto nxt) {
#pragma unroll
    for (int ks = 0; ks < BK / 16; ++ks) {
      const half8 a = ra[0][ks];
#pragma unroll
      for (int l = 0; l < L; ++l)
        acc[l] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, opnd(l * 32 + lr, ks), acc[l], 0, 0, 0);
      if constexpr (decltype(nxt)::value == 1) aload1(s + 1, ks);
      if constexpr (decltype(nxt)::value == 2) aload2(0, ks);
    }
  };
  auto mult2 = [&](int kst, auto nxt) {
#pragma unroll
    for (int ks = 0; ks < BK / 16; ++ks) {
      const half8 bb = opnd(lr, ks);
#pragma unroll
      for (int l = 0; l < L; ++l) acc[l] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ra[l][ks], bb, acc[l], 0, 0, 0);
      if constexpr (decltype(nxt)::value == 2) aload2(kst + 1, ks);
    }
  };
  // the H and P loads of the epilogue overlap the last stage's multiplications
  auto epi_loads = [&]() {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int64_t off = (int64_t)((r >> 2) * 8 + (r & 3)) * ldx;
      hv[r] = __builtin_nontemporal_load(Hu + off + lo);
    }
  };
  // P is loaded once the accumulators are being consumed (holding both H and P
  // through the last stage would raise the register count)
  auto p_loads = [&]() {
    if constexpr (EPI == EPI_BA) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t off = (int64_t)((r >> 2) * 8 + (r & 3)) * ldx;
        pv[r] = __builtin_nontemporal_load(Pu + off + lo);
      }
    }
  };
  auto stage_in = [&](int s) {
    if (s > 0) __syncthreads();  // the previous stage has been read
    sstore(s);
    __syncthreads();
  };
  // one loop per phase (a phase branch inside one loop makes the compiler copy
  // every accumulator where the two paths merge); the last stage is peeled so
  // the epilogue loads can join it
  using N0 = std::integral_constant<int, 0>;
  using N1 = std::integral_constant<int, 1>;
  using N2c = std::integral_constant<int, 2>;
  gload1(0);
#pragma unroll
  for (int ks = 0; ks < BK / 16; ++ks) aload1(0, ks);
  if (nst2 > 0) {
#pragma unroll 1
    for (int s = 0; s + 1 < nst1; ++s) {
      stage_in(s);
      gload1(s + 1);
      mult1(s, N1{});
    }
    stage_in(nst1 - 1);
    gload2(0);
    mult1(nst1 - 1, N2c{});
#pragma unroll 1
    for (int kst = 0; kst + 1 < nst2; ++kst) {
      stage_in(nst1 + kst);
      gload2(kst + 1);
      mult2(kst, N2c{});
    }
    stage_in(S - 1);
    epi_loads();
    mult2(nst2 - 1, N0{});
  } else {
#pragma unroll 1
    for (int s = 0; s + 1 < S; ++s) {
      stage_in(s);
      gload1(s + 1);
      mult1(s, N1{});
    }
    stage_in(S - 1);
    epi_loads();
    mult1(S - 1, N0{});
  }
  __syncthreads();  // etab visible

  // ---- epilogue, one element at a time: S from the limb sums (plus the
  // residuals of rounded counts), D, and for the BA step the update.
  const int ev = (int)ea.scal[kScalVScale];
  double sc[L];
#pragma unroll
  for (int l = 0; l < L; ++l) sc[l] = ldexp(1.0, ev - bits * l);
  const double mu = ea.scal[kScalMu], dref = ea.scal[kScalDRef];
  double s0 = 0.0, m1 = -INFINITY, s2 = 0.0, wmax = 0.0;
  double dv[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    double S = 0.0;
#pragma unroll
    for (int l = 0; l < L; ++l) S = fma((double)acc[l][r], sc[l], S);
    dv[r] = S;
  }
  // residual terms of rounded counts (every tile of a residual launch; in the
  // quarter launch, the residual tiles listed first), added to S in the same
  // order either way
  if (RES == 1 || (RES == 2 && (int)blockIdx.x < go.nres_first)) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int x1 = m0 + wave * 32 + mfma32_row(lane, r), x2 = n0 + lr;
      double S = dv[r];
      for (int e = rs.rowptrA[x1]; e < rs.rowptrA[x1 + 1]; ++e)
        S += (double)rs.entA[2 * e + 1] * rs.V64[(int64_t)rs.entA[2 * e] * N2 + x2];
      for (int e = rs.rowptrB[x2]; e < rs.rowptrB[x2 + 1]; ++e)
        S += rs.W64[(int64_t)x1 * rs.KRp + rs.entB[2 * e]] * (double)rs.entB[2 * e + 1];
      dv[r] = S;
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) dv[r] = fma(-dv[r], ea.inv_c, hv[r] - mu);
  // the accumulators and H are dead now: room for the P tile
  p_loads();
  const int Nq = N2, halfq = N2 / 2;  // quarter mode: N1 = N2
  // quarter mode: the tile's mirror values (-1: none) go to LDS rows of 33
  // doubles; the operand buffer is free since the barrier after the last stage
  double *tl = reinterpret_cast<double *>(sm);
  // canonical: row <= lim, its own mirror: row == lim (lim = c left, N - 1 - c right)
  const int qlim = (n0 + lr < halfq) ? n0 + lr : Nq - 1 - (n0 + lr);
  const int qrow0 = m0 + wave * 32 + (lane >> 5) * 4;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int64_t off = (int64_t)((r >> 2) * 8 + (r & 3)) * ldx;
    const double D = dv[r];
    if constexpr (QTR) {
      const int row = qrow0 + (r >> 2) * 8 + (r & 3);
      const bool canon = row <= qlim, self = row == qlim;
      const double wt = canon ? (self ? 1.0 : 2.0) : 0.0;
      const double p = pv[r];
      const double w = p * exp_tang(D - dref, etab);
      if (canon) {
        __builtin_nontemporal_store(w, Pu + off + lo);
        m1 = fmax(m1, D);
        wmax = fmax(wmax, w);
      }
      s0 += wt * w;
      s2 = fma(wt * p, D, s2);
      if (go.mirror) tl[(wave * 32 + mfma32_row(lane, r)) * 33 + lr] = canon && !self ? w : -1.0;
    } else {
      m1 = fmax(m1, D);
      if constexpr (EPI == EPI_BA) {
        const double p = pv[r];
        const double w = p * exp_tang(D - dref, etab);
        __builtin_nontemporal_store(w, Pu + off + lo);
        s0 += w;
        s2 = fma(p, D, s2);
        wmax = fmax(wmax, w);
      } else {
        Pu[off + lo] = D;
      }
    }
  }
  block_reduce3(s0, m1, s2, red);
  if (tid == 0) {
    double *pp = ea.part + (int64_t)(QTR ? go.part_off + (int)blockIdx.x : ty * TX + tx) * kPartStrideDev;
    pp[0] = s0;
    pp[1] = m1;
    pp[2] = s2;
  }
  // ---- BA step: exponent of the new P tile for the next output law (whose
  // contraction splits P into limbs as it reads it)
  if constexpr (EPI == EPI_BA) {
    double wmax_w = wmax;  // this wave's 32 rows
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wmax_w = fmax(wmax_w, __shfl_xor(wmax_w, o, 64));
    const double wmax_b = block_max(wmax, red + 16);
    wmax = wmax_b;
    if constexpr (QTR) {
      // (a tile or wave whose canonical p are all zero enters no exponent)
      if (tid == 0 && wmax_b > 0.0) atomicMax(&go.texp[(m0 / 128) * go.tcols + n0 / 32], tile_exponent(wmax_b, go.bits_p));
      if (!go.mirror && lane == 0 && wmax_w > 0.0) {
        // the output law reads this wave's 32 rows of the tile transposed, as
        // part of the strip of columns m0 + 32 wave (left of the middle) or
        // N - 32 - m0 - 32 wave (right), at the 128-row chunk holding the
        // tile's columns: they enter that chunk's exponent
        const bool lt = n0 < halfq;
        const int c0 = lt ? (n0 & ~127) : ((Nq - 32 - n0) & ~127);
        const int s0 = lt ? m0 + 32 * wave : Nq - 32 - m0 - 32 * wave;
        if (c0 + 127 > (lt ? s0 : Nq - 32 - s0))
          atomicMax(&go.texp[(c0 / 128) * go.tcols + s0 / 32], tile_exponent(wmax_w, go.bits_p));
      }
    } else {
      if (tid == 0) go.texp[(m0 / 128) * go.tcols + n0 / 32] = tile_exponent(wmax, go.bits_p);
    }
  }
  if (QTR && go.mirror) {
    // mirror image of the tile (canonical, non-self elements) from LDS: wave w
    // writes mirror rows 8w..8w+7 (tile columns), each in two passes of 64
    // consecutive entries (tile rows), so a store instruction covers 512
    // contiguous bytes (descending addresses right of the middle column)
    __syncthreads();
    const bool left = n0 < halfq;
    double vlo = -1.0, vhi = -1.0;  // maxima over this lane's entries in tile rows < 64 / >= 64
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int mr = wave * 8 + (i >> 1), mc = (i & 1) * 64 + lane;
      const double v = tl[mc * 33 + mr];
      if (v >= 0.0) {
        const int64_t grow = left ? n0 + mr : Nq - 1 - (n0 + mr);
        const int64_t gcol = left ? m0 + mc : Nq - 1 - (m0 + mc);
        __builtin_nontemporal_store(v, ea.P + grow * ldx + gcol);
        if (i & 1)
          vhi = fmax(vhi, v);
        else
          vlo = fmax(vlo, v);
      }
    }
    // exponent tiles (32 mirror columns each): tile rows mc in [32g, 32g + 32)
    // are lanes 32(g & 1).. of the pass g >> 1; reduce within 32-lane halves
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
      vlo = fmax(vlo, __shfl_xor(vlo, o, 64));
      vhi = fmax(vhi, __shfl_xor(vhi, o, 64));
    }
    if ((lane & 31) == 0) {
      const int64_t grow = left ? n0 + wave * 8 : Nq - 1 - (n0 + wave * 8);
      const int trow = (int)(grow / 128);
      for (int pass = 0; pass < 2; ++pass) {
        const double vm = pass ? vhi : vlo;
        if (vm < 0.0) continue;
        const int g = pass * 2 + (lane >> 5);
        const int tcol = left ? (m0 + 32 * g) / 32 : (Nq - m0 - 32 * g - 32) / 32;
        atomicMax(&go.texp[trow * go.tcols + tcol], tile_exponent(vm, go.bits_p));
      }
    }
  }
}

// reset of the P tile exponents before a quarter-mode step
__global__ void k_fill_i32(int32_t *__restrict__ a, int len, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) a[i] = v;
}

// Completes the half space (rows r < N/2 of N columns) of a quarter-layout P
// after BA steps that updated only its canonical elements: every other element
// takes its partner's value, P[r][c] = P[c][r] left of the middle column and
// P[N-1-c][N-1-r] right of it (the partner is canonical in both cases).  One
// 64 x 64 tile per block; the partner tile comes through LDS so that reads and
// writes both run along rows.  Entries of the partner tile that are not
// partners of this tile's elements are read but unused.
__global__ void __launch_bounds__(256) k_fill_mirror(double *__restrict__ P, int N) {
  __shared__ double t[64][65];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const bool left = c0 < N / 2;
  if (left ? r0 + 63 <= c0 : r0 + c0 + 126 <= N - 1) return;  // canonical throughout
  const int pr0 = left ? c0 : N - 64 - c0, pc0 = left ? r0 : N - 64 - r0;
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int rr = i >> 6, cc = i & 63;
    t[rr][cc] = P[(int64_t)(pr0 + rr) * N + pc0 + cc];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int rr = i >> 6, cc = i & 63, r = r0 + rr, c = c0 + cc;
    if (left ? r > c : r + c > N - 1) P[(int64_t)r * N + c] = left ? t[cc][rr] : t[63 - cc][63 - rr];
  }
}

int launch_fill_mirror(double *P, int N, hipStream_t s) {
  if (N < 128 || N % 128) return BDC_ERR_ARG;
  hipLaunchKernelGGL(k_fill_mirror, dim3(N / 64, N / 128), dim3(256), 0, s, P, N);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// ----------------------------------------------------------------- launchers
#define BDC_SPLIT_L(L, call)     \
  switch (L) {                   \
    case 2: call(2); break;      \
    case 3: call(3); break;      \
    case 4: call(4); break;      \
    default: return BDC_ERR_ARG; \
  }

int launch_tile_exp(const double *P, int M, int N2, int bits, int32_t *texp, hipStream_t s) {
  if (M % 128 || N2 % 64) return BDC_ERR_ARG;
  hipLaunchKernelGGL(k_tile_exp, dim3(N2 / 64, M / 128), dim3(256), 0, s, P, N2, bits, N2 / 32, texp);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

int launch_split_vw(const double *V, int KL, int KLp, int N2, const double *W, int M, int KR, int KRp, int limbs,
                    int bits, _Float16 *vT, _Float16 *wl, double *part, int np_pre, double *scal, int32_t *fill,
                    int fill_n, hipStream_t s) {
  int np = np_pre;
  if (np <= 0) {
    np = 512;
    hipLaunchKernelGGL(k_absmax_vw, dim3(np), dim3(256), 0, s, V, (int64_t)KL * N2, W, M, KR, KRp, part);
  }
  if (N2 % 64 || KLp % 64) return BDC_ERR_ARG;
  dim3 gv(N2 / 64, KLp / 64);
#define LV(LL) \
  hipLaunchKernelGGL(k_split_v<LL>, gv, dim3(256), 0, s, V, KL, KLp, N2, bits, part, np, scal, vT, fill, fill_n)
  BDC_SPLIT_L(limbs, LV)
#undef LV
  if (KR > 0) {
    const int64_t len = (int64_t)M * KRp;
#define LW(LL) \
  hipLaunchKernelGGL(k_split_w<LL>, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, W, len, KR, KRp, bits, scal, wl)
    BDC_SPLIT_L(limbs, LW)
#undef LW
  }
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

int launch_vw_limbs(const double *V, int KL, int KLp, int N2, const MirrorW &mw, const double *Vm, int n2, double *W,
                    int M, int KR, int KRp, int limbs, int bits, _Float16 *vT, _Float16 *wl, double *part, int np,
                    double *scal, int32_t *fill, int fill_n, hipStream_t s) {
  int b = 0;
  while (b < 31 && (1 << b) < M) ++b;
  const bool tiled = (1 << b) == M && M >= 32 && b <= n2 && N2 == (1 << n2) && KR > 0;
  if (!tiled || np <= 0 || N2 % 64 || KLp % 64) {  // the separate launches
    const int rc = launch_w_from_v(mw, Vm, N2, n2, W, M, KR, KRp, s);
    if (rc != BDC_OK) return rc;
    return launch_split_vw(V, KL, KLp, N2, W, M, KR, KRp, limbs, bits, vT, wl, part, np, scal, fill, fill_n, s);
  }
  const int nbv = (N2 / 64) * (KLp / 64), nbw = ((KR + 63) / 64) * (M / 32);
#define LVW(LL)                                                                                                       \
  hipLaunchKernelGGL(k_vw_limbs<LL>, dim3((unsigned)(nbv + nbw)), dim3(256), 0, s, V, KL, KLp, N2, bits, part, np, scal, \
                     vT, fill, fill_n, mw, Vm, n2, n2 - b, W, wl, M, KR, KRp, nbv)
  BDC_SPLIT_L(limbs, LVW)
#undef LVW
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

int launch_split_g1(int KLp, int Kdim, int lda, int ncols, const _Float16 *aT, const double *X, const int32_t *texp,
                    int limbs, int bits, int ks, double *Tpart, int tri, int merge, int fields, int uniform_loops,
                    hipStream_t s) {
  if (ks < 1 || Kdim % (ks * 128) || ncols % 64 || lda < Kdim || lda % 16) return BDC_ERR_ARG;
  if (tri && (ncols != 2 * Kdim || ncols < 256)) return BDC_ERR_ARG;
  if (reinterpret_cast<uintptr_t>(aT) % 16) return BDC_ERR_ARG;
  const half8 *aO = reinterpret_cast<const half8 *>(aT);
  const int blocks = ks * (ncols / 32) * ((KLp + 255) / 256);
#define LG4(LL, TT, MM, IB)                                                                                          \
  hipLaunchKernelGGL((k_split_g1<LL, TT, MM, IB>), dim3(blocks), dim3(512), 0, s, KLp, Kdim, lda, ncols, aO, X, texp, \
                     bits, ks, Tpart, uniform_loops)
#define LG3(LL, TT, MM) LG4(LL, TT, MM, 0)
#define LG(LL)                                              \
  if (merge && LL >= 4) {                                   \
    if (tri) LG3(LL, true, (LL >= 4)); else LG3(LL, false, (LL >= 4)); \
  } else {                                                  \
    if (tri) LG3(LL, true, false); else LG3(LL, false, false); \
  }
#define LFB(IB)                                  \
  if (tri) LG4(5, true, true, IB); else LG4(5, false, true, IB);
  // five limbs for the narrow limbs of half length >= 15 (7 bits: 35 bits of p per tile)
  if (limbs == 5 && merge && fields && bits >= 7 && bits <= 9) {
    // limbs from bit fields (7 to 9 bits: half lengths 14 to 16)
    if (bits == 9) {
      LFB(9);
    } else if (bits == 8) {
      LFB(8);
    } else {
      LFB(7);
    }
  } else if (limbs == 5) {
    LG(5);
  } else {
    BDC_SPLIT_L(limbs, LG)
  }
#undef LFB
#undef LG4
#undef LG3
#undef LG
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

int launch_resid_g1(const int32_t *res, int nres, const double *X, int rows, int N2, int tri, double *T,
                    hipStream_t s) {
  if (nres == 0) return BDC_OK;
  if (tri && N2 != 2 * rows) return BDC_ERR_ARG;
  hipLaunchKernelGGL(k_resid_g1, dim3((unsigned)((N2 + 255) / 256)), dim3(256), 0, s, res, nres, X, rows, N2, tri, T);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

int launch_split_g2(const SplitG2Args &a, const EpiArgs &ea, hipStream_t s, int *nb) {
  if (a.M % 128 || a.N2 % 64 || a.KLp % 64 || a.KRp % 64) return BDC_ERR_ARG;
  G2Resid rs{a.blkA, a.blkB, a.rowptrA, a.entA, a.rowptrB, a.entB, a.V64, a.W64, a.KRp};
  const int TX = a.N2 / 32, TY = a.M / 128;
  G2Out go{};
  go.texp = a.texp_next;
  go.bits_p = a.bits_p;
  go.tcols = a.N2 / 32;
  go.patches = (TX % 8 == 0 && TY % 8 == 0 && ((TX / 8) * (TY / 8)) % 8 == 0) ? 1 : 0;
  go.M = a.M;
  go.exptab = a.exptab;
  go.mirror = a.mirror;
  if (a.epi == EPI_BA && !a.texp_next) return BDC_ERR_ARG;
  const int blocks = TX * TY;
  if (nb && !a.quarter) *nb = blocks;
  if (a.quarter) {  // the BA step over the quarter's tiles (lists), partials by list position
    if (a.epi != EPI_BA || a.N1 != a.N2 || a.M != a.N2 / 2) return BDC_ERR_ARG;
    const int nt = (a.M / 128) * (a.N2 / 32);
    if (!a.texp_reset)  // (else the V limb split has reset the exponents)
      hipLaunchKernelGGL(k_fill_i32, dim3((nt + 255) / 256), dim3(256), 0, s, a.texp_next, nt, (int32_t)INT32_MIN);
    if (nb) *nb = a.nq_tiles + a.nqres_tiles;
    // one launch over the residual tiles, then the rest (a.qall_tiles)
    go.part_off = 0;
    go.nres_first = a.nqres_tiles;
#define LQ(LL)                                                                                                      \
  hipLaunchKernelGGL((k_split_g2<LL, EPI_BA, 2, true>), dim3(a.nq_tiles + a.nqres_tiles), dim3(256), 0, s, a.N2, \
                     a.KLp, a.KRp, a.A16, a.vT, a.wl, a.B16, a.bits, ea, rs, go, a.qall_tiles)
    BDC_SPLIT_L(a.limbs, LQ)
#undef LQ
    return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
  }
#define LG2(LL, EE)                                                                                               \
  do {                                                                                                            \
    hipLaunchKernelGGL((k_split_g2<LL, EE, 0, false>), dim3(blocks), dim3(256), 0, s, a.N2, a.KLp, a.KRp,     \
                       a.A16, a.vT, a.wl, a.B16, a.bits, ea, rs, go, nullptr);                                    \
    if (a.nres_tiles)                                                                                             \
      hipLaunchKernelGGL((k_split_g2<LL, EE, 1, false>), dim3(a.nres_tiles), dim3(256), 0, s, a.N2, a.KLp,     \
                         a.KRp, a.A16, a.vT, a.wl, a.B16, a.bits, ea, rs, go, a.res_tiles);                        \
  } while (0)
#define LG2_E(EE)                \
  switch (a.limbs) {             \
    case 2: LG2(2, EE); break;   \
    case 3: LG2(3, EE); break;   \
    case 4: LG2(4, EE); break;   \
    default: return BDC_ERR_ARG; \
  }
  if (a.epi == EPI_BA) {
    LG2_E(EPI_BA)
  } else if (a.epi == EPI_D) {
    LG2_E(EPI_D)
  } else {
    return BDC_ERR_ARG;
  }
#undef LG2_E
#undef LG2
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

}  // namespace bdc
