// FP64 contractions on CDNA4 FP64 MFMA (v_mfma_f64_16x16x4f64).
//
// One kernel template serves every FP64 contraction of a BA step: the output
// law G1 = Abar^T P, the cross term G2 = [Abar | Wbar][Vbar ; Bbar^T] (two K
// phases with different operand types), and the small stage GEMMs that turn
// G1's slabs into q and ln q into G2's operands.  Operands are read through
// (row, col) strides so transposed views cost nothing but coalescing; count
// matrices stay u16 in HBM and are widened to FP64 on the way into LDS.
//
// The G2 instantiation carries the BA update in its epilogue
// (backend/BAA.cu:487-530 log W, :590-639 update): for each x it forms
// D(x||q) = H(x)/C - ln C - mu - S(x)/C, writes p~(x) exp(D - dref) in place,
// and emits per-block partials (sum of the new p~, max D, sum p~ D).
#include "device_common.hpp"
#include "fast_exp.hpp"
#include "kernels.hpp"

namespace bdc {

typedef double double4_ __attribute__((ext_vector_type(4)));

template <int T>
__device__ inline double ld_elem(const void *p, int64_t i) {
  if constexpr (T == 0)
    return static_cast<const double *>(p)[i];
  else
    return (double)static_cast<const uint16_t *>(p)[i];
}

// Row of the FP64 16x16 MFMA result held by `lane` in register r.
__device__ inline int mfma64_row(int lane, int r) { return (lane >> 4) + 4 * r; }

template <int BM, int BN, int WM, int WN, int TA1, int TB1, int TA2, int TB2, int EPI>
__global__ void __launch_bounds__(256) k_gemm64(int M, int N, Opnd A1, Opnd B1, int K1, Opnd A2, Opnd B2, int K2,
                                                double *__restrict__ C, int64_t crs, int64_t ccs, int kchunk,
                                                int64_t split_stride, EpiArgs ea) {
  constexpr int BK = 16;
  constexpr int TM = BM / WM, TN = BN / WN, SM = TM / 16, SN = TN / 16;
  static_assert(WM * WN == 4, "four waves");
  static_assert(SM >= 1 && SN >= 1, "wave tile at least 16x16");
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];
  __shared__ double red[48];
  __shared__ double etab[32];  // exp table in LDS (a global one could alias the P stores)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (EPI == EPI_BA && tid < 32) etab[tid] = ea.exptab[tid];
  const int wm = wave / WN, wn = wave % WN;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int z = blockIdx.z;

  double4_ acc[SM][SN];
#pragma unroll
  for (int i = 0; i < SM; ++i)
#pragma unroll
    for (int j = 0; j < SN; ++j) acc[i][j] = (double4_){0.0, 0.0, 0.0, 0.0};

  constexpr int AE = BM * BK / 256, BE = BN * BK / 256;  // elements per thread per tile
  double ra[AE], rb[BE];
#pragma unroll
  for (int ph = 0; ph < 2; ++ph) {
    const Opnd A = ph ? A2 : A1;
    const Opnd B = ph ? B2 : B1;
    int kb, ke;
    if (ph == 0) {
      kb = z * kchunk;
      ke = min(K1, kb + kchunk);
    } else {
      kb = 0;
      ke = K2;
    }
    if (kb >= ke) continue;
    // global -> registers for the tile at k0 (zero outside the operand)
    auto gload = [&](int k0) {
#pragma unroll
      for (int q = 0; q < AE; ++q) {
        const int e = tid + 256 * q;
        int m, kk;
        if (A.cs == 1) {
          kk = e % BK;
          m = e / BK;
        } else {
          m = e % BM;
          kk = e / BM;
        }
        const int gm = m0 + m, gk = k0 + kk;
        ra[q] = (gm < M && gk < ke) ? (ph ? ld_elem<TA2>(A.p, gm * A.rs + gk * A.cs) : ld_elem<TA1>(A.p, gm * A.rs + gk * A.cs)) : 0.0;
      }
#pragma unroll
      for (int q = 0; q < BE; ++q) {
        const int e = tid + 256 * q;
        int nn, kk;
        if (B.cs == 1) {
          nn = e % BN;
          kk = e / BN;
        } else {
          kk = e % BK;
          nn = e / BK;
        }
        const int gn = n0 + nn, gk = k0 + kk;
        rb[q] = (gn < N && gk < ke) ? (ph ? ld_elem<TB2>(B.p, gk * B.rs + gn * B.cs) : ld_elem<TB1>(B.p, gk * B.rs + gn * B.cs)) : 0.0;
      }
    };
    auto sstore = [&]() {
#pragma unroll
      for (int q = 0; q < AE; ++q) {
        const int e = tid + 256 * q;
        if (A.cs == 1)
          As[e % BK][e / BK] = ra[q];
        else
          As[e / BM][e % BM] = ra[q];
      }
#pragma unroll
      for (int q = 0; q < BE; ++q) {
        const int e = tid + 256 * q;
        if (B.cs == 1)
          Bs[e / BN][e % BN] = rb[q];
        else
          Bs[e % BK][e / BK] = rb[q];
      }
    };
    gload(kb);
    for (int k0 = kb; k0 < ke; k0 += BK) {
      __syncthreads();  // previous tile fully consumed
      sstore();
      __syncthreads();
      if (k0 + BK < ke) gload(k0 + BK);  // prefetch overlaps the MFMAs below
#pragma unroll
      for (int ks = 0; ks < BK / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        double a[SM], b[SN];
#pragma unroll
        for (int i = 0; i < SM; ++i) a[i] = As[kr][wm * TM + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < SN; ++j) b[j] = Bs[kr][wn * TN + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
          for (int j = 0; j < SN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }

  if constexpr (EPI == EPI_STORE) {
    double *Cz = C + (int64_t)z * split_stride;
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * TM + i * 16 + mfma64_row(lane, r);
          const int col = n0 + wn * TN + j * 16 + (lane & 15);
          if (row < M && col < N) Cz[row * crs + col * ccs] = acc[i][j][r];
        }
  } else {
    __syncthreads();  // etab visible even when no K loop ran
    const double mu = ea.scal[kScalMu], dref = ea.scal[kScalDRef];
    double s0 = 0.0, m1 = -INFINITY, s2 = 0.0;
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      // issue the H (and P) loads of 16 elements before any store to P, so
      // their latencies overlap instead of ordering behind the stores
      double hv[SN][4], pv[SN][4];
#pragma unroll
      for (int j = 0; j < SN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * TM + i * 16 + mfma64_row(lane, r);
          const int col = n0 + wn * TN + j * 16 + (lane & 15);
          const int64_t x = (int64_t)row * ea.ldx + col;
          const bool in = row < M && col < N;
          hv[j][r] = in ? ea.Hc[x] : 0.0;
          if constexpr (EPI == EPI_BA) pv[j][r] = in ? ea.P[x] : 0.0;
        }
#pragma unroll
      for (int j = 0; j < SN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * TM + i * 16 + mfma64_row(lane, r);
          const int col = n0 + wn * TN + j * 16 + (lane & 15);
          if (row < M && col < N) {
            const int64_t x = (int64_t)row * ea.ldx + col;
            const double D = hv[j][r] - mu - acc[i][j][r] * ea.inv_c;
            m1 = fmax(m1, D);
            if constexpr (EPI == EPI_BA) {
              const double p = pv[j][r];
              const double w = p * exp_tang(D - dref, etab);
              ea.P[x] = w;
              s0 += w;
              s2 += p * D;
            } else {
              ea.D[x] = D;
            }
          }
        }
    }
    block_reduce3(s0, m1, s2, red);
    if (tid == 0) {
      double *pp = ea.part + (int64_t)(blockIdx.y * gridDim.x + blockIdx.x) * kPartStrideDev;
      pp[0] = s0;
      pp[1] = m1;
      pp[2] = s2;
    }
  }
}

template <int BM, int BN, int WM, int WN, int TA1, int TB1, int TA2, int TB2, int EPI>
static int run(int M, int N, const Opnd &A1, const Opnd &B1, int K1, const Opnd &A2, const Opnd &B2, int K2, double *C,
               int64_t crs, int64_t ccs, int splits, int64_t split_stride, const EpiArgs &ea, hipStream_t s, int *nb) {
  const int kchunk = ((K1 + splits - 1) / splits + 15) / 16 * 16;
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, splits);
  if (nb) *nb = (int)(grid.x * grid.y);
  hipLaunchKernelGGL((k_gemm64<BM, BN, WM, WN, TA1, TB1, TA2, TB2, EPI>), grid, dim3(256), 0, s, M, N, A1, B1, K1, A2,
                     B2, K2, C, crs, ccs, kchunk, split_stride, ea);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

template <int TA1, int TB1, int TA2, int TB2, int EPI>
static int dispatch_shape(int M, int N, const Opnd &A1, const Opnd &B1, int K1, const Opnd &A2, const Opnd &B2, int K2,
                          double *C, int64_t crs, int64_t ccs, int splits, int64_t split_stride, const EpiArgs &ea,
                          hipStream_t s, int *nb) {
  if (M <= 16 && EPI == EPI_STORE)
    return run<16, 64, 1, 4, TA1, TB1, TA2, TB2, EPI>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, splits, split_stride,
                                                      ea, s, nb);
  if (N <= 16 && EPI == EPI_STORE)
    return run<64, 16, 4, 1, TA1, TB1, TA2, TB2, EPI>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, splits, split_stride,
                                                      ea, s, nb);
  if (M >= 128 && N >= 128 && (int64_t)M * N >= (int64_t)128 * 128 * 256)
    return run<128, 128, 2, 2, TA1, TB1, TA2, TB2, EPI>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, splits, split_stride,
                                                        ea, s, nb);
  return run<64, 64, 2, 2, TA1, TB1, TA2, TB2, EPI>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, splits, split_stride,
                                                    ea, s, nb);
}

int gemm64(int M, int N, const Opnd &A1, const Opnd &B1, int K1, const Opnd &A2, const Opnd &B2, int K2, double *C,
           int64_t crs, int64_t ccs, int splits, int64_t split_stride, int epi, const EpiArgs *ea_in, hipStream_t s,
           int *nb) {
  EpiArgs ea{};
  if (ea_in) ea = *ea_in;
  if (splits < 1) splits = 1;
  if (splits > 1 && K2 != 0) return BDC_ERR_ARG;
  const int code = (A1.type << 3) | (B1.type << 2) | (K2 ? ((A2.type << 1) | B2.type) : 0);
  if (epi == EPI_STORE) {
    switch (code) {
      case 0b0100:  // f64 x u16
        return dispatch_shape<0, 1, 0, 0, EPI_STORE>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, splits, split_stride,
                                                     ea, s, nb);
      case 0b1000:  // u16 x f64
        return dispatch_shape<1, 0, 0, 0, EPI_STORE>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, splits, split_stride,
                                                     ea, s, nb);
      case 0b0000:  // f64 x f64
        return dispatch_shape<0, 0, 0, 0, EPI_STORE>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, splits, split_stride,
                                                     ea, s, nb);
      case 0b1001:  // G2 phases: (u16 x f64) then (f64 x u16)
        return dispatch_shape<1, 0, 0, 1, EPI_STORE>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, splits, split_stride,
                                                     ea, s, nb);
      default:
        return BDC_ERR_ARG;
    }
  }
  if (code != 0b1001 && code != 0b1000) return BDC_ERR_ARG;
  if (epi == EPI_BA) {
    if (code == 0b1001) return dispatch_shape<1, 0, 0, 1, EPI_BA>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, 1, 0, ea, s, nb);
    return dispatch_shape<1, 0, 0, 0, EPI_BA>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, 1, 0, ea, s, nb);
  }
  if (epi == EPI_D) {
    if (code == 0b1001) return dispatch_shape<1, 0, 0, 1, EPI_D>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, 1, 0, ea, s, nb);
    return dispatch_shape<1, 0, 0, 0, EPI_D>(M, N, A1, B1, K1, A2, B2, K2, C, crs, ccs, 1, 0, ea, s, nb);
  }
  return BDC_ERR_ARG;
}

// ---------------------------------------------------------------------------
// The two large FP64 products of a BA step on tile-aligned operands (every
// dimension a multiple of the tile once K is padded to the operands' zero
// columns): 128 x 128 tiles of four waves (64 x 64 each, 4 x 4 MFMA tiles),
// K stages of 16 through two LDS buffers (one barrier per stage; the next
// stage's global loads are issued before the current stage's MFMAs), and
// 16-byte global loads.  Operand layouts:
//   A code 0: u16 A(m, k) = a[k * lda + m]    (Abar^T of the output law)
//   A code 1: u16 A(m, k) = a[m * lda + k]    (Abar of the cross term)
//   A code 2: f64 A(m, k) = a[m * lda + k]    (Wbar of the cross term)
//   B code 0: f64 B(k, n) = b[k * ldb + n]    (p, Vbar)
//   B code 1: u16 B(k, n) = b[n * ldb + k]    (Bbar^T of the cross term)
constexpr int kG64Tile = 128, kG64BK = 16, kG64Row = kG64Tile + 2;

template <int ACODE, int BCODE>
__device__ inline void g64_load(const void *a, int64_t lda, const void *b, int64_t ldb, int m0, int n0, int k0,
                                int tid, double (&ra)[8], double (&rb)[8]) {
  if constexpr (ACODE == 0) {
    const int k = tid >> 4, m = (tid & 15) * 8;
    const uint4 v = *reinterpret_cast<const uint4 *>(static_cast<const uint16_t *>(a) + (int64_t)(k0 + k) * lda + m0 + m);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) ra[i] = (double)((w[i >> 1] >> (16 * (i & 1))) & 0xffffu);
  } else if constexpr (ACODE == 1) {
    const int m = tid >> 1, k = (tid & 1) * 8;
    const uint4 v = *reinterpret_cast<const uint4 *>(static_cast<const uint16_t *>(a) + (int64_t)(m0 + m) * lda + k0 + k);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) ra[i] = (double)((w[i >> 1] >> (16 * (i & 1))) & 0xffffu);
  } else {
    const int m = tid >> 1, k = (tid & 1) * 8;
    const double2 *p = reinterpret_cast<const double2 *>(static_cast<const double *>(a) + (int64_t)(m0 + m) * lda + k0 + k);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double2 v = p[i];
      ra[2 * i] = v.x;
      ra[2 * i + 1] = v.y;
    }
  }
  if constexpr (BCODE == 0) {
    const int k = tid >> 4, n = (tid & 15) * 8;
    const double2 *p = reinterpret_cast<const double2 *>(static_cast<const double *>(b) + (int64_t)(k0 + k) * ldb + n0 + n);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double2 v = p[i];
      rb[2 * i] = v.x;
      rb[2 * i + 1] = v.y;
    }
  } else {
    const int n = tid >> 1, k = (tid & 1) * 8;
    const uint4 v = *reinterpret_cast<const uint4 *>(static_cast<const uint16_t *>(b) + (int64_t)(n0 + n) * ldb + k0 + k);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) rb[i] = (double)((w[i >> 1] >> (16 * (i & 1))) & 0xffffu);
  }
}

template <int ACODE, int BCODE>
__device__ inline void g64_store(double (*As)[kG64Row], double (*Bs)[kG64Row], int tid, const double (&ra)[8],
                                 const double (&rb)[8]) {
  if constexpr (ACODE == 0) {
    const int k = tid >> 4, m = (tid & 15) * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) As[k][m + i] = ra[i];
  } else {
    const int m = tid >> 1, k = (tid & 1) * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) As[k + i][m] = ra[i];
  }
  if constexpr (BCODE == 0) {
    const int k = tid >> 4, n = (tid & 15) * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) Bs[k][n + i] = rb[i];
  } else {
    const int n = tid >> 1, k = (tid & 1) * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) Bs[k + i][n] = rb[i];
  }
}

// acc += A[m0.., kb..ke) B[kb..ke), n0..]; leaves both LDS buffers free (ends on a barrier)
template <int ACODE, int BCODE>
__device__ inline void g64_phase(const void *a, int64_t lda, const void *b, int64_t ldb, int m0, int n0, int kb,
                                 int ke, double (*As)[kG64BK][kG64Row], double (*Bs)[kG64BK][kG64Row],
                                 double4_ (&acc)[4][4]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int nst = (ke - kb) / kG64BK;
  if (nst <= 0) return;
  double ra[8], rb[8];
  g64_load<ACODE, BCODE>(a, lda, b, ldb, m0, n0, kb, tid, ra, rb);
  g64_store<ACODE, BCODE>(As[0], Bs[0], tid, ra, rb);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int cur = st & 1;
    const bool more = st + 1 < nst;
    if (more) g64_load<ACODE, BCODE>(a, lda, b, ldb, m0, n0, kb + (st + 1) * kG64BK, tid, ra, rb);
#pragma unroll
    for (int ks = 0; ks < kG64BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      double av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = As[cur][kr][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[cur][kr][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    if (more) g64_store<ACODE, BCODE>(As[cur ^ 1], Bs[cur ^ 1], tid, ra, rb);
    __syncthreads();
  }
}

// acc += A1[m0.., 0..K1) B1[0..K1), n0..] + A2[m0.., 0..K2) B2[0..K2), n0..]
// (the cross term's two K phases, operand codes 1/0 then 2/1) in one loop:
// only the stage loads and LDS stores switch on the phase (a uniform branch),
// so the accumulators live in one register set (two loops, one per phase,
// made the compiler keep a second set and copy between them)
__device__ inline void g64_two_phase(const void *a1, int64_t lda1, const void *b1, int64_t ldb1, int K1,
                                     const void *a2, int64_t lda2, const void *b2, int64_t ldb2, int K2, int m0,
                                     int n0, double (*As)[kG64BK][kG64Row], double (*Bs)[kG64BK][kG64Row],
                                     double4_ (&acc)[4][4]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int nst1 = K1 / kG64BK, nst = nst1 + K2 / kG64BK;
  if (nst <= 0) return;
  double ra[8], rb[8];
  auto load = [&](int st) {
    if (st < nst1)
      g64_load<1, 0>(a1, lda1, b1, ldb1, m0, n0, st * kG64BK, tid, ra, rb);
    else
      g64_load<2, 1>(a2, lda2, b2, ldb2, m0, n0, (st - nst1) * kG64BK, tid, ra, rb);
  };
  auto store = [&](int buf, int st) {
    if (st < nst1)
      g64_store<1, 0>(As[buf], Bs[buf], tid, ra, rb);
    else
      g64_store<2, 1>(As[buf], Bs[buf], tid, ra, rb);
  };
  load(0);
  store(0, 0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int cur = st & 1;
    const bool more = st + 1 < nst;
    if (more) load(st + 1);
#pragma unroll
    for (int ks = 0; ks < kG64BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      double av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = As[cur][kr][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[cur][kr][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    if (more) store(cur ^ 1, st + 1);
    __syncthreads();
  }
}

// Output law: C[z] = Abar^T X over the K range of split z (C + z * split_stride)
__global__ void __launch_bounds__(256) k_g64_g1(const uint16_t *__restrict__ A, int64_t lda,
                                                const double *__restrict__ X, int64_t ldb, int kchunk,
                                                double *__restrict__ C, int64_t ldc, int64_t split_stride) {
  __shared__ double As[2][kG64BK][kG64Row];
  __shared__ double Bs[2][kG64BK][kG64Row];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * kG64Tile, n0 = blockIdx.x * kG64Tile;
  double4_ acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (double4_){0.0, 0.0, 0.0, 0.0};
  const int kb = blockIdx.z * kchunk;
  g64_phase<0, 0>(A, lda, X, ldb, m0, n0, kb, kb + kchunk, As, Bs, acc);
  double *Cz = C + (int64_t)blockIdx.z * split_stride;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + mfma64_row(lane, r);
        const int col = n0 + wn * 64 + j * 16 + (lane & 15);
        Cz[(int64_t)row * ldc + col] = acc[i][j][r];
      }
}

// Cross term with the fused BA update (EPI_BA) or the D pass (EPI_D):
// S = Abar[rows] Vbar + Wbar[rows] Bbar^T, then the epilogue of k_gemm64
template <int EPI>
__global__ void __launch_bounds__(256, 2) k_g64_g2(const uint16_t *__restrict__ A1, int64_t lda1,
                                                const double *__restrict__ B1, int64_t ldb1, int K1,
                                                const double *__restrict__ A2, int64_t lda2,
                                                const uint16_t *__restrict__ B2, int64_t ldb2, int K2, EpiArgs ea) {
  __shared__ double As[2][kG64BK][kG64Row];
  __shared__ double Bs[2][kG64BK][kG64Row];
  __shared__ double red[48];
  __shared__ double etab[32];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wm = wave >> 1, wn = wave & 1;
  if (EPI == EPI_BA && tid < 32) etab[tid] = ea.exptab[tid];
  const int m0 = blockIdx.y * kG64Tile, n0 = blockIdx.x * kG64Tile;
  double4_ acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (double4_){0.0, 0.0, 0.0, 0.0};
  g64_two_phase(A1, lda1, B1, ldb1, K1, A2, lda2, B2, ldb2, K2, m0, n0, As, Bs, acc);
  __syncthreads();  // etab visible even when no K stage ran
  const double mu = ea.scal[kScalMu], dref = ea.scal[kScalDRef];
  double s0 = 0.0, m1 = -INFINITY, s2 = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double hv[4][4], pv[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t x = (int64_t)(m0 + wm * 64 + i * 16 + mfma64_row(lane, r)) * ea.ldx + n0 + wn * 64 + j * 16 +
                          (lane & 15);
        hv[j][r] = ea.Hc[x];
        if constexpr (EPI == EPI_BA) pv[j][r] = ea.P[x];
      }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t x = (int64_t)(m0 + wm * 64 + i * 16 + mfma64_row(lane, r)) * ea.ldx + n0 + wn * 64 + j * 16 +
                          (lane & 15);
        const double D = hv[j][r] - mu - acc[i][j][r] * ea.inv_c;
        m1 = fmax(m1, D);
        if constexpr (EPI == EPI_BA) {
          const double p = pv[j][r];
          const double w = p * exp_tang(D - dref, etab);
          ea.P[x] = w;
          s0 += w;
          s2 += p * D;
        } else {
          ea.D[x] = D;
        }
      }
  }
  block_reduce3(s0, m1, s2, red);
  if (threadIdx.x == 0) {
    double *pp = ea.part + (int64_t)(blockIdx.y * gridDim.x + blockIdx.x) * kPartStrideDev;
    pp[0] = s0;
    pp[1] = m1;
    pp[2] = s2;
  }
}

bool g64_g1_fits(int Mpad, int N, int K, int splits) {
  return Mpad % kG64Tile == 0 && N % kG64Tile == 0 && splits >= 1 && K % (kG64BK * splits) == 0;
}
int launch_g64_g1(int Mpad, int N, int K, const uint16_t *A, int64_t lda, const double *X, int64_t ldb, double *C,
                  int64_t ldc, int splits, int64_t split_stride, hipStream_t s) {
  if (!g64_g1_fits(Mpad, N, K, splits) || lda % 8 || ldb % 2) return BDC_ERR_ARG;
  dim3 grid(N / kG64Tile, Mpad / kG64Tile, splits);
  hipLaunchKernelGGL(k_g64_g1, grid, dim3(256), 0, s, A, lda, X, ldb, K / splits, C, ldc, split_stride);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}
bool g64_g2_fits(int M, int N, int K1, int K2) {
  return M % kG64Tile == 0 && N % kG64Tile == 0 && K1 % kG64BK == 0 && K2 % kG64BK == 0;
}
int launch_g64_g2(int M, int N, const uint16_t *A1, int64_t lda1, const double *B1, int64_t ldb1, int K1,
                  const double *A2, int64_t lda2, const uint16_t *B2, int64_t ldb2, int K2, int epi,
                  const EpiArgs &ea, hipStream_t s, int *nb) {
  if (!g64_g2_fits(M, N, K1, K2) || lda1 % 8 || ldb1 % 2 || lda2 % 2 || ldb2 % 8) return BDC_ERR_ARG;
  dim3 grid(N / kG64Tile, M / kG64Tile);
  if (nb) *nb = (int)(grid.x * grid.y);
  if (epi == EPI_BA)
    hipLaunchKernelGGL(k_g64_g2<EPI_BA>, grid, dim3(256), 0, s, A1, lda1, B1, ldb1, K1, A2, lda2, B2, ldb2, K2, ea);
  else if (epi == EPI_D)
    hipLaunchKernelGGL(k_g64_g2<EPI_D>, grid, dim3(256), 0, s, A1, lda1, B1, ldb1, K1, A2, lda2, B2, ldb2, K2, ea);
  else
    return BDC_ERR_ARG;
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

// C[i] = sum_z part[z * stride + i]: four lanes per output (each a quarter of
// the splits, combined in a fixed order), so small outputs still fill the GPU
__global__ void k_reduce_splits(const double *__restrict__ part, int splits, int64_t stride, uint64_t len,
                                double *__restrict__ C) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t i = t >> 2;
  const int q = (int)(t & 3);
  double s = 0.0;
  if (i < len)
    for (int z = q; z < splits; z += 4) s += part[(int64_t)z * stride + i];
  s += __shfl_xor(s, 1);
  s += __shfl_xor(s, 2);
  if (i < len && q == 0) C[i] = s;
}
int launch_reduce_splits(const double *part, int splits, int64_t stride, uint64_t len, double *C, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_splits, dim3((unsigned)((4 * len + 255) / 256)), dim3(256), 0, s, part, splits, stride,
                     len, C);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

}  // namespace bdc
