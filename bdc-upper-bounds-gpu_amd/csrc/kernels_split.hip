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
// integers exactly, whatever order the matrix core uses.  The only error left
// is the truncation of the FP64 operand to L*b bits.
#include "device_common.hpp"
#include "kernels.hpp"

namespace bdc {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16_ __attribute__((ext_vector_type(16)));

// row of the 32x32 MFMA result held by `lane` in register r (verified layout)
__device__ inline int mfma32_row(int lane, int r) { return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); }

__device__ inline int tile_exponent(double mx, int bits) {
  // smallest e with mx <= 2^(e + bits - 1)
  if (!(mx > 0.0)) return 0;
  return ilogb(mx) + 1 - (bits - 1);
}

// ----------------------------------------------------------------- P limbs
// One 128x128 tile of P per block: tile max -> exponent, limbs written
// transposed ([l][x2][x1]) so that the output-law contraction reads K = x1
// contiguously.
template <int L>
__global__ void __launch_bounds__(1024) k_split_p(const double *__restrict__ P, const double *__restrict__ scal, int N1,
                                                  int N2, int bits, _Float16 *__restrict__ limbT,
                                                  int32_t *__restrict__ texp) {
  constexpr int PITCH = 136;  // halves per LDS row: 16-byte aligned rows
  __shared__ __attribute__((aligned(16))) _Float16 tile[128][PITCH];  // [x2][x1]
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int m0 = blockIdx.y * 128, n0 = blockIdx.x * 128;
  double v[16];
  double mx = 0.0;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int e = tid + 1024 * r, row = e >> 7, col = e & 127;
    v[r] = P[(int64_t)(m0 + row) * N2 + n0 + col];
    mx = fmax(mx, fabs(v[r]));
  }
  mx = block_max(mx, red);
  const int et = tile_exponent(mx, bits);
  if (tid == 0) texp[blockIdx.y * gridDim.x + blockIdx.x] = et;
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = ldexp(v[r], -et);
  const double radix = ldexp(1.0, bits);
  for (int l = 0; l < L; ++l) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int e = tid + 1024 * r, row = e >> 7, col = e & 127;
      const double li = rint(v[r]);
      tile[col][row] = (_Float16)(float)li;
      v[r] = (v[r] - li) * radix;
    }
    __syncthreads();
    _Float16 *dst = limbT + (int64_t)l * N1 * N2;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int ch = tid + 1024 * q, c2 = ch >> 4, g = ch & 15;
      *reinterpret_cast<uint4 *>(dst + (int64_t)(n0 + c2) * N1 + m0 + 8 * g) =
          *reinterpret_cast<const uint4 *>(&tile[c2][8 * g]);
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------- V/W limbs
// max |V| over the KL x N2 block and |W| over the N1 x KR block
__global__ void __launch_bounds__(256) k_absmax_vw(const double *__restrict__ V, int64_t vlen, const double *__restrict__ W,
                                                   int N1, int KR, int KRp, double *__restrict__ part) {
  __shared__ double red[16];
  double m = 0.0;
  const int64_t wlen = (int64_t)N1 * KR;
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
__global__ void __launch_bounds__(256) k_vw_exponent(const double *__restrict__ part, int np, int bits,
                                                     double *__restrict__ scal) {
  __shared__ double red[16];
  double m = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) m = fmax(m, part[i]);
  m = block_max(m, red);
  if (threadIdx.x == 0) scal[kScalVScale] = (double)tile_exponent(m, bits);
}
// V (KLp x N2, row kappa) -> vT[l][x2][kappa] (N2 x KLp), 64x64 tiles through LDS
template <int L>
__global__ void __launch_bounds__(256) k_split_v(const double *__restrict__ V, int KL, int KLp, int N2, int bits,
                                                 const double *__restrict__ scal, _Float16 *__restrict__ vT) {
  __shared__ double t[64][65];
  const int k0 = blockIdx.y * 64, x0 = blockIdx.x * 64;
  const int e = (int)scal[kScalVScale];
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
      vT[(int64_t)l * N2 * KLp + (int64_t)(x0 + xx) * KLp + k0 + kk] = (_Float16)(float)li;
      v = (v - li) * radix;
    }
  }
}
// W (N1 x KRp) -> wl[l][x1][kappa']
template <int L>
__global__ void k_split_w(const double *__restrict__ W, int64_t len, int KR, int KRp, int bits,
                          const double *__restrict__ scal, _Float16 *__restrict__ wl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const int e = (int)scal[kScalVScale];
  const int c = (int)(i % KRp);
  double v = c < KR ? ldexp(W[i], -e) : 0.0;
  const double radix = ldexp(1.0, bits);
  for (int l = 0; l < L; ++l) {
    const double li = rint(v);
    wl[(int64_t)l * len + i] = (_Float16)(float)li;
    v = (v - li) * radix;
  }
}

// ----------------------------------------------------------------- G1
// Tbar[kappa][x2] = sum_x1 Abar[x1][kappa] P[x1][x2] with P in L limbs.
// Block: 256 (kappa) x 32 (x2); 8 waves, each 32 kappa x 32 x2 and all L limbs
// (64 FP32 accumulators + 16 FP64 per lane).  K step 32 through LDS,
// double-buffered; FP32 sums promoted to FP64 every 128 rows of x1 (one P tile
// row -> one exponent).  blockIdx.z splits K.
template <int L>
__global__ void __launch_bounds__(512) k_split_g1(int KLp, int N1, int N2, const _Float16 *__restrict__ aT,
                                                  const _Float16 *__restrict__ pT, const int32_t *__restrict__ texp,
                                                  int bits, int krows, double *__restrict__ T) {
  constexpr int BM = 256, BN = 32, BK = 32, PADK = BK + 8;
  __shared__ __attribute__((aligned(16))) _Float16 As[2][BM][PADK];
  __shared__ __attribute__((aligned(16))) _Float16 Bs[2][L][BN][PADK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int ntile2 = N2 / 128;

  float16_ acc[L];
  double acc64[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc64[r] = 0.0;
#pragma unroll
  for (int l = 0; l < L; ++l) acc[l] = (float16_){0};

  constexpr int ACH = BM * BK / 8 / 512;              // 2 chunks of 16 B per thread
  constexpr int BTOT = L * BN * BK / 8;               // L * 128 chunks
  constexpr int BCH = (BTOT + 511) / 512;             // 1
  uint4 ra[ACH], rb[BCH];
  auto gload = [&](int k0) {
#pragma unroll
    for (int c = 0; c < ACH; ++c) {
      const int ch = tid + 512 * c, row = ch >> 2, kq = ch & 3;
      const int gr = m0 + row;
      ra[c] = gr < KLp ? *reinterpret_cast<const uint4 *>(aT + (int64_t)gr * N1 + k0 + kq * 8) : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int c = 0; c < BCH; ++c) {
      const int ch = tid + 512 * c;
      if (ch < BTOT) {
        const int l = ch / (BN * BK / 8), rem = ch % (BN * BK / 8), row = rem >> 2, kq = rem & 3;
        rb[c] = *reinterpret_cast<const uint4 *>(pT + (int64_t)l * N1 * N2 + (int64_t)(n0 + row) * N1 + k0 + kq * 8);
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int c = 0; c < ACH; ++c) {
      const int ch = tid + 512 * c, row = ch >> 2, kq = ch & 3;
      *reinterpret_cast<uint4 *>(&As[buf][row][kq * 8]) = ra[c];
    }
#pragma unroll
    for (int c = 0; c < BCH; ++c) {
      const int ch = tid + 512 * c;
      if (ch < BTOT) {
        const int l = ch / (BN * BK / 8), rem = ch % (BN * BK / 8), row = rem >> 2, kq = rem & 3;
        *reinterpret_cast<uint4 *>(&Bs[buf][l][row][kq * 8]) = rb[c];
      }
    }
  };

  const int kbase = blockIdx.z * krows;
  T += (int64_t)blockIdx.z * KLp * N2;
  const int nsteps = krows / BK;
  gload(kbase);
  sstore(0);
  __syncthreads();
  for (int s = 0; s < nsteps; ++s) {
    const int buf = s & 1;
    if (s + 1 < nsteps) gload(kbase + (s + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK / 16; ++kk) {
      const int kof = kk * 16 + 8 * (lane >> 5);
      const half8 a = *reinterpret_cast<const half8 *>(&As[buf][wave * 32 + (lane & 31)][kof]);
#pragma unroll
      for (int l = 0; l < L; ++l) {
        const half8 b = *reinterpret_cast<const half8 *>(&Bs[buf][l][lane & 31][kof]);
        acc[l] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[l], 0, 0, 0);
      }
    }
    // promote after every 128 rows of x1: exact integers -> FP64
    if (((s + 1) * BK) % 128 == 0) {
      const int et = texp[((kbase + s * BK) / 128) * ntile2 + n0 / 128];
      double sc[L];
#pragma unroll
      for (int l = 0; l < L; ++l) sc[l] = ldexp(1.0, et - bits * l);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        double t = 0.0;
#pragma unroll
        for (int l = 0; l < L; ++l) t += (double)acc[l][r] * sc[l];
        acc64[r] += t;
      }
#pragma unroll
      for (int l = 0; l < L; ++l) acc[l] = (float16_){0};
    }
    if (s + 1 < nsteps) {
      sstore(buf ^ 1);
      __syncthreads();
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + wave * 32 + mfma32_row(lane, r);
    const int col = n0 + (lane & 31);
    if (row < KLp) T[(int64_t)row * N2 + col] = acc64[r];
  }
}

// T[kappa][x2] += r * P[x1][x2] for the residuals of the rounded counts; one
// thread owns a column x2 and walks every residual, so updates never race.
__global__ void k_resid_g1(const int32_t *__restrict__ res, int nres, const double *__restrict__ P, int N2,
                           double *__restrict__ T) {
  const int x2 = blockIdx.x * blockDim.x + threadIdx.x;
  if (x2 >= N2) return;
  for (int e = 0; e < nres; ++e) {
    const int x1 = res[3 * e], kap = res[3 * e + 1], r = res[3 * e + 2];
    T[(int64_t)kap * N2 + x2] += (double)r * P[(int64_t)x1 * N2 + x2];
  }
}

// ----------------------------------------------------------------- G2
struct G2Resid {
  const uint8_t *blkA;     // N1/128: block of x1 rows has a residual
  const uint8_t *blkB;     // N2/128: block of x2 rows (of Bbar) has a residual
  const int32_t *rowptrA;  // N1 + 1
  const int32_t *entA;     // (kappa, r) pairs
  const int32_t *rowptrB;  // N2 + 1
  const int32_t *entB;     // (kappa', r) pairs
  const double *V64;       // KLp x N2
  const double *W64;       // N1 x KRp
  int KRp;
};

// S[x1][x2] = sum_kappa A16[x1][kappa] V[kappa][x2] + sum_kappa' W[x1][kappa'] B16[x2][kappa']
// with V, W in L limbs of one global scale.  Block 128 x 128 (one P tile),
// 8 waves of 32 x 64.  Blocks are mapped to tiles XCD by XCD: the 8 XCDs each
// sweep their own 8 x 8 patches of tiles so the operand rows a patch needs
// (about 3.75 MiB) stay in that XCD's L2.
// Epilogue: the BA update (or D), and with fuse_limbs the next step's P limbs
// and tile exponent, written transposed through LDS.
struct G2Out {
  _Float16 *limbT;  // next P limbs [L][N2][N1] (fused BA epilogue) or null
  int32_t *texp;    // tile exponents
  int limbs_p, bits_p;
  int patches;      // 1: XCD patch mapping
};
template <int L, int EPI>
__global__ void __launch_bounds__(512) k_split_g2(int N1, int N2, int KL, int KLp, int KR, int KRp,
                                                  const _Float16 *__restrict__ A16, const _Float16 *__restrict__ vT,
                                                  const _Float16 *__restrict__ wl, const _Float16 *__restrict__ B16,
                                                  int bits, EpiArgs ea, G2Resid rs, G2Out go) {
  constexpr int BM = 128, BN = 128, BK = 32, PADK = BK + 8;
  static_assert(BM == BN, "slabs are shared between the A and B roles");
  // L + 1 slabs of 128 x 32 per stage: phase 1 uses slab 0 for A16 and slabs
  // 1..L for the V limbs; phase 2 uses slabs 0..L-1 for the W limbs and slab L for B16.
  __shared__ __attribute__((aligned(16))) _Float16 Sl[2][L + 1][BM][PADK];
  __shared__ double red[48];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int TX = N2 / BN;
  int ty, tx;
  if (go.patches) {
    const int b = blockIdx.x, xcd = b & 7, sq = b >> 3;
    const int patch = (sq >> 6) * 8 + xcd, w = sq & 63;
    ty = (patch / (TX >> 3)) * 8 + (w >> 3);
    tx = (patch % (TX >> 3)) * 8 + (w & 7);
  } else {
    ty = blockIdx.x / TX;
    tx = blockIdx.x % TX;
  }
  const int m0 = ty * BM, n0 = tx * BN;

  float16_ acc[L][2];
#pragma unroll
  for (int l = 0; l < L; ++l)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[l][j] = (float16_){0};

  // one 16-byte chunk per thread covers a 128 x 32 slab
  auto load_slab = [&](const _Float16 *base, int64_t ld, int row0, int k0, int kmax) -> uint4 {
    const int row = tid >> 2, kq = tid & 3;
    const int kk = k0 + kq * 8;
    if (kk >= kmax) return make_uint4(0, 0, 0, 0);
    return *reinterpret_cast<const uint4 *>(base + (int64_t)(row0 + row) * ld + kk);
  };
  auto store_slab = [&](int buf, int slab, uint4 v) {
    const int row = tid >> 2, kq = tid & 3;
    *reinterpret_cast<uint4 *>(&Sl[buf][slab][row][kq * 8]) = v;
  };

  // ---- phase 1: A16 (shared) x V limbs
  {
    const int nst = (KL + BK - 1) / BK;
    uint4 r[L + 1];
    auto gl = [&](int k0) {
      r[0] = load_slab(A16, KLp, m0, k0, KLp);
#pragma unroll
      for (int l = 0; l < L; ++l) r[l + 1] = load_slab(vT + (int64_t)l * N2 * KLp, KLp, n0, k0, KLp);
    };
    auto st = [&](int buf) {
#pragma unroll
      for (int q = 0; q <= L; ++q) store_slab(buf, q, r[q]);
    };
    if (nst > 0) {
      gl(0);
      st(0);
      __syncthreads();
    }
    for (int s = 0; s < nst; ++s) {
      const int buf = s & 1;
      if (s + 1 < nst) gl((s + 1) * BK);
#pragma unroll
      for (int kk = 0; kk < BK / 16; ++kk) {
        const int kof = kk * 16 + 8 * (lane >> 5);
        const half8 a = *reinterpret_cast<const half8 *>(&Sl[buf][0][wm * 32 + (lane & 31)][kof]);
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int l = 0; l < L; ++l) {
            const half8 b = *reinterpret_cast<const half8 *>(&Sl[buf][l + 1][wn * 64 + j * 32 + (lane & 31)][kof]);
            acc[l][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[l][j], 0, 0, 0);
          }
      }
      if (s + 1 < nst) {
        st(buf ^ 1);
        __syncthreads();
      }
    }
    __syncthreads();
  }
  // ---- phase 2: W limbs x B16 (shared)
  if (KR > 0) {
    const int nst = (KR + BK - 1) / BK;
    uint4 r[L + 1];
    auto gl = [&](int k0) {
#pragma unroll
      for (int l = 0; l < L; ++l) r[l] = load_slab(wl + (int64_t)l * N1 * KRp, KRp, m0, k0, KRp);
      r[L] = load_slab(B16, KRp, n0, k0, KRp);
    };
    auto st = [&](int buf) {
#pragma unroll
      for (int q = 0; q <= L; ++q) store_slab(buf, q, r[q]);
    };
    gl(0);
    st(0);
    __syncthreads();
    for (int s = 0; s < nst; ++s) {
      const int buf = s & 1;
      if (s + 1 < nst) gl((s + 1) * BK);
#pragma unroll
      for (int kk = 0; kk < BK / 16; ++kk) {
        const int kof = kk * 16 + 8 * (lane >> 5);
        half8 b[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const half8 *>(&Sl[buf][L][wn * 64 + j * 32 + (lane & 31)][kof]);
#pragma unroll
        for (int l = 0; l < L; ++l) {
          const half8 a = *reinterpret_cast<const half8 *>(&Sl[buf][l][wm * 32 + (lane & 31)][kof]);
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[l][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b[j], acc[l][j], 0, 0, 0);
        }
      }
      if (s + 1 < nst) {
        st(buf ^ 1);
        __syncthreads();
      }
    }
    __syncthreads();
  }

  // ---- epilogue pass 1: combine limbs into S (FP64) and apply residuals
  const int ev = (int)ea.scal[kScalVScale];
  double sc[L];
#pragma unroll
  for (int l = 0; l < L; ++l) sc[l] = ldexp(1.0, ev - bits * l);
  double Sv[2][16];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      double S = 0.0;
#pragma unroll
      for (int l = 0; l < L; ++l) S += (double)acc[l][j][r] * sc[l];
      Sv[j][r] = S;
    }
  if (rs.blkA[ty] || rs.blkB[tx]) {
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int x1 = m0 + wm * 32 + mfma32_row(lane, r);
        const int x2 = n0 + wn * 64 + j * 32 + (lane & 31);
        for (int e = rs.rowptrA[x1]; e < rs.rowptrA[x1 + 1]; ++e)
          Sv[j][r] += (double)rs.entA[2 * e + 1] * rs.V64[(int64_t)rs.entA[2 * e] * N2 + x2];
        for (int e = rs.rowptrB[x2]; e < rs.rowptrB[x2 + 1]; ++e)
          Sv[j][r] += rs.W64[(int64_t)x1 * rs.KRp + rs.entB[2 * e]] * (double)rs.entB[2 * e + 1];
      }
  }
  // ---- epilogue pass 2: D, the update, block partials
  const double mu = ea.scal[kScalMu], dref = ea.scal[kScalDRef];
  double s0 = 0.0, m1 = -INFINITY, s2 = 0.0, wmax = 0.0;
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int x1 = m0 + wm * 32 + mfma32_row(lane, r);
      const int x2 = n0 + wn * 64 + j * 32 + (lane & 31);
      const int64_t x = (int64_t)x1 * ea.ldx + x2;
      const double D = ea.Hc[x] - mu - Sv[j][r] * ea.inv_c;
      m1 = fmax(m1, D);
      if constexpr (EPI == EPI_BA) {
        const double p = ea.P[x];
        const double w = p * exp(D - dref);
        ea.P[x] = w;
        s0 += w;
        s2 += p * D;
        Sv[j][r] = w;
        wmax = fmax(wmax, w);
      } else {
        ea.D[x] = D;
      }
    }
  block_reduce3(s0, m1, s2, red);
  if (tid == 0) {
    double *pp = ea.part + (int64_t)(ty * TX + tx) * kPartStrideDev;
    pp[0] = s0;
    pp[1] = m1;
    pp[2] = s2;
  }
  // ---- epilogue pass 3 (BA, fused): limbs of the new P tile for the next output law
  if constexpr (EPI == EPI_BA) {
    if (go.limbT) {
      wmax = block_max(wmax, red + 16);
      const int et = tile_exponent(wmax, go.bits_p);
      if (tid == 0) go.texp[ty * TX + tx] = et;
      constexpr int PITCH = 136;
      _Float16(*tileT)[PITCH] = reinterpret_cast<_Float16(*)[PITCH]>(&Sl[0][0][0][0]);  // 128 x 136, 34 KB
      const double radix = ldexp(1.0, go.bits_p);
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) Sv[j][r] = ldexp(Sv[j][r], -et);
      for (int l = 0; l < go.limbs_p; ++l) {
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int xl1 = wm * 32 + mfma32_row(lane, r);
            const int xl2 = wn * 64 + j * 32 + (lane & 31);
            const double li = rint(Sv[j][r]);
            tileT[xl2][xl1] = (_Float16)(float)li;
            Sv[j][r] = (Sv[j][r] - li) * radix;
          }
        __syncthreads();
        _Float16 *dst = go.limbT + (int64_t)l * N1 * N2;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int ch = tid + 512 * q, c2 = ch >> 4, g = ch & 15;
          *reinterpret_cast<uint4 *>(dst + (int64_t)(n0 + c2) * N1 + m0 + 8 * g) =
              *reinterpret_cast<const uint4 *>(&tileT[c2][8 * g]);
        }
        __syncthreads();
      }
    }
  }
}

// ----------------------------------------------------------------- launchers
#define BDC_SPLIT_L(L, call) \
  switch (L) {               \
    case 2: call(2); break;  \
    case 3: call(3); break;  \
    case 4: call(4); break;  \
    case 5: call(5); break;  \
    default: return BDC_ERR_ARG; \
  }

int launch_split_p(const double *P, const double *scal, int N1, int N2, int limbs, int bits, _Float16 *limbT,
                   int32_t *texp, hipStream_t s) {
  if (N1 % 128 || N2 % 128) return BDC_ERR_ARG;
  dim3 grid(N2 / 128, N1 / 128);
#define LP(LL) hipLaunchKernelGGL(k_split_p<LL>, grid, dim3(1024), 0, s, P, scal, N1, N2, bits, limbT, texp)
  BDC_SPLIT_L(limbs, LP)
#undef LP
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

int launch_split_vw(const double *V, int KL, int KLp, int N2, const double *W, int N1, int KR, int KRp, int limbs,
                    int bits, _Float16 *vT, _Float16 *wl, double *part, double *scal, hipStream_t s) {
  const int np = 512;
  hipLaunchKernelGGL(k_absmax_vw, dim3(np), dim3(256), 0, s, V, (int64_t)KL * N2, W, N1, KR, KRp, part);
  hipLaunchKernelGGL(k_vw_exponent, dim3(1), dim3(256), 0, s, part, np, bits, scal);
  if (N2 % 64 || KLp % 64) return BDC_ERR_ARG;
  dim3 gv(N2 / 64, KLp / 64);
#define LV(LL) hipLaunchKernelGGL(k_split_v<LL>, gv, dim3(256), 0, s, V, KL, KLp, N2, bits, scal, vT)
  BDC_SPLIT_L(limbs, LV)
#undef LV
  if (KR > 0) {
    const int64_t len = (int64_t)N1 * KRp;
#define LW(LL) hipLaunchKernelGGL(k_split_w<LL>, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, W, len, KR, KRp, bits, scal, wl)
    BDC_SPLIT_L(limbs, LW)
#undef LW
  }
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

int launch_split_g1(int KLp, int N1, int N2, const _Float16 *aT, const _Float16 *pT, const int32_t *texp, int limbs,
                    int bits, int ksplit, double *T, hipStream_t s) {
  if (N1 % 128 || N2 % 128 || ksplit < 1 || (N1 / 128) % ksplit) return BDC_ERR_ARG;
  const int krows = N1 / ksplit;
  dim3 grid(N2 / 32, (KLp + 255) / 256, ksplit);
#define LG(LL) hipLaunchKernelGGL(k_split_g1<LL>, grid, dim3(512), 0, s, KLp, N1, N2, aT, pT, texp, bits, krows, T)
  switch (limbs) {
    case 2: LG(2); break;
    case 3: LG(3); break;
    case 4: LG(4); break;
    default: return BDC_ERR_ARG;
  }
#undef LG
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

int launch_resid_g1(const int32_t *res, int nres, const double *P, int N2, double *T, hipStream_t s) {
  if (nres == 0) return BDC_OK;
  hipLaunchKernelGGL(k_resid_g1, dim3((unsigned)((N2 + 255) / 256)), dim3(256), 0, s, res, nres, P, N2, T);
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

int launch_split_g2(int N1, int N2, int KL, int KLp, int KR, int KRp, const _Float16 *A16, const _Float16 *vT,
                    const _Float16 *wl, const _Float16 *B16, int limbs, int bits, int epi, const EpiArgs &ea,
                    const uint8_t *blkA, const uint8_t *blkB, const int32_t *rowptrA, const int32_t *entA,
                    const int32_t *rowptrB, const int32_t *entB, const double *V64, const double *W64,
                    _Float16 *limbT_next, int32_t *texp_next, int limbs_p, int bits_p, hipStream_t s, int *nb) {
  if (N1 % 128 || N2 % 128) return BDC_ERR_ARG;
  G2Resid rs{blkA, blkB, rowptrA, entA, rowptrB, entB, V64, W64, KRp};
  const int TX = N2 / 128, TY = N1 / 128;
  G2Out go{limbT_next, texp_next, limbs_p, bits_p, (TX % 8 == 0 && TY % 8 == 0 && ((TX / 8) * (TY / 8)) % 8 == 0) ? 1 : 0};
  if (limbT_next && (limbs_p < 1 || limbs_p > 4)) return BDC_ERR_ARG;
  const int blocks = TX * TY;
  if (nb) *nb = blocks;
#define LG2(LL, EE) hipLaunchKernelGGL((k_split_g2<LL, EE>), dim3(blocks), dim3(512), 0, s, N1, N2, KL, KLp, KR, KRp, A16, vT, wl, B16, bits, ea, rs, go)
  if (epi == EPI_BA) {
    switch (limbs) {
      case 2: LG2(2, EPI_BA); break;
      case 3: LG2(3, EPI_BA); break;
      case 4: LG2(4, EPI_BA); break;
      default: return BDC_ERR_ARG;
    }
  } else if (epi == EPI_D) {
    switch (limbs) {
      case 2: LG2(2, EPI_D); break;
      case 3: LG2(3, EPI_D); break;
      case 4: LG2(4, EPI_D); break;
      default: return BDC_ERR_ARG;
    }
  } else {
    return BDC_ERR_ARG;
  }
#undef LG2
  return hipGetLastError() == hipSuccess ? BDC_OK : BDC_ERR_HIP;
}

}  // namespace bdc
