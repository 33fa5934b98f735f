// gemm.hip -- the complex FP64 matrix product on the matrix cores.
//
// dotProduct / dotProduct(dims) of two large operands (the reference's tiled
// shared-memory form is CuDotProductSharedProcedures.h:95-166): out(x, y) =
// sum_k A[y][k] B[k][x] over a W x H window, complex as four real products
// (Cr = Ar Br - Ai Bi, Ci = Ar Bi + Ai Br), each on v_mfma_f64_16x16x4_f64.
//
// Tiling for CDNA4: a 16-wave workgroup owns a 128 x 128 output tile, each
// wave a 32 x 32 part as 2 x 2 MFMA tiles with separate real and imaginary
// accumulators (32 doubles per lane), so four waves share each SIMD and hide
// one another's LDS and load waits (measured at 4096^3 complex on MI355X:
// 4 waves of 64 x 64 0.70 of the FP64 peak, 8 waves 0.80, 16 waves 0.84; K
// slices of 16 instead of 8 were slower).  K advances in slices of 8 staged
// through LDS, double-buffered: the next slice's global loads are in flight
// while the current one feeds the matrix cores.  LDS keeps A as [k][row] and
// B as [k][col], so the 16 lanes that share one k of an MFMA operand read 16
// consecutive doubles.
//
// Operand layout of v_mfma_f64_16x16x4_f64 (lane l): A(l % 16, l / 16),
// B(l / 16, l % 16); result C(l / 16 + 4 i, l % 16), i = 0..3.
#include "internal.h"

namespace oaph {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 8;
constexpr int kPad = 2;  // LDS row padding (doubles) against bank conflicts of the transposed A stores
constexpr int WM = 4, WN = 4;          // waves along rows and columns of the tile
constexpr int kThreads = 64 * WM * WN;
constexpr int TI = BM / WM / 16, TJ = BN / WN / 16;  // 16 x 16 MFMA tiles per wave

// MODE 0: complex; 1: real part only (Ar Br); 2: imaginary plane only with the
// dims semantics of the reference (-Ai Bi).
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_zgemm_mfma(
    uint32_t W, uint32_t H, uint32_t K, const double* __restrict__ are,
    const double* __restrict__ aim, uint32_t lda, const double* __restrict__ bre,
    const double* __restrict__ bim, uint32_t ldb, double* __restrict__ ore,
    double* __restrict__ oim, uint32_t ldo) {
  constexpr bool RE = MODE != 2, IM = MODE != 1;  // which operand planes are read
  __shared__ double sA[2][2][BK][BM + kPad];  // [buffer][plane][k][row]
  __shared__ double sB[2][2][BK][BN + kPad];  // [buffer][plane][k][col]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t x0 = blockIdx.x * BN, y0 = blockIdx.y * BM;
  const int wr = (wave / WN) * (16 * TI), wc = (wave % WN) * (16 * TJ);  // the wave's part

  // global -> register staging of one K slice: A BM rows x BK k (thread: one
  // row, Q consecutive k), B BK k x BN columns (thread: one k, Q consecutive
  // columns)
  constexpr int Q = BM * BK / kThreads;
  const int ar = tid / (BK / Q), ak = (tid % (BK / Q)) * Q;
  const int bk = tid / (BN / Q), bc = (tid % (BN / Q)) * Q;
  double ra[2][Q], rb[2][Q];
  auto load_slice = [&](uint32_t k0) {
    const uint32_t gy = y0 + ar, gx = x0 + bc, gkb = k0 + bk;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const uint32_t gk = k0 + ak + q;
      const bool oka = gy < H && gk < K;
      const size_t ia = (size_t)gy * lda + gk;
      // an absent plane reads as zero (a real left operand times a complex right one)
      ra[0][q] = RE && oka && are ? are[ia] : 0.0;
      ra[1][q] = IM && oka && aim ? aim[ia] : 0.0;
      const bool okb = gkb < K && gx + q < W;
      const size_t ib = (size_t)gkb * ldb + gx + q;
      rb[0][q] = RE && okb && bre ? bre[ib] : 0.0;
      rb[1][q] = IM && okb && bim ? bim[ib] : 0.0;
    }
  };
  auto store_slice = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        sA[buf][p][ak + q][ar] = ra[p][q];
        sB[buf][p][bk][bc + q] = rb[p][q];
      }
    }
  };

  v4d cr[TI][TJ], ci[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      cr[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
      ci[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  const uint32_t nslices = (K + BK - 1) / BK;
  load_slice(0);
  store_slice(0);
  __syncthreads();
  const int fr = lane & 15, fk = lane >> 4;  // operand position of this lane
  for (uint32_t t = 0; t < nslices; ++t) {
    const int buf = t & 1;
    if (t + 1 < nslices) load_slice((t + 1) * BK);  // in flight during the products
#pragma unroll
    for (int kq = 0; kq < BK / 4; ++kq) {
      const int kk = kq * 4 + fk;
      double a_re[TI], a_im[TI], b_re[TJ], b_im[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) {
        a_re[i] = sA[buf][0][kk][wr + 16 * i + fr];
        a_im[i] = sA[buf][1][kk][wr + 16 * i + fr];
      }
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        b_re[j] = sB[buf][0][kk][wc + 16 * j + fr];
        b_im[j] = sB[buf][1][kk][wc + 16 * j + fr];
      }
#pragma unroll
      for (int i = 0; i < TI; ++i) {
        const double na_im = -a_im[i];
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
          if (MODE != 2) cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_re[i], b_re[j], cr[i][j], 0, 0, 0);
          if (MODE == 0) cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(na_im, b_im[j], cr[i][j], 0, 0, 0);
          if (MODE == 0) {
            ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_re[i], b_im[j], ci[i][j], 0, 0, 0);
            ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_im[i], b_re[j], ci[i][j], 0, 0, 0);
          }
          if (MODE == 2) ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(na_im, b_im[j], ci[i][j], 0, 0, 0);
        }
      }
    }
    if (t + 1 < nslices) store_slice(buf ^ 1);  // the other buffer: last read one slice ago
    __syncthreads();
  }

  // C(l / 16 + 4 e, l % 16) of each 16 x 16 tile
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t gy = y0 + wr + 16 * i + fk + 4 * e, gx = x0 + wc + 16 * j + fr;
        if (gy < H && gx < W) {
          const size_t o = (size_t)gy * ldo + gx;
          if (MODE != 2) ore[o] = cr[i][j][e];
          if (MODE != 1) oim[o] = ci[i][j][e];
        }
      }
}

}  // namespace

bool gemm_mfma_eligible(uint32_t W, uint32_t H, uint32_t K) {
  return W >= 256 && H >= 256 && K >= 64;
}

oaph_status launch_gemm_mfma(hipStream_t s, int mode, uint32_t W, uint32_t H, uint32_t K,
                             const double* are, const double* aim, uint32_t lda,
                             const double* bre, const double* bim, uint32_t ldb, double* ore,
                             double* oim, uint32_t ldo) {
  const dim3 grid((W + BN - 1) / BN, (H + BM - 1) / BM);
  OAPH_CHECK(grid.y <= 65535, OAPH_ERR_UNSUPPORTED);
  if (mode == 0)
    hipLaunchKernelGGL(k_zgemm_mfma<0>, grid, dim3(kThreads), 0, s, W, H, K, are, aim, lda, bre,
                       bim, ldb, ore, oim, ldo);
  else if (mode == 1)
    hipLaunchKernelGGL(k_zgemm_mfma<1>, grid, dim3(kThreads), 0, s, W, H, K, are, nullptr, lda,
                       bre, nullptr, ldb, ore, nullptr, ldo);
  else
    hipLaunchKernelGGL(k_zgemm_mfma<2>, grid, dim3(kThreads), 0, s, W, H, K, nullptr, aim, lda,
                       nullptr, bim, ldb, nullptr, oim, ldo);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

}  // namespace oaph
