// gemv.hip -- complex dotProduct on gfx950: the streaming matvec (the hot
// path: the user matvec of the eigensolver and dotProduct with a one-column
// right operand) and a tiled general product for everything else.
//
// Reference: CUDAKernel_DotProduct -> CUDA_specific_dotProduct
// (oapMatrixCuda/CuProcedures/CuDotProductSpecificProcedures.h:64-141), one
// thread per output that walks a whole row of A, so the 32 lanes of a block
// read addresses n*8 bytes apart.  Here a row of A (row-major plane, element
// (c, r) at c + r*columns) is contiguous, so a wave64 streams R rows with
// 16-byte loads per lane (64 lanes x 16 B = 1 KiB per row per plane per
// step), shares each x chunk across its R rows, and reduces the R complex
// partial sums across lanes at the end.  Algorithmic bytes per call:
// 16*rows*cols (A) + 16*cols (x) + 16*rows (y) for a complex FP64 operand.
#include "internal.h"

namespace oaph {

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// The projection epilogue (MvProj; ncols <= 64): lane c sums conj(V[c][r]) y[r]
// over the rows r = wave, wave + 4, ... of the block, and the four waves' sums
// are added in wave order.  The V entries are loaded into registers when the
// FP32-operand kernel starts (project_prefetch), so their latency hides
// behind the operator stream; the block's rows of y are in ys.
template <int RB>
struct ProjRegs {
  double a[RB / 4], b[RB / 4];
};

template <int RB>
__device__ __forceinline__ void project_prefetch(const MvProj& pj, uint32_t rows,
                                                 ProjRegs<RB>& v) {
  const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
  const uint32_t rb = blockIdx.x * RB;
  const bool col = (uint32_t)lane < pj.ncols;
  const double* vr = pj.Vre + (size_t)(col ? lane : 0) * pj.ldv + rb;
  const double* vi = pj.Vim + (size_t)(col ? lane : 0) * pj.ldv + rb;
#pragma unroll
  for (int i = 0; i < RB / 4; ++i) {
    const uint32_t r = wib + 4 * i;
    const bool in = col && rb + r < rows;
    v.a[i] = in ? vr[r] : 0.0;
    v.b[i] = in ? vi[r] : 0.0;
  }
}

template <int RB>
__device__ void block_project(const MvProj& pj, uint32_t rows, const double2* ys,
                              const ProjRegs<RB>& v) {
  __shared__ double2 acc[4][64];
  const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
  const uint32_t rb = blockIdx.x * RB;
  double sr = 0.0, si = 0.0;
#pragma unroll
  for (int i = 0; i < RB / 4; ++i) {
    const uint32_t r = wib + 4 * i;
    if (rb + r < rows) {
      const double2 y = ys[r];
      sr = fma(v.a[i], y.x, fma(v.b[i], y.y, sr));
      si = fma(v.a[i], y.y, fma(-v.b[i], y.x, si));
    }
  }
  acc[wib][lane] = make_double2(sr, si);
  __syncthreads();
  if (wib == 0 && (uint32_t)lane < pj.ncols) {
    const double2 p0 = acc[0][lane], p1 = acc[1][lane], p2 = acc[2][lane], p3 = acc[3][lane];
    pj.part[(size_t)blockIdx.x * pj.ldp + lane] =
        make_double2((p0.x + p1.x) + (p2.x + p3.x), (p0.y + p1.y) + (p2.y + p3.y));
  }
}

// ---------------------------------------------------------------- complex FP64 GEMV
// y[r] = sum_c A[r][c] x[c] for r < rows.  Requires cols even and 16-byte
// aligned rows (lda even); the launcher falls back to k_zgemv_scalar otherwise.
// G groups of R rows per wave, one group after another.
template <int R, int U, int MODE, int G = 1>
__global__ __launch_bounds__(256) void k_zgemv(uint32_t rows, uint32_t cols, uint32_t lda,
                                               const double* __restrict__ are,
                                               const double* __restrict__ aim,
                                               const double* __restrict__ xre,
                                               const double* __restrict__ xim,
                                               double* __restrict__ yre,
                                               double* __restrict__ yim, uint32_t rotate,
                                               MvProj pj) {
  __shared__ double2 ys[4 * G * R];  // the block's rows of y (projection epilogue)
  const int lane = threadIdx.x & 63;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  constexpr bool CPLX = MODE == 0;      // complex A (and x)
  constexpr bool XIM = MODE != 1;       // x has an imaginary plane
  const uint32_t npair = cols >> 1;
  for (int g = 0; g < G; ++g) {
  const uint32_t r0 = (wave * G + g) * R;
  if (r0 >= rows) break;
  double sr[R], si[R];
#pragma unroll
  for (int q = 0; q < R; ++q) { sr[q] = 0.0; si[q] = 0.0; }
  const v2d* xr2 = reinterpret_cast<const v2d*>(xre);
  const v2d* xi2 = reinterpret_cast<const v2d*>(xim);
  const v2d* ar2[R];
  const v2d* ai2[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    uint32_t r = r0 + q < rows ? r0 + q : rows - 1;  // clamp; results of clamped rows dropped
    ar2[q] = reinterpret_cast<const v2d*>(are + (size_t)r * lda);
    ai2[q] = CPLX ? reinterpret_cast<const v2d*>(aim + (size_t)r * lda) : nullptr;
  }
  // main loop: nch full chunks of U steps of 64 pairs, all loads of a chunk
  // issued first; with `rotate` each wave starts at chunk (wave mod nch) and
  // wraps, so that waves in step read different column ranges (the same
  // offsets in every row at once concentrate the traffic on few channels)
  const uint32_t nch = npair / (64 * U);
  uint32_t c = rotate && nch ? wave % nch : 0;
  for (uint32_t it = 0; it < nch; ++it, c = c + 1 == nch ? 0 : c + 1) {
    const uint32_t p = c * 64 * U + lane;
    v2d xa[U], xb[U];
    v2d a[R][U], b[R][U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      xa[u] = xr2[p + 64 * u];
      if (XIM) xb[u] = xi2[p + 64 * u];
    }
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        a[q][u] = __builtin_nontemporal_load(ar2[q] + p + 64 * u);
        if (CPLX) b[q][u] = __builtin_nontemporal_load(ai2[q] + p + 64 * u);
      }
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        sr[q] = fma(a[q][u].x, xa[u].x, sr[q]);
        sr[q] = fma(a[q][u].y, xa[u].y, sr[q]);
        if (CPLX) {
          sr[q] = fma(-b[q][u].x, xb[u].x, sr[q]);
          sr[q] = fma(-b[q][u].y, xb[u].y, sr[q]);
          si[q] = fma(b[q][u].x, xa[u].x, si[q]);
          si[q] = fma(b[q][u].y, xa[u].y, si[q]);
        }
        if (XIM) {
          si[q] = fma(a[q][u].x, xb[u].x, si[q]);
          si[q] = fma(a[q][u].y, xb[u].y, si[q]);
        }
      }
  }
  for (uint32_t p = nch * 64 * U + lane; p < npair; p += 64) {
    v2d xa = xr2[p], xb = XIM ? xi2[p] : v2d{0, 0};
#pragma unroll
    for (int q = 0; q < R; ++q) {
      v2d a = __builtin_nontemporal_load(ar2[q] + p);
      sr[q] = fma(a.x, xa.x, sr[q]);
      sr[q] = fma(a.y, xa.y, sr[q]);
      if (XIM) {
        si[q] = fma(a.x, xb.x, si[q]);
        si[q] = fma(a.y, xb.y, si[q]);
      }
      if (CPLX) {
        v2d b = __builtin_nontemporal_load(ai2[q] + p);
        sr[q] = fma(-b.x, xb.x, sr[q]);
        sr[q] = fma(-b.y, xb.y, sr[q]);
        si[q] = fma(b.x, xa.x, si[q]);
        si[q] = fma(b.y, xa.y, si[q]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    sr[q] = wave_sum(sr[q]);
    if (XIM) si[q] = wave_sum(si[q]);
  }
  if (lane < R && r0 + lane < rows) {
    double vr = sr[0], vi = si[0];
#pragma unroll
    for (int q = 1; q < R; ++q)
      if (lane == q) { vr = sr[q]; vi = si[q]; }
    yre[r0 + lane] = vr;
    if (XIM) yim[r0 + lane] = vi;
    ys[((threadIdx.x >> 6) * G + g) * R + lane] = make_double2(vr, vi);
  }
  }
  if constexpr (XIM) {
    if (pj.part) {  // uniform over the grid
      // V loaded here, not at the start: registers held across the FP64
      // stream slowed it (171 vs 163 us per launch at n = 8192)
      ProjRegs<4 * G * R> pv;
      project_prefetch(pj, rows, pv);
      __syncthreads();
      block_project<4 * G * R>(pj, rows, ys, pv);
    }
  }
}

// Y[:, t] = A X[:, t] for t < K in one pass over A (the mixed solver's FP64
// check of its K wanted Ritz vectors: one read of the operator instead of
// K).  A wave takes R rows, a lane one 16-byte pair of each row per step and
// the same pair of every x vector (x from L2: the K vectors are a few MiB);
// the R K sums are reduced across the lanes at the end.  MODE 0: complex A;
// 2: real A (complex x and y).  Requires cols even and 16-byte aligned rows
// and vectors.
template <int R, int K, int MODE>
__global__ __launch_bounds__(256) void k_zgemv_multi(uint32_t rows, uint32_t cols, uint32_t lda,
                                                     const double* __restrict__ are,
                                                     const double* __restrict__ aim,
                                                     const double* __restrict__ xre,
                                                     const double* __restrict__ xim,
                                                     uint32_t ldx, uint32_t nvec,
                                                     double* __restrict__ yre,
                                                     double* __restrict__ yim, uint32_t ldy) {
  constexpr bool CPLX = MODE == 0;
  const int lane = threadIdx.x & 63;
  const uint32_t r0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * R;
  if (r0 >= rows) return;  // no barrier below
  double sr[R][K], si[R][K];
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int t = 0; t < K; ++t) { sr[q][t] = 0.0; si[q][t] = 0.0; }
  const v2d* ar2[R];
  const v2d* ai2[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const uint32_t r = r0 + q < rows ? r0 + q : rows - 1;  // clamped rows are dropped
    ar2[q] = reinterpret_cast<const v2d*>(are + (size_t)r * lda);
    ai2[q] = CPLX ? reinterpret_cast<const v2d*>(aim + (size_t)r * lda) : nullptr;
  }
  const uint32_t npair = cols >> 1;
  for (uint32_t p = lane; p < npair; p += 64) {
    v2d a[R], b[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      a[q] = __builtin_nontemporal_load(ar2[q] + p);
      if (CPLX) b[q] = __builtin_nontemporal_load(ai2[q] + p);
    }
#pragma unroll
    for (int t = 0; t < K; ++t) {
      if ((uint32_t)t < nvec) {  // uniform
        const v2d xa = reinterpret_cast<const v2d*>(xre + (size_t)t * ldx)[p];
        const v2d xb = reinterpret_cast<const v2d*>(xim + (size_t)t * ldx)[p];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          sr[q][t] = fma(a[q].x, xa.x, sr[q][t]);
          sr[q][t] = fma(a[q].y, xa.y, sr[q][t]);
          si[q][t] = fma(a[q].x, xb.x, si[q][t]);
          si[q][t] = fma(a[q].y, xb.y, si[q][t]);
          if (CPLX) {
            sr[q][t] = fma(-b[q].x, xb.x, sr[q][t]);
            sr[q][t] = fma(-b[q].y, xb.y, sr[q][t]);
            si[q][t] = fma(b[q].x, xa.x, si[q][t]);
            si[q][t] = fma(b[q].y, xa.y, si[q][t]);
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int t = 0; t < K; ++t) {
      if ((uint32_t)t < nvec) {
        const double vr = wave_sum(sr[q][t]), vi = wave_sum(si[q][t]);
        if (lane == 0 && r0 + q < rows) {
          yre[(size_t)t * ldy + r0 + q] = vr;
          yim[(size_t)t * ldy + r0 + q] = vi;
        }
      }
    }
}

// any alignment: one wave per row, scalar loads
template <int MODE>
__global__ __launch_bounds__(256) void k_zgemv_scalar(uint32_t rows, uint32_t cols, uint32_t lda,
                                                      const double* __restrict__ are,
                                                      const double* __restrict__ aim,
                                                      const double* __restrict__ xre,
                                                      const double* __restrict__ xim,
                                                      double* __restrict__ yre,
                                                      double* __restrict__ yim) {
  const int lane = threadIdx.x & 63;
  const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= rows) return;
  constexpr bool CPLX = MODE == 0, XIM = MODE != 1;
  double sr = 0, si = 0;
  const double* ar = are + (size_t)r * lda;
  const double* ai = CPLX ? aim + (size_t)r * lda : nullptr;
  for (uint32_t c = lane; c < cols; c += 64) {
    double a = ar[c], xa = xre[c];
    sr = fma(a, xa, sr);
    if (XIM) si = fma(a, xim[c], si);
    if (CPLX) {
      double b = ai[c], xb = xim[c];
      sr = fma(-b, xb, sr);
      si = fma(b, xa, si);
    }
  }
  sr = wave_sum(sr);
  if (XIM) si = wave_sum(si);
  if (lane == 0) {
    yre[r] = sr;
    if (XIM) yim[r] = si;
  }
}

// ---------------------------------------------------------------- reduced-precision-operand GEMV
// A stored in a reduced-precision copy (the mixed-precision solver's FP32
// copy of the operator), x and y FP64, products and sums in FP64: the only
// rounding beyond FP64 is the storage of A.  One load per lane brings 4
// consecutive row elements (16 B of FP32); OP::at widens element i of a load
// to FP64 exactly.
typedef float v4f __attribute__((ext_vector_type(4)));
struct OpF32 {
  using elem = float;
  using vec = v4f;
  static __device__ __forceinline__ double at(const vec& v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
  }
};
// CPLX = false: real A (aim unused), complex x -- half the operator bytes
// G groups of R rows per wave, one group after another.
template <class OP, int R, int U, bool CPLX, int G>
__device__ __forceinline__ void lowp_gemv(uint32_t rows, uint32_t cols, uint32_t lda,
                                          const typename OP::elem* __restrict__ are,
                                          const typename OP::elem* __restrict__ aim,
                                          const double* __restrict__ xre,
                                          const double* __restrict__ xim,
                                          double* __restrict__ yre, double* __restrict__ yim,
                                          uint32_t rotate, const MvProj& pj, double2* ys) {
  using vec = typename OP::vec;
  const int lane = threadIdx.x & 63;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t nquad = cols >> 2;
  ProjRegs<4 * G * R> pv;
  if (pj.part) project_prefetch(pj, rows, pv);
  for (int g = 0; g < G; ++g) {
  const uint32_t r0 = (wave * G + g) * R;
  if (r0 >= rows) break;
  double sr[R], si[R];
#pragma unroll
  for (int q = 0; q < R; ++q) { sr[q] = 0.0; si[q] = 0.0; }
  const v2d* xr2 = reinterpret_cast<const v2d*>(xre);
  const v2d* xi2 = reinterpret_cast<const v2d*>(xim);
  const vec* ar4[R];
  const vec* ai4[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    uint32_t r = r0 + q < rows ? r0 + q : rows - 1;
    ar4[q] = reinterpret_cast<const vec*>(are + (size_t)r * lda);
    ai4[q] = CPLX ? reinterpret_cast<const vec*>(aim + (size_t)r * lda) : nullptr;
  }
  // with `rotate` the wave walks its rows from quad `off` (a multiple of 64
  // that differs between neighbouring waves) and wraps at the end, so that
  // waves in step read different column ranges
  const uint32_t nq64 = nquad / 64;
  const uint32_t off = rotate && nq64 ? (wave * U % nq64) * 64 : 0;
  for (uint32_t p = lane; p < nquad; p += 64 * U) {
    v2d xa[U][2], xb[U][2];
    vec a[R][U], b[R][U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uint32_t pp = p + 64 * u;
      if (pp < nquad) {
        const uint32_t wq = pp + off < nquad ? pp + off : pp + off - nquad;
        xa[u][0] = xr2[2 * wq]; xa[u][1] = xr2[2 * wq + 1];
        xb[u][0] = xi2[2 * wq]; xb[u][1] = xi2[2 * wq + 1];
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        uint32_t pp = p + 64 * u;
        if (pp < nquad) {
          const uint32_t wq = pp + off < nquad ? pp + off : pp + off - nquad;
          a[q][u] = __builtin_nontemporal_load(ar4[q] + wq);
          if (CPLX) b[q][u] = __builtin_nontemporal_load(ai4[q] + wq);
        }
      }
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (p + 64 * u < nquad) {
          double a0 = OP::at(a[q][u], 0), a1 = OP::at(a[q][u], 1), a2 = OP::at(a[q][u], 2),
                 a3 = OP::at(a[q][u], 3);
          sr[q] = fma(a0, xa[u][0].x, sr[q]); sr[q] = fma(a1, xa[u][0].y, sr[q]);
          sr[q] = fma(a2, xa[u][1].x, sr[q]); sr[q] = fma(a3, xa[u][1].y, sr[q]);
          if (CPLX) {
            double b0 = OP::at(b[q][u], 0), b1 = OP::at(b[q][u], 1), b2 = OP::at(b[q][u], 2),
                   b3 = OP::at(b[q][u], 3);
            sr[q] = fma(-b0, xb[u][0].x, sr[q]); sr[q] = fma(-b1, xb[u][0].y, sr[q]);
            sr[q] = fma(-b2, xb[u][1].x, sr[q]); sr[q] = fma(-b3, xb[u][1].y, sr[q]);
          }
          si[q] = fma(a0, xb[u][0].x, si[q]); si[q] = fma(a1, xb[u][0].y, si[q]);
          si[q] = fma(a2, xb[u][1].x, si[q]); si[q] = fma(a3, xb[u][1].y, si[q]);
          if (CPLX) {
            double b0 = OP::at(b[q][u], 0), b1 = OP::at(b[q][u], 1), b2 = OP::at(b[q][u], 2),
                   b3 = OP::at(b[q][u], 3);
            si[q] = fma(b0, xa[u][0].x, si[q]); si[q] = fma(b1, xa[u][0].y, si[q]);
            si[q] = fma(b2, xa[u][1].x, si[q]); si[q] = fma(b3, xa[u][1].y, si[q]);
          }
        }
      }
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    sr[q] = wave_sum(sr[q]);
    si[q] = wave_sum(si[q]);
  }
  if (lane < R && r0 + lane < rows) {
    double vr = sr[0], vi = si[0];
#pragma unroll
    for (int q = 1; q < R; ++q)
      if (lane == q) { vr = sr[q]; vi = si[q]; }
    yre[r0 + lane] = vr;
    yim[r0 + lane] = vi;
    ys[((threadIdx.x >> 6) * G + g) * R + lane] = make_double2(vr, vi);
  }
  }
  if (pj.part) {  // uniform over the grid
    __syncthreads();
    block_project<4 * G * R>(pj, rows, ys, pv);
  }
}

// FP32 operand (the mixed-precision solver's FP32 copy)
template <int R, int U, bool CPLX = true, int G = 1>
__global__ __launch_bounds__(256) void k_cgemv32(uint32_t rows, uint32_t cols, uint32_t lda,
                                                 const float* __restrict__ are,
                                                 const float* __restrict__ aim,
                                                 const double* __restrict__ xre,
                                                 const double* __restrict__ xim,
                                                 double* __restrict__ yre,
                                                 double* __restrict__ yim, uint32_t rotate,
                                                 MvProj pj) {
  __shared__ double2 ys[4 * G * R];  // the block's rows of y (projection epilogue)
  lowp_gemv<OpF32, R, U, CPLX, G>(rows, cols, lda, are, aim, xre, xim, yre, yim, rotate, pj, ys);
}

// ---------------------------------------------------------------- FP16-operand GEMV
// A stored as FP16 values of 2^s A (the mixed-precision solver's 16-bit copy,
// 2 bytes per value; the power of two keeps the largest entry below 2^15),
// x rounded to FP32 as it is loaded, products and per-lane sums in FP32 --
// v_fma_mix_f32 takes the FP16 halves of a load as they are, one instruction
// per product -- then the cross-lane sums in FP64 and y = 2^-s times them.
// 8 elements (16 B) per lane per load, so rows need cols % 8 == 0.
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
// V octets (8 columns each) of R rows at octet index p, p + 64, ...: the A
// loads (hgemv_load) and the products (hgemv_fma, x read at p: XL FP32 x
// staged in LDS, else FP64 x from global memory converted on arrival).  No
// octet index may be out of range (no guards, so no branch splits the loads).
typedef float f4v __attribute__((ext_vector_type(4)));
template <int R, int V, bool CPLX>
struct HOct {
  h8 a[R][V], b[R][V];
};
template <int R, int V, bool CPLX>
__device__ __forceinline__ void hgemv_load(HOct<R, V, CPLX>& o, uint32_t p,
                                           const h8* const (&ar8)[R], const h8* const (&ai8)[R]) {
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int u = 0; u < V; ++u) {
      o.a[q][u] = __builtin_nontemporal_load(ar8[q] + p + 64 * u);
      if (CPLX) o.b[q][u] = __builtin_nontemporal_load(ai8[q] + p + 64 * u);
    }
}
template <int R, int V, bool CPLX, bool XL>
__device__ __forceinline__ void hgemv_fma(const HOct<R, V, CPLX>& o, uint32_t p, const v2d* xr2,
                                          const v2d* xi2, const float* xs, uint32_t xsim,
                                          float (&sr)[R], float (&si)[R]) {
  float xa[V][8], xb[V][8];
  if (XL) {
#pragma unroll
    for (int u = 0; u < V; ++u)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const f4v r4 = *reinterpret_cast<const f4v*>(xs + 8 * (p + 64 * u) + 4 * h);
        const f4v i4 = *reinterpret_cast<const f4v*>(xs + xsim + 8 * (p + 64 * u) + 4 * h);
        xa[u][4 * h] = r4.x; xa[u][4 * h + 1] = r4.y; xa[u][4 * h + 2] = r4.z; xa[u][4 * h + 3] = r4.w;
        xb[u][4 * h] = i4.x; xb[u][4 * h + 1] = i4.y; xb[u][4 * h + 2] = i4.z; xb[u][4 * h + 3] = i4.w;
      }
  } else {
    v2d tr[V][4], ti[V][4];
#pragma unroll
    for (int u = 0; u < V; ++u)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        tr[u][e] = xr2[4 * (p + 64 * u) + e];
        ti[u][e] = xi2[4 * (p + 64 * u) + e];
      }
#pragma unroll
    for (int u = 0; u < V; ++u)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        xa[u][2 * e] = (float)tr[u][e].x; xa[u][2 * e + 1] = (float)tr[u][e].y;
        xb[u][2 * e] = (float)ti[u][e].x; xb[u][2 * e + 1] = (float)ti[u][e].y;
      }
  }
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int u = 0; u < V; ++u)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float av = (float)o.a[q][u][e];
        sr[q] = __builtin_fmaf(av, xa[u][e], sr[q]);
        si[q] = __builtin_fmaf(av, xb[u][e], si[q]);
        if (CPLX) {
          const float bv = (float)o.b[q][u][e];
          sr[q] = __builtin_fmaf(-bv, xb[u][e], sr[q]);
          si[q] = __builtin_fmaf(bv, xa[u][e], si[q]);
        }
      }
}

// x staged in LDS as FP32 when the row length is at most kXLdsCols (64 KiB
// per workgroup): one coalesced pass over x per workgroup instead of every
// wave reading FP64 x from L2 and converting it each step
constexpr uint32_t kXLdsCols = 8192;
template <int R, int U, bool CPLX = true, int G = 1, bool XL = false>
__global__ __launch_bounds__(256) void k_hgemv16(uint32_t rows, uint32_t cols, uint32_t lda,
                                                 const _Float16* __restrict__ are,
                                                 const _Float16* __restrict__ aim,
                                                 const double* __restrict__ xre,
                                                 const double* __restrict__ xim,
                                                 double* __restrict__ yre,
                                                 double* __restrict__ yim, double unscale,
                                                 uint32_t rotate, MvProj pj) {
  __shared__ double2 ys[4 * G * R];  // the block's rows of y (projection epilogue)
  __shared__ float xs[XL ? 2 * kXLdsCols : 1];
  const int lane = threadIdx.x & 63;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t noct = cols >> 3;
  ProjRegs<4 * G * R> pv;
  if (pj.part) project_prefetch(pj, rows, pv);
  const v2d* xr2 = reinterpret_cast<const v2d*>(xre);
  const v2d* xi2 = reinterpret_cast<const v2d*>(xim);
  if (XL) {
    for (uint32_t j = threadIdx.x; j < cols; j += 256) {
      xs[j] = (float)xre[j];
      xs[kXLdsCols + j] = (float)xim[j];
    }
    __syncthreads();
  }
  for (int g = 0; g < G; ++g) {
    const uint32_t r0 = (wave * G + g) * R;
    if (r0 >= rows) break;
    float sr[R], si[R];
#pragma unroll
    for (int q = 0; q < R; ++q) { sr[q] = 0.0f; si[q] = 0.0f; }
    const h8* ar8[R];
    const h8* ai8[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      uint32_t r = r0 + q < rows ? r0 + q : rows - 1;
      ar8[q] = reinterpret_cast<const h8*>(are + (size_t)r * lda);
      ai8[q] = CPLX ? reinterpret_cast<const h8*>(aim + (size_t)r * lda) : ar8[q];
    }
    // software pipeline: the next U octets' loads are in flight while the
    // current ones are multiplied (one wave per SIMD at 8192 rows, so no
    // other wave hides the load latency).  The row is walked in nch chunks
    // of kStep octets; with `rotate` each wave starts at chunk (wave mod nch)
    // and wraps, so that waves in step read different column ranges
    constexpr uint32_t kStep = 64 * U;
    const uint32_t nch = noct / kStep;
    if (nch) {
      uint32_t c = rotate ? wave % nch : 0;
      HOct<R, U, CPLX> cur;
      hgemv_load(cur, c * kStep + lane, ar8, ai8);
      for (uint32_t i = 1; i < nch; ++i) {
        const uint32_t cn = c + 1 == nch ? 0 : c + 1;
        HOct<R, U, CPLX> nxt;
        hgemv_load(nxt, cn * kStep + lane, ar8, ai8);
        hgemv_fma<R, U, CPLX, XL>(cur, c * kStep + lane, xr2, xi2, xs, kXLdsCols, sr, si);
        cur = nxt;
        c = cn;
      }
      hgemv_fma<R, U, CPLX, XL>(cur, c * kStep + lane, xr2, xi2, xs, kXLdsCols, sr, si);
    }
    for (uint32_t p = nch * kStep + lane; p < noct; p += 64) {
      HOct<R, 1, CPLX> one;
      hgemv_load(one, p, ar8, ai8);
      hgemv_fma<R, 1, CPLX, XL>(one, p, xr2, xi2, xs, kXLdsCols, sr, si);
    }
    double dr[R], di[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      dr[q] = wave_sum((double)sr[q]) * unscale;
      di[q] = wave_sum((double)si[q]) * unscale;
    }
    if (lane < R && r0 + lane < rows) {
      double vr = dr[0], vi = di[0];
#pragma unroll
      for (int q = 1; q < R; ++q)
        if (lane == q) { vr = dr[q]; vi = di[q]; }
      yre[r0 + lane] = vr;
      yim[r0 + lane] = vi;
      ys[((threadIdx.x >> 6) * G + g) * R + lane] = make_double2(vr, vi);
    }
  }
  if (pj.part) {  // uniform over the grid
    __syncthreads();
    block_project<4 * G * R>(pj, rows, ys, pv);
  }
}

// ---------------------------------------------------------------- general complex product
// out(x, y) = sum_k A[y][k] B[k][x] over a W x H window, K terms, with the
// matrices' own row strides (covers dotProduct and dotProduct(dims),
// CuDotProductDimProcedures.h:30-48).  64x64 output tile per 256 threads
// (4x4 outputs per thread), K staged through LDS in slices of 16.
// MODE 0: complex, 1: re only, 2: im only (dims semantics: im = -sum a_im b_im).
// TB: the right operand is read transposed, B[k][x] = b[x * ldb + k] (A * B^T
// without forming B^T).
constexpr int GT = 64, GK = 16;
template <int MODE, bool TB = false>
__global__ __launch_bounds__(256) void k_gemm(uint32_t W, uint32_t H, uint32_t K,
                                              const double* __restrict__ are,
                                              const double* __restrict__ aim, uint32_t lda,
                                              const double* __restrict__ bre,
                                              const double* __restrict__ bim, uint32_t ldb,
                                              double* __restrict__ ore, double* __restrict__ oim,
                                              uint32_t ldo) {
  constexpr bool HASRE = MODE != 2, HASIM = MODE != 1;
  __shared__ double sAr[GK][GT + 1], sAi[GK][GT + 1];  // A tile transposed: [k][row]
  __shared__ double sBr[GK][GT + 1], sBi[GK][GT + 1];  // B tile: [k][col]
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const uint32_t x0 = blockIdx.x * GT, y0 = blockIdx.y * GT;
  double cr[4][4], ci[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) { cr[i][j] = 0; ci[i][j] = 0; }
  for (uint32_t k0 = 0; k0 < K; k0 += GK) {
    // load A[y0 .. y0+64)[k0 .. k0+16) and B[k0 .. k0+16)[x0 .. x0+64): 1024 elements each
    for (int e = threadIdx.x; e < GT * GK; e += 256) {
      int kk = e & (GK - 1), rr = e / GK;  // A: consecutive threads walk k (row-contiguous)
      uint32_t gy = y0 + rr, gk = k0 + kk;
      bool ok = gy < H && gk < K;
      size_t ia = (size_t)gy * lda + gk;
      sAr[kk][rr] = (HASRE || MODE == 0) && ok && are ? are[ia] : 0.0;
      sAi[kk][rr] = HASIM && ok && aim ? aim[ia] : 0.0;
      // B: consecutive threads walk columns (or k, when B is read transposed)
      int cc = TB ? rr : (e & (GT - 1)), kb = TB ? kk : e / GT;
      uint32_t gx = x0 + cc, gk2 = k0 + kb;
      bool ok2 = gx < W && gk2 < K;
      size_t ib = TB ? (size_t)gx * ldb + gk2 : (size_t)gk2 * ldb + gx;
      sBr[kb][cc] = (HASRE || MODE == 0) && ok2 && bre ? bre[ib] : 0.0;
      sBi[kb][cc] = HASIM && ok2 && bim ? bim[ib] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < GK; ++kk) {
      double ar[4], ai[4], br[4], bi[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ar[i] = sAr[kk][ty + 16 * i];
        ai[i] = sAi[kk][ty + 16 * i];
        br[i] = sBr[kk][tx + 16 * i];
        bi[i] = sBi[kk][tx + 16 * i];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (MODE == 0) {
            cr[i][j] = fma(ar[i], br[j], cr[i][j]);
            cr[i][j] = fma(-ai[i], bi[j], cr[i][j]);
            ci[i][j] = fma(ar[i], bi[j], ci[i][j]);
            ci[i][j] = fma(ai[i], br[j], ci[i][j]);
          } else if (MODE == 1) {
            cr[i][j] = fma(ar[i], br[j], cr[i][j]);
          } else {
            ci[i][j] = fma(-ai[i], bi[j], ci[i][j]);
          }
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      uint32_t gy = y0 + ty + 16 * i, gx = x0 + tx + 16 * j;
      if (gy < H && gx < W) {
        size_t o = (size_t)gy * ldo + gx;
        if (HASRE) ore[o] = cr[i][j];
        if (HASIM) oim[o] = ci[i][j];
      }
    }
}

}  // namespace

namespace {

// Launch shapes of the streaming matvecs: R rows per wave read together, U
// loads in flight per row and plane, G groups of R rows per wave one after
// another.  More rows per wave re-read x (from L2) less often per byte of A;
// more loads in flight hide HBM latency; groups give fewer, longer-lived
// waves (one per SIMD at n = 8192 with R G = 8).  Selected per process with
// oaph_set_matvec_shape.  Entry 0 is the default, the fastest measured on
// MI355X at n = 8192 with rotated column chunks (averages of 50 launches,
// operator, x, y and the 48 basis vectors of the projection epilogue):
// FP64 (2, 4, 2) 155.3 us = 6.95 TB/s vs (2, 3, 4) 156.9 us (161.7 us
// without rotation); FP32 operand (4, 2, 2) 79.4 us = 6.85 TB/s vs
// (4, 3, 2) 82.3 us (83.1 us without rotation).
struct Shape {
  int R, U, G;
};
constexpr Shape kZShapes[] = {{2, 4, 2}, {4, 2, 1}, {2, 4, 4}, {4, 3, 2}, {2, 6, 4},
                              {2, 3, 4}, {4, 6, 2}, {8, 2, 1}};
constexpr Shape kCShapes[] = {{4, 2, 2}, {4, 3, 1}, {4, 3, 2}, {4, 4, 2}, {2, 2, 4},
                              {2, 3, 4}, {8, 2, 1}, {4, 2, 1}};
// FP16 operand: 16 B per load as for FP32, but twice the elements; all keep
// R G = 8 (the fused projection epilogue's 256 workgroups at 8192 rows).
// Entry 0 measured fastest at n = 8192 with x in LDS and rotated column
// chunks: 45.4 us (0.76 of HBM; 53.5 us without rotation; (4, 3, 2) 46.0
// us, (8, 2, 1) 51.6 us).
constexpr Shape kHShapes[] = {{2, 4, 4}, {8, 1, 1}, {8, 2, 1}, {4, 2, 2}, {4, 1, 2},
                              {4, 3, 2}, {2, 2, 4}, {8, 3, 1}};
constexpr int kNumZ = sizeof(kZShapes) / sizeof(kZShapes[0]);
constexpr int kNumC = sizeof(kCShapes) / sizeof(kCShapes[0]);
constexpr int kNumH = sizeof(kHShapes) / sizeof(kHShapes[0]);
// g_rot: waves of the streaming matvecs start at rotated column chunks
int g_zshape = 0, g_cshape = 0, g_hshape = 0, g_hxlds = 1, g_rot = 1;
// fallback (G = 1) shapes per operand precision
int g_z1 = 0, g_c1 = 0, g_h1 = 0;
// dotProducts on the matrix cores: 1 large ones (gemm_mfma_eligible), 2 all,
// 0 none (the vector-FMA tiles)
int g_gemm_mfma = 1;

using ZFn = void (*)(uint32_t, uint32_t, uint32_t, const double*, const double*, const double*,
                     const double*, double*, double*, uint32_t, MvProj);
using CFn = void (*)(uint32_t, uint32_t, uint32_t, const float*, const float*, const double*,
                     const double*, double*, double*, uint32_t, MvProj);
using HFn = void (*)(uint32_t, uint32_t, uint32_t, const _Float16*, const _Float16*,
                     const double*, const double*, double*, double*, double, uint32_t, MvProj);

// kZ / kC / kH: each shape as listed; kZ1 / kC1 / kH1 (below): what runs
// when the grouped grid would leave SIMDs without a wave.
template <int MODE>
constexpr ZFn kZ[] = {k_zgemv<2, 4, MODE, 2>, k_zgemv<4, 2, MODE, 1>, k_zgemv<2, 4, MODE, 4>,
                      k_zgemv<4, 3, MODE, 2>, k_zgemv<2, 6, MODE, 4>, k_zgemv<2, 3, MODE, 4>,
                      k_zgemv<4, 6, MODE, 2>, k_zgemv<8, 2, MODE, 1>};
// The G = 1 shapes that run where row groups would leave SIMDs without a
// wave (rows < 4 R G simd_count, e.g. 4096 rows), selected independently of
// the grouped shape with oaph_set_matvec_fallback_shape; entry 0 the default
// (at 4096 rows, 50 launches: FP32-operand real 14.2 us (0.59 of HBM for the
// operator) vs 14.5-20.8 us for the others, complex 23.7 us vs 23.5-34.1 us;
// FP16-operand complex 17.6 us vs 17.5-26.1 us).
struct Shape1 {
  int R, U;
};
constexpr Shape1 kS1Shapes[] = {{4, 2}, {4, 3}, {4, 4}, {4, 6}, {8, 2}, {8, 3}, {4, 8}, {8, 4}};
constexpr int kNumS1 = sizeof(kS1Shapes) / sizeof(kS1Shapes[0]);
template <int MODE>
constexpr ZFn kZ1[] = {k_zgemv<4, 2, MODE, 1>, k_zgemv<4, 3, MODE, 1>, k_zgemv<4, 4, MODE, 1>,
                       k_zgemv<4, 6, MODE, 1>, k_zgemv<8, 2, MODE, 1>, k_zgemv<8, 3, MODE, 1>,
                       k_zgemv<4, 8, MODE, 1>, k_zgemv<8, 4, MODE, 1>};
template <bool CPLX>
constexpr CFn kC[] = {k_cgemv32<4, 2, CPLX, 2>, k_cgemv32<4, 3, CPLX, 1>, k_cgemv32<4, 3, CPLX, 2>,
                      k_cgemv32<4, 4, CPLX, 2>, k_cgemv32<2, 2, CPLX, 4>, k_cgemv32<2, 3, CPLX, 4>,
                      k_cgemv32<8, 2, CPLX, 1>, k_cgemv32<4, 2, CPLX, 1>};
template <bool CPLX>
constexpr CFn kC1[] = {k_cgemv32<4, 2, CPLX, 1>, k_cgemv32<4, 3, CPLX, 1>, k_cgemv32<4, 4, CPLX, 1>,
                       k_cgemv32<4, 6, CPLX, 1>, k_cgemv32<8, 2, CPLX, 1>, k_cgemv32<8, 3, CPLX, 1>,
                       k_cgemv32<4, 8, CPLX, 1>, k_cgemv32<8, 4, CPLX, 1>};
template <bool CPLX, bool XL>
constexpr HFn kH[] = {k_hgemv16<2, 4, CPLX, 4, XL>, k_hgemv16<8, 1, CPLX, 1, XL>,
                      k_hgemv16<8, 2, CPLX, 1, XL>, k_hgemv16<4, 2, CPLX, 2, XL>,
                      k_hgemv16<4, 1, CPLX, 2, XL>, k_hgemv16<4, 3, CPLX, 2, XL>,
                      k_hgemv16<2, 2, CPLX, 4, XL>, k_hgemv16<8, 3, CPLX, 1, XL>};
template <bool CPLX>
constexpr HFn kH1[] = {k_hgemv16<4, 2, CPLX, 1>, k_hgemv16<4, 3, CPLX, 1>, k_hgemv16<4, 4, CPLX, 1>,
                       k_hgemv16<4, 6, CPLX, 1>, k_hgemv16<8, 2, CPLX, 1>, k_hgemv16<8, 3, CPLX, 1>,
                       k_hgemv16<4, 8, CPLX, 1>, k_hgemv16<8, 4, CPLX, 1>};
static_assert(sizeof(kH<true, false>) / sizeof(HFn) == kNumH, "one kernel per FP16 shape");
static_assert(sizeof(kH1<true>) / sizeof(HFn) == kNumS1, "one kernel per fallback shape");
static_assert(sizeof(kZ1<0>) / sizeof(ZFn) == kNumS1, "one kernel per fallback shape");
static_assert(sizeof(kC1<true>) / sizeof(CFn) == kNumS1, "one kernel per fallback shape");
// At most kSmallRows rows without row groups: two rows per wave, six loads in
// flight (n = 2048 complex: FP64 18.8 -> 12.8 us, FP32 operand 11.1 -> 8.4
// us); at 4096 rows this grid would exceed the projection epilogue's 256
// workgroups, so four rows per wave stay there.
constexpr uint32_t kSmallRows = 2048;
template <int MODE>
constexpr ZFn kZs = k_zgemv<2, 6, MODE, 1>;
template <bool CPLX>
constexpr CFn kCs = k_cgemv32<2, 6, CPLX, 1>;
template <bool CPLX>
constexpr HFn kHs = k_hgemv16<2, 4, CPLX, 1>;

// Waves the device runs at once with one per SIMD (4 SIMDs per CU).
unsigned simd_count() {
  static unsigned n = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    return 4u * (unsigned)cus;
  }();
  return n;
}

// Row groups pay off only while the grouped grid still gives every SIMD a
// wave; smaller operators take the kZ1 / kC1 / kH1 fallback shapes.
bool use_groups(const Shape& sh, uint32_t rows) {
  return rows / (uint32_t)(sh.R * sh.G) >= simd_count();
}

static_assert(sizeof(kZ<0>) / sizeof(ZFn) == kNumZ, "one kernel per FP64 shape");
static_assert(sizeof(kC<true>) / sizeof(CFn) == kNumC, "one kernel per FP32 shape");

// The epilogue arguments a launch of grid blocks passes: the caller's when it
// asked for the projection, y is complex and the partial sums fit (blocks set
// to the grid size), else none (blocks = 0).
MvProj fused_projection(MvProj* pj, bool complex_y, uint32_t grid) {
  if (!pj) return MvProj{};
  pj->blocks = (complex_y && pj->part && pj->ncols <= 64 && grid <= pj->maxparts) ? grid : 0;
  return pj->blocks ? *pj : MvProj{};
}

}  // namespace

oaph_status launch_zgemv(hipStream_t s, uint32_t rows, uint32_t cols, uint32_t lda,
                         const double* are, const double* aim, const double* xre,
                         const double* xim, double* yre, double* yim, MvProj* pj) {
  // MODE 0: complex A, complex x; 1: real A, real x; 2: real A, complex x
  const int mode = aim ? 0 : (xim ? 2 : 1);
  const bool vec_ok = (cols % 2 == 0) && (lda % 2 == 0) &&
                      ((((uintptr_t)are) | ((uintptr_t)aim) | ((uintptr_t)xre) |
                        ((uintptr_t)xim)) & 15) == 0 && cols >= 128;
  if (vec_ok) {
    const Shape& sh = kZShapes[g_zshape];
    const bool grp = sh.G == 1 || use_groups(sh, rows);
    const bool small = !grp && rows <= kSmallRows;
    const int RG = grp ? sh.R * sh.G : (small ? 2 : kS1Shapes[g_z1].R);
    const unsigned waves = (rows + RG - 1) / RG;
    const dim3 grid((waves + 3) / 4);
    const ZFn fn = grp ? (mode == 0 ? kZ<0>[g_zshape] : mode == 1 ? kZ<1>[g_zshape] : kZ<2>[g_zshape])
                 : small ? (mode == 0 ? kZs<0> : mode == 1 ? kZs<1> : kZs<2>)
                       : (mode == 0 ? kZ1<0>[g_z1] : mode == 1 ? kZ1<1>[g_z1] : kZ1<2>[g_z1]);
    const MvProj kp = fused_projection(pj, mode != 1, grid.x);
    hipLaunchKernelGGL(fn, grid, dim3(256), 0, s, rows, cols, lda, are, aim, xre, xim, yre, yim,
                       (uint32_t)g_rot, kp);
  } else {
    if (pj) pj->blocks = 0;
    dim3 grid((rows + 3) / 4);
    if (mode == 0)
      hipLaunchKernelGGL(k_zgemv_scalar<0>, grid, dim3(256), 0, s, rows, cols, lda, are, aim, xre, xim, yre, yim);
    else if (mode == 1)
      hipLaunchKernelGGL(k_zgemv_scalar<1>, grid, dim3(256), 0, s, rows, cols, lda, are, aim, xre, xim, yre, yim);
    else
      hipLaunchKernelGGL(k_zgemv_scalar<2>, grid, dim3(256), 0, s, rows, cols, lda, are, aim, xre, xim, yre, yim);
  }
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_zgemv_multi(hipStream_t s, uint32_t rows, uint32_t cols, uint32_t lda,
                               const double* are, const double* aim, const double* xre,
                               const double* xim, uint32_t ldx, uint32_t nvec, double* yre,
                               double* yim, uint32_t ldy) {
  OAPH_CHECK(nvec >= 1 && nvec <= kMultiMaxVectors, OAPH_ERR_UNSUPPORTED);
  OAPH_CHECK(xim != nullptr && yim != nullptr, OAPH_ERR_PLANE);
  const bool ok = (cols % 2 == 0) && (lda % 2 == 0) && (ldx % 2 == 0) &&
                  ((((uintptr_t)are) | ((uintptr_t)aim) | ((uintptr_t)xre) | ((uintptr_t)xim)) &
                   15) == 0;
  OAPH_CHECK(ok, OAPH_ERR_UNSUPPORTED);
  constexpr int R = 2;
  const dim3 grid((rows + 4 * R - 1) / (4 * R));
  if (aim)
    hipLaunchKernelGGL((k_zgemv_multi<R, kMultiMaxVectors, 0>), grid, dim3(256), 0, s, rows, cols,
                       lda, are, aim, xre, xim, ldx, nvec, yre, yim, ldy);
  else
    hipLaunchKernelGGL((k_zgemv_multi<R, kMultiMaxVectors, 2>), grid, dim3(256), 0, s, rows, cols,
                       lda, are, nullptr, xre, xim, ldx, nvec, yre, yim, ldy);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_cgemv32(hipStream_t s, uint32_t rows, uint32_t cols, uint32_t lda,
                           const float* are, const float* aim, const double* xre,
                           const double* xim, double* yre, double* yim, MvProj* pj) {
  OAPH_CHECK(cols % 4 == 0 && lda % 4 == 0, OAPH_ERR_UNSUPPORTED);
  OAPH_CHECK(((((uintptr_t)are) | ((uintptr_t)aim) | ((uintptr_t)xre) | ((uintptr_t)xim)) & 15) == 0,
             OAPH_ERR_UNSUPPORTED);
  OAPH_CHECK(xim != nullptr && yim != nullptr, OAPH_ERR_PLANE);
  const Shape& sh = kCShapes[g_cshape];
  const bool grp = sh.G == 1 || use_groups(sh, rows);
  const bool small = !grp && rows <= kSmallRows;
  const int RG = grp ? sh.R * sh.G : (small ? 2 : kS1Shapes[g_c1].R);
  const unsigned waves = (rows + RG - 1) / RG;
  const dim3 grid((waves + 3) / 4);
  const CFn fn = grp ? (aim ? kC<true>[g_cshape] : kC<false>[g_cshape])
                 : small ? (aim ? kCs<true> : kCs<false>)
                     : (aim ? kC1<true>[g_c1] : kC1<false>[g_c1]);
  const MvProj kp = fused_projection(pj, true, grid.x);
  hipLaunchKernelGGL(fn, grid, dim3(256), 0, s, rows, cols, lda,
                     are, aim, xre, xim, yre, yim, (uint32_t)g_rot, kp);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_hgemv16(hipStream_t s, uint32_t rows, uint32_t cols, uint32_t lda,
                           const _Float16* are, const _Float16* aim, const double* xre,
                           const double* xim, double* yre, double* yim, double unscale,
                           MvProj* pj) {
  OAPH_CHECK(cols % 8 == 0 && lda % 8 == 0, OAPH_ERR_UNSUPPORTED);
  OAPH_CHECK(((((uintptr_t)are) | ((uintptr_t)aim) | ((uintptr_t)xre) | ((uintptr_t)xim)) & 15) == 0,
             OAPH_ERR_UNSUPPORTED);
  OAPH_CHECK(xim != nullptr && yim != nullptr, OAPH_ERR_PLANE);
  const Shape& sh = kHShapes[g_hshape];
  const bool grp = sh.G == 1 || use_groups(sh, rows);
  const bool small = !grp && rows <= kSmallRows;
  const int RG = grp ? sh.R * sh.G : (small ? 2 : kS1Shapes[g_h1].R);
  const unsigned waves = (rows + RG - 1) / RG;
  const dim3 grid((waves + 3) / 4);
  // row groups (the large operators) with rows of at most kXLdsCols take x
  // from LDS (g_hxlds = 0 turns that off)
  const bool xl = g_hxlds && cols <= kXLdsCols;
  const HFn fn = grp ? (aim ? (xl ? kH<true, true>[g_hshape] : kH<true, false>[g_hshape])
                            : (xl ? kH<false, true>[g_hshape] : kH<false, false>[g_hshape]))
                 : small ? (aim ? kHs<true> : kHs<false>)
                     : (aim ? kH1<true>[g_h1] : kH1<false>[g_h1]);
  const MvProj kp = fused_projection(pj, true, grid.x);
  hipLaunchKernelGGL(fn, grid, dim3(256), 0, s, rows, cols, lda, are, aim, xre, xim, yre, yim,
                     unscale, (uint32_t)g_rot, kp);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_gemm(hipStream_t s, oaph_matrix* out, const oaph_matrix* a,
                        const oaph_matrix* b, uint32_t W, uint32_t H, uint32_t K,
                        bool dims_mode) {
  const bool ore = out->reValues != nullptr, oim = out->imValues != nullptr;
  int mode = ore && oim ? 0 : ore ? 1 : 2;
  (void)dims_mode;
  // one-column right operand over full rows of A: the streaming GEMV
  if (W == 1 && b->columns == 1 && mode != 2 && K == a->columns && out->columns == 1) {
    return launch_zgemv(s, H, K, a->columns, a->reValues, mode == 0 ? a->imValues : nullptr,
                        b->reValues, mode == 0 ? b->imValues : nullptr, out->reValues,
                        mode == 0 ? out->imValues : nullptr);
  }
  // large products on the matrix cores (gemm.hip)
  if (g_gemm_mfma == 2 || (g_gemm_mfma == 1 && gemm_mfma_eligible(W, H, K)))
    return launch_gemm_mfma(s, mode, W, H, K, a->reValues, a->imValues, a->columns, b->reValues,
                            b->imValues, b->columns, out->reValues, out->imValues, out->columns);
  dim3 grid((W + GT - 1) / GT, (H + GT - 1) / GT);
  OAPH_CHECK(grid.y <= 65535, OAPH_ERR_UNSUPPORTED);
  if (mode == 0)
    hipLaunchKernelGGL(k_gemm<0>, grid, dim3(256), 0, s, W, H, K, a->reValues, a->imValues,
                       a->columns, b->reValues, b->imValues, b->columns, out->reValues,
                       out->imValues, out->columns);
  else if (mode == 1)
    hipLaunchKernelGGL(k_gemm<1>, grid, dim3(256), 0, s, W, H, K, a->reValues, nullptr,
                       a->columns, b->reValues, nullptr, b->columns, out->reValues, nullptr,
                       out->columns);
  else
    hipLaunchKernelGGL(k_gemm<2>, grid, dim3(256), 0, s, W, H, K, nullptr, a->imValues,
                       a->columns, nullptr, b->imValues, b->columns, nullptr, out->imValues,
                       out->columns);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_gemm_nt(hipStream_t s, uint32_t W, uint32_t H, uint32_t K, const double* are,
                           const double* aim, uint32_t lda, const double* bre, const double* bim,
                           uint32_t ldb, double* ore, double* oim, uint32_t ldo) {
  dim3 grid((W + GT - 1) / GT, (H + GT - 1) / GT);
  OAPH_CHECK(grid.y <= 65535, OAPH_ERR_UNSUPPORTED);
  if (ore && oim)
    hipLaunchKernelGGL((k_gemm<0, true>), grid, dim3(256), 0, s, W, H, K, are, aim, lda, bre, bim,
                       ldb, ore, oim, ldo);
  else if (ore)
    hipLaunchKernelGGL((k_gemm<1, true>), grid, dim3(256), 0, s, W, H, K, are, nullptr, lda, bre,
                       nullptr, ldb, ore, nullptr, ldo);
  else
    hipLaunchKernelGGL((k_gemm<2, true>), grid, dim3(256), 0, s, W, H, K, nullptr, aim, lda,
                       nullptr, bim, ldb, nullptr, oim, ldo);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

}  // namespace oaph

extern "C" {

oaph_status oaph_set_gemm_path(int mfma) {
  OAPH_CHECK(mfma >= 0 && mfma <= 2, OAPH_ERR_ARG);
  oaph::g_gemm_mfma = mfma;
  return OAPH_OK;
}

oaph_status oaph_set_matvec_shape(int fp32, int variant) {
  using namespace oaph;
  if (fp32 == 4) {  // every streaming matvec: waves start at rotated column chunks (1) or at 0
    OAPH_CHECK(variant == 0 || variant == 1, OAPH_ERR_ARG);
    g_rot = variant;
  } else if (fp32 == 3) {  // FP16 operand: x staged in LDS (1) or read from global memory (0)
    OAPH_CHECK(variant == 0 || variant == 1, OAPH_ERR_ARG);
    g_hxlds = variant;
  } else if (fp32 == 2) {
    OAPH_CHECK(variant >= 0 && variant < kNumH, OAPH_ERR_ARG);
    g_hshape = variant;
  } else if (fp32) {
    OAPH_CHECK(variant >= 0 && variant < kNumC, OAPH_ERR_ARG);
    g_cshape = variant;
  } else {
    OAPH_CHECK(variant >= 0 && variant < kNumZ, OAPH_ERR_ARG);
    g_zshape = variant;
  }
  return OAPH_OK;
}

oaph_status oaph_set_matvec_fallback_shape(int fp32, int variant) {
  using namespace oaph;
  OAPH_CHECK(fp32 >= 0 && fp32 <= 2 && variant >= 0 && variant < kNumS1, OAPH_ERR_ARG);
  (fp32 == 2 ? g_h1 : fp32 == 1 ? g_c1 : g_z1) = variant;
  return OAPH_OK;
}

oaph_status oaph_matvec_shape_info(int fp32, int variant, int* rows_per_wave, int* loads_in_flight,
                                   int* row_groups) {
  using namespace oaph;
  const int n = fp32 == 2 ? kNumH : fp32 ? kNumC : kNumZ;
  OAPH_CHECK(variant >= 0 && variant < n, OAPH_ERR_ARG);
  const Shape sh = fp32 == 2 ? kHShapes[variant] : fp32 ? kCShapes[variant] : kZShapes[variant];
  if (rows_per_wave) *rows_per_wave = sh.R;
  if (loads_in_flight) *loads_in_flight = sh.U;
  if (row_groups) *row_groups = sh.G;
  return OAPH_OK;
}

}  // extern "C"
