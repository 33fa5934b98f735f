// rts.hip -- the square operator S = M M^T of a rectangular device matrix M
// (R rows x C columns, S is R x R): explicit row blocks of S through the
// tiled GEMM reading M transposed in place, and the implicit product
// S v = M (M^T v), which streams M twice and never forms S.
//
// Reference: oapMatrixCuda/RecToSquareApi.cpp:57-141 (blocks of S; a matrix
// that is already square is used as it is), oapMatrixCuda/SquareMatrixApi.cpp
// :143-181 (a block is subM * M^T by dotProduct against a cached transposed
// copy), oap2dt3dDevice/MainAPExecutor.cpp:42-127 (callbacks that multiply S
// by v, whole or block by block).
#include <new>

#include "internal.h"

struct oaph_rts {
  oaph_api* api = nullptr;
  const oaph_matrix* rec = nullptr;
  oaph::DevBuf part;  // chunk partial sums of M^T v: 2 planes x chunks x C
  oaph::DevBuf u;     // M^T v: 2 planes x C
  oaph::DevBuf wpart; // column-slice partial products of M u: 2 planes x slices x R
};

namespace oaph {
namespace {

// part[chunk][c] = sum over the chunk's rows r of M[r][c] v[r] (plain
// transpose, complex product), VW adjacent columns per thread.  Rows in a
// chunk are taken in order, four loads per column in flight.
template <bool AIM, bool XIM, int VW>
__global__ __launch_bounds__(256) void k_gemv_t(uint32_t rows, uint32_t cols,
                                                uint32_t rows_per_chunk,
                                                const double* __restrict__ are,
                                                const double* __restrict__ aim,
                                                const double* __restrict__ vre,
                                                const double* __restrict__ vim,
                                                double* __restrict__ ore, double* __restrict__ oim,
                                                size_t ostride) {
  constexpr bool CPLX = AIM || XIM;
  const uint32_t c0 = (blockIdx.x * 256 + threadIdx.x) * VW;
  if (c0 >= cols) return;
  const uint32_t r0 = blockIdx.y * rows_per_chunk;
  const uint32_t r1 = min(rows, r0 + rows_per_chunk);
  double sr[VW], si[VW];
#pragma unroll
  for (int q = 0; q < VW; ++q) sr[q] = si[q] = 0.0;
  auto load = [&](const double* p, size_t i, double* d) {
    if constexpr (VW == 2) {
      const double2 t = *reinterpret_cast<const double2*>(p + i);
      d[0] = t.x;
      d[1] = t.y;
    } else {
      d[0] = p[i];
    }
  };
  auto acc = [&](const double* ar, const double* ai, double xr, double xi) {
#pragma unroll
    for (int q = 0; q < VW; ++q) {
      sr[q] = fma(ar[q], xr, sr[q]);
      if (AIM && XIM) sr[q] = fma(-ai[q], xi, sr[q]);
      if (XIM) si[q] = fma(ar[q], xi, si[q]);
      if (AIM) si[q] = fma(ai[q], xr, si[q]);
    }
  };
  uint32_t r = r0;
  for (; r + 4 <= r1; r += 4) {
    double ar[4][VW], ai[4][VW], xr[4], xi[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t i = (size_t)(r + u) * cols + c0;
      load(are, i, ar[u]);
      if (AIM) load(aim, i, ai[u]);
      xr[u] = vre[r + u];
      xi[u] = XIM ? vim[r + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) acc(ar[u], ai[u], xr[u], xi[u]);
  }
  for (; r < r1; ++r) {
    double ar[VW], ai[VW];
    const size_t i = (size_t)r * cols + c0;
    load(are, i, ar);
    if (AIM) load(aim, i, ai);
    acc(ar, ai, vre[r], XIM ? vim[r] : 0.0);
  }
  const size_t o = (size_t)blockIdx.y * ostride + c0;
#pragma unroll
  for (int q = 0; q < VW; ++q) {
    ore[o + q] = sr[q];
    if (CPLX) oim[o + q] = si[q];
  }
}

// u[c] = sum over chunks of part[chunk][c]: 32 columns per workgroup, eight
// interleaved chunk sets per column summed in order, then combined in slice
// order (the result does not depend on the launch)
__global__ __launch_bounds__(256) void k_sum_chunks(uint32_t cols, uint32_t chunks,
                                                    const double* __restrict__ pre,
                                                    const double* __restrict__ pim,
                                                    double* __restrict__ ure,
                                                    double* __restrict__ uim) {
  __shared__ double lr[8][33], li[8][33];
  const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const uint32_t c = blockIdx.x * 32 + tx;
  double sr = 0, si = 0;
  if (c < cols) {
    uint32_t k = ty;
    for (; k + 24 < chunks; k += 32) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = pre[(size_t)(k + 8 * u) * cols + c];
        b[u] = pim ? pim[(size_t)(k + 8 * u) * cols + c] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        sr += a[u];
        si += b[u];
      }
    }
    for (; k < chunks; k += 8) {
      sr += pre[(size_t)k * cols + c];
      if (pim) si += pim[(size_t)k * cols + c];
    }
  }
  lr[ty][tx] = sr;
  li[ty][tx] = si;
  __syncthreads();
  if (ty == 0 && c < cols) {
    double tr = 0, ti = 0;
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      tr += lr[g][tx];
      ti += li[g][tx];
    }
    ure[c] = tr;
    if (uim) uim[c] = ti;
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// part[q][r] = sum over slice q's columns [q cw, min(C, (q+1) cw)) of
// A[r][c] x[c] (complex product), four rows per wave, VW adjacent columns
// per lane, two column steps of loads in flight; the slice is blockIdx.y.
template <bool AIM, bool XIM, int VW>
__global__ __launch_bounds__(256) void k_gemv_split(uint32_t rows, uint32_t cols, uint32_t cw,
                                                    const double* __restrict__ are,
                                                    const double* __restrict__ aim,
                                                    const double* __restrict__ xre,
                                                    const double* __restrict__ xim,
                                                    double* __restrict__ pre,
                                                    double* __restrict__ pim) {
  constexpr int RW = 4, STEP = 64 * VW;
  const int lane = threadIdx.x & 63;
  const uint32_t r0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * RW;
  if (r0 >= rows) return;
  const uint32_t c0 = blockIdx.y * cw, c1 = min(cols, c0 + cw);
  const double* ar[RW];
  const double* ai[RW];
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    const uint32_t r = r0 + q < rows ? r0 + q : rows - 1;  // clamped rows are dropped
    ar[q] = are + (size_t)r * cols;
    ai[q] = AIM ? aim + (size_t)r * cols : nullptr;
  }
  double sr[RW], si[RW];
#pragma unroll
  for (int q = 0; q < RW; ++q) sr[q] = si[q] = 0.0;
  auto ld = [&](const double* p, uint32_t c, double* d) {
    if constexpr (VW == 2) {
      const double2 t = *reinterpret_cast<const double2*>(p + c);
      d[0] = t.x;
      d[1] = t.y;
    } else {
      d[0] = p[c];
    }
  };
  auto step = [&](const double (*a)[VW], const double (*b)[VW], const double* x,
                  const double* y) {
#pragma unroll
    for (int q = 0; q < RW; ++q)
#pragma unroll
      for (int e = 0; e < VW; ++e) {
        sr[q] = fma(a[q][e], x[e], sr[q]);
        if (AIM && XIM) sr[q] = fma(-b[q][e], y[e], sr[q]);
        if (XIM) si[q] = fma(a[q][e], y[e], si[q]);
        if (AIM) si[q] = fma(b[q][e], x[e], si[q]);
      }
  };
  uint32_t c = c0 + VW * lane;
  for (; c + STEP < c1; c += 2 * STEP) {
    double a0[RW][VW], b0[RW][VW], a1[RW][VW], b1[RW][VW];
    double x0[VW], y0[VW], x1[VW], y1[VW];
    ld(xre, c, x0);
    ld(xre, c + STEP, x1);
    if (XIM) {
      ld(xim, c, y0);
      ld(xim, c + STEP, y1);
    }
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      ld(ar[q], c, a0[q]);
      ld(ar[q], c + STEP, a1[q]);
      if (AIM) {
        ld(ai[q], c, b0[q]);
        ld(ai[q], c + STEP, b1[q]);
      }
    }
    step(a0, b0, x0, y0);
    step(a1, b1, x1, y1);
  }
  if (c < c1) {
    double a0[RW][VW], b0[RW][VW], x0[VW], y0[VW];
    ld(xre, c, x0);
    if (XIM) ld(xim, c, y0);
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      ld(ar[q], c, a0[q]);
      if (AIM) ld(ai[q], c, b0[q]);
    }
    step(a0, b0, x0, y0);
  }
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    sr[q] = wave_sum(sr[q]);
    if (AIM || XIM) si[q] = wave_sum(si[q]);
  }
  if (lane < RW && r0 + lane < rows) {
    double vr = sr[0], vi = si[0];
#pragma unroll
    for (int q = 1; q < RW; ++q)
      if (lane == q) {
        vr = sr[q];
        vi = si[q];
      }
    pre[(size_t)blockIdx.y * rows + r0 + lane] = vr;
    if (AIM || XIM) pim[(size_t)blockIdx.y * rows + r0 + lane] = vi;
  }
}

__global__ void k_fill_zero(double* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

oaph_status zero_plane(hipStream_t s, double* p, size_t n) {
  if (!p || !n) return OAPH_OK;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

// u = M^T v over R rows and C columns; returns whether u is complex
oaph_status gemv_t(oaph_rts* t, hipStream_t s, const double* vre, const double* vim, bool* cplx) {
  const oaph_matrix* M = t->rec;
  const uint32_t R = M->rows, C = M->columns;
  const bool aim = M->imValues != nullptr, xim = vim != nullptr;
  *cplx = aim || xim;
  const bool aligned =
      ((((uintptr_t)M->reValues) | ((uintptr_t)M->imValues)) & 15) == 0;
  const int vw = (C % 2 == 0 && aligned) ? 2 : 1;
  const uint32_t col_blocks = (C + 256 * vw - 1) / (256 * vw);
  // enough chunks of rows for ~2048 workgroups, each chunk at least 16 rows
  uint32_t chunks = (2048 + col_blocks - 1) / col_blocks;
  const uint32_t max_chunks = (R + 15) / 16;
  if (chunks > max_chunks) chunks = max_chunks;
  if (chunks < 1) chunks = 1;
  const uint32_t rpc = (R + chunks - 1) / chunks;
  chunks = (R + rpc - 1) / rpc;
  OAPH_CHECK(chunks <= 65535, OAPH_ERR_UNSUPPORTED);
  OAPH_TRY(t->u.reserve(2 * (size_t)C * sizeof(double)));
  double* ure = t->u.as<double>();
  double* uim = ure + C;
  double *ore = ure, *oim = uim;
  if (chunks > 1) {
    OAPH_TRY(t->part.reserve(2 * (size_t)chunks * C * sizeof(double)));
    ore = t->part.as<double>();
    oim = ore + (size_t)chunks * C;
  }
  const dim3 grid(col_blocks, chunks);
#define OAPH_GT(A, X, W)                                                                      \
  hipLaunchKernelGGL((k_gemv_t<A, X, W>), grid, dim3(256), 0, s, R, C, rpc, M->reValues,     \
                     M->imValues, vre, vim, ore, oim, (size_t)C)
  if (vw == 2) {
    if (aim && xim) OAPH_GT(true, true, 2);
    else if (aim) OAPH_GT(true, false, 2);
    else if (xim) OAPH_GT(false, true, 2);
    else OAPH_GT(false, false, 2);
  } else {
    if (aim && xim) OAPH_GT(true, true, 1);
    else if (aim) OAPH_GT(true, false, 1);
    else if (xim) OAPH_GT(false, true, 1);
    else OAPH_GT(false, false, 1);
  }
#undef OAPH_GT
  OAPH_HIP(hipGetLastError());
  if (chunks > 1) {
    hipLaunchKernelGGL(k_sum_chunks, dim3((C + 31) / 32), dim3(256), 0, s, C, chunks, ore,
                       *cplx ? oim : nullptr, ure, *cplx ? uim : nullptr);
    OAPH_HIP(hipGetLastError());
  }
  return OAPH_OK;
}

// w = A x for A (rows x cols, row stride cols) with any plane combination;
// writes zero into an imaginary plane of w that the product leaves real.
// With few rows the columns are split into slices (one launch, a slice per
// grid row, into a workspace, then an ordered sum), so that the launch still
// has thousands of waves.
oaph_status gemv_any(oaph_rts* t, hipStream_t s, const oaph_matrix* A, const double* xre,
                     const double* xim, double* wre, double* wim) {
  const uint32_t R = A->rows, C = A->columns;
  const bool aim = A->imValues != nullptr, cplx = aim || xim;
  if (cplx) OAPH_CHECK(wim != nullptr, OAPH_ERR_PLANE);
  const uint32_t waves = (R + 3) / 4;
  uint32_t S = 1;
  if (waves < 2048 && C >= 4096) {
    S = (2048 + waves - 1) / waves;
    if (S > 16) S = 16;
    if (S > C / 2048) S = C / 2048;
  }
  if (S == 1) {
    if (aim && !xim) {
      // complex A times real x: two real products
      OAPH_TRY(launch_zgemv(s, R, C, C, A->reValues, nullptr, xre, nullptr, wre, nullptr));
      OAPH_TRY(launch_zgemv(s, R, C, C, A->imValues, nullptr, xre, nullptr, wim, nullptr));
    } else {
      OAPH_TRY(launch_zgemv(s, R, C, C, A->reValues, aim ? A->imValues : nullptr, xre, xim, wre,
                            xim ? wim : nullptr));
    }
  } else {
    const uint32_t cw = ((C + S - 1) / S + 255) / 256 * 256;
    S = (C + cw - 1) / cw;
    OAPH_TRY(t->wpart.reserve(2 * (size_t)S * R * sizeof(double)));
    double* pre = t->wpart.as<double>();
    double* pim = pre + (size_t)S * R;
    const bool aligned = C % 2 == 0 &&
                         ((((uintptr_t)A->reValues) | ((uintptr_t)A->imValues) |
                           ((uintptr_t)xre) | ((uintptr_t)xim)) & 15) == 0;
    const dim3 grid((R + 15) / 16, S);
#define OAPH_GS(AI, XI, W)                                                                    \
  hipLaunchKernelGGL((k_gemv_split<AI, XI, W>), grid, dim3(256), 0, s, R, C, cw, A->reValues, \
                     A->imValues, xre, xim, pre, pim)
    if (aligned) {
      if (aim && xim) OAPH_GS(true, true, 2);
      else if (aim) OAPH_GS(true, false, 2);
      else if (xim) OAPH_GS(false, true, 2);
      else OAPH_GS(false, false, 2);
    } else {
      if (aim && xim) OAPH_GS(true, true, 1);
      else if (aim) OAPH_GS(true, false, 1);
      else if (xim) OAPH_GS(false, true, 1);
      else OAPH_GS(false, false, 1);
    }
#undef OAPH_GS
    OAPH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_chunks, dim3((R + 31) / 32), dim3(256), 0, s, R, S, pre,
                       cplx ? pim : nullptr, wre, cplx ? wim : nullptr);
    OAPH_HIP(hipGetLastError());
  }
  if (!cplx) OAPH_TRY(zero_plane(s, wim, R));
  return OAPH_OK;
}

// copies rows [r0, r0 + n) of M into out (the already-square case)
oaph_status copy_rows(hipStream_t s, oaph_matrix* out, const oaph_matrix* M, uint32_t r0,
                      uint32_t n) {
  const size_t off = (size_t)r0 * M->columns, len = (size_t)n * M->columns;
  if (out->reValues) {
    if (M->reValues)
      OAPH_HIP(hipMemcpyAsync(out->reValues, M->reValues + off, len * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
    else
      OAPH_TRY(zero_plane(s, out->reValues, len));
  }
  if (out->imValues) {
    if (M->imValues)
      OAPH_HIP(hipMemcpyAsync(out->imValues, M->imValues + off, len * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
    else
      OAPH_TRY(zero_plane(s, out->imValues, len));
  }
  return OAPH_OK;
}

// gives out exactly `rows` rows, reallocating its planes when the count
// differs (SquareMatrixApi.cpp:162-168 reallocates the output the same way)
oaph_status resize_rows(oaph_matrix* out, uint32_t rows) {
  if (out->rows == rows) return OAPH_OK;
  const size_t bytes = (size_t)out->columns * rows * sizeof(double);
  double *re = nullptr, *im = nullptr;
  if (out->reValues && hipMalloc((void**)&re, bytes) != hipSuccess) return OAPH_ERR_NOMEM;
  if (out->imValues && hipMalloc((void**)&im, bytes) != hipSuccess) {
    if (re) (void)hipFree(re);
    return OAPH_ERR_NOMEM;
  }
  OAPH_HIP(hipDeviceSynchronize());
  if (out->reValues) (void)hipFree(out->reValues);
  if (out->imValues) (void)hipFree(out->imValues);
  out->reValues = re;
  out->imValues = im;
  out->rows = out->realRows = rows;
  return OAPH_OK;
}

}  // namespace
}  // namespace oaph

using namespace oaph;

extern "C" {

oaph_status oaph_rts_create(oaph_rts** out, oaph_api* api, const oaph_matrix* rec) {
  OAPH_CHECK(out && api && rec, OAPH_ERR_ARG);
  OAPH_CHECK(rec->reValues || rec->imValues, OAPH_ERR_PLANE);
  OAPH_CHECK(rec->columns > 0 && rec->rows > 0, OAPH_ERR_SHAPE);
  OAPH_TRY(ensure_init());
  oaph_rts* t = new (std::nothrow) oaph_rts;
  OAPH_CHECK(t, OAPH_ERR_NOMEM);
  t->api = api;
  t->rec = rec;
  *out = t;
  return OAPH_OK;
}

oaph_status oaph_rts_destroy(oaph_rts* t) {
  if (!t) return OAPH_OK;
  t->part.release();
  t->u.release();
  t->wpart.release();
  delete t;
  return OAPH_OK;
}

oaph_status oaph_rts_info(const oaph_rts* t, oaph_matrix_info* info) {
  OAPH_CHECK(t && info, OAPH_ERR_ARG);
  info->isRe = t->rec->reValues != nullptr;
  info->isIm = t->rec->imValues != nullptr;
  info->columns = t->rec->rows;
  info->rows = t->rec->rows;
  return OAPH_OK;
}

oaph_status oaph_rts_sub_matrix(oaph_rts* t, uint32_t rindex, uint32_t rlength, oaph_matrix* out) {
  OAPH_CHECK(t && out, OAPH_ERR_ARG);
  const oaph_matrix* M = t->rec;
  const uint32_t R = M->rows, C = M->columns;
  OAPH_CHECK(rindex < R && rlength > 0, OAPH_ERR_ARG);
  OAPH_CHECK(out->reValues || out->imValues, OAPH_ERR_PLANE);
  OAPH_CHECK(out->columns == R, OAPH_ERR_SHAPE);
  const uint32_t n = rindex + rlength < R ? rlength : R - rindex;
  OAPH_TRY(resize_rows(out, n));
  const hipStream_t s = t->api->stream;
  if (C == R) return copy_rows(s, out, M, rindex, n);
  // out(x, y) = sum_k M[rindex + y][k] M[x][k]: dotProduct(subM, M^T) with
  // the output's planes selecting the product's planes as dotProduct does
  const size_t off = (size_t)rindex * C;
  const bool want_re = out->reValues != nullptr, want_im = out->imValues != nullptr;
  if (want_re && want_im) {
    if (!M->imValues) {
      OAPH_TRY(launch_gemm_nt(s, R, n, C, M->reValues + off, nullptr, C, M->reValues, nullptr, C,
                              out->reValues, nullptr, R));
      return zero_plane(s, out->imValues, (size_t)R * n);
    }
    OAPH_CHECK(M->reValues, OAPH_ERR_PLANE);
    return launch_gemm_nt(s, R, n, C, M->reValues + off, M->imValues + off, C, M->reValues,
                          M->imValues, C, out->reValues, out->imValues, R);
  }
  if (want_re) {
    OAPH_CHECK(M->reValues, OAPH_ERR_PLANE);
    return launch_gemm_nt(s, R, n, C, M->reValues + off, nullptr, C, M->reValues, nullptr, C,
                          out->reValues, nullptr, R);
  }
  OAPH_CHECK(M->imValues, OAPH_ERR_PLANE);
  return launch_gemm_nt(s, R, n, C, nullptr, M->imValues + off, C, nullptr, M->imValues, C,
                        nullptr, out->imValues, R);
}

oaph_status oaph_rts_square(oaph_rts* t, oaph_matrix* out) {
  OAPH_CHECK(t && out, OAPH_ERR_ARG);
  OAPH_CHECK(out->rows == t->rec->rows, OAPH_ERR_SHAPE);
  return oaph_rts_sub_matrix(t, 0, t->rec->rows, out);
}

oaph_status oaph_rts_row_vector(oaph_rts* t, uint32_t index, oaph_matrix* out) {
  OAPH_CHECK(t && out, OAPH_ERR_ARG);
  OAPH_CHECK(out->rows == 1, OAPH_ERR_SHAPE);
  return oaph_rts_sub_matrix(t, index, 1, out);
}

oaph_status oaph_rts_multiply(oaph_rts* t, oaph_api* api, oaph_matrix* w, const oaph_matrix* v) {
  OAPH_CHECK(t && w && v, OAPH_ERR_ARG);
  const oaph_matrix* M = t->rec;
  const uint32_t R = M->rows, C = M->columns;
  OAPH_CHECK(v->columns == 1 && v->rows == R && w->columns == 1 && w->rows == R, OAPH_ERR_SHAPE);
  OAPH_CHECK(v->reValues && w->reValues && M->reValues, OAPH_ERR_PLANE);
  const hipStream_t s = (api ? api : t->api)->stream;
  if (C == R) return gemv_any(t, s, M, v->reValues, v->imValues, w->reValues, w->imValues);
  bool cplx = false;
  OAPH_TRY(gemv_t(t, s, v->reValues, v->imValues, &cplx));
  const double* ure = t->u.as<double>();
  return gemv_any(t, s, M, ure, cplx ? ure + C : nullptr, w->reValues, w->imValues);
}

}  // extern "C"
