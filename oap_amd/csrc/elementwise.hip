// elementwise.hip -- bandwidth-bound per-plane procedures on gfx950.
//
// Per-plane semantics follow the reference kernels they replace (cited per
// function); the implementation is wave64 grid-stride loops with 16-byte
// accesses where the layout allows, and an LDS-tiled transpose.
#include "internal.h"

namespace oaph {

namespace {

constexpr int kThreads = 256;

inline unsigned grid_for(size_t n, int per_thread = 1) {
  size_t blocks = (n + (size_t)kThreads * per_thread - 1) / ((size_t)kThreads * per_thread);
  if (blocks < 1) blocks = 1;
  if (blocks > 65535u * 16u) blocks = 65535u * 16u;
  return (unsigned)blocks;
}

// ------------------------------------------------------------ transpose
// out(c, r) = in(r, c) (CuTransposeProcedures.h:106-158); conj negates im
// (CuConjugateTransposeProcedures.h:106-158).  32x32 tiles staged in LDS with
// one column of padding so that the column reads of the tile hit distinct
// banks; reads and writes of HBM are both row-contiguous.
constexpr int TT = 32;
__global__ __launch_bounds__(256) void k_transpose(double* __restrict__ ore, double* __restrict__ oim,
                                                   const double* __restrict__ are,
                                                   const double* __restrict__ aim,
                                                   uint32_t out_cols, uint32_t out_rows,
                                                   double imsign) {
  __shared__ double tre[TT][TT + 1];
  __shared__ double tim[TT][TT + 1];
  // input is out_cols rows x out_rows columns
  const uint32_t in_cols = out_rows;
  const uint32_t bx = blockIdx.x * TT, by = blockIdx.y * TT;  // tile origin in output (c, r)
  const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;  // 32 x 8
  // load input tile: input rows = output columns (bx..), input cols = output rows (by..)
  for (int j = ty; j < TT; j += 8) {
    uint32_t in_r = bx + j, in_c = by + tx;
    if (in_r < out_cols && in_c < out_rows) {
      size_t i = (size_t)in_r * in_cols + in_c;
      if (are) tre[j][tx] = are[i];
      if (aim) tim[j][tx] = aim[i];
    }
  }
  __syncthreads();
  for (int j = ty; j < TT; j += 8) {
    uint32_t oc = bx + tx, orow = by + j;
    if (oc < out_cols && orow < out_rows) {
      size_t o = (size_t)orow * out_cols + oc;
      if (ore) ore[o] = tre[tx][j];
      if (oim) oim[o] = imsign * tim[tx][j];
    }
  }
}

// ------------------------------------------------------------ set diagonal / identity
// CuSetMatrixProcedures.h:31-97: diag entries = (re, im), all others 0.
__global__ void k_set_diagonal(double* __restrict__ re, double* __restrict__ im, uint32_t cols,
                               uint32_t rows, double dre, double dim) {
  size_t n = (size_t)cols * rows;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    uint32_t r = (uint32_t)(i / cols), c = (uint32_t)(i - (size_t)r * cols);
    bool d = r == c;
    if (re) re[i] = d ? dre : 0.0;
    if (im) im[i] = d ? dim : 0.0;
  }
}

// ------------------------------------------------------------ binary per-plane ops
// add (CuAdditionProcedures.h:88-107), substract (CuSubstractionProcedures.h:92-110),
// hadamard (CuHadamardProductProcedures.h:84-106), hadamardVec
// (CuPartialHadamardProductProcedures.h:87-109: b is broadcast by row).
template <int OP>
__global__ void k_binary(double* __restrict__ ore, double* __restrict__ oim,
                         const double* __restrict__ are, const double* __restrict__ aim,
                         const double* __restrict__ bre, const double* __restrict__ bim,
                         size_t n, uint32_t cols) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    size_t j = OP == 3 ? i / cols : i;
    if (ore) {
      double a = are[i], b = bre[j];
      ore[i] = OP == 0 ? a + b : OP == 1 ? a - b : a * b;
    }
    if (oim) {
      double a = aim[i], b = bim[j];
      oim[i] = OP == 0 ? a + b : OP == 1 ? a - b : a * b;
    }
  }
}

// ------------------------------------------------------------ constant scaling
// multiplyConstant: re*alpha_re, im*alpha_im per plane (CuMultiplicationProcedures.h:56-116);
// multiplyReConstant: the re plane only (CuMatrixKernels.cu:119-123).
__global__ void k_scale(double* __restrict__ ore, double* __restrict__ oim,
                        const double* __restrict__ are, const double* __restrict__ aim, size_t n,
                        double sre, double sim) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    if (ore) ore[i] = are[i] * sre;
    if (oim) oim[i] = aim[i] * sim;
  }
}

// ------------------------------------------------------------ tensor (Kronecker) product
// CuTensorProductDimProcedures.h:26-48: out(x, y) = A(x / bc, y / br) * B(x % bc, y % br),
// re plane only.
__global__ void k_tensor(double* __restrict__ ore, const double* __restrict__ are,
                         const double* __restrict__ bre, uint32_t W, uint32_t H, uint32_t ocols,
                         uint32_t acols, uint32_t bcols, uint32_t bc, uint32_t br) {
  size_t n = (size_t)W * H;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    uint32_t y = (uint32_t)(i / W), x = (uint32_t)(i - (size_t)y * W);
    double v0 = are[(size_t)(x / bc) + (size_t)acols * (y / br)];
    double v1 = bre[(size_t)(x % bc) + (size_t)bcols * (y % br)];
    ore[(size_t)x + (size_t)ocols * y] = v0 * v1;
  }
}

// ------------------------------------------------------------ set / get vector
// CuVectorUtils.h:27-69 over the grid of v (GenericProceduresApi.h:610-641):
// V[y * V.columns + column + x] <-> v[y * v.columns + x] for y < length.
template <bool SET>
__global__ void k_vector(double* __restrict__ Vre, double* __restrict__ Vim, uint32_t Vcols,
                         double* __restrict__ vre, double* __restrict__ vim, uint32_t vcols,
                         uint32_t vrows, uint32_t column, uint32_t length) {
  size_t n = (size_t)vcols * (vrows < length ? vrows : length);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    uint32_t y = (uint32_t)(i / vcols), x = (uint32_t)(i - (size_t)y * vcols);
    size_t i1 = (size_t)y * Vcols + column + x, i2 = (size_t)y * vcols + x;
    if (Vre && vre) {
      if (SET) Vre[i1] = vre[i2];
      else vre[i2] = Vre[i1];
    }
    if (Vim && vim) {
      if (SET) Vim[i1] = vim[i2];
      else vim[i2] = Vim[i1];
    }
  }
}

// ------------------------------------------------------------ region forms
// add / substract / addSubstract over the columns x rows region at the top
// left of the matrices (CuProceduresApi.cpp:309-325 launch their kernels on
// a columns x rows grid; CuAddSubstractionProcedures.h:48-62: out += a - b).
// OP 0 add, 1 substract, 4 addSubstract.
template <int OP>
__global__ void k_binary_region(double* __restrict__ ore, double* __restrict__ oim,
                                const double* __restrict__ are, const double* __restrict__ aim,
                                const double* __restrict__ bre, const double* __restrict__ bim,
                                uint32_t ocols, uint32_t acols, uint32_t bcols, uint32_t W,
                                uint32_t H) {
  const size_t n = (size_t)W * H;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t y = (uint32_t)(i / W), x = (uint32_t)(i - (size_t)y * W);
    const size_t io = (size_t)y * ocols + x, ia = (size_t)y * acols + x, ib = (size_t)y * bcols + x;
    if (ore) {
      const double a = are[ia], b = bre[ib];
      ore[io] = OP == 0 ? a + b : OP == 1 ? a - b : ore[io] + (a - b);
    }
    if (oim) {
      const double a = aim[ia], b = bim[ib];
      oim[io] = OP == 0 ? a + b : OP == 1 ? a - b : oim[io] + (a - b);
    }
  }
}

// out(x, y) = a(y, x) for x < W, y < H (transposeEx: the region the
// reference's launch grid and its MatrixEx bound select,
// CuTransposeProcedures.h:27-80)
__global__ void k_transpose_region(double* __restrict__ ore, double* __restrict__ oim,
                                   const double* __restrict__ are, const double* __restrict__ aim,
                                   uint32_t ocols, uint32_t acols, uint32_t W, uint32_t H) {
  const size_t n = (size_t)W * H;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t y = (uint32_t)(i / W), x = (uint32_t)(i - (size_t)y * W);
    const size_t io = (size_t)y * ocols + x, ia = (size_t)x * acols + y;
    if (ore) ore[io] = are[ia];
    if (oim) oim[io] = aim[ia];
  }
}

// General row-mapped product for the periodic and region forms of
// dotProduct: for x in [x0, x0 + W) and output rows y in [y0, y0 + H),
//   out(x, y) = sum_{k < K} A(k, arow(y)) * B(x, brow(y) + k)
// with arow(y) = (y - y0) % P + ay0 and brow(y) = ((y - y0) / P) * Q + by0
// (a plain region: P = H, Q = 0); rows whose (y - y0) % P is not below S are
// left alone (the periodic form's rows outside each period's block).  The k order and the per-plane
// accumulation follow cuda_generic_dotProductRealExOffset
// (CuDotProductGenericProcedures.h:86-104).  MODE 0: complex, 1: re only.
template <int MODE>
__global__ void k_dot_mapped(double* __restrict__ ore, double* __restrict__ oim,
                             const double* __restrict__ are, const double* __restrict__ aim,
                             const double* __restrict__ bre, const double* __restrict__ bim,
                             uint32_t ocols, uint32_t acols, uint32_t bcols, uint32_t x0,
                             uint32_t y0, uint32_t W, uint32_t H, uint32_t K, uint32_t P,
                             uint32_t Q, uint32_t ay0, uint32_t by0, uint32_t S) {
  const size_t n = (size_t)W * H;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t dy = (uint32_t)(i / W), dx = (uint32_t)(i - (size_t)dy * W);
    const uint32_t x = x0 + dx, y = y0 + dy;
    if (dy % P >= S) continue;
    const size_t ra = (size_t)(dy % P + ay0) * acols;
    const size_t rb = (size_t)((dy / P) * Q + by0);
    double sr = 0.0, si = 0.0;
    for (uint32_t k = 0; k < K; ++k) {
      const size_t ia = ra + k, ib = (rb + k) * bcols + x;
      if (MODE == 0) {
        sr += are[ia] * bre[ib];
        sr -= aim[ia] * bim[ib];
        si += are[ia] * bim[ib];
        si += aim[ia] * bre[ib];
      } else {
        sr += are[ia] * bre[ib];
      }
    }
    ore[(size_t)y * ocols + x] = sr;
    if (MODE == 0) oim[(size_t)y * ocols + x] = si;
  }
}

}  // namespace

oaph_status launch_binary_region(hipStream_t s, int op, oaph_matrix* out, const oaph_matrix* a,
                                 const oaph_matrix* b, uint32_t W, uint32_t H) {
  const size_t n = (size_t)W * H;
  if (n == 0) return OAPH_OK;
  dim3 g(grid_for(n, 4)), t(kThreads);
  double* ore = out->reValues;
  double* oim = out->imValues;
  if (op == 0)
    hipLaunchKernelGGL(k_binary_region<0>, g, t, 0, s, ore, oim, a->reValues, a->imValues,
                       b->reValues, b->imValues, out->columns, a->columns, b->columns, W, H);
  else if (op == 1)
    hipLaunchKernelGGL(k_binary_region<1>, g, t, 0, s, ore, oim, a->reValues, a->imValues,
                       b->reValues, b->imValues, out->columns, a->columns, b->columns, W, H);
  else
    hipLaunchKernelGGL(k_binary_region<4>, g, t, 0, s, ore, oim, a->reValues, a->imValues,
                       b->reValues, b->imValues, out->columns, a->columns, b->columns, W, H);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_transpose_region(hipStream_t s, oaph_matrix* out, const oaph_matrix* a,
                                    uint32_t W, uint32_t H) {
  const size_t n = (size_t)W * H;
  if (n == 0) return OAPH_OK;
  hipLaunchKernelGGL(k_transpose_region, dim3(grid_for(n, 4)), dim3(kThreads), 0, s,
                     out->reValues, a->imValues ? out->imValues : nullptr, a->reValues,
                     a->imValues, out->columns, a->columns, W, H);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_dot_mapped(hipStream_t s, oaph_matrix* out, const oaph_matrix* a,
                              const oaph_matrix* b, const DotMap& mp) {
  const size_t n = (size_t)mp.W * mp.H;
  if (n == 0) return OAPH_OK;
  dim3 g(grid_for(n)), t(kThreads);
  if (out->imValues)
    hipLaunchKernelGGL(k_dot_mapped<0>, g, t, 0, s, out->reValues, out->imValues, a->reValues,
                       a->imValues, b->reValues, b->imValues, out->columns, a->columns,
                       b->columns, mp.x0, mp.y0, mp.W, mp.H, mp.K, mp.P, mp.Q, mp.ay0, mp.by0, mp.S);
  else
    hipLaunchKernelGGL(k_dot_mapped<1>, g, t, 0, s, out->reValues, nullptr, a->reValues, nullptr,
                       b->reValues, nullptr, out->columns, a->columns, b->columns, mp.x0, mp.y0,
                       mp.W, mp.H, mp.K, mp.P, mp.Q, mp.ay0, mp.by0, mp.S);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_transpose(hipStream_t s, oaph_matrix* out, const oaph_matrix* a, bool conj) {
  dim3 grid((out->columns + TT - 1) / TT, (out->rows + TT - 1) / TT);
  OAPH_CHECK(grid.y <= 65535, OAPH_ERR_UNSUPPORTED);
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, s, out->reValues,
                     a->imValues ? out->imValues : nullptr, a->reValues, a->imValues,
                     out->columns, out->rows, conj ? -1.0 : 1.0);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_set_diagonal(hipStream_t s, oaph_matrix* m, double re, double im) {
  size_t n = numel(m);
  hipLaunchKernelGGL(k_set_diagonal, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, m->reValues,
                     m->imValues, m->columns, m->rows, re, im);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_binary(hipStream_t s, int op, oaph_matrix* out, const oaph_matrix* a,
                          const oaph_matrix* b) {
  size_t n = numel(out);
  dim3 g(grid_for(n, 4)), t(kThreads);
  double* ore = out->reValues;
  double* oim = out->imValues;
  switch (op) {
    case 0: hipLaunchKernelGGL(k_binary<0>, g, t, 0, s, ore, oim, a->reValues, a->imValues, b->reValues, b->imValues, n, out->columns); break;
    case 1: hipLaunchKernelGGL(k_binary<1>, g, t, 0, s, ore, oim, a->reValues, a->imValues, b->reValues, b->imValues, n, out->columns); break;
    case 2: hipLaunchKernelGGL(k_binary<2>, g, t, 0, s, ore, oim, a->reValues, a->imValues, b->reValues, b->imValues, n, out->columns); break;
    default: hipLaunchKernelGGL(k_binary<3>, g, t, 0, s, ore, oim, a->reValues, a->imValues, b->reValues, b->imValues, n, out->columns); break;
  }
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_scale(hipStream_t s, oaph_matrix* out, const oaph_matrix* a, double re,
                         double im, bool re_only) {
  size_t n = numel(out);
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, out->reValues,
                     re_only ? nullptr : out->imValues, a->reValues, a->imValues, n, re, im);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_tensor(hipStream_t s, oaph_matrix* out, const oaph_matrix* a,
                          const oaph_matrix* b, const uint32_t dims[3][2]) {
  uint32_t W = dims[0][0], H = dims[0][1];
  size_t n = (size_t)W * H;
  hipLaunchKernelGGL(k_tensor, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, out->reValues,
                     a->reValues, b->reValues, W, H, out->columns, a->columns, b->columns,
                     dims[2][0], dims[2][1]);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_set_vector(hipStream_t s, oaph_matrix* V, uint32_t column, const oaph_matrix* v,
                              uint32_t length) {
  size_t n = numel(v);
  hipLaunchKernelGGL(k_vector<true>, dim3(grid_for(n)), dim3(kThreads), 0, s, V->reValues,
                     V->imValues, V->columns, v->reValues, v->imValues, v->columns, v->rows,
                     column, length);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_get_vector(hipStream_t s, oaph_matrix* v, uint32_t length, const oaph_matrix* V,
                              uint32_t column) {
  size_t n = numel(v);
  hipLaunchKernelGGL(k_vector<false>, dim3(grid_for(n)), dim3(kThreads), 0, s, V->reValues,
                     V->imValues, V->columns, v->reValues, v->imValues, v->columns, v->rows,
                     column, length);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

}  // namespace oaph
