// procedures.hip -- C ABI of oap::CuProceduresApi (oapMatrixCuda/CuProceduresApi.h:45-309).
//
// Every entry point validates shapes and planes on the host before it
// enqueues a kernel: the reference only debug-asserts
// (GenericValidationApi.h:51-81), so a bad column index there writes out of
// bounds; here it is OAPH_ERR_SHAPE.
#include "internal.h"

using namespace oaph;

namespace {

inline bool same_shape(const oaph_matrix* a, const oaph_matrix* b) {
  return a->columns == b->columns && a->rows == b->rows;
}

// every plane of `out` must exist in `in`
inline bool planes_cover(const oaph_matrix* out, const oaph_matrix* in) {
  return (!out->reValues || in->reValues) && (!out->imValues || in->imValues);
}

oaph_status check_api(oaph_api* api) {
  if (!api) return OAPH_ERR_ARG;
  return ensure_init();
}

}  // namespace

namespace oaph {
oaph_status enqueue_reduction(oaph_api* api, const oaph_matrix* m, bool sumsq) {
  size_t n = numel(m);
  int blocks = reduce_blocks_for(n);
  OAPH_TRY(api->partials.reserve((size_t)blocks * 2 * sizeof(double)));
  double* result = api->scalars.as<double>();
  if (sumsq)
    return launch_sumsq(api->stream, m->reValues, m->imValues, n, api->partials.as<double>(),
                        result);
  return launch_sum2(api->stream, m->reValues, m->imValues, n, api->partials.as<double>(), result);
}
}  // namespace oaph

extern "C" {

oaph_status oaph_dot_product(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                             const oaph_matrix* b) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a && b, OAPH_ERR_ARG);
  // check_dotProduct (GenericValidationApi.h:51-81)
  OAPH_CHECK(a->columns == b->rows && out->columns == b->columns && out->rows == a->rows,
             OAPH_ERR_SHAPE);
  // Im-only output: the reference writes into the absent re plane
  // (CuDotProductSpecificProcedures.h:45-62); rejected here.
  OAPH_CHECK(out->reValues != nullptr, OAPH_ERR_PLANE);
  if (out->imValues) {
    // a real left operand (no im plane) with a complex right operand is
    // accepted as A + 0i: the solver's vectors are complex even for a real
    // operator (an extension; the reference dereferences the missing plane)
    OAPH_CHECK(a->reValues && b->reValues && b->imValues, OAPH_ERR_PLANE);
  } else {
    OAPH_CHECK(a->reValues && b->reValues, OAPH_ERR_PLANE);
  }
  return launch_gemm(api->stream, out, a, b, out->columns, out->rows, a->columns, false);
}

oaph_status oaph_dot_product_shared(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                    const oaph_matrix* b) {
  return oaph_dot_product(api, out, a, b);
}

oaph_status oaph_dot_product_dims(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                  const oaph_matrix* b, const uint32_t dims[3][2]) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a && b && dims, OAPH_ERR_ARG);
  // check_dotProduct(dims) (GenericValidationApi.h:93-118)
  OAPH_CHECK(dims[1][0] == dims[2][1] && dims[0][0] == dims[2][0] && dims[0][1] == dims[1][1],
             OAPH_ERR_SHAPE);
  OAPH_CHECK(dims[0][0] <= out->columns && dims[0][1] <= out->rows && dims[1][0] <= a->columns &&
                 dims[1][1] <= a->rows && dims[2][0] <= b->columns && dims[2][1] <= b->rows,
             OAPH_ERR_SHAPE);
  if (out->reValues && out->imValues)
    OAPH_CHECK(a->reValues && a->imValues && b->reValues && b->imValues, OAPH_ERR_PLANE);
  else if (out->reValues)
    OAPH_CHECK(a->reValues && b->reValues, OAPH_ERR_PLANE);
  else
    OAPH_CHECK(a->imValues && b->imValues, OAPH_ERR_PLANE);
  if (dims[0][0] == 0 || dims[0][1] == 0) return OAPH_OK;
  return launch_gemm(api->stream, out, a, b, dims[0][0], dims[0][1], dims[1][0], true);
}

static oaph_status transpose_common(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                    bool conj) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a, OAPH_ERR_ARG);
  OAPH_CHECK(out->rows == a->columns && out->columns == a->rows, OAPH_ERR_SHAPE);
  OAPH_CHECK(planes_cover(out, a), OAPH_ERR_PLANE);
  // vector <-> row: plain device copy in the reference (CuProceduresApi.cpp:277-280),
  // so the conjugate variant does not negate there either.
  if (out->columns == 1 || out->rows == 1) {
    size_t bytes = numel(out) * sizeof(double);
    if (out->reValues)
      OAPH_HIP(hipMemcpyAsync(out->reValues, a->reValues, bytes, hipMemcpyDeviceToDevice,
                              api->stream));
    if (out->imValues)
      OAPH_HIP(hipMemcpyAsync(out->imValues, a->imValues, bytes, hipMemcpyDeviceToDevice,
                              api->stream));
    return OAPH_OK;
  }
  return launch_transpose(api->stream, out, a, conj);
}

oaph_status oaph_transpose(oaph_api* api, oaph_matrix* out, const oaph_matrix* a) {
  return transpose_common(api, out, a, false);
}

oaph_status oaph_conjugate_transpose(oaph_api* api, oaph_matrix* out, const oaph_matrix* a) {
  return transpose_common(api, out, a, true);
}

oaph_status oaph_tensor_product(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                const oaph_matrix* b, const uint32_t dims[3][2]) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a && b && dims, OAPH_ERR_ARG);
  // check_tensorProduct (GenericValidationApi.h:205-230)
  OAPH_CHECK(dims[0][0] == dims[1][0] * dims[2][0] && dims[0][1] == dims[1][1] * dims[2][1],
             OAPH_ERR_SHAPE);
  OAPH_CHECK(dims[0][0] <= out->columns && dims[0][1] <= out->rows && dims[1][0] <= a->columns &&
                 dims[1][1] <= a->rows && dims[2][0] <= b->columns && dims[2][1] <= b->rows,
             OAPH_ERR_SHAPE);
  // the reference computes the re plane only and only for a re-only output
  // (CuTensorProductDimProcedures.h:51-63, 98-112); other plane layouts are a no-op
  if (!out->reValues || out->imValues || !a->reValues || !b->reValues) return OAPH_OK;
  if (dims[0][0] == 0 || dims[0][1] == 0) return OAPH_OK;
  return launch_tensor(api->stream, out, a, b, dims);
}

static oaph_status binary_common(oaph_api* api, int op, oaph_matrix* out, const oaph_matrix* a,
                                 const oaph_matrix* b) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a && b, OAPH_ERR_ARG);
  OAPH_CHECK(same_shape(out, a), OAPH_ERR_SHAPE);
  if (op == 3) {
    OAPH_CHECK(b->rows == out->rows && b->columns >= 1, OAPH_ERR_SHAPE);
  } else {
    OAPH_CHECK(same_shape(out, b), OAPH_ERR_SHAPE);
  }
  OAPH_CHECK(planes_cover(out, a) && planes_cover(out, b), OAPH_ERR_PLANE);
  if (op == 3 && b->columns != 1) return OAPH_ERR_SHAPE;
  return launch_binary(api->stream, op, out, a, b);
}

oaph_status oaph_add(oaph_api* api, oaph_matrix* out, const oaph_matrix* a, const oaph_matrix* b) {
  return binary_common(api, 0, out, a, b);
}
oaph_status oaph_substract(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                           const oaph_matrix* b) {
  return binary_common(api, 1, out, a, b);
}
oaph_status oaph_hadamard_product(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                  const oaph_matrix* b) {
  return binary_common(api, 2, out, a, b);
}
oaph_status oaph_hadamard_product_vec(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                      const oaph_matrix* b) {
  return binary_common(api, 3, out, a, b);
}

oaph_status oaph_multiply_re_constant(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                      double re) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a, OAPH_ERR_ARG);
  OAPH_CHECK(same_shape(out, a), OAPH_ERR_SHAPE);
  if (!out->reValues) return OAPH_OK;
  OAPH_CHECK(a->reValues, OAPH_ERR_PLANE);
  return launch_scale(api->stream, out, a, re, 0.0, true);
}

oaph_status oaph_multiply_constant(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                   double re, double im) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a, OAPH_ERR_ARG);
  OAPH_CHECK(same_shape(out, a), OAPH_ERR_SHAPE);
  OAPH_CHECK(planes_cover(out, a), OAPH_ERR_PLANE);
  return launch_scale(api->stream, out, a, re, im, false);
}

oaph_status oaph_set_diagonal(oaph_api* api, oaph_matrix* m, double re, double im) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(m, OAPH_ERR_ARG);
  return launch_set_diagonal(api->stream, m, re, im);
}
oaph_status oaph_set_identity(oaph_api* api, oaph_matrix* m) {
  return oaph_set_diagonal(api, m, 1.0, 0.0);
}
oaph_status oaph_set_zero_matrix(oaph_api* api, oaph_matrix* m) {
  return oaph_set_diagonal(api, m, 0.0, 0.0);
}

oaph_status oaph_set_vector(oaph_api* api, oaph_matrix* V, uint32_t column, const oaph_matrix* v,
                            uint32_t length) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(V && v, OAPH_ERR_ARG);
  uint32_t rows = v->rows < length ? v->rows : length;
  OAPH_CHECK((uint64_t)column + v->columns <= V->columns && rows <= V->rows, OAPH_ERR_SHAPE);
  if (rows == 0) return OAPH_OK;
  return launch_set_vector(api->stream, V, column, v, length);
}

oaph_status oaph_get_vector(oaph_api* api, oaph_matrix* v, uint32_t length, const oaph_matrix* V,
                            uint32_t column) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(V && v, OAPH_ERR_ARG);
  uint32_t rows = v->rows < length ? v->rows : length;
  OAPH_CHECK((uint64_t)column + v->columns <= V->columns && rows <= V->rows, OAPH_ERR_SHAPE);
  if (rows == 0) return OAPH_OK;
  return launch_get_vector(api->stream, v, length, V, column);
}

static oaph_status reduce_common(oaph_api* api, const oaph_matrix* m, bool sumsq) {
  OAPH_TRY(enqueue_reduction(api, m, sumsq));
  double* result = api->scalars.as<double>();
  OAPH_HIP(hipMemcpyAsync(api->host_scalars, result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          api->stream));
  OAPH_HIP(hipStreamSynchronize(api->stream));
  return OAPH_OK;
}

oaph_status oaph_magnitude(oaph_api* api, double* out, const oaph_matrix* m) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && m, OAPH_ERR_ARG);
  OAPH_TRY(reduce_common(api, m, true));
  *out = std::sqrt(api->host_scalars[0]);
  return OAPH_OK;
}

oaph_status oaph_sum(oaph_api* api, double* re, double* im, const oaph_matrix* m) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(m, OAPH_ERR_ARG);
  OAPH_TRY(reduce_common(api, m, false));
  if (re) *re = m->reValues ? api->host_scalars[0] : 0.0;
  if (im) *im = m->imValues ? api->host_scalars[1] : 0.0;
  return OAPH_OK;
}

oaph_status oaph_qrgr(oaph_api* api, oaph_matrix* Q, oaph_matrix* R, const oaph_matrix* H,
                      oaph_matrix*, oaph_matrix*, oaph_matrix*, oaph_matrix*) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(Q && R && H, OAPH_ERR_ARG);
  OAPH_CHECK(H->columns == H->rows && same_shape(Q, H) && same_shape(R, H), OAPH_ERR_SHAPE);
  OAPH_CHECK(H->reValues && Q->reValues && R->reValues, OAPH_ERR_PLANE);
  if (H->imValues) OAPH_CHECK(Q->imValues && R->imValues, OAPH_ERR_PLANE);
  OAPH_CHECK(R->reValues != H->reValues && Q->reValues != H->reValues, OAPH_ERR_ARG);
  return launch_qrgr(api->stream, Q, R, H);
}

oaph_status oaph_is_upper_triangular(oaph_api* api, const oaph_matrix* m, int* result) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(m && result, OAPH_ERR_ARG);
  OAPH_CHECK(m->columns == m->rows, OAPH_ERR_SHAPE);
  int* dflag = reinterpret_cast<int*>(api->scalars.as<double>() + 8);
  OAPH_TRY(launch_is_upper_triangular(api->stream, m, dflag));
  OAPH_HIP(hipMemcpyAsync(api->host_scalars, dflag, sizeof(int), hipMemcpyDeviceToHost,
                          api->stream));
  OAPH_HIP(hipStreamSynchronize(api->stream));
  *result = *reinterpret_cast<int*>(api->host_scalars);
  return OAPH_OK;
}

oaph_status oaph_calc_triangular_h(oaph_api* api, oaph_matrix* H, oaph_matrix* Q, oaph_matrix* R,
                                   oaph_matrix*, oaph_matrix*, oaph_matrix*, oaph_matrix*,
                                   oaph_matrix*, oaph_matrix*) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(H && Q, OAPH_ERR_ARG);
  OAPH_CHECK(H->columns == H->rows && same_shape(Q, H), OAPH_ERR_SHAPE);
  OAPH_CHECK(H->reValues && Q->reValues, OAPH_ERR_PLANE);
  if (H->imValues) OAPH_CHECK(Q->imValues, OAPH_ERR_PLANE);
  (void)R;
  size_t mm = numel(H);
  OAPH_TRY(api->work.reserve(6 * mm * sizeof(double)));
  // CUDA_calcUTMatrix_STEPS = 10000 (CuTriangularH.h:29)
  return launch_calc_triangular_h(api->stream, H, Q, api->work.as<double>(), 10000);
}

// ------------------------------------------------------------------ dims
// forms and MatrixEx procedures

// the (columns, rows) forms of hadamardProduct(Vec) and tensorProduct check
// the given output size and run over the whole matrices
// (CuProceduresApi.cpp:136-192, GenericValidationApi.h:181-270)
oaph_status oaph_hadamard_product_cr(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                     const oaph_matrix* b, uint32_t columns, uint32_t rows) {
  OAPH_CHECK(out && a && b, OAPH_ERR_ARG);
  OAPH_CHECK(columns == out->columns && rows == out->rows, OAPH_ERR_SHAPE);
  return oaph_hadamard_product(api, out, a, b);
}
oaph_status oaph_hadamard_product_vec_cr(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                         const oaph_matrix* b, uint32_t columns, uint32_t rows) {
  OAPH_CHECK(out && a && b, OAPH_ERR_ARG);
  OAPH_CHECK(columns == out->columns && rows == out->rows, OAPH_ERR_SHAPE);
  return oaph_hadamard_product_vec(api, out, a, b);
}
oaph_status oaph_tensor_product_cr(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                   const oaph_matrix* b, uint32_t columns, uint32_t rows) {
  OAPH_CHECK(out && a && b, OAPH_ERR_ARG);
  OAPH_CHECK(columns <= out->columns && rows <= out->rows, OAPH_ERR_SHAPE);
  const uint32_t dims[3][2] = {{columns, rows}, {a->columns, a->rows}, {b->columns, b->rows}};
  return oaph_tensor_product(api, out, a, b, dims);
}

// add / substract / addSubstract over the columns x rows region
// (CuProceduresApi.cpp:309-325; addSubstract: out += a - b,
// CuAddSubstractionProcedures.h:48-62)
static oaph_status region_common(oaph_api* api, int op, oaph_matrix* out, const oaph_matrix* a,
                                 const oaph_matrix* b, uint32_t columns, uint32_t rows) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a && b, OAPH_ERR_ARG);
  OAPH_CHECK(columns <= out->columns && columns <= a->columns && columns <= b->columns &&
                 rows <= out->rows && rows <= a->rows && rows <= b->rows,
             OAPH_ERR_SHAPE);
  OAPH_CHECK(planes_cover(out, a) && planes_cover(out, b), OAPH_ERR_PLANE);
  return launch_binary_region(api->stream, op, out, a, b, columns, rows);
}
oaph_status oaph_add_cr(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                        const oaph_matrix* b, uint32_t columns, uint32_t rows) {
  return region_common(api, 0, out, a, b, columns, rows);
}
oaph_status oaph_substract_cr(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                              const oaph_matrix* b, uint32_t columns, uint32_t rows) {
  return region_common(api, 1, out, a, b, columns, rows);
}
oaph_status oaph_add_substract(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                               const oaph_matrix* b, uint32_t columns, uint32_t rows) {
  return region_common(api, 4, out, a, b, columns, rows);
}

// planes of a product output: complex, or re only (im-only outputs are
// rejected as in oaph_dot_product)
static oaph_status product_planes(const oaph_matrix* out, const oaph_matrix* a,
                                  const oaph_matrix* b) {
  OAPH_CHECK(out->reValues != nullptr, OAPH_ERR_PLANE);
  if (out->imValues)
    OAPH_CHECK(a->reValues && a->imValues && b->reValues && b->imValues, OAPH_ERR_PLANE);
  else
    OAPH_CHECK(a->reValues && b->reValues, OAPH_ERR_PLANE);
  return OAPH_OK;
}

// addDotProduct(output, a, b[, columns, rows]): the reference kernel writes
// output = a b (CuAddDotProductProcedures.h:26-84: "=", not "+=")
oaph_status oaph_add_dot_product(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                 const oaph_matrix* b) {
  return oaph_dot_product(api, out, a, b);
}

// dotProductPeriodic (CuDotProductPeriodicProcedures.h:65-85): with
// K = a.columns, out(x, y) = sum_k a(k, y % K) b(x, (y / K) K + k)
oaph_status oaph_dot_product_periodic(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                      const oaph_matrix* b) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a && b, OAPH_ERR_ARG);
  const uint32_t K = a->columns;
  // check_dotProductPeriodic (GenericValidationApi.h:120-141); rows of a
  // read: y % K < a.rows
  OAPH_CHECK(K > 0 && b->rows % K == 0 && out->columns == b->columns && a->rows >= K &&
                 out->rows % a->rows == 0 && out->rows <= b->rows,
             OAPH_ERR_SHAPE);
  OAPH_TRY(product_planes(out, a, b));
  DotMap mp;
  mp.W = out->columns;
  mp.H = out->rows;
  mp.K = K;
  mp.P = K;
  mp.Q = K;
  mp.S = K;
  return launch_dot_mapped(api->stream, out, a, b, mp);
}

// dotProductDimPeriodic (CuDotProductDimPeriodicProcedures.h:27-45): with
// W = dims[0][0], H = dims[0][1], K = dims[1][0] and P = periodic_rows, for
// every output row y with y % P < H and x < W:
// out(x, y) = sum_k a(k, y % P) b(x, (y / P) K + k)
oaph_status oaph_dot_product_dim_periodic(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                          const oaph_matrix* b, const uint32_t dims[3][2],
                                          uint32_t periodic_rows) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a && b && dims && periodic_rows > 0, OAPH_ERR_ARG);
  const uint32_t W = dims[0][0], H = dims[0][1], K = dims[1][0], P = periodic_rows;
  OAPH_CHECK(W <= out->columns && W <= b->columns && K <= a->columns &&
                 (H < P ? H : P) <= a->rows,
             OAPH_ERR_SHAPE);
  // rows of b read: ((out.rows - 1) / P) K + K
  OAPH_CHECK(out->rows == 0 || (uint64_t)((out->rows - 1) / P) * K + K <= b->rows,
             OAPH_ERR_SHAPE);
  OAPH_TRY(product_planes(out, a, b));
  DotMap mp;
  mp.W = W;
  mp.H = out->rows;
  mp.K = K;
  mp.P = P;
  mp.Q = K;
  mp.S = H;
  return launch_dot_mapped(api->stream, out, a, b, mp);
}

// dotProductEx(output, a, b, matrixEx[, columns, rows]): the reference
// launches a kernel "CUDAKernel_DotProductEx" that its kernel file does not
// define (CuProceduresApi.cpp:201-209), so the call has no defined result
// there.  Here: the product restricted to the output region of matrixEx,
// out(x, y) = sum_{k < a.columns} a(k, y) b(x, k) for
// column <= x < column + columns, row <= y < row + rows.
oaph_status oaph_dot_product_ex(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                const oaph_matrix* b, const oaph_matrix_ex* ex) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a && b && ex, OAPH_ERR_ARG);
  OAPH_CHECK(a->columns == b->rows, OAPH_ERR_SHAPE);
  OAPH_CHECK((uint64_t)ex->column + ex->columns <= out->columns &&
                 (uint64_t)ex->column + ex->columns <= b->columns &&
                 (uint64_t)ex->row + ex->rows <= out->rows &&
                 (uint64_t)ex->row + ex->rows <= a->rows,
             OAPH_ERR_SHAPE);
  OAPH_TRY(product_planes(out, a, b));
  DotMap mp;
  mp.x0 = ex->column;
  mp.y0 = ex->row;
  mp.W = ex->columns;
  mp.H = ex->rows;
  mp.K = a->columns;
  mp.P = ex->rows ? ex->rows : 1;
  mp.Q = 0;
  mp.ay0 = ex->row;
  mp.S = mp.P;
  return launch_dot_mapped(api->stream, out, a, b, mp);
}

// transposeEx(output, a, matrixEx) (CuProceduresApi.cpp:260-266,
// CuTransposeProcedures.h:27-80): the kernel runs on the
// SetThreadsBlocks(columns, rows) grid and a thread (x, y) writes
// out(x, y) = a(y, x) when y < row + rows and x < column + columns -- the
// region [0, min(grid width, column + columns)) x [0, min(grid height,
// row + rows)).
oaph_status oaph_transpose_ex(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                              const oaph_matrix_ex* ex) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && a && ex, OAPH_ERR_ARG);
  OAPH_CHECK(planes_cover(out, a), OAPH_ERR_PLANE);
  if (ex->columns == 0 || ex->rows == 0) return OAPH_OK;
  uint32_t blocks[2], threads[2];
  OAPH_TRY(oaph_set_threads_blocks(blocks, threads, ex->columns, ex->rows, 1024));
  const uint64_t gw = (uint64_t)blocks[0] * threads[0], gh = (uint64_t)blocks[1] * threads[1];
  const uint64_t ex_w = (uint64_t)ex->column + ex->columns, ex_h = (uint64_t)ex->row + ex->rows;
  const uint32_t W = (uint32_t)(gw < ex_w ? gw : ex_w), H = (uint32_t)(gh < ex_h ? gh : ex_h);
  OAPH_CHECK(W <= out->columns && H <= out->rows && W <= a->rows && H <= a->columns,
             OAPH_ERR_SHAPE);
  return launch_transpose_region(api->stream, out, a, W, H);
}

// magnitude2(out, m) = sum re^2 + im^2 (CuProceduresApi.cpp:401-420; the Opt
// and OptVer2 forms compute the same sum with other kernels)
oaph_status oaph_magnitude2(oaph_api* api, double* out, const oaph_matrix* m) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(out && m, OAPH_ERR_ARG);
  OAPH_TRY(enqueue_reduction(api, m, true));
  OAPH_HIP(hipMemcpyAsync(api->host_scalars, api->scalars.as<double>(), sizeof(double),
                          hipMemcpyDeviceToHost, api->stream));
  OAPH_HIP(hipStreamSynchronize(api->stream));
  *out = api->host_scalars[0];
  return OAPH_OK;
}

// calcTriangularHStep (CuProceduresApi.cpp:482-490, CuTriangularH.h:61-77):
// one unshifted QR step H <- R Q with Q the step's factor; an H that is
// already upper triangular is left alone and Q = I (the reference leaves Q
// with the previous contents of an aux matrix there)
oaph_status oaph_calc_triangular_h_step(oaph_api* api, oaph_matrix* H, oaph_matrix* Q) {
  OAPH_TRY(check_api(api));
  OAPH_CHECK(H && Q, OAPH_ERR_ARG);
  OAPH_CHECK(H->columns == H->rows && same_shape(Q, H), OAPH_ERR_SHAPE);
  OAPH_CHECK(H->reValues && Q->reValues, OAPH_ERR_PLANE);
  if (H->imValues) OAPH_CHECK(Q->imValues, OAPH_ERR_PLANE);
  OAPH_TRY(api->work.reserve(6 * numel(H) * sizeof(double)));
  return launch_calc_triangular_h(api->stream, H, Q, api->work.as<double>(), 1);
}

}  // extern "C"
