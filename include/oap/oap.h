// oap.h -- header-only C++ facade with the reference's names over the C ABI
// (include/oaph.h), so that code written against oap's procedures API and
// Arnoldi package compiles unchanged against the MI355X library:
//
//   math::Matrix, math::MatrixInfo        oapMatrix/Matrix.h:31-44, MatrixInfo.h:16-67
//   oap::cuda::NewDeviceMatrix & co.      oapMatrixCuda/oapCudaMatrixUtils.h:38-74
//   oap::CuProceduresApi                  oapMatrixCuda/CuProceduresApi.h:45-309
//   oap::QRType, VecMultiplicationType,
//   InitVVectorType                       ArnoldiPackage/oapGenericArnoldiApi.h:31-48
//   EigenPair, ArnUtils::*                ArnoldiPackage/ArnoldiUtils.h:28-87
//   CuHArnoldi, CuHArnoldiDefault,
//   CuHArnoldiCallback                    ArnoldiPackage/ArnoldiProcedures.h:33-241,
//                                         ArnoldiProceduresImpl.h:26-74
//
// Errors: the reference aborts on driver errors; here every failing call
// throws oap::Error carrying the oaph_status.
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../oaph.h"

typedef double floatt;
typedef uint32_t uintt;

namespace math {
// same field order and types as oaph_matrix (a reinterpret_cast is valid)
struct Matrix {
  uintt columns;
  uintt rows;
  uintt realColumns;
  uintt realRows;
  floatt* reValues;
  floatt* imValues;
};
static_assert(sizeof(Matrix) == sizeof(oaph_matrix), "math::Matrix must match oaph_matrix");

struct MatrixDim {
  uintt columns;
  uintt rows;
};

struct MatrixInfo {
  bool isRe = true;
  bool isIm = true;
  MatrixDim m_matrixDim{0, 0};
  MatrixInfo() = default;
  MatrixInfo(bool re, bool im, uintt columns, uintt rows) : isRe(re), isIm(im), m_matrixDim{columns, rows} {}
  uintt columns() const { return m_matrixDim.columns; }
  uintt rows() const { return m_matrixDim.rows; }
};
}  // namespace math

// MatrixEx (oapMatrix/MatrixEx.h:26-55): a region of a matrix.  Here a host
// struct passed by pointer (the reference allocates it on the device).
struct MatrixEx {
  uintt column;
  uintt columns;
  uintt row;
  uintt rows;
};
static_assert(sizeof(MatrixEx) == sizeof(oaph_matrix_ex), "MatrixEx must match oaph_matrix_ex");
#define mex_erow(matrixex) matrixex.row + matrixex.rows
#define mex_ecolumn(matrixex) matrixex.column + matrixex.columns
inline bool adjustColumns(MatrixEx& ex, uintt val) {
  if (ex.column > val) return false;
  if (ex.column + ex.columns > val) ex.columns = val - ex.column;
  return true;
}
inline bool adjustRows(MatrixEx& ex, uintt val) {
  if (ex.row > val) return false;
  if (ex.row + ex.rows > val) ex.rows = val - ex.row;
  return true;
}

namespace oap {

class Error : public std::runtime_error {
 public:
  Error(oaph_status s, const std::string& where)
      : std::runtime_error(where + ": " + oaph_status_string(s) + " " + oaph_last_error()),
        status(s) {}
  oaph_status status;
};

inline void check(oaph_status s, const char* where) {
  if (s != OAPH_OK) throw Error(s, where);
}

inline oaph_matrix* c(math::Matrix* m) { return reinterpret_cast<oaph_matrix*>(m); }
inline const oaph_matrix* c(const math::Matrix* m) { return reinterpret_cast<const oaph_matrix*>(m); }

namespace cuda {
inline math::Matrix* NewDeviceMatrix(bool isRe, bool isIm, uintt columns, uintt rows) {
  oaph_matrix* m = nullptr;
  check(oaph_matrix_new(&m, isRe, isIm, columns, rows), "NewDeviceMatrix");
  return reinterpret_cast<math::Matrix*>(m);
}
inline math::Matrix* NewDeviceMatrix(const math::MatrixInfo& info) {
  return NewDeviceMatrix(info.isRe, info.isIm, info.columns(), info.rows());
}
inline math::Matrix* NewDeviceReMatrix(uintt columns, uintt rows) {
  return NewDeviceMatrix(true, false, columns, rows);
}
inline void DeleteDeviceMatrix(math::Matrix* m) { check(oaph_matrix_free(c(m)), "DeleteDeviceMatrix"); }
inline math::MatrixInfo GetMatrixInfo(const math::Matrix* m) {
  oaph_matrix_info i;
  check(oaph_matrix_info_get(c(m), &i), "GetMatrixInfo");
  return math::MatrixInfo(i.isRe != 0, i.isIm != 0, i.columns, i.rows);
}
inline uintt GetColumns(const math::Matrix* m) { return m->columns; }
inline uintt GetRows(const math::Matrix* m) { return m->rows; }
// host arrays (row-major planes) <-> device matrix
inline void CopyHostArraysToDeviceMatrix(math::Matrix* dst, const floatt* re, const floatt* im) {
  check(oaph_matrix_copy_h2d(c(dst), re, im), "CopyHostArraysToDeviceMatrix");
}
inline void CopyDeviceMatrixToHostArrays(floatt* re, floatt* im, const math::Matrix* src) {
  check(oaph_matrix_copy_d2h(re, im, c(src)), "CopyDeviceMatrixToHostArrays");
}
inline void CopyDeviceMatrixToDeviceMatrix(math::Matrix* dst, const math::Matrix* src) {
  check(oaph_matrix_copy_d2d(c(dst), c(src)), "CopyDeviceMatrixToDeviceMatrix");
}
// dims[0] = {{column, columns}, {row, rows}} of dst, dims[1] of src
// (oapCudaMatrixUtils.h:102-116); a host matrix is a math::Matrix whose
// planes are host pointers
inline void CopyDeviceMatrixToHostMatrixDims(math::Matrix* dst, const math::Matrix* src,
                                             uintt dims[2][2][2]) {
  check(oaph_matrix_copy_d2h_dims(c(dst), c(src), dims), "CopyDeviceMatrixToHostMatrixDims");
}
inline void CopyHostMatrixToDeviceMatrixDims(math::Matrix* dst, const math::Matrix* src,
                                             uintt dims[2][2][2]) {
  check(oaph_matrix_copy_h2d_dims(c(dst), c(src), dims), "CopyHostMatrixToDeviceMatrixDims");
}
inline void CopyDeviceMatrixToDeviceMatrixDims(math::Matrix* dst, const math::Matrix* src,
                                               uintt dims[2][2][2]) {
  check(oaph_matrix_copy_d2d_dims(c(dst), c(src), dims), "CopyDeviceMatrixToDeviceMatrixDims");
}
// SetMatrix / SetReMatrix / SetImMatrix (oapCudaMatrixUtils.h:142-146)
inline void SetMatrix(math::Matrix* m, math::Matrix* m1, uintt column, uintt row) {
  check(oaph_matrix_set_submatrix(c(m), c(m1), column, row, 3), "SetMatrix");
}
inline void SetReMatrix(math::Matrix* m, math::Matrix* m1, uintt column, uintt row) {
  check(oaph_matrix_set_submatrix(c(m), c(m1), column, row, 1), "SetReMatrix");
}
inline void SetImMatrix(math::Matrix* m, math::Matrix* m1, uintt column, uintt row) {
  check(oaph_matrix_set_submatrix(c(m), c(m1), column, row, 2), "SetImMatrix");
}
// values by index or (column, row) (oapCudaMatrixUtils.h:182-200)
inline void GetValue(floatt& re, floatt& im, const math::Matrix* m, uintt index) {
  check(oaph_matrix_get_value(c(m), index, &re, &im), "GetValue");
}
inline floatt GetReValue(const math::Matrix* m, uintt index) {
  floatt re = 0, im = 0;
  GetValue(re, im, m, index);
  return re;
}
inline floatt GetReValue(const math::Matrix* m, uintt column, uintt row) {
  return GetReValue(m, column + row * m->columns);
}
inline floatt GetImValue(const math::Matrix* m, uintt index) {
  floatt re = 0, im = 0;
  GetValue(re, im, m, index);
  return im;
}
inline floatt GetImValue(const math::Matrix* m, uintt column, uintt row) {
  return GetImValue(m, column + row * m->columns);
}
inline void SetValue(math::Matrix* m, floatt re, floatt im, uintt column, uintt row) {
  check(oaph_matrix_set_value(c(m), column + row * m->columns, re, im), "SetValue");
}
inline void SetReValue(math::Matrix* m, floatt value, uintt index) {
  floatt re = 0, im = 0;
  GetValue(re, im, m, index);
  check(oaph_matrix_set_value(c(m), index, value, im), "SetReValue");
}
inline void SetReValue(math::Matrix* m, floatt value, uintt column, uintt row) {
  SetReValue(m, value, column + row * m->columns);
}
inline void SetImValue(math::Matrix* m, floatt value, uintt index) {
  floatt re = 0, im = 0;
  GetValue(re, im, m, index);
  check(oaph_matrix_set_value(c(m), index, re, value), "SetImValue");
}
inline void SetImValue(math::Matrix* m, floatt value, uintt column, uintt row) {
  SetImValue(m, value, column + row * m->columns);
}
inline void SetZeroRow(math::Matrix* m, uintt index, bool re = true, bool im = true) {
  check(oaph_matrix_set_zero_row(c(m), index, re ? 1 : 0, im ? 1 : 0), "SetZeroRow");
}
// MatrixEx objects (oapCudaMatrixUtils.h:148-167): host structs here
inline MatrixEx* NewDeviceMatrixEx() { return new MatrixEx{0, 0, 0, 0}; }
inline MatrixEx* NewDeviceMatrixExCopy(const MatrixEx& ex) { return new MatrixEx(ex); }
inline void SetMatrixEx(MatrixEx* dst, const MatrixEx* src) { *dst = *src; }
inline void GetMatrixEx(MatrixEx* dst, const MatrixEx* src) { *dst = *src; }
inline void DeleteDeviceMatrixEx(MatrixEx* ex) { delete ex; }
}  // namespace cuda

class CuProceduresApi {
 public:
  CuProceduresApi() { check(oaph_api_create(&m_api), "CuProceduresApi"); }
  explicit CuProceduresApi(oaph_api* borrowed) : m_api(borrowed), m_owned(false) {}
  virtual ~CuProceduresApi() {
    if (m_owned) oaph_api_destroy(m_api);
  }
  CuProceduresApi(const CuProceduresApi&) = delete;
  CuProceduresApi(CuProceduresApi&&) = delete;
  CuProceduresApi& operator=(const CuProceduresApi&) = delete;
  CuProceduresApi& operator=(CuProceduresApi&&) = delete;

  void dotProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b) {
    check(oaph_dot_product(m_api, c(o), c(a), c(b)), "dotProduct");
  }
  void dotProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt dims[3][2]) {
    check(oaph_dot_product_dims(m_api, c(o), c(a), c(b), dims), "dotProduct(dims)");
  }
  void dotProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt od[2], uintt ad[2],
                  uintt bd[2]) {
    uintt dims[3][2] = {{od[0], od[1]}, {ad[0], ad[1]}, {bd[0], bd[1]}};
    dotProduct(o, a, b, dims);
  }
  void dotProductShared(math::Matrix* o, math::Matrix* a, math::Matrix* b) {
    check(oaph_dot_product_shared(m_api, c(o), c(a), c(b)), "dotProductShared");
  }
  void tensorProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt dims[3][2]) {
    check(oaph_tensor_product(m_api, c(o), c(a), c(b), dims), "tensorProduct");
  }
  void tensorProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b) {
    uintt dims[3][2] = {{o->columns, o->rows}, {a->columns, a->rows}, {b->columns, b->rows}};
    tensorProduct(o, a, b, dims);
  }
  void tensorProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt od[2], uintt ad[2],
                     uintt bd[2]) {
    uintt dims[3][2] = {{od[0], od[1]}, {ad[0], ad[1]}, {bd[0], bd[1]}};
    tensorProduct(o, a, b, dims);
  }
  void tensorProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt columns, uintt rows) {
    check(oaph_tensor_product_cr(m_api, c(o), c(a), c(b), columns, rows), "tensorProduct");
  }
  void hadamardProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b) {
    check(oaph_hadamard_product(m_api, c(o), c(a), c(b)), "hadamardProduct");
  }
  void hadamardProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt columns, uintt rows) {
    check(oaph_hadamard_product_cr(m_api, c(o), c(a), c(b), columns, rows), "hadamardProduct");
  }
  void hadamardProductVec(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt columns,
                          uintt rows) {
    check(oaph_hadamard_product_vec_cr(m_api, c(o), c(a), c(b), columns, rows),
          "hadamardProductVec");
  }
  // the reference kernel writes output = a b (CuAddDotProductProcedures.h:26-84)
  void addDotProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b) {
    check(oaph_add_dot_product(m_api, c(o), c(a), c(b)), "addDotProduct");
  }
  void addDotProduct(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt, uintt) {
    addDotProduct(o, a, b);
  }
  void dotProductPeriodic(math::Matrix* o, math::Matrix* a, math::Matrix* b) {
    check(oaph_dot_product_periodic(m_api, c(o), c(a), c(b)), "dotProductPeriodic");
  }
  void dotProductDimPeriodic(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt dims[3][2],
                             uintt periodicRows) {
    check(oaph_dot_product_dim_periodic(m_api, c(o), c(a), c(b), dims, periodicRows),
          "dotProductDimPeriodic");
  }
  void dotProductDimPeriodic(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt dims[3][2]) {
    dotProductDimPeriodic(o, a, b, dims, a->rows);
  }
  void dotProductEx(math::Matrix* o, math::Matrix* a, math::Matrix* b, MatrixEx* ex, uintt, uintt) {
    check(oaph_dot_product_ex(m_api, c(o), c(a), c(b), reinterpret_cast<oaph_matrix_ex*>(ex)),
          "dotProductEx");
  }
  void dotProductEx(math::Matrix* o, math::Matrix* a, math::Matrix* b, MatrixEx* ex) {
    dotProductEx(o, a, b, ex, ex->columns, ex->rows);
  }
  void dotProductExOpt(math::Matrix* o, math::Matrix* a, math::Matrix* b, MatrixEx* ex) {
    dotProductEx(o, a, b, ex);
  }
  void dotProductOpt(math::Matrix* o, math::Matrix* a, math::Matrix* b) { dotProduct(o, a, b); }
  void dotProductOpt(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt, uintt, uintt,
                     uintt) {
    dotProduct(o, a, b);
  }
  void transposeEx(math::Matrix* o, math::Matrix* a, MatrixEx* ex) {
    check(oaph_transpose_ex(m_api, c(o), c(a), reinterpret_cast<oaph_matrix_ex*>(ex)),
          "transposeEx");
  }
  // calculateQTHQ (CuProceduresApi.cpp:194-199): output = Q H Q^T by way of aux
  void calculateQTHQ(math::Matrix* o, math::Matrix* H, math::Matrix* Q, math::Matrix* aux) {
    transpose(o, Q);
    dotProduct(aux, H, o);
    dotProduct(o, Q, aux);
  }
  void hadamardProductVec(math::Matrix* o, math::Matrix* a, math::Matrix* b) {
    check(oaph_hadamard_product_vec(m_api, c(o), c(a), c(b)), "hadamardProductVec");
  }
  void transpose(math::Matrix* o, math::Matrix* a) { check(oaph_transpose(m_api, c(o), c(a)), "transpose"); }
  void conjugateTranspose(math::Matrix* o, math::Matrix* a) {
    check(oaph_conjugate_transpose(m_api, c(o), c(a)), "conjugateTranspose");
  }
  void add(math::Matrix* o, math::Matrix* a, math::Matrix* b) { check(oaph_add(m_api, c(o), c(a), c(b)), "add"); }
  void substract(math::Matrix* o, math::Matrix* a, math::Matrix* b) {
    check(oaph_substract(m_api, c(o), c(a), c(b)), "substract");
  }
  void add(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt columns, uintt rows) {
    check(oaph_add_cr(m_api, c(o), c(a), c(b), columns, rows), "add");
  }
  void substract(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt columns, uintt rows) {
    check(oaph_substract_cr(m_api, c(o), c(a), c(b), columns, rows), "substract");
  }
  // output += a - b (CuAddSubstractionProcedures.h:48-62)
  void addSubstract(math::Matrix* o, math::Matrix* a, math::Matrix* b, uintt columns, uintt rows) {
    check(oaph_add_substract(m_api, c(o), c(a), c(b), columns, rows), "addSubstract");
  }
  void addSubstract(math::Matrix* o, math::Matrix* a, math::Matrix* b) {
    addSubstract(o, a, b, o->columns, o->rows);
  }
  void multiplyReConstant(math::Matrix* o, math::Matrix* a, floatt re) {
    check(oaph_multiply_re_constant(m_api, c(o), c(a), re), "multiplyReConstant");
  }
  void multiplyConstant(math::Matrix* o, math::Matrix* a, floatt re, floatt im) {
    check(oaph_multiply_constant(m_api, c(o), c(a), re, im), "multiplyConstant");
  }
  void setDiagonal(math::Matrix* m, floatt re, floatt im) {
    check(oaph_set_diagonal(m_api, c(m), re, im), "setDiagonal");
  }
  void setIdentity(math::Matrix* m) { check(oaph_set_identity(m_api, c(m)), "setIdentity"); }
  void setZeroMatrix(math::Matrix* m) { check(oaph_set_zero_matrix(m_api, c(m)), "setZeroMatrix"); }
  void setVector(math::Matrix* V, uintt column, math::Matrix* v, uintt length) {
    check(oaph_set_vector(m_api, c(V), column, c(v), length), "setVector");
  }
  void getVector(math::Matrix* v, uintt length, math::Matrix* V, uintt column) {
    check(oaph_get_vector(m_api, c(v), length, c(V), column), "getVector");
  }
  void getVector(math::Matrix* v, math::Matrix* V, uintt column) {
    getVector(v, v->columns * v->rows, V, column);
  }
  void magnitude(floatt& out, math::Matrix* m) { check(oaph_magnitude(m_api, &out, c(m)), "magnitude"); }
  void magnitudeOpt(floatt& out, math::Matrix* m) { magnitude(out, m); }
  void magnitudeOptVer2(floatt& out, math::Matrix* m) { magnitude(out, m); }
  void magnitude2(floatt& out, math::Matrix* m) { check(oaph_magnitude2(m_api, &out, c(m)), "magnitude2"); }
  void magnitude2Opt(floatt& out, math::Matrix* m) { magnitude2(out, m); }
  void magnitude2OptVer2(floatt& out, math::Matrix* m) { magnitude2(out, m); }
  void sum(floatt& re, floatt& im, math::Matrix* m) { check(oaph_sum(m_api, &re, &im, c(m)), "sum"); }
  void sum(floatt& re, math::Matrix* m) {
    floatt im = 0;
    sum(re, im, m);
  }
  void QRGR(math::Matrix* Q, math::Matrix* R, math::Matrix* H, math::Matrix* a0, math::Matrix* a1,
            math::Matrix* a2, math::Matrix* a3) {
    check(oaph_qrgr(m_api, c(Q), c(R), c(H), c(a0), c(a1), c(a2), c(a3)), "QRGR");
  }
  bool isUpperTriangular(math::Matrix* m) {
    int r = 0;
    check(oaph_is_upper_triangular(m_api, c(m), &r), "isUpperTriangular");
    return r == 1;
  }
  void calcTriangularH(math::Matrix* H, math::Matrix* Q, math::Matrix* R, math::Matrix* a1,
                       math::Matrix* a2, math::Matrix* a3, math::Matrix* a4, math::Matrix* a5,
                       math::Matrix* a6) {
    check(oaph_calc_triangular_h(m_api, c(H), c(Q), c(R), c(a1), c(a2), c(a3), c(a4), c(a5), c(a6)),
          "calcTriangularH");
  }
  void calcTriangularHStep(math::Matrix* H, math::Matrix* Q, math::Matrix*, math::Matrix*,
                           math::Matrix*, math::Matrix*, math::Matrix*, math::Matrix*,
                           math::Matrix*) {
    check(oaph_calc_triangular_h_step(m_api, c(H), c(Q)), "calcTriangularHStep");
  }
  oaph_api* handle() const { return m_api; }

 private:
  oaph_api* m_api = nullptr;
  bool m_owned = true;
};

// oap::RecToSquareApi (oapMatrixCuda/RecToSquareApi.h:33-58) over a DEVICE
// matrix M (the reference takes a host matrix and copies blocks of it to the
// device on demand; here M stays resident in HBM).  With deallocate = true
// the object frees M when it is destroyed, as the reference frees its host
// matrix.  S = M M^T, or M itself when M is square.
class RecToSquareApi final {
 public:
  RecToSquareApi(CuProceduresApi& api, const math::Matrix* rec, bool deallocate)
      : m_rec(rec), m_deallocate(deallocate) {
    check(oaph_rts_create(&m_rts, api.handle(), c(rec)), "RecToSquareApi");
  }
  ~RecToSquareApi() {
    oaph_rts_destroy(m_rts);
    if (m_deallocate) oaph_matrix_free(const_cast<oaph_matrix*>(c(m_rec)));
  }
  RecToSquareApi(const RecToSquareApi&) = delete;
  RecToSquareApi(RecToSquareApi&&) = delete;
  RecToSquareApi& operator=(const RecToSquareApi&) = delete;
  RecToSquareApi& operator=(RecToSquareApi&&) = delete;

  math::MatrixInfo getMatrixInfo() const {
    oaph_matrix_info i{};
    check(oaph_rts_info(m_rts, &i), "RecToSquareApi::getMatrixInfo");
    return math::MatrixInfo(i.isRe != 0, i.isIm != 0, i.columns, i.rows);
  }
  math::Matrix* createDeviceMatrix() {
    const math::MatrixInfo i = getMatrixInfo();
    math::Matrix* m = cuda::NewDeviceReMatrix(i.columns(), i.rows());
    return getDeviceMatrix(m);
  }
  math::Matrix* getDeviceMatrix(math::Matrix* dmatrix) {
    check(oaph_rts_square(m_rts, c(dmatrix)), "RecToSquareApi::getDeviceMatrix");
    return dmatrix;
  }
  math::Matrix* createDeviceSubMatrix(uintt rindex, uintt rlength) {
    const math::MatrixInfo i = getMatrixInfo();
    const uintt rows = rindex < i.rows() ? std::min(rlength, i.rows() - rindex) : 1;
    math::Matrix* m = cuda::NewDeviceMatrix(i.isRe, i.isIm, i.columns(), rows ? rows : 1);
    return getDeviceSubMatrix(rindex, rlength, m);
  }
  math::Matrix* getDeviceSubMatrix(uintt rindex, uintt rlength, math::Matrix* dmatrix) {
    if (dmatrix == nullptr) return createDeviceSubMatrix(rindex, rlength);
    check(oaph_rts_sub_matrix(m_rts, rindex, rlength, c(dmatrix)),
          "RecToSquareApi::getDeviceSubMatrix");
    return dmatrix;
  }
  math::Matrix* createDeviceRowVector(uintt index) {
    const math::MatrixInfo i = getMatrixInfo();
    return getDeviceRowVector(index, cuda::NewDeviceReMatrix(i.columns(), 1));
  }
  math::Matrix* getDeviceRowVector(uintt index, math::Matrix* dmatrix) {
    if (dmatrix == nullptr) return createDeviceRowVector(index);
    check(oaph_rts_row_vector(m_rts, index, c(dmatrix)), "RecToSquareApi::getDeviceRowVector");
    return dmatrix;
  }
  // w = S v without forming S (the product of the reference's multiply
  // callbacks, oap2dt3dDevice/MainAPExecutor.cpp:42-127), enqueued on the
  // stream of `api` -- in a multiply callback, the api the callback receives
  void multiply(math::Matrix* w, const math::Matrix* v, CuProceduresApi& api) {
    check(oaph_rts_multiply(m_rts, api.handle(), c(w), c(v)), "RecToSquareApi::multiply");
  }
  void multiply(math::Matrix* w, const math::Matrix* v) {
    check(oaph_rts_multiply(m_rts, nullptr, c(w), c(v)), "RecToSquareApi::multiply");
  }
  oaph_rts* handle() const { return m_rts; }

 private:
  oaph_rts* m_rts = nullptr;
  const math::Matrix* m_rec;
  bool m_deallocate;
};

enum class QRType { NONE = -1, QRGR, QRHT };
enum class VecMultiplicationType { TYPE_EIGENVECTOR, TYPE_WV };
enum class InitVVectorType { RANDOM, FIRST_VALUE_IS_ONE };
}  // namespace oap

struct Complex {
  floatt re = 0, im = 0;
  Complex() = default;
  Complex(floatt r, floatt i = 0) : re(r), im(i) {}
};

class EigenPair {
 public:
  EigenPair(floatt re, const math::Matrix* m = nullptr, uintt index = 0) : m_v(re), m_m(m), m_i(index) {}
  EigenPair(const Complex& v, uintt index) : m_v(v), m_m(nullptr), m_i(index) {}
  Complex getEValue() const { return m_v; }
  floatt re() const { return m_v.re; }
  floatt im() const { return m_v.im; }
  uintt getIndex() const { return m_i; }

 private:
  Complex m_v;
  const math::Matrix* m_m;
  uintt m_i;
};

namespace ArnUtils {
// Comparators (ArnoldiUtils.h:63-72, ArnoldiUtils.cpp:24-50): true when i
// sorts before j.  setSortType maps these six to the same orders applied on
// the device; any other comparator orders the Ritz values on the host.
inline bool SortLargestValues(const EigenPair& i, const EigenPair& j) {
  return i.re() * i.re() + i.im() * i.im() > j.re() * j.re() + j.im() * j.im();
}
inline bool SortLargestReValues(const EigenPair& i, const EigenPair& j) { return i.re() > j.re(); }
inline bool SortLargestImValues(const EigenPair& i, const EigenPair& j) { return i.im() > j.im(); }
inline bool SortSmallestValues(const EigenPair& i, const EigenPair& j) {
  return i.re() * i.re() + i.im() * i.im() < j.re() * j.re() + j.im() * j.im();
}
inline bool SortSmallestReValues(const EigenPair& i, const EigenPair& j) { return i.re() < j.re(); }
inline bool SortSmallestImValues(const EigenPair& i, const EigenPair& j) { return i.im() < j.im(); }
typedef bool (*SortType)(const EigenPair& i, const EigenPair& j);
// the device-side code of a built-in comparator, -1 for any other
inline int builtinSortCode(SortType t) {
  if (t == &SortLargestValues) return OAPH_SORT_LARGEST_VALUES;
  if (t == &SortLargestReValues) return OAPH_SORT_LARGEST_RE;
  if (t == &SortLargestImValues) return OAPH_SORT_LARGEST_IM;
  if (t == &SortSmallestValues) return OAPH_SORT_SMALLEST_VALUES;
  if (t == &SortSmallestReValues) return OAPH_SORT_SMALLEST_RE;
  if (t == &SortSmallestImValues) return OAPH_SORT_SMALLEST_IM;
  return -1;
}
enum CheckType { CHECK_INTERNAL, CHECK_FIRST_STOP, CHECK_COUNTER };
enum TriangularHProcedureType { CALC_IN_DEVICE, CALC_IN_HOST, CALC_IN_DEVICE_STEP };
enum Type { UNDEFINED, DEVICE, HOST };
}  // namespace ArnUtils

class CuHArnoldi {
 public:
  CuHArnoldi() { oap::check(oaph_arnoldi_create(&m_a), "CuHArnoldi"); }
  virtual ~CuHArnoldi() { oaph_arnoldi_destroy(m_a); }
  CuHArnoldi(const CuHArnoldi&) = delete;
  CuHArnoldi& operator=(const CuHArnoldi&) = delete;

  void setRho(floatt rho = 1. / 3.14) { oap::check(oaph_arnoldi_set_rho(m_a, rho), "setRho"); }
  void setBLimit(floatt b) { oap::check(oaph_arnoldi_set_blimit(m_a, b), "setBLimit"); }
  void setQRType(oap::QRType t = oap::QRType::QRGR) {
    oap::check(oaph_arnoldi_set_qr_type(m_a, (int)t), "setQRType");
  }
  void setVecInitType(oap::InitVVectorType t) {
    oap::check(oaph_arnoldi_set_vec_init_type(m_a, (int)t), "setVecInitType");
  }
  void setSortType(ArnUtils::SortType t) {
    const int code = ArnUtils::builtinSortCode(t);
    if (code >= 0) {
      oap::check(oaph_arnoldi_set_sort_type(m_a, code), "setSortType");
    } else {
      m_sort = t;
      oap::check(oaph_arnoldi_set_sort_fn(m_a, &CuHArnoldi::sortTrampoline, this), "setSortType");
    }
  }
  void setCheckType(ArnUtils::CheckType t) {
    oap::check(oaph_arnoldi_set_check_type(m_a, (int)t), "setCheckType");
  }
  void setCheckCounts(uintt n) { oap::check(oaph_arnoldi_set_check_counts(m_a, n), "setCheckCounts"); }
  void setCheckTolerance(floatt t) {
    oap::check(oaph_arnoldi_set_check_tolerance(m_a, t), "setCheckTolerance");
  }
  void setOutputsEigenvalues(floatt* re, floatt* im) {
    oap::check(oaph_arnoldi_set_outputs_eigenvalues(m_a, re, im), "setOutputsEigenvalues");
  }
  void setOutputsEigenvectors(math::Matrix** v) {
    oap::check(oaph_arnoldi_set_outputs_eigenvectors(m_a, reinterpret_cast<oaph_matrix**>(v)),
               "setOutputsEigenvectors");
  }
  void setOutputType(ArnUtils::Type t) { oap::check(oaph_arnoldi_set_output_type(m_a, (int)t), "setOutputType"); }
  void setCalcTraingularHType(ArnUtils::TriangularHProcedureType t) {
    oap::check(oaph_arnoldi_set_calc_triangular_h_type(m_a, (int)t), "setCalcTraingularHType");
  }
  // extensions
  void setSeed(uint64_t s) { oap::check(oaph_arnoldi_set_seed(m_a, s), "setSeed"); }
  void setNormA(floatt n) { oap::check(oaph_arnoldi_set_norm_a(m_a, n), "setNormA"); }
  void setMaxRestarts(uintt n) { oap::check(oaph_arnoldi_set_max_restarts(m_a, n), "setMaxRestarts"); }
  void setPrecision(int p) { oap::check(oaph_arnoldi_set_precision(m_a, p), "setPrecision"); }
  void setProjection(int mode) { oap::check(oaph_arnoldi_set_projection(m_a, mode), "setProjection"); }
  void setMixedTolerance(double t) {
    oap::check(oaph_arnoldi_set_mixed_tolerance(m_a, t), "setMixedTolerance");
  }
  void setF16Phase(bool on, double tol = 2e-3) {
    oap::check(oaph_arnoldi_set_f16_phase(m_a, on ? 1 : 0, tol), "setF16Phase");
  }
  void setMixedCheck(bool on, double fraction = 0.5) {
    oap::check(oaph_arnoldi_set_mixed_check(m_a, on ? 1 : 0, fraction), "setMixedCheck");
  }
  void setRestartStages(const std::vector<double>& tols) {
    oap::check(oaph_arnoldi_set_restart_stages(m_a, tols.empty() ? nullptr : tols.data(),
                                               (uint32_t)tols.size()),
               "setRestartStages");
  }
  void setRestartGuard(uintt g) {
    oap::check(oaph_arnoldi_set_restart_guard(m_a, (uint32_t)g), "setRestartGuard");
  }
  void setRestartGuardWeight(double w) {
    oap::check(oaph_arnoldi_set_restart_guard_weight(m_a, w), "setRestartGuardWeight");
  }
  void setRestartKeep(uintt keep) {
    oap::check(oaph_arnoldi_set_restart_keep(m_a, (uint32_t)keep), "setRestartKeep");
  }
  /* extension: 1 perfect-shift Schur form where possible (default), 0 Wilkinson */
  void setSchurMethod(int method) {
    oap::check(oaph_arnoldi_set_schur_method(m_a, method), "setSchurMethod");
  }
  oaph_arnoldi_stats stats() const {
    oaph_arnoldi_stats s;
    oap::check(oaph_arnoldi_get_stats(m_a, &s), "stats");
    return s;
  }

  floatt testOutcome(size_t index) {
    floatt out = 0;
    oap::check(oaph_arnoldi_test_outcome(m_a, (uint32_t)index, &out), "testOutcome");
    return out;
  }

 protected:
  void execute(uintt hdim, uintt wanted, const math::MatrixInfo& info,
               ArnUtils::Type type = ArnUtils::DEVICE) {
    prepare();
    oaph_matrix_info ci{info.isRe, info.isIm, info.columns(), info.rows()};
    oaph_status s = oaph_arnoldi_execute(m_a, hdim, wanted, &ci, (int)type);
    if (m_pending) std::rethrow_exception(m_pending);
    oap::check(s, "execute");
  }
  virtual void multiply(math::Matrix* w, math::Matrix* v, oap::CuProceduresApi& api,
                        oap::VecMultiplicationType mt) = 0;
  virtual bool checkEigenspair(floatt re, floatt im, math::Matrix* v, uintt index, uintt max) = 0;
  // called before execute: binds the virtual multiply to the C callback
  virtual void prepare() {
    oap::check(oaph_arnoldi_set_callback(m_a, &CuHArnoldi::tramp, this), "set_callback");
    oap::check(oaph_arnoldi_set_check_callback(m_a, &CuHArnoldi::ctramp, this), "set_check");
  }
  oaph_arnoldi* m_a = nullptr;
  ArnUtils::SortType m_sort = nullptr;
  static int sortTrampoline(double ri, double ii, double rj, double ij, void* user) {
    const CuHArnoldi* self = static_cast<const CuHArnoldi*>(user);
    return self->m_sort(EigenPair(Complex(ri, ii), 0u), EigenPair(Complex(rj, ij), 1u)) ? 1 : 0;
  }
  std::exception_ptr m_pending;

 private:
  static void tramp(oaph_matrix* w, oaph_matrix* v, oaph_api* api, void* user, int type) {
    auto* self = static_cast<CuHArnoldi*>(user);
    try {
      oap::CuProceduresApi borrowed(api);
      self->multiply(reinterpret_cast<math::Matrix*>(w), reinterpret_cast<math::Matrix*>(v), borrowed,
                     type == OAPH_TYPE_WV ? oap::VecMultiplicationType::TYPE_WV
                                          : oap::VecMultiplicationType::TYPE_EIGENVECTOR);
    } catch (...) {
      self->m_pending = std::current_exception();
    }
  }
  static int ctramp(double re, double im, oaph_matrix* v, uint32_t index, uint32_t max, void* user) {
    auto* self = static_cast<CuHArnoldi*>(user);
    return self->checkEigenspair(re, im, reinterpret_cast<math::Matrix*>(v), index, max) ? 1 : 0;
  }
};

// CuHArnoldiDefault (ArnoldiProceduresImpl.h:26-43): w = A v with the
// solver's own streaming matvec kernel
class CuHArnoldiDefault : public CuHArnoldi {
 public:
  void setMatrix(math::Matrix* A) {
    m_A = A;
    oap::check(oaph_arnoldi_set_matrix(m_a, oap::c(A)), "setMatrix");
  }
  void execute(uintt hdim, uintt wanted, const math::MatrixInfo& info,
               ArnUtils::Type type = ArnUtils::DEVICE) {
    CuHArnoldi::execute(hdim, wanted, info, type);
  }

 protected:
  void prepare() override {}
  void multiply(math::Matrix* w, math::Matrix* v, oap::CuProceduresApi& api,
                oap::VecMultiplicationType) override {
    api.dotProduct(w, m_A, v);
  }
  bool checkEigenspair(floatt, floatt, math::Matrix*, uintt, uintt) override { return true; }

 private:
  math::Matrix* m_A = nullptr;
};

// CuHArnoldiCallbackBase / CuHArnoldiCallback (ArnoldiProceduresImpl.h:45-74)
class CuHArnoldiCallbackBase : public CuHArnoldi {
 public:
  typedef void (*MultiplyFunc)(math::Matrix* w, math::Matrix* v, oap::CuProceduresApi& api,
                               void* userData, oap::VecMultiplicationType mt);
  typedef bool (*CheckFunc)(floatt re, floatt im, math::Matrix* vector, uintt index, uintt max,
                            void* userData);
  void setCallback(MultiplyFunc f, void* user) {
    m_mf = f;
    m_mu = user;
  }
  void setCheckCallback(CheckFunc f, void* user) {
    m_cf = f;
    m_cu = user;
  }

 protected:
  void multiply(math::Matrix* w, math::Matrix* v, oap::CuProceduresApi& api,
                oap::VecMultiplicationType mt) override {
    m_mf(w, v, api, m_mu, mt);
  }
  bool checkEigenspair(floatt re, floatt im, math::Matrix* v, uintt index, uintt max) override {
    return m_cf ? m_cf(re, im, v, index, max, m_cu) : true;
  }

 private:
  MultiplyFunc m_mf = nullptr;
  void* m_mu = nullptr;
  CheckFunc m_cf = nullptr;
  void* m_cu = nullptr;
};

class CuHArnoldiCallback : public CuHArnoldiCallbackBase {
 public:
  void execute(uintt hdim, uintt wanted, const math::MatrixInfo& info,
               ArnUtils::Type type = ArnUtils::DEVICE) {
    CuHArnoldi::execute(hdim, wanted, info, type);
  }
};
