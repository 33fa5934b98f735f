/*
 * oaph.h -- C ABI of the MI355X (gfx950) implementation of oap's complex
 * matrix procedures and implicitly restarted Arnoldi eigensolver.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an oaph_status (0 = success).  Errors never abort the process.
 * Each declaration cites the reference interface it replaces
 * (paths relative to the marcinmatula/oap source tree).
 *
 * Threading: an oaph_api owns one HIP stream; calls on one oaph_api must not
 * overlap.  Procedures are enqueued on that stream; calls that return a host
 * scalar (magnitude, sum, is_upper_triangular, get_value) synchronise it.
 */
#ifndef OAPH_H
#define OAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* status                                                              */
/* ------------------------------------------------------------------ */
typedef int oaph_status;
#define OAPH_OK 0
#define OAPH_ERR_ARG 1           /* NULL or inconsistent argument */
#define OAPH_ERR_SHAPE 2         /* matrix dimensions do not match */
#define OAPH_ERR_PLANE 3         /* required re/im plane missing */
#define OAPH_ERR_HIP 4           /* HIP runtime error */
#define OAPH_ERR_NOMEM 5         /* device allocation failed */
#define OAPH_ERR_UNSUPPORTED 6   /* option or size not supported */
#define OAPH_ERR_NOT_CONVERGED 7 /* solver stopped before all pairs converged */
#define OAPH_ERR_STATE 8         /* call out of order (e.g. no callback set) */

/* Human-readable text for a status; replaces the reference's abort() on
 * driver errors (oapCuda/KernelExecutor.h:34-41). */
const char* oaph_status_string(oaph_status s);
/* Text of the last HIP error seen by the library (thread-local). */
const char* oaph_last_error(void);

/* ------------------------------------------------------------------ */
/* data model                                                          */
/* ------------------------------------------------------------------ */
/* math::Matrix (oapMatrix/Matrix.h:31-44): element (c, r) at c + r*columns
 * of each plane; either plane may be NULL.  For device matrices the struct
 * lives in host memory and the planes in HBM (the reference keeps the
 * struct itself on the device, oapMatrixCuda/CudaUtils.cpp:175-248). */
typedef struct oaph_matrix {
  uint32_t columns;
  uint32_t rows;
  uint32_t realColumns;
  uint32_t realRows;
  double* reValues;
  double* imValues;
} oaph_matrix;

/* MatrixEx (oapMatrix/MatrixEx.h:26-32): a region, columns [column,
 * column + columns) and rows [row, row + rows).  Passed by pointer to host
 * memory (the reference allocates it on the device). */
typedef struct oaph_matrix_ex {
  uint32_t column;
  uint32_t columns;
  uint32_t row;
  uint32_t rows;
} oaph_matrix_ex;

/* math::MatrixInfo (oapMatrix/MatrixInfo.h:16-67) */
typedef struct oaph_matrix_info {
  int32_t isRe;
  int32_t isIm;
  uint32_t columns;
  uint32_t rows;
} oaph_matrix_info;

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
/* oap::cuda::Context::create (oapCuda/KernelExecutor.cpp:147-187): selects
 * the device.  Optional; the first call of any entry point initialises
 * device 0. */
oaph_status oaph_init(int device);
oaph_status oaph_device_synchronize(void);
/* number of visible devices (0 when no GPU) */
int oaph_device_count(void);

/* ------------------------------------------------------------------ */
/* device matrices (oapMatrixCuda/oapCudaMatrixUtils.h:38-74,         */
/* oapMatrixCuda/CudaUtils.cpp:175-359)                               */
/* ------------------------------------------------------------------ */
/* oap::cuda::NewDeviceMatrix(isRe, isIm, columns, rows) */
oaph_status oaph_matrix_new(oaph_matrix** out, int isRe, int isIm, uint32_t columns,
                            uint32_t rows);
/* oap::cuda::DeleteDeviceMatrix; NULL is accepted */
oaph_status oaph_matrix_free(oaph_matrix* m);
/* oap::cuda::GetMatrixInfo */
oaph_status oaph_matrix_info_get(const oaph_matrix* m, oaph_matrix_info* info);
/* oap::cuda::CopyHostMatrixToDeviceMatrix with plain host arrays (NULL skips a plane) */
oaph_status oaph_matrix_copy_h2d(oaph_matrix* dst, const double* re, const double* im);
/* oap::cuda::CopyDeviceMatrixToHostMatrix */
oaph_status oaph_matrix_copy_d2h(double* re, double* im, const oaph_matrix* src);
/* oap::cuda::CopyDeviceMatrixToDeviceMatrix (same element count) */
oaph_status oaph_matrix_copy_d2d(oaph_matrix* dst, const oaph_matrix* src);
/* CudaUtils::GetReValue/GetImValue, SetReValue/SetImValue (index = c + r*columns) */
oaph_status oaph_matrix_get_value(const oaph_matrix* m, uint32_t index, double* re, double* im);
oaph_status oaph_matrix_set_value(oaph_matrix* m, uint32_t index, double re, double im);
/* CudaUtils::SetZeroRow (oapMatrixCuda/CudaUtils.cpp:332-359) */
oaph_status oaph_matrix_set_zero_row(oaph_matrix* m, uint32_t row, int re, int im);
/* CopyDeviceMatrixToDeviceMatrixDims / CopyHostMatrixToDeviceMatrixDims /
 * CopyDeviceMatrixToHostMatrixDims (oapCudaMatrixUtils.h:102-116,
 * oapHostMatrixUtils.h:803-892): dims[0] = {{column, columns}, {row, rows}}
 * of dst, dims[1] the same of src (equal sizes); the block is copied plane
 * by plane (planes present in both).  A host matrix is an oaph_matrix whose
 * planes are host pointers. */
oaph_status oaph_matrix_copy_d2d_dims(oaph_matrix* dst, const oaph_matrix* src,
                                      const uint32_t dims[2][2][2]);
oaph_status oaph_matrix_copy_h2d_dims(oaph_matrix* dst, const oaph_matrix* src_host,
                                      const uint32_t dims[2][2][2]);
oaph_status oaph_matrix_copy_d2h_dims(oaph_matrix* dst_host, const oaph_matrix* src,
                                      const uint32_t dims[2][2][2]);
/* SetMatrix / SetReMatrix / SetImMatrix(matrix, matrix1, column, row)
 * (oapCudaMatrixUtils.h:142-146): all of src written into dst at
 * (column, row); planes: 1 re, 2 im, 3 both (those present in both). */
oaph_status oaph_matrix_set_submatrix(oaph_matrix* dst, const oaph_matrix* src, uint32_t column,
                                      uint32_t row, int planes);

/* ------------------------------------------------------------------ */
/* matrix formats (host code; device matrices are copied once)         */
/* ------------------------------------------------------------------ */
/* Text, as oap::host::ToString prints it (oapMatrixCpu/oapHostMatrixUtils.cpp
 * :576-588 -> oapMatrix/MatrixPrinter.h:28-112 with the default PrintArgs,
 * oapMatrix/MatrixUtils.h:104-176): "[v, v, ...|\n v ...]" per plane, the im
 * plane after " + i * ", fixed notation with 9 digits, runs of equal values
 * as "v <repeats n times>".  *len receives the length without the NUL; buf
 * may be NULL to query it, else cap must exceed it.  A NULL matrix prints
 * "nullptr". */
oaph_status oaph_matrix_to_string(const oaph_matrix* m, char* buf, size_t cap, size_t* len);
oaph_status oaph_host_matrix_to_string(uint32_t columns, uint32_t rows, const double* re,
                                       const double* im, char* buf, size_t cap, size_t* len);
/* oap::host::NewMatrix(text) (oapMatrixCpu/oapHostMatrixUtils.cpp:223-274,
 * oapMatrix/MatrixParser.cpp): "columns=<c>" and "rows=<r>" anywhere in the
 * text (else a square of side floor(sqrt(n)) of the first array's length
 * n), the first "[...]" is the re plane and the second the im plane,
 * elements split at ',' or '|', "v <repeats n times>" expands; values past
 * c x r are ignored.  A text without an array -> OAPH_ERR_ARG; an array
 * shorter than c x r -> OAPH_ERR_SHAPE (the reference reads past it). */
oaph_status oaph_matrix_parse(const char* text, oaph_matrix** out);
/* the same into host arrays (re/im may be NULL; cap = elements available) */
oaph_status oaph_host_parse_matrix(const char* text, oaph_matrix_info* info, double* re,
                                   double* im, size_t cap);
/* matrixUtils::CreateArray(text, which) (MatrixParser.cpp:199-210) */
oaph_status oaph_host_parse_array(const char* text, uint32_t which, double* out, size_t cap,
                                  size_t* len);
/* Binary record of oap::cuda::SaveMatrix / oap::host::SaveMatrix
 * (oapMatrixCuda/oapCudaMatrixUtils.cpp:562-589, oapMatrixCpu/oapHostMatrixUtils
 * .cpp:1043-1070): u8 isMatrix, u8 isRe, u8 isIm, u32 columns, u32 rows, the
 * re doubles, the im doubles (little endian, unaligned); a NULL matrix is
 * the single byte 0.  *len receives the record size (buf NULL: query). */
oaph_status oaph_matrix_save(const oaph_matrix* m, uint8_t* buf, size_t cap, size_t* len);
oaph_status oaph_host_matrix_save(int present, uint32_t columns, uint32_t rows, const double* re,
                                  const double* im, uint8_t* buf, size_t cap, size_t* len);
/* LoadMatrix (oapMatrixCuda/oapCudaMatrixUtils.cpp:591-603): one record into
 * a new device matrix (*out = NULL for a NULL record); *consumed = bytes used */
oaph_status oaph_matrix_load(oaph_matrix** out, const uint8_t* buf, size_t len, size_t* consumed);
/* the same into host arrays (re/im may be NULL: info only) */
oaph_status oaph_host_matrix_load(const uint8_t* buf, size_t len, size_t* consumed, int* present,
                                  oaph_matrix_info* info, double* re, double* im, size_t cap);
/* utils::ByteBuffer::fwrite / fread (oapUtils/Buffer.h): size_t 1, size_t
 * length, the bytes.  read: buf NULL queries *len. */
oaph_status oaph_buffer_write_file(const char* path, const uint8_t* buf, size_t len);
oaph_status oaph_buffer_read_file(const char* path, uint8_t* buf, size_t cap, size_t* len);

/* ------------------------------------------------------------------ */
/* procedures API: oap::CuProceduresApi (oapMatrixCuda/CuProceduresApi.h:45-309) */
/* ------------------------------------------------------------------ */
typedef struct oaph_api oaph_api;

/* CuProceduresApi::CuProceduresApi / ~CuProceduresApi (CuProceduresApi.cpp:64-89) */
oaph_status oaph_api_create(oaph_api** out);
oaph_status oaph_api_destroy(oaph_api* api);
/* waits for all work enqueued on the api's stream */
oaph_status oaph_api_synchronize(oaph_api* api);
/* the api's hipStream_t, as void* */
void* oaph_api_stream(oaph_api* api);

/* dotProduct(out, A, B): C = A*B complex (CuProceduresApi.cpp:91-105) */
oaph_status oaph_dot_product(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                             const oaph_matrix* b);
/* dotProduct(out, A, B, dims[3][2]) (CuProceduresApi.cpp:221-225) */
oaph_status oaph_dot_product_dims(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                  const oaph_matrix* b, const uint32_t dims[3][2]);
/* dotProductShared (CuProceduresApi.cpp:107-117): same result as dotProduct */
oaph_status oaph_dot_product_shared(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                    const oaph_matrix* b);
/* transpose / conjugateTranspose (CuProceduresApi.cpp:268-307) */
oaph_status oaph_transpose(oaph_api* api, oaph_matrix* out, const oaph_matrix* a);
oaph_status oaph_conjugate_transpose(oaph_api* api, oaph_matrix* out, const oaph_matrix* a);
/* tensorProduct(out, A, B, dims[3][2]) (CuProceduresApi.cpp:153-156); re plane only */
oaph_status oaph_tensor_product(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                const oaph_matrix* b, const uint32_t dims[3][2]);
/* hadamardProduct / hadamardProductVec (CuProceduresApi.cpp:158-192); per plane */
oaph_status oaph_hadamard_product(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                  const oaph_matrix* b);
oaph_status oaph_hadamard_product_vec(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                      const oaph_matrix* b);
/* add / substract (CuProceduresApi.cpp:309-325) */
oaph_status oaph_add(oaph_api* api, oaph_matrix* out, const oaph_matrix* a, const oaph_matrix* b);
oaph_status oaph_substract(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                           const oaph_matrix* b);
/* multiplyReConstant (re plane only) / multiplyConstant (per plane)
 * (CuProceduresApi.cpp:492-507) */
oaph_status oaph_multiply_re_constant(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                      double re);
oaph_status oaph_multiply_constant(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                   double re, double im);
/* setDiagonal / setIdentity / setZeroMatrix (CuProceduresApi.cpp:422-437) */
oaph_status oaph_set_diagonal(oaph_api* api, oaph_matrix* m, double re, double im);
oaph_status oaph_set_identity(oaph_api* api, oaph_matrix* m);
oaph_status oaph_set_zero_matrix(oaph_api* api, oaph_matrix* m);
/* setVector / getVector (CuProceduresApi.cpp:327-340) */
oaph_status oaph_set_vector(oaph_api* api, oaph_matrix* V, uint32_t column, const oaph_matrix* v,
                            uint32_t length);
oaph_status oaph_get_vector(oaph_api* api, oaph_matrix* v, uint32_t length, const oaph_matrix* V,
                            uint32_t column);
/* magnitude2(floatt&, m) = sum re^2 + im^2 (CuProceduresApi.cpp:401-420;
 * magnitude2Opt and magnitude2OptVer2 are the same value) */
oaph_status oaph_magnitude2(oaph_api* api, double* out, const oaph_matrix* m);
/* magnitude(floatt&, m) = sqrt(sum re^2 + im^2) (CuProceduresApi.cpp:342-345) */
oaph_status oaph_magnitude(oaph_api* api, double* out, const oaph_matrix* m);
/* Allocation-free form of oaph_magnitude; no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops): calls no hipMalloc/hipFree.  `out` is one
 * double in DEVICE memory that the caller owns; sqrt(sum re^2 + im^2) is
 * written there in the order of the api's stream (read it after
 * oaph_api_synchronize).  The block partial sums live in a static device
 * array of the library (2048 doubles, 16 KiB per GPU, held until the process
 * ends); any matrix size is accepted (the reduction grid has at most 2048
 * workgroups; a larger one would get OAPH_ERR_UNSUPPORTED).  Calls that overlap on one GPU share
 * that array and race: serialise them (one stream, or events). */
oaph_status oaph_magnitude_noalloc(oaph_api* api, double* out, const oaph_matrix* m);
/* sum(re&, im&, m) (CuProceduresApi.cpp:358-388) */
oaph_status oaph_sum(oaph_api* api, double* re, double* im, const oaph_matrix* m);
/* QRGR(Q, R, H, aux0..aux3) (CuProceduresApi.cpp:439-455); the aux matrices
 * are accepted for signature parity and may be NULL (not used). */
oaph_status oaph_qrgr(oaph_api* api, oaph_matrix* Q, oaph_matrix* R, const oaph_matrix* H,
                      oaph_matrix* aux0, oaph_matrix* aux1, oaph_matrix* aux2,
                      oaph_matrix* aux3);
/* isUpperTriangular (CuProceduresApi.cpp:462-470) */
oaph_status oaph_is_upper_triangular(oaph_api* api, const oaph_matrix* m, int* result);
/* calcTriangularH(H, Q, R, aux1..aux6) (CuProceduresApi.cpp:472-480); aux may be NULL */
oaph_status oaph_calc_triangular_h(oaph_api* api, oaph_matrix* H, oaph_matrix* Q, oaph_matrix* R,
                                   oaph_matrix* aux1, oaph_matrix* aux2, oaph_matrix* aux3,
                                   oaph_matrix* aux4, oaph_matrix* aux5, oaph_matrix* aux6);
/* calcTriangularHStep (CuProceduresApi.cpp:482-490, CuTriangularH.h:61-77):
 * one unshifted QR step H <- R Q, Q the step's factor (Q = I when H is
 * already upper triangular; the reference leaves an aux matrix's previous
 * contents there) */
oaph_status oaph_calc_triangular_h_step(oaph_api* api, oaph_matrix* H, oaph_matrix* Q);

/* dims forms and MatrixEx procedures */
/* hadamardProduct / hadamardProductVec / tensorProduct (output, a, b, columns,
 * rows) (CuProceduresApi.cpp:136-192): the given output size is checked; the
 * operation is the one without the size (tensor: over columns x rows) */
oaph_status oaph_hadamard_product_cr(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                     const oaph_matrix* b, uint32_t columns, uint32_t rows);
oaph_status oaph_hadamard_product_vec_cr(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                         const oaph_matrix* b, uint32_t columns, uint32_t rows);
oaph_status oaph_tensor_product_cr(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                   const oaph_matrix* b, uint32_t columns, uint32_t rows);
/* add / substract (output, a, b, columns, rows) and addSubstract (out += a - b)
 * over the columns x rows region at the top left (CuProceduresApi.cpp:309-325,
 * CuAddSubstractionProcedures.h:48-62) */
oaph_status oaph_add_cr(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                        const oaph_matrix* b, uint32_t columns, uint32_t rows);
oaph_status oaph_substract_cr(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                              const oaph_matrix* b, uint32_t columns, uint32_t rows);
oaph_status oaph_add_substract(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                               const oaph_matrix* b, uint32_t columns, uint32_t rows);
/* addDotProduct (CuProceduresApi.cpp:119-134): the reference kernel writes
 * output = a b (CuAddDotProductProcedures.h:26-84), not +=; kept */
oaph_status oaph_add_dot_product(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                 const oaph_matrix* b);
/* dotProductPeriodic (CuProceduresApi.cpp:211-214, CuDotProductPeriodicProcedures.h
 * :65-85): with K = a.columns, out(x, y) = sum_k a(k, y % K) b(x, (y / K) K + k) */
oaph_status oaph_dot_product_periodic(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                      const oaph_matrix* b);
/* dotProductDimPeriodic (CuProceduresApi.cpp:216-219, CuDotProductDimPeriodic
 * Procedures.h:27-45): W = dims[0][0], H = dims[0][1], K = dims[1][0],
 * P = periodic_rows; for output rows y with y % P < H and x < W,
 * out(x, y) = sum_k a(k, y % P) b(x, (y / P) K + k) */
oaph_status oaph_dot_product_dim_periodic(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                          const oaph_matrix* b, const uint32_t dims[3][2],
                                          uint32_t periodic_rows);
/* dotProductEx / dotProductExOpt (output, a, b, matrixEx) (CuProceduresApi.cpp
 * :201-209, 242-258): the product on the output region of matrixEx,
 * out(x, y) = sum_{k < a.columns} a(k, y) b(x, k).  The reference launches a
 * kernel its kernel file does not define; this is the region product. */
oaph_status oaph_dot_product_ex(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                const oaph_matrix* b, const oaph_matrix_ex* ex);
/* transposeEx(output, a, matrixEx) (CuProceduresApi.cpp:260-266,
 * CuTransposeProcedures.h:27-80): out(x, y) = a(y, x) for
 * x < min(W, column + columns), y < min(H, row + rows), (W, H) the extent of
 * the SetThreadsBlocks(columns, rows) grid the reference launches */
oaph_status oaph_transpose_ex(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                              const oaph_matrix_ex* ex);

/* utils::mapper::SetThreadsBlocks (oapUtils/ThreadsMapper.cpp:36-74): the
 * reference's launch-geometry map, kept bit-exact for callers that size
 * their own grids with it. */
oaph_status oaph_set_threads_blocks(uint32_t blocks[2], uint32_t threads[2], uint32_t w,
                                    uint32_t h, uint32_t threads_limit);

/* ------------------------------------------------------------------ */
/* Arnoldi: CuHArnoldi / CuHArnoldiCallback                            */
/* (ArnoldiPackage/ArnoldiProcedures.h:33-241, ArnoldiProceduresImpl.h:45-74) */
/* ------------------------------------------------------------------ */
typedef struct oaph_arnoldi oaph_arnoldi;

/* oap::VecMultiplicationType (ArnoldiPackage/oapGenericArnoldiApi.h:38-42) */
#define OAPH_TYPE_EIGENVECTOR 0
#define OAPH_TYPE_WV 1
/* oap::QRType (oapGenericArnoldiApi.h:31-36) */
#define OAPH_QRTYPE_NONE (-1)
#define OAPH_QRTYPE_QRGR 0
#define OAPH_QRTYPE_QRHT 1
/* oap::InitVVectorType (oapGenericArnoldiApi.h:44-48) */
#define OAPH_INIT_RANDOM 0
#define OAPH_INIT_FIRST_VALUE_IS_ONE 1
/* ArnUtils sort functions (ArnoldiPackage/ArnoldiUtils.h:60-72) */
#define OAPH_SORT_LARGEST_VALUES 0
#define OAPH_SORT_LARGEST_RE 1
#define OAPH_SORT_LARGEST_IM 2
#define OAPH_SORT_SMALLEST_VALUES 3
#define OAPH_SORT_SMALLEST_RE 4
#define OAPH_SORT_SMALLEST_IM 5
/* ArnUtils::CheckType (ArnoldiUtils.h:74-78) */
#define OAPH_CHECK_INTERNAL 0
#define OAPH_CHECK_FIRST_STOP 1
#define OAPH_CHECK_COUNTER 2
/* ArnUtils::TriangularHProcedureType (ArnoldiUtils.h:80-84) */
#define OAPH_CALC_IN_DEVICE 0
#define OAPH_CALC_IN_HOST 1
#define OAPH_CALC_IN_DEVICE_STEP 2
/* ArnUtils::Type (ArnoldiUtils.h:86) */
#define OAPH_UNDEFINED 0
#define OAPH_DEVICE 1
#define OAPH_HOST 2
/* matvec precision (extension, no reference counterpart) */
#define OAPH_PREC_FP64 0
#define OAPH_PREC_MIXED_FP32 1 /* FP32 matvec bulk, FP64 polish */
/* first Gram-Schmidt pass of each Arnoldi step (extension) */
#define OAPH_PROJ_FP64 0   /* FP64, in the dense operator's matvec epilogue */
#define OAPH_PROJ_BF16X3 1 /* BF16-split (hi + lo) basis, hi.hi + hi.lo + lo.hi on
                              the matrix cores with FP32 accumulation */

/* CuHArnoldiCallbackBase::MultiplyFunc (ArnoldiProceduresImpl.h:51):
 * w = A v.  w and v are 1-column device matrices owned by the solver; the
 * callback may use any oaph_* procedure on `api` (the solver's stream) and
 * must leave the result enqueued on that stream. */
typedef void (*oaph_multiply_fn)(oaph_matrix* w, oaph_matrix* v, oaph_api* api, void* user,
                                 int type);
/* ArnUtils::SortType (ArnoldiUtils.h:63-72) as a C comparator: nonzero when
 * Ritz value i = (re_i, im_i) sorts before j (the wanted ones first). */
typedef int (*oaph_sort_fn)(double re_i, double im_i, double re_j, double im_j, void* user);
/* CuHArnoldiCallbackBase::CheckFunc (ArnoldiProceduresImpl.h:53) */
typedef int (*oaph_check_fn)(double re, double im, oaph_matrix* vector, uint32_t index,
                             uint32_t max, void* user);

/* CuHArnoldi::CuHArnoldi / ~CuHArnoldi (ArnoldiProcedures.cpp:35-66) */
oaph_status oaph_arnoldi_create(oaph_arnoldi** out);
oaph_status oaph_arnoldi_destroy(oaph_arnoldi* a);
/* the procedures api the solver runs on (CuHArnoldi::m_cuApi, ArnoldiProcedures.h:81) */
oaph_api* oaph_arnoldi_api(oaph_arnoldi* a);
/* setters (ArnoldiProcedures.h:41-62, ArnoldiProcedures.cpp:68-126) */
oaph_status oaph_arnoldi_set_rho(oaph_arnoldi* a, double rho);
oaph_status oaph_arnoldi_set_blimit(oaph_arnoldi* a, double blimit);
oaph_status oaph_arnoldi_set_qr_type(oaph_arnoldi* a, int qrtype);
oaph_status oaph_arnoldi_set_vec_init_type(oaph_arnoldi* a, int init_type);
oaph_status oaph_arnoldi_set_sort_type(oaph_arnoldi* a, int sort_type);
oaph_status oaph_arnoldi_set_check_type(oaph_arnoldi* a, int check_type);
/* CuHArnoldi::setSortType with a caller's comparator (ArnoldiProcedures.cpp
 * :87-90, 370-380): the Ritz values of each restart are ordered on the host
 * by a stable sort with `fn` (one more launch and host round trip per
 * restart); oaph_arnoldi_set_sort_type (the built-in orders, applied on the
 * device) replaces it, and fn = NULL removes it. */
oaph_status oaph_arnoldi_set_sort_fn(oaph_arnoldi* a, oaph_sort_fn fn, void* user);
oaph_status oaph_arnoldi_set_check_counts(oaph_arnoldi* a, uint32_t count);
/* residual tolerance: a pair is converged when ||A x - l x|| <= tol * norm_a
 * (the reference stores this value but never reads it, ArnoldiProcedures.cpp:102-105) */
oaph_status oaph_arnoldi_set_check_tolerance(oaph_arnoldi* a, double tolerance);
oaph_status oaph_arnoldi_set_outputs_eigenvalues(oaph_arnoldi* a, double* re, double* im);
oaph_status oaph_arnoldi_set_outputs_eigenvectors(oaph_arnoldi* a, oaph_matrix** vectors);
oaph_status oaph_arnoldi_set_output_type(oaph_arnoldi* a, int type);
oaph_status oaph_arnoldi_set_calc_triangular_h_type(oaph_arnoldi* a, int type);
/* CuHArnoldiCallbackBase::setCallback / setCheckCallback (ArnoldiProceduresImpl.cpp:42-52) */
oaph_status oaph_arnoldi_set_callback(oaph_arnoldi* a, oaph_multiply_fn fn, void* user);
oaph_status oaph_arnoldi_set_check_callback(oaph_arnoldi* a, oaph_check_fn fn, void* user);
/* CuHArnoldiDefault::setMatrix (ArnoldiProceduresImpl.h:32): a dense device
 * operator used instead of the callback, with the solver's own matvec kernel */
oaph_status oaph_arnoldi_set_matrix(oaph_arnoldi* a, const oaph_matrix* A);
/* CuHArnoldiCallback::execute(hdim, wanted, info, type) (ArnoldiProceduresImpl.cpp:54-57,
 * ArnoldiProcedures.cpp:129-140).  Returns OAPH_ERR_NOT_CONVERGED when the
 * restart cap is hit; outputs then hold the best available pairs. */
oaph_status oaph_arnoldi_execute(oaph_arnoldi* a, uint32_t hdim, uint32_t wanted,
                                 const oaph_matrix_info* info, int type);
/* CuHArnoldi::testOutcome(index) (ArnoldiProcedures.cpp:642-649, 521-542):
 * ||A x - lambda x||_2 with the full complex lambda */
oaph_status oaph_arnoldi_test_outcome(oaph_arnoldi* a, uint32_t index, double* out);

/* extensions (no reference counterpart) */
oaph_status oaph_arnoldi_set_seed(oaph_arnoldi* a, uint64_t seed);
oaph_status oaph_arnoldi_set_norm_a(oaph_arnoldi* a, double norm_a);
oaph_status oaph_arnoldi_set_max_restarts(oaph_arnoldi* a, uint32_t max_restarts);
/* OAPH_PREC_MIXED_FP32 (dense operator only): the first execute after
 * oaph_arnoldi_set_matrix makes an FP32 copy of the operator; call set_matrix
 * again after changing the operator's values.  When every entry is an FP32
 * value the copy IS the operator and all cycles use it (no FP64 restart);
 * otherwise FP32 cycles to the mixed tolerance, then an FP64 restart from the
 * sum of the wanted and guard Ritz vectors.  Residuals are explicit FP64
 * either way.  An operator whose row length is not a multiple of 4, or with
 * an entry beyond the FP32 range, has no usable FP32 copy: the solve then
 * runs FP64 cycles only (stats.mixed_declined = 1). */
oaph_status oaph_arnoldi_set_precision(oaph_arnoldi* a, int precision);
/* The first Gram-Schmidt pass V^H w of every Arnoldi step: OAPH_PROJ_FP64
 * (default) or OAPH_PROJ_BF16X3 -- a BF16 hi + lo copy of the basis (about
 * 16 significant bits, kept current column by column) and
 * v_mfma_f32_16x16x16_bf16 products hi.hi + hi.lo + lo.hi with FP32
 * accumulation per block; the second pass (DGKS re-orthogonalisation), H,
 * the restart and the residuals stay FP64.  Used for n % 64 == 0 and
 * Krylov dimensions up to 64, else FP64.  No reference counterpart. */
oaph_status oaph_arnoldi_set_projection(oaph_arnoldi* a, int mode);
/* Staged solve (opt-in, default none): at each tolerance tols[0] > tols[1]
 * > ... (relative to the largest |Ritz value|) the cycles stop and the solve
 * restarts from the sum of the wanted (and guard) Ritz vectors (in the FP32
 * phase with OAPH_PREC_MIXED_FP32, stages at or below the mixed tolerance
 * ignored).  n = 0: no stages.  Such a restart discards every direction
 * outside the current wanted set: when the wanted eigenvalues are close in
 * the sort order to the next ones it can settle on a wrong set (see
 * oaph_arnoldi_set_restart_guard).  No reference counterpart. */
oaph_status oaph_arnoldi_set_restart_stages(oaph_arnoldi* a, const double* tols, uint32_t n);
/* Ritz vectors after the k wanted ones (next in the wanted order) that a
 * Ritz-sum restart (the mixed flow's FP64 start, stages) also adds into its
 * sum, so that a wanted eigenvalue whose Ritz value is still ranked just
 * below the k-th is kept in the new start vector (default 4; at most m - k
 * are used).  No reference counterpart. */
oaph_status oaph_arnoldi_set_restart_guard(oaph_arnoldi* a, uint32_t guard);
/* Weight of the guard Ritz vectors in a Ritz-sum restart (0..1, default
 * 1e-3; the wanted ones have weight 1).  No reference counterpart. */
oaph_status oaph_arnoldi_set_restart_guard_weight(oaph_arnoldi* a, double weight);
/* Smallest number of Ritz vectors an exact-shift restart keeps (the rest,
 * m minus that many, are the shifts).  0: the reference's k, plus
 * min(nconv, (m - k) / 2) as ARPACK adjusts it; a larger value keeps more of
 * the basis per restart (fewer matvecs on clustered spectra, more restarts);
 * capped at m - 1.  OAPH_KEEP_AUTO (default): m / 2 in the mixed flow (the
 * Krylov-Schur practice), 0 in FP64 and for an operator whose FP32 copy is
 * exact (all cycles FP32) -- those restart as the reference does.  No
 * reference counterpart. */
#define OAPH_KEEP_AUTO 0xFFFFFFFFu
oaph_status oaph_arnoldi_set_restart_keep(oaph_arnoldi* a, uint32_t keep);
/* How the restart finds the eigenpairs of H (extension, no reference
 * counterpart; the reference triangularises H by unshifted QR steps,
 * CuTriangularH.h).  For hdim <= 64 the eigenvalues of H come first
 * (Ehrlich-Aberth iteration on det(H - zI) by Hyman's recurrence).
 * 2 (default): the wanted eigenvectors of H by inverse iteration (LU of
 * H - lambda I with partial pivoting, LAPACK zlaein's method), each accepted
 * on its growth and an FP64 residual <= 64 hdim eps ||H||_F; no Schur form is
 * formed.  1: the Aberth values serve as perfect shifts for the Schur form,
 * eight per bulge train, smallest modulus first.  0: Wilkinson shifts only.
 * Where a method does not succeed (no convergence, a vector not accepted, a
 * split that does not happen) the next lower one takes over; with a
 * comparator (oaph_arnoldi_set_sort_fn) method 2 runs as 1.  Same Ritz pairs
 * to rounding either way. */
oaph_status oaph_arnoldi_set_schur_method(oaph_arnoldi* a, int method);
/* OAPH_PREC_MIXED_FP32: residual tolerance of the FP32 phase, relative to
 * the largest Ritz value (default 1e-7) */
oaph_status oaph_arnoldi_set_mixed_tolerance(oaph_arnoldi* a, double tol_low);
/* OAPH_PREC_MIXED_FP32, what follows the FP32 phase (no reference
 * counterpart).  1 (default): FP32-operand cycles go on from the Ritz-vector
 * sum down to `fraction` (0 < fraction <= 1, default 0.5) of the requested
 * tolerance; then each wanted Ritz vector gets its FP64 Rayleigh quotient
 * and explicit FP64 residual against the FP64 operator, and FP64 cycles run
 * only when some residual exceeds the tolerance (stats.fp64_polish).  The
 * FP32-operand matvec accumulates in FP64, so those cycles are Arnoldi on
 * the FP32-rounded operator to FP64 rounding; whether its eigenvectors meet
 * the tolerance for the FP64 operator depends on the operator and is decided
 * by the check alone.  0: FP64 cycles always follow the FP32 phase. */
oaph_status oaph_arnoldi_set_mixed_check(oaph_arnoldi* a, int on, double fraction);
/* OAPH_PREC_MIXED_FP32: a first phase on an FP16 copy of the operator (no
 * reference counterpart).  on = 1: cycles with the FP16-operand matvec (the
 * copy holds the FP16 rounding of 2^s A, the largest entry in [2^14, 2^15);
 * x rounded to FP32,
 * FP32 products and per-lane sums, 2 bytes per operator value) until the
 * Ritz estimates reach tol (0 < tol < 1, relative to the largest Ritz value,
 * default 2e-3); the FP32 phase then starts from the sum of their wanted and
 * guard Ritz vectors.  The copy is made (and kept) on the first solve that
 * uses it: 2 bytes per operator value per plane of device memory; an
 * operator whose row length is not a multiple of 8 or with a non-finite
 * entry skips the phase, as does one whose entries are all FP32 values (its
 * cycles run on the FP32 copy alone).  On by default. */
oaph_status oaph_arnoldi_set_f16_phase(oaph_arnoldi* a, int on, double tol);
/* start vector for the next execute (n complex values, NULL im = real) */
oaph_status oaph_arnoldi_set_start_vector(oaph_arnoldi* a, const double* re, const double* im,
                                          uint32_t n);
typedef struct oaph_arnoldi_stats {
  uint32_t restarts;
  uint32_t matvecs;
  uint32_t matvecs_low; /* matvecs done in reduced precision */
  uint32_t nconv;
  double seconds;       /* wall time of execute */
  double max_resid;     /* max explicit ||A x - l x|| over the wanted pairs */
  uint32_t fallbacks;   /* mixed flow: times the FP64 cycles met a wanted
                           eigenvalue the FP32 phase had not found and went
                           back to the FP32 operand */
  uint32_t mixed_declined; /* 1: OAPH_PREC_MIXED_FP32 was asked for but the
                              FP32 copy is unusable (row length not a multiple
                              of 4, or an entry beyond the FP32 range); the
                              solve ran FP64 cycles only */
  uint32_t restarts_perfect_shift; /* restarts whose Schur form came from the
                                      perfect shifts (oaph_arnoldi_set_schur_method) */
  uint32_t restarts_inverse_iteration; /* restarts that took the Aberth eigenvalues
                                          and inverse-iteration eigenvectors of H
                                          (method 2) without a Schur form */
  uint32_t fp64_polish; /* mixed flow with the FP64 check (the default): 1 when
                           the Ritz pairs of the FP32 cycles failed the explicit
                           FP64 residual check and FP64 cycles polished them */
  uint32_t matvecs_f16; /* of matvecs_low, those on the FP16 copy */
  double check_resid;   /* mixed flow: max explicit FP64 residual of the FP32
                           cycles' pairs at the FP64 check (0: no check ran) */
} oaph_arnoldi_stats;
oaph_status oaph_arnoldi_get_stats(oaph_arnoldi* a, oaph_arnoldi_stats* stats);

/* Average duration (HIP events on the solver's stream) of the solver's own
 * operator application for the dense operator set with
 * oaph_arnoldi_set_matrix: FP64 (low = 0), the FP32 copy (low = 1) or the
 * FP16 copy (low = 2), as the
 * last Arnoldi step of a cycle launches it (the matvec kernel also forms the
 * per-block partial sums of V^H y against the m basis vectors, the first
 * Gram-Schmidt pass's reduction, when its grid has at most 256 blocks). */
oaph_status oaph_arnoldi_time_matvec(oaph_arnoldi* a, int low, uint32_t iters, double* avg_ms);
/* y = A x with the solver's dense operator (oaph_arnoldi_set_matrix) at the
 * given precision: 0 the FP64 planes, 1 the FP32 copy (FP64 products and
 * sums: y is the FP64 product of the rounded operator with x), 2 the FP16
 * copy (x rounded to FP32, FP32 products and per-lane sums, FP64 cross-lane
 * sums); copies are made on first use; x and y complex device vectors of n
 * elements.  No reference counterpart (testing and measurement). */
oaph_status oaph_arnoldi_apply_operator(oaph_arnoldi* a, int low, const oaph_matrix* x,
                                        oaph_matrix* y);
/* Shader-clock stamps of the phases of the last restart kernel: load+Schur,
 * eigenvectors+estimates, ordering, wanted Ritz vectors, exact shifts
 * (6 values; consecutive differences are the phase durations in cycles). */
oaph_status oaph_arnoldi_restart_profile(oaph_arnoldi* a, long long* clk6);

/* ------------------------------------------------------------------ */
/* measurement support (no reference counterpart)                      */
/* ------------------------------------------------------------------ */
/* Seeded synthetic square operator written in place (counter-based
 * SplitMix64 + Box-Muller):
 *   kind 0: A = diag(1/(1+i)) + p1 G / sqrt(n), real (im plane zeroed if present)
 *   kind 1: A = diag(exp(-k/p2) e^{2 pi i u_k}) + p1 (G_re + i G_im) / sqrt(n) */
oaph_status oaph_synth_matrix(oaph_matrix* A, int kind, uint64_t seed, double p1, double p2);
/* Convection-diffusion-reaction operator -(alpha u_xx + u_yy) + pe (u_x + u_y)
 * + c u on the grid x grid interior of the unit square (h = 1/(grid+1),
 * Dirichlet, central differences), unknown (ix, iy) at ix + grid * iy,
 * written dense into the re plane (im zeroed if present) with every
 * coefficient rounded to FP32.  A must be grid^2 x grid^2. */
oaph_status oaph_synth_cdr(oaph_matrix* A, uint32_t grid, double pe, double c, double alpha);
/* Launch shape of the streaming matvec (fp32 = 1: the FP32-operand form used
 * by OAPH_PREC_MIXED_FP32, 2: its FP16-operand form, 0: FP64; fp32 = 3 sets
 * whether the FP16-operand form stages x in LDS, variant 1 (default) or 0,
 * fp32 = 4 whether the waves of every streaming matvec start at rotated
 * column chunks, 1 (default) or 0): variant index
 * into a table of (rows per wave read
 * together, loads in flight per row and plane, groups of such rows per wave
 * taken one after another); process-wide.  No reference counterpart (tuning
 * knob). */
oaph_status oaph_set_matvec_shape(int fp32, int variant);
/* Launch shape (variant index into a table of (rows per wave, loads in
 * flight), one group) of the matvec that runs where row groups would leave
 * SIMDs without a wave (the operators of a few thousand rows): fp32 = 0
 * FP64, 1 FP32 operand, 2 FP16 operand; process-wide.  No reference
 * counterpart (tuning knob). */
oaph_status oaph_set_matvec_fallback_shape(int fp32, int variant);
/* extension (no reference counterpart): which kernel forms dotProduct.
 * 1 (default): products with W, H >= 256 and K >= 64 on the FP64 matrix
 * cores (v_mfma_f64_16x16x4_f64, 128 x 128 tiles), smaller ones on the
 * vector-FMA tiles; 2: every product on the matrix cores; 0: every product on
 * the vector-FMA tiles.  Both are FP64 throughout and agree to rounding (they
 * sum the K terms in different orders). */
oaph_status oaph_set_gemm_path(int mfma);
oaph_status oaph_matvec_shape_info(int fp32, int variant, int* rows_per_wave,
                                   int* loads_in_flight, int* row_groups);
/* Average duration of `iters` back-to-back dotProduct calls (HIP events on
 * the api's stream), after one warm-up call. */
oaph_status oaph_time_dot_product(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                  const oaph_matrix* b, uint32_t iters, double* avg_ms);
/* Average duration of `iters` back-to-back magnitude (sumsq != 0) or sum
 * reductions of m on the api's stream (the kernel alone: the result stays on
 * the device), after one warm-up launch. */
oaph_status oaph_time_reduction(oaph_api* api, int sumsq, const oaph_matrix* m, uint32_t iters,
                                double* avg_ms);
/* Workgroups of the magnitude/sum reductions at most (1..2048, default 256:
 * one wave per SIMD); process-wide tuning knob. */
oaph_status oaph_set_reduce_blocks(int blocks);

/* ------------------------------------------------------------------ */
/* Activation procedures (CuProceduresApi.h:239-290): every plane of the
 * output on its own; OAPH_FN_MULTIPLY_* multiply the output in place by the
 * derivative at the input.                                               */
/* ------------------------------------------------------------------ */
#define OAPH_FN_SIGMOID 0
#define OAPH_FN_DSIGMOID 1
#define OAPH_FN_MULTIPLY_DSIGMOID 2
#define OAPH_FN_LINEAR 3
#define OAPH_FN_DLINEAR 4
#define OAPH_FN_TANH 5
#define OAPH_FN_DTANH 6
#define OAPH_FN_MULTIPLY_DTANH 7
#define OAPH_FN_SIN 8
#define OAPH_FN_DSIN 9 /* cos */
#define OAPH_FN_MULTIPLY_DSIN 10

/* sigmoid(out, in) ... dsin(out, in) over the whole matrix */
oaph_status oaph_activation(oaph_api* api, int fn, oaph_matrix* out, const oaph_matrix* in);
/* ... (out, in, dims[2]): elements with column < dims[0] and row < dims[1] */
oaph_status oaph_activation_dims(oaph_api* api, int fn, oaph_matrix* out, const oaph_matrix* in,
                                 const uint32_t dims[2]);
/* ... (out, in, dims[2][2]): column < dims[0][0] and (row mod dims[1][1]) <
 * dims[0][1] (the periodic form, CuFuncProcedures.h cuda_inRangePD) */
oaph_status oaph_activation_dims_periodic(oaph_api* api, int fn, oaph_matrix* out,
                                          const oaph_matrix* in, const uint32_t dims[2][2]);

/* ------------------------------------------------------------------ */
/* FP64 feed-forward network with oap's training semantics
 * (oapNeuralDevice/oapNetwork.h:33-120, oapNeuralHost/oapGenericNeuralApi.h
 * :195-364): per-sample forward / accumulateErrors (MEAN_SQUARE_ERROR) /
 * backPropagation, updateWeights after a batch.  Activations: OAPH_FN_SIGMOID,
 * _LINEAR, _TANH, _SIN.                                                   */
/* ------------------------------------------------------------------ */
typedef struct oaph_network oaph_network;

/* api NULL: the network owns a stream of its own */
oaph_status oaph_network_create(oaph_network** out, oaph_api* api);
oaph_status oaph_network_destroy(oaph_network* n);
/* Network::createLayer(neurons, addBias, activation); connects the previous layer */
oaph_status oaph_network_add_layer(oaph_network* n, uint32_t neurons, int has_bias,
                                   int activation);
oaph_status oaph_network_set_learning_rate(oaph_network* n, double lr);
/* weights layer -> layer+1: (neurons+bias of layer) columns x (neurons of
 * layer+1) rows, element (c, r) at c + r * columns */
oaph_status oaph_network_set_host_weights(oaph_network* n, uint32_t layer, const double* w);
oaph_status oaph_network_get_host_weights(oaph_network* n, uint32_t layer, double* w);
oaph_status oaph_network_set_inputs(oaph_network* n, const double* x);
oaph_status oaph_network_set_expected(oaph_network* n, const double* t);
oaph_status oaph_network_forward(oaph_network* n);
oaph_status oaph_network_get_outputs(oaph_network* n, double* y);
oaph_status oaph_network_accumulate_errors(oaph_network* n);
oaph_status oaph_network_backward(oaph_network* n);
/* W <- W - (lr / N) sum of nabla W (N accumulated errors), then post_step */
oaph_status oaph_network_update_weights(oaph_network* n);
oaph_status oaph_network_post_step(oaph_network* n);
/* mean of 0.5 e^2 over the accumulated errors */
oaph_status oaph_network_calculate_error(oaph_network* n, double* out);

/* ------------------------------------------------------------------ */
/* Learned-start MLP: tanh hidden layers, linear output, BF16 operands with
 * FP32 accumulation on the matrix cores, FP32 master weights, Adam.  The
 * batched form of oap's neural routines (oapNeuralHost/oapGenericNeuralApi.h
 * :195-354).  Loss of a batch: 1/(2B) sum_b sum_j w_j (y_bj - t_bj)^2.
 * Widths d_l <= 512.  Parameters flat: per layer W (d_{l+1} x d_l,
 * row-major), then b.  Host arrays are row-major (n x d).                 */
/* ------------------------------------------------------------------ */
typedef struct oaph_mlp oaph_mlp;

oaph_status oaph_mlp_create(oaph_mlp** out, const uint32_t* dims, uint32_t ndims,
                            uint32_t max_batch);
oaph_status oaph_mlp_destroy(oaph_mlp* m);
oaph_status oaph_mlp_num_params(oaph_mlp* m, size_t* n);
/* sets the parameters and resets the optimiser state */
oaph_status oaph_mlp_set_params(oaph_mlp* m, const float* params);
oaph_status oaph_mlp_get_params(oaph_mlp* m, float* params);
/* per-output loss weights w_j (default 1) */
oaph_status oaph_mlp_set_output_weights(oaph_mlp* m, const float* w);
oaph_status oaph_mlp_predict(oaph_mlp* m, const float* X, uint32_t n, float* Y);
/* loss and flat gradients of one batch (B <= max_batch) */
oaph_status oaph_mlp_gradients(oaph_mlp* m, const float* X, const float* T, uint32_t B,
                               float* grads, float* loss);
/* training set kept in HBM for oaph_mlp_train */
oaph_status oaph_mlp_set_dataset(oaph_mlp* m, const float* X, const float* T, uint32_t n);
/* `epochs` shuffled passes of Adam (lr, betas 0.9 / 0.999) in batches of B;
 * losses (may be NULL) receives the mean batch loss of each epoch */
oaph_status oaph_mlp_train(oaph_mlp* m, uint32_t B, uint32_t epochs, float lr, uint64_t seed,
                           float* losses);

/* ------------------------------------------------------------------ */
/* Square operator of a rectangular matrix: oap::RecToSquareApi           */
/* (oapMatrixCuda/RecToSquareApi.h:33-58, RecToSquareApi.cpp:57-141;      */
/* oapMatrixCuda/SquareMatrixApi.cpp:143-181).  For a device matrix M of  */
/* R rows and C columns the operator is S = M M^T (R x R, plain transpose */
/* as in the reference); when C == R, S is M itself (RecToSquareApi.cpp   */
/* :57-63,94-97).  M stays owned by the caller and must outlive the       */
/* handle; work is enqueued on the api's stream.                          */
/* ------------------------------------------------------------------ */
typedef struct oaph_rts oaph_rts;

/* RecToSquareApi(api, recMatrix, deallocate = false) over a device matrix */
oaph_status oaph_rts_create(oaph_rts** out, oaph_api* api, const oaph_matrix* rec);
oaph_status oaph_rts_destroy(oaph_rts* t);
/* getMatrixInfo: R x R with M's planes (RecToSquareApi.cpp:31-43) */
oaph_status oaph_rts_info(const oaph_rts* t, oaph_matrix_info* info);
/* getDeviceMatrix / createDeviceMatrix: out (R x R) = S; out's planes
 * select the product's planes as in oaph_dot_product */
oaph_status oaph_rts_square(oaph_rts* t, oaph_matrix* out);
/* getDeviceSubMatrix(rindex, rlength, dmatrix): rows [rindex, rindex + n)
 * of S with n = min(rlength, R - rindex); out must have R columns and is
 * given exactly n rows (its planes are reallocated when its row count
 * differs, as the reference reallocates, SquareMatrixApi.cpp:162-168).
 * rindex >= R or rlength == 0 -> OAPH_ERR_ARG (the reference throws,
 * RecToSquareApi.cpp:143-165). */
oaph_status oaph_rts_sub_matrix(oaph_rts* t, uint32_t rindex, uint32_t rlength, oaph_matrix* out);
/* getDeviceRowVector(index, dmatrix): row `index` of S into out (R x 1 row) */
oaph_status oaph_rts_row_vector(oaph_rts* t, uint32_t index, oaph_matrix* out);
/* w = S v without forming S: u = M^T v, then w = M u (the product the
 * reference's multiply callbacks form block by block,
 * oap2dt3dDevice/MainAPExecutor.cpp:42-127).  v, w are R x 1 columns; w
 * needs an imaginary plane when M or v has one.  Enqueued on api's stream
 * (NULL: the handle's api); inside a solver multiply callback pass the api
 * the callback receives, so the product is ordered with the solver's work.
 * Calls on one handle must not overlap (shared workspace). */
oaph_status oaph_rts_multiply(oaph_rts* t, oaph_api* api, oaph_matrix* w, const oaph_matrix* v);

/* ------------------------------------------------------------------ */
/* PCA data path of the oap2dt3d application (host code): PNG images ->  */
/* cropped pixel columns -> the data matrix that RecToSquareApi squares. */
/* ------------------------------------------------------------------ */
/* 8-bit non-interlaced PNG, colour types grey / RGB / palette (raw index
 * bytes) / grey+alpha / RGBA: raw rows (height x width*colors bytes), as
 * PngFile reads them (oapAppUtils/PngFile.cpp:107-140).  pixels NULL: size
 * query.  Decoder of this library's own (zlib). */
oaph_status oaph_png_read(const char* path, uint32_t* width, uint32_t* height, uint32_t* colors,
                          uint8_t* pixels, size_t cap);
/* writes such rows as a PNG (colors 1, 2, 3, 4), for synthetic image stacks */
oaph_status oaph_png_write(const char* path, uint32_t width, uint32_t height, uint32_t colors,
                           const uint8_t* pixels);
/* oap::GetOptWidth / GetOptHeight (oapAppUtils/GraphicsUtilsImpl.h):
 * out = {width, begin column, height, begin row} of the non-constant window */
oaph_status oaph_image_opt_size(const uint8_t* bitmap, uint32_t width, uint32_t height,
                                uint32_t colors, uint32_t out[4]);
/* DataLoader's load process (oapAppUtils/DataLoader.cpp:162-230) on given
 * windows sizes[4i..4i+3] = {w, wbegin, h, hbegin}: out = the windows used,
 * loads (may be NULL) = bitmap loads the reference performs per image */
oaph_status oaph_dataloader_plan(const uint32_t* sizes, uint32_t count, uint32_t* out,
                                 uint32_t* loads);
typedef struct oaph_dataloader oaph_dataloader;
/* DataLoader(images) over PNG files (truncate: crop to the optimal window,
 * PngFile's default); every image gets one column of pixel / 0xFFFFFF
 * values (Image.cpp:134-178).  Windows that do not fit an image or differ
 * in area -> OAPH_ERR_SHAPE (the reference reads past the image, or throws
 * NotIdenticalLengths). */
oaph_status oaph_dataloader_create(oaph_dataloader** out, const char* const* paths,
                                   uint32_t count, int truncate);
oaph_status oaph_dataloader_destroy(oaph_dataloader* d);
/* getMatrixInfo: columns = images, rows = window pixels, re plane only */
oaph_status oaph_dataloader_info(const oaph_dataloader* d, oaph_matrix_info* info);
oaph_status oaph_dataloader_window(const oaph_dataloader* d, uint32_t image, uint32_t out[4],
                                   uint32_t* loads);
/* createColumnVector(index) into host memory (rows values) */
oaph_status oaph_dataloader_column(const oaph_dataloader* d, uint32_t image, double* out);
/* createMatrix() into a device matrix of images columns x pixels rows */
oaph_status oaph_dataloader_matrix(const oaph_dataloader* d, oaph_matrix* m);

/* ------------------------------------------------------------------ */
/* Batched covariance eigenproblems.  Problem i is C_i = X_i^T X_i / s with
 * X_i an s x n_i FP32 row-major sample matrix: the M M^T operator the
 * reference forms with RecToSquareApi / SquareMatrixApi
 * (oapMatrixCuda/RecToSquareApi.cpp:57-141, oapMatrixCuda/SquareMatrixApi.cpp)
 * applied implicitly, solved for the `wanted` extreme eigenpairs by the same
 * implicitly restarted Arnoldi iteration as oaph_arnoldi_execute
 * (ArnoldiPackage/ArnoldiProcedures.cpp:129-140), all problems advancing in
 * one persistent launch (one workgroup per problem).  Limits: n_i in
 * [2, 1024], s <= 1024, hdim <= 24, n_i >= hdim.                           */
/* ------------------------------------------------------------------ */
typedef struct oaph_batch oaph_batch;

/* Sizes n_i = 64 + floor(961 u_i) of the seeded synthetic batch (u_i the
 * first uniform of problem i's SplitMix64 stream). */
oaph_status oaph_batch_c3_sizes(uint32_t count, uint64_t seed, uint32_t* n_out);
oaph_status oaph_batch_create(oaph_batch** out, uint32_t count, const uint32_t* n, uint32_t s);
oaph_status oaph_batch_destroy(oaph_batch* b);
/* X_i as s x n_i row-major floats (host memory, synchronous copy). */
oaph_status oaph_batch_upload_x(oaph_batch* b, uint32_t i, const float* X);
oaph_status oaph_batch_download_x(oaph_batch* b, uint32_t i, float* X);
/* Synthetic 16-bit grey patches (k/65535, k uniform) plus three planted
 * rank-1 components of amplitude 5, generated in HBM. */
oaph_status oaph_batch_synth_c3(oaph_batch* b, uint64_t seed);
/* Launch variant (threads per workgroup x rows per wave step x steps of
 * rows in flight); see oaph_batch_variant_shape.  Default 1. */
oaph_status oaph_batch_set_variant(oaph_batch* b, int variant);
oaph_status oaph_batch_variant_info(int variant, int* threads, int* rows);
/* The variant's shape with the depth of its row ring (register buffers of
 * rows: depth - 1 steps of loads in flight while one is computed). */
oaph_status oaph_batch_variant_shape(int variant, int* threads, int* rows, int* depth);
/* Capture the solve launch in a HIP graph and replay it while its
 * arguments are unchanged (default on). */
oaph_status oaph_batch_set_graph(oaph_batch* b, int enable);
/* Solves every problem: residual test beta |e_m^T y| <= tol * max |Ritz|,
 * sort_type as OAPH_SORT_*.  Synchronous. */
oaph_status oaph_batch_solve(oaph_batch* b, uint32_t hdim, uint32_t wanted, int sort_type,
                             double tol, uint32_t max_restarts, uint64_t seed);
/* Times every variant (`reps` solves each after a warm-up), selects the
 * fastest and appends it to the text cache at cache_path (NULL: no cache).
 * With times_ms == NULL a cached choice for the same batch shape and device
 * is taken without timing.  times_ms, if given, receives one time per
 * variant. */
oaph_status oaph_batch_autotune(oaph_batch* b, uint32_t hdim, uint32_t wanted, int sort_type,
                                double tol, uint32_t max_restarts, uint64_t seed, uint32_t reps,
                                const char* cache_path, int* chosen, double* times_ms);
/* Per problem (count x wanted, wanted order): Rayleigh quotients of the
 * normalised Ritz vectors, explicit FP64 residuals ||C x - lambda x||_2,
 * converged count, restarts, operator applications; and the kernel time of
 * the last solve.  Any pointer may be NULL. */
oaph_status oaph_batch_results(oaph_batch* b, double* evals, double* resid, uint32_t* nconv,
                               uint32_t* restarts, uint32_t* matvecs, double* kernel_ms);
/* Per problem (count x 5): shader-clock cycles of the last solve spent in the
 * operator, the orthogonalisation, the restarts, the basis update and the
 * final Ritz vectors (profiling). */
oaph_status oaph_batch_phase_cycles(oaph_batch* b, long long* cycles);
/* Ritz vectors of problem i: wanted x n_i doubles, vector t at x + t * n_i. */
oaph_status oaph_batch_eigenvectors(oaph_batch* b, uint32_t i, double* x);

#ifdef __cplusplus
}
#endif

#endif /* OAPH_H */
