// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.  A C ABI over the reference's own
// host code, compiled from the reference sources where they lie (see
// oracle/Makefile) into oracle/_ref/libref.so.  It lets the tests check the
// C restatement (oracle.c) against the reference itself: HostProcedures
// (oapMatrixCpu/HostProcedures.h:30-102) on the reference's CPU thread
// emulator, the reference's device functions run through its HostKernel
// emulator (oapMatrixCpu/HostKernel.h:31-70, the pattern of
// oapHostTests/oapQRTests.cpp:55-102), and utils::mapper::SetThreadsBlocks.
//
// Matrices cross this ABI as (isRe, isIm, columns, rows, re*, im*) with
// row-major planes; each call copies into freshly allocated reference host
// matrices (oap::host::NewMatrix) and back.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>

#include <string>

#include "ByteBuffer.h"
#include <vector>

#include "CuMatrixProcedures.h"
#include "GraphicUtils.h"
#include "HostKernel.h"
#include "MatrixParser.h"
#include "HostProcedures.h"
#include "ThreadsMapper.h"
#include "oapHostMatrixUtils.h"

namespace {

math::Matrix* make(int isre, int isim, uint32_t c, uint32_t r, const double* re,
                   const double* im) {
  math::Matrix* m = oap::host::NewMatrix(isre != 0, isim != 0, c, r, 0);
  size_t n = (size_t)c * r;
  if (isre && re) std::memcpy(m->reValues, re, n * sizeof(double));
  if (isim && im) std::memcpy(m->imValues, im, n * sizeof(double));
  return m;
}

void take(math::Matrix* m, double* re, double* im) {
  size_t n = (size_t)m->columns * m->rows;
  if (m->reValues && re) std::memcpy(re, m->reValues, n * sizeof(double));
  if (m->imValues && im) std::memcpy(im, m->imValues, n * sizeof(double));
  oap::host::DeleteMatrix(m);
}

// Runs a reference device function on the reference emulator with the
// reference's own launch geometry for a (w, h) output.
class Stub : public HostKernel {
 public:
  std::function<void()> body;
  Stub(uint32_t w, uint32_t h, std::function<void()> f) : body(std::move(f)) {
    calculateDims(w, h);
  }
  Stub(dim3 grid, dim3 block, std::function<void()> f) : body(std::move(f)) {
    setDims(grid, block);
  }
  void execute(const dim3&, const dim3&) override { body(); }
};

}  // namespace

#define M_ARGS(p) int p##_isre, int p##_isim, uint32_t p##_c, uint32_t p##_r, double *p##_re, double *p##_im
#define M_MAKE(p) make(p##_isre, p##_isim, p##_c, p##_r, p##_re, p##_im)

extern "C" {

void ref_set_threads_blocks(uint32_t blocks[2], uint32_t threads[2], uint32_t w, uint32_t h,
                            uint32_t limit) {
  uint b[2], t[2];
  utils::mapper::SetThreadsBlocks(b, t, w, h, limit);
  blocks[0] = b[0]; blocks[1] = b[1];
  threads[0] = t[0]; threads[1] = t[1];
}

// ---- HostProcedures
void ref_dot_product(M_ARGS(o), M_ARGS(a), M_ARGS(b)) {
  HostProcedures hp;
  math::Matrix *O = M_MAKE(o), *A = M_MAKE(a), *B = M_MAKE(b);
  hp.dotProduct(O, A, B);
  take(O, o_re, o_im); take(A, nullptr, nullptr); take(B, nullptr, nullptr);
}

void ref_dot_product_dims(M_ARGS(o), M_ARGS(a), M_ARGS(b), const uint32_t* dims6) {
  HostProcedures hp;
  math::Matrix *O = M_MAKE(o), *A = M_MAKE(a), *B = M_MAKE(b);
  uintt dims[3][2] = {{dims6[0], dims6[1]}, {dims6[2], dims6[3]}, {dims6[4], dims6[5]}};
  hp.dotProduct(O, A, B, dims);
  take(O, o_re, o_im); take(A, nullptr, nullptr); take(B, nullptr, nullptr);
}

void ref_transpose(M_ARGS(o), M_ARGS(a)) {
  HostProcedures hp;
  math::Matrix *O = M_MAKE(o), *A = M_MAKE(a);
  hp.transpose(O, A);
  take(O, o_re, o_im); take(A, nullptr, nullptr);
}

void ref_tensor_product(M_ARGS(o), M_ARGS(a), M_ARGS(b), const uint32_t* dims6) {
  HostProcedures hp;
  math::Matrix *O = M_MAKE(o), *A = M_MAKE(a), *B = M_MAKE(b);
  uintt dims[3][2] = {{dims6[0], dims6[1]}, {dims6[2], dims6[3]}, {dims6[4], dims6[5]}};
  hp.tensorProduct(O, A, B, dims);
  take(O, o_re, o_im); take(A, nullptr, nullptr); take(B, nullptr, nullptr);
}

void ref_sum(M_ARGS(a), double* sre, double* sim) {
  HostProcedures hp;
  math::Matrix* A = M_MAKE(a);
  floatt r = 0, i = 0;
  hp.sum(r, i, A);
  *sre = r; *sim = i;
  take(A, nullptr, nullptr);
}

void ref_set_vector(M_ARGS(V), uint32_t column, M_ARGS(v), uint32_t length) {
  HostProcedures hp;
  math::Matrix *MV = M_MAKE(V), *mv = M_MAKE(v);
  hp.setVector(MV, column, mv, length);
  take(MV, V_re, V_im); take(mv, nullptr, nullptr);
}

void ref_get_vector(M_ARGS(v), uint32_t length, M_ARGS(V), uint32_t column) {
  HostProcedures hp;
  math::Matrix *MV = M_MAKE(V), *mv = M_MAKE(v);
  hp.getVector(mv, length, MV, column);
  take(mv, v_re, v_im); take(MV, nullptr, nullptr);
}

void ref_set_identity(M_ARGS(o)) {
  HostProcedures hp;
  math::Matrix* O = M_MAKE(o);
  hp.setIdentity(O);
  take(O, o_re, o_im);
}

// ---- device functions on the emulator
void ref_conjugate_transpose(M_ARGS(o), M_ARGS(a)) {
  math::Matrix *O = M_MAKE(o), *A = M_MAKE(a);
  Stub s(O->columns, O->rows, [&] { CUDA_conjugateTransposeMatrix(O, A); });
  s.executeKernelAsync();
  take(O, o_re, o_im); take(A, nullptr, nullptr);
}

void ref_binary(int op, M_ARGS(o), M_ARGS(a), M_ARGS(b)) {
  math::Matrix *O = M_MAKE(o), *A = M_MAKE(a), *B = M_MAKE(b);
  Stub s(O->columns, O->rows, [&] {
    switch (op) {
      case 0: CUDA_addMatrix(O, A, B); break;
      case 1: CUDA_substractMatrices(O, A, B); break;
      case 2: CUDA_hadamardProduct(O, A, B); break;
      default: CUDA_phadamardProduct(O, A, B); break;
    }
  });
  s.executeKernelAsync();
  take(O, o_re, o_im); take(A, nullptr, nullptr); take(B, nullptr, nullptr);
}

void ref_multiply_constant(int re_only, M_ARGS(o), M_ARGS(a), double re, double im) {
  math::Matrix *O = M_MAKE(o), *A = M_MAKE(a);
  Stub s(O->columns, O->rows, [&] {
    if (re_only) CUDA_multiplyConstantReMatrix(O, A, re);
    else CUDA_multiplyConstantMatrix(O, A, re, im);
  });
  s.executeKernelSync();
  take(O, o_re, o_im); take(A, nullptr, nullptr);
}

void ref_set_diagonal(M_ARGS(o), double re, double im) {
  math::Matrix* O = M_MAKE(o);
  Stub s(O->columns, O->rows, [&] { CUDA_setDiagonalMatrix(O, re, im); });
  s.executeKernelAsync();
  take(O, o_re, o_im);
}

// CuProceduresApi::magnitude2Procedure (CuProceduresApi.cpp:765-803): per-block
// CUDA_magnitudeOpt partials, then host sum and sqrt
double ref_magnitude(M_ARGS(a)) {
  math::Matrix* A = M_MAKE(a);
  uint b[2], t[2];
  utils::mapper::SetThreadsBlocks(b, t, A->columns, A->rows, 1024);
  std::vector<floatt> sums((size_t)b[0] * b[1], 0.0), buffer((size_t)t[0] * t[1], 0.0);
  Stub s(dim3(b[0], b[1]), dim3(t[0], t[1]),
         [&] { CUDA_magnitudeOpt(sums.data(), A, buffer.data()); });
  s.executeKernelAsync();
  double out = 0;
  for (double v : sums) out += v;
  take(A, nullptr, nullptr);
  return std::sqrt(out);
}

// CUDA_QRGR (CuQRProcedures_GR.h:106-150), one block of m x m threads
void ref_qrgr(M_ARGS(q), M_ARGS(r), M_ARGS(a)) {
  math::Matrix *Q = M_MAKE(q), *R = M_MAKE(r), *A = M_MAKE(a);
  math::Matrix* t1 = oap::host::NewMatrixRef(A);
  math::Matrix* t2 = oap::host::NewMatrixRef(A);
  math::Matrix* t3 = oap::host::NewMatrixRef(A);
  math::Matrix* t4 = oap::host::NewMatrixRef(A);
  Stub s(A->columns, A->rows, [&] { CUDA_QRGR(Q, R, A, t1, t2, t3, t4); });
  s.executeKernelAsync();
  take(Q, q_re, q_im); take(R, r_re, r_im); take(A, nullptr, nullptr);
  oap::host::DeleteMatrix(t1); oap::host::DeleteMatrix(t2);
  oap::host::DeleteMatrix(t3); oap::host::DeleteMatrix(t4);
}

int ref_is_upper_triangular(M_ARGS(a)) {
  math::Matrix* A = M_MAKE(a);
  int res = CUDA_isUpperTriangular(A);
  take(A, nullptr, nullptr);
  return res;
}

// ---- matrix formats: oap::host::ToString, NewMatrix(text), CreateArray,
// SaveMatrix / LoadMatrix into utils::ByteBuffer and its file framing ----
size_t ref_to_string(M_ARGS(a), char* buf, size_t cap) {
  math::Matrix* A = M_MAKE(a);
  std::string s;
  oap::host::ToString(s, A);
  take(A, nullptr, nullptr);
  if (buf && cap > s.size()) {
    std::memcpy(buf, s.data(), s.size());
    buf[s.size()] = 0;
  }
  return s.size();
}

// returns 0 when the text gives no matrix; else writes info and planes (cap elements)
int ref_parse_matrix(const char* text, int* isre, int* isim, uint32_t* c, uint32_t* r, double* re,
                     double* im, size_t cap) {
  math::Matrix* m = oap::host::NewMatrix(std::string(text));
  if (!m) return 0;
  *isre = m->reValues != nullptr;
  *isim = m->imValues != nullptr;
  *c = m->columns;
  *r = m->rows;
  size_t n = (size_t)m->columns * m->rows;
  if (n <= cap) take(m, re, im);
  else oap::host::DeleteMatrix(m);
  return 1;
}

// length of CreateArray(text, which), -1 when it fails; values into out (cap)
long ref_create_array(const char* text, unsigned which, double* out, size_t cap) {
  std::pair<floatt*, size_t> p = matrixUtils::CreateArray(std::string(text), which);
  if (!p.first) return -1;
  if (p.second <= cap) std::memcpy(out, p.first, p.second * sizeof(double));
  delete[] p.first;
  return (long)p.second;
}

// SaveMatrix of a host matrix (is_null: of nullptr) into a ByteBuffer
size_t ref_save_matrix(int is_null, M_ARGS(a), uint8_t* buf, size_t cap) {
  utils::ByteBuffer b;
  math::Matrix* A = is_null ? nullptr : M_MAKE(a);
  oap::host::SaveMatrix(A, b);
  if (A) take(A, nullptr, nullptr);
  size_t n = b.getUsedLength();
  if (buf && n <= cap) std::memcpy(buf, b.m_buffer, n);
  return n;
}

// LoadMatrix from bytes; returns 0 for a null record
int ref_load_matrix(const uint8_t* bytes, size_t len, int* isre, int* isim, uint32_t* c,
                    uint32_t* r, double* re, double* im, size_t cap) {
  utils::ByteBuffer b;
  b.push_back(reinterpret_cast<const char*>(bytes), len);
  math::Matrix* m = oap::host::LoadMatrix(b);
  if (!m) return 0;
  *isre = m->reValues != nullptr;
  *isim = m->imValues != nullptr;
  *c = m->columns;
  *r = m->rows;
  size_t n = (size_t)m->columns * m->rows;
  if (n <= cap) take(m, re, im);
  else oap::host::DeleteMatrix(m);
  return 1;
}

// oap::GetOptWidth / GetOptHeight over a row-major byte bitmap
void ref_opt_size(const uint8_t* bm, uint32_t w, uint32_t h, uint32_t colors, uint32_t* out) {
  std::vector<unsigned char*> rows(h);
  for (uint32_t y = 0; y < h; ++y) rows[y] = const_cast<unsigned char*>(bm + (size_t)y * w * colors);
  oap::OptSize ow = oap::GetOptWidth<unsigned char**, unsigned char>(rows.data(), w, h, colors);
  oap::OptSize oh = oap::GetOptHeight<unsigned char**, unsigned char>(rows.data(), w, h, colors);
  out[0] = (uint32_t)ow.optSize;
  out[1] = (uint32_t)ow.begin;
  out[2] = (uint32_t)oh.optSize;
  out[3] = (uint32_t)oh.begin;
}

// ByteBuffer::fwrite of the given bytes
void ref_buffer_fwrite(const char* path, const uint8_t* bytes, size_t len) {
  utils::ByteBuffer b;
  b.push_back(reinterpret_cast<const char*>(bytes), len);
  b.fwrite(std::string(path));
}

}  // extern "C"
