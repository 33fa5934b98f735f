// internal.h -- shared declarations of the gfx950 implementation behind oaph.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/oaph.h"

namespace oaph {

// Thread-local text of the last HIP failure (oaph_last_error()).
void set_last_error(const char* where, hipError_t e);
void set_last_message(const std::string& msg);

#define OAPH_HIP(expr)                                  \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) {                             \
      ::oaph::set_last_error(#expr, e_);                \
      return OAPH_ERR_HIP;                              \
    }                                                   \
  } while (0)

#define OAPH_CHECK(cond, code) \
  do {                         \
    if (!(cond)) return (code);\
  } while (0)

#define OAPH_TRY(expr)              \
  do {                              \
    oaph_status s_ = (expr);        \
    if (s_ != OAPH_OK) return s_;   \
  } while (0)

oaph_status ensure_init();

// A device scratch buffer that only grows.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  oaph_status reserve(size_t n);
  void release();
  template <typename T>
  T* as() const { return static_cast<T*>(ptr); }
};

}  // namespace oaph

// oaph_api: one HIP stream plus scratch for reductions (the reference's
// CuProceduresApi owns two device scalars and growable TBuffers,
// oapMatrixCuda/CuProceduresApi.cpp:64-89).
struct oaph_api {
  hipStream_t stream = nullptr;
  oaph::DevBuf partials;   // per-block partial sums
  oaph::DevBuf counters;   // last-block-done counters (kept zero between launches)
  oaph::DevBuf scalars;    // device-side results
  oaph::DevBuf pinned_host;  // not used for device memory; see host_scalars
  double* host_scalars = nullptr;  // pinned, 16 doubles
  oaph::DevBuf work;       // general device workspace (QR, triangular-H)
};

namespace oaph {

inline size_t numel(const oaph_matrix* m) { return (size_t)m->columns * m->rows; }

// ---- kernel launchers (each enqueues on `s`, returns OAPH_OK or HIP error) ----
// elementwise.hip
oaph_status launch_transpose(hipStream_t s, oaph_matrix* out, const oaph_matrix* a, bool conj);
oaph_status launch_set_diagonal(hipStream_t s, oaph_matrix* m, double re, double im);
oaph_status launch_binary(hipStream_t s, int op, oaph_matrix* out, const oaph_matrix* a,
                          const oaph_matrix* b);  // 0 add, 1 sub, 2 hadamard, 3 hadamard_vec
oaph_status launch_scale(hipStream_t s, oaph_matrix* out, const oaph_matrix* a, double re,
                         double im, bool re_only);
oaph_status launch_tensor(hipStream_t s, oaph_matrix* out, const oaph_matrix* a,
                          const oaph_matrix* b, const uint32_t dims[3][2]);
oaph_status launch_set_vector(hipStream_t s, oaph_matrix* V, uint32_t column, const oaph_matrix* v,
                              uint32_t length);
oaph_status launch_get_vector(hipStream_t s, oaph_matrix* v, uint32_t length, const oaph_matrix* V,
                              uint32_t column);
// region forms: op 0 add, 1 substract, 4 addSubstract (out += a - b) over
// the W x H region at the top left
oaph_status launch_binary_region(hipStream_t s, int op, oaph_matrix* out, const oaph_matrix* a,
                                 const oaph_matrix* b, uint32_t W, uint32_t H);
// out(x, y) = a(y, x) for x < W, y < H
oaph_status launch_transpose_region(hipStream_t s, oaph_matrix* out, const oaph_matrix* a,
                                    uint32_t W, uint32_t H);
// Row-mapped product: out(x, y) = sum_{k<K} a(k, (y-y0)%P + ay0) b(x, ((y-y0)/P) Q + by0 + k)
// for x0 <= x < x0 + W, y0 <= y < y0 + H with (y-y0)%P < S; complex when out
// has an im plane, else re only.
struct DotMap {
  uint32_t x0 = 0, y0 = 0, W = 0, H = 0, K = 0, P = 1, Q = 0, ay0 = 0, by0 = 0, S = 1;
};
oaph_status launch_dot_mapped(hipStream_t s, oaph_matrix* out, const oaph_matrix* a,
                              const oaph_matrix* b, const DotMap& mp);
// reduce.hip
oaph_status launch_sumsq(hipStream_t s, const double* re, const double* im, size_t n,
                         double* partials, double* result);
oaph_status launch_sum2(hipStream_t s, const double* re, const double* im, size_t n,
                        double* partials, double* result);
int reduce_blocks_for(size_t n);
// magnitude (sumsq) or sum of m into api->scalars, enqueued on api->stream
oaph_status enqueue_reduction(oaph_api* api, const oaph_matrix* m, bool sumsq);
// gemv.hip
oaph_status launch_gemm(hipStream_t s, oaph_matrix* out, const oaph_matrix* a,
                        const oaph_matrix* b, uint32_t W, uint32_t H, uint32_t K, bool dims_mode);
// Optional epilogue of a matvec y = A x with complex x, y: each workgroup
// also writes part[block * ldp + c] = sum over its rows r of conj(V[c][r]) y[r]
// for c < ncols (V column c at Vre + c * ldv), the per-block partial sums of
// V^H y that the first Gram-Schmidt pass of an Arnoldi step needs.  The
// launcher fuses it when ncols <= 64 and the grid has at most maxparts
// workgroups, and sets blocks to the grid size; blocks = 0: not fused, y only.
struct MvProj {
  const double* Vre = nullptr;
  const double* Vim = nullptr;
  uint32_t ldv = 0, ncols = 0, ldp = 0, maxparts = 0;
  double2* part = nullptr;
  uint32_t blocks = 0;
};
oaph_status launch_zgemv(hipStream_t s, uint32_t rows, uint32_t cols, uint32_t lda,
                         const double* are, const double* aim, const double* xre,
                         const double* xim, double* yre, double* yim, MvProj* pj = nullptr);
// Y[:, t] = A X[:, t], t < nvec <= kMultiMaxVectors, in one pass over A (FP64;
// A complex or real, x and y complex; vector t at x + t ldx, y + t ldy)
constexpr uint32_t kMultiMaxVectors = 8;
oaph_status launch_zgemv_multi(hipStream_t s, uint32_t rows, uint32_t cols, uint32_t lda,
                               const double* are, const double* aim, const double* xre,
                               const double* xim, uint32_t ldx, uint32_t nvec, double* yre,
                               double* yim, uint32_t ldy);
// out(x, y) = sum_k A[y][k] B[x][k] (A * B^T, B read in place), W x H, K terms
// the matrix-core complex product (gemm.hip): 128 x 128 tiles, for W, H >= 256
bool gemm_mfma_eligible(uint32_t W, uint32_t H, uint32_t K);
oaph_status launch_gemm_mfma(hipStream_t s, int mode, uint32_t W, uint32_t H, uint32_t K,
                             const double* are, const double* aim, uint32_t lda,
                             const double* bre, const double* bim, uint32_t ldb, double* ore,
                             double* oim, uint32_t ldo);
oaph_status launch_gemm_nt(hipStream_t s, uint32_t W, uint32_t H, uint32_t K, const double* are,
                           const double* aim, uint32_t lda, const double* bre, const double* bim,
                           uint32_t ldb, double* ore, double* oim, uint32_t ldo);
// FP32 operand (aim NULL: real operator), complex FP64 x and y
oaph_status launch_cgemv32(hipStream_t s, uint32_t rows, uint32_t cols, uint32_t lda,
                           const float* are, const float* aim, const double* xre,
                           const double* xim, double* yre, double* yim, MvProj* pj = nullptr);
// FP16 operand holding 2^s A (aim NULL: real operator), complex FP64 x and
// y: products and per-lane sums in FP32, y = unscale (= 2^-s) times the sums
oaph_status launch_hgemv16(hipStream_t s, uint32_t rows, uint32_t cols, uint32_t lda,
                           const _Float16* are, const _Float16* aim, const double* xre,
                           const double* xim, double* yre, double* yim, double unscale,
                           MvProj* pj = nullptr);
// small.hip
oaph_status launch_qrgr(hipStream_t s, oaph_matrix* Q, oaph_matrix* R, const oaph_matrix* A);
oaph_status launch_is_upper_triangular(hipStream_t s, const oaph_matrix* m, int* dflag);
oaph_status launch_calc_triangular_h(hipStream_t s, oaph_matrix* H, oaph_matrix* Q,
                                     double* work, uint32_t max_steps);

}  // namespace oaph
