/*
 * oracle.h -- CPU restatement of oap's complex procedures and of the
 * implicitly restarted Arnoldi method they drive.
 *
 * TEST INFRASTRUCTURE ONLY.  This file and oracle.c are the checker for the
 * HIP library: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load liboracle.so.  Nothing in the product (oap_amd/,
 * include/) links or calls it.
 *
 * Data model (reference oapMatrix/Matrix.h:31-44, oapMatrix/MatrixAPI.h:130-140):
 * element (c, r) of a matrix with `columns` columns sits at c + r*columns in
 * each of two independent planes (re, im); either plane may be NULL.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - set_threads_blocks ........ golden table of oapThreadsMapperTest.cpp and
 *                                 bit-exact vs the reference's own ThreadsMapper.cpp
 *                                 compiled into oracle/_ref.
 *   - transpose / conj-transpose / identity / diagonal / set/get vector /
 *     hadamard / tensor / add / substract / multiply-constant ... bit-exact vs
 *     oracle/_ref (HostProcedures + the reference's device functions on its
 *     CPU thread emulator) and closed forms.
 *   - dot_product / dot_product_dims .... golden arrays in oapTestsData, closed
 *     forms (constant matrices) and bit-exact vs oracle/_ref (same k order).
 *   - magnitude / sum ........... closed forms (sqrt(385), integer sums) and
 *     vs oracle/_ref within gamma_N (block-tree order differs, ulp level).
 *   - qrgr ...................... host_qrtest1-4 and qrtest1-9 fixtures through
 *                                 A = QR, Q^T Q = I, R upper; vs oracle/_ref
 *                                 (reference CUDA_QRGR on its emulator) on
 *                                 inputs with an odd rotation count.
 *   - calc_triangular_h ......... schur1-3 diagonal fixtures.
 *   - hess_eig (shifted complex Hessenberg QR, not in the reference) ... numpy
 *     LAPACK eigvals and A x = lambda x residuals.
 *   - iram ...................... dense LAPACK eigenvalues of the same matrix
 *     (committed fixtures) and explicit FP64 residuals.
 */
#ifndef OAP_ORACLE_H
#define OAP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_matrix {
  uint32_t columns;
  uint32_t rows;
  double* re; /* may be NULL */
  double* im; /* may be NULL */
} orc_matrix;

/* ---- launch geometry (oapUtils/ThreadsMapper.cpp:29-74) ---- */
void orc_set_threads_blocks(uint32_t blocks[2], uint32_t threads[2], uint32_t w, uint32_t h,
                            uint32_t limit);

/* ---- procedures; return 0 on success, nonzero on a shape/plane error ---- */
int orc_dot_product(orc_matrix* out, const orc_matrix* a, const orc_matrix* b);
int orc_dot_product_dims(orc_matrix* out, const orc_matrix* a, const orc_matrix* b,
                         const uint32_t dims[3][2]);
int orc_transpose(orc_matrix* out, const orc_matrix* a);
int orc_conjugate_transpose(orc_matrix* out, const orc_matrix* a);
int orc_tensor_product(orc_matrix* out, const orc_matrix* a, const orc_matrix* b,
                       const uint32_t dims[3][2]);
int orc_hadamard_product(orc_matrix* out, const orc_matrix* a, const orc_matrix* b);
int orc_hadamard_product_vec(orc_matrix* out, const orc_matrix* a, const orc_matrix* b);
int orc_add(orc_matrix* out, const orc_matrix* a, const orc_matrix* b);
int orc_substract(orc_matrix* out, const orc_matrix* a, const orc_matrix* b);
int orc_multiply_re_constant(orc_matrix* out, const orc_matrix* a, double re);
int orc_multiply_constant(orc_matrix* out, const orc_matrix* a, double re, double im);
int orc_set_diagonal(orc_matrix* m, double re, double im);
int orc_set_identity(orc_matrix* m);
int orc_set_zero(orc_matrix* m);
int orc_set_vector(orc_matrix* V, uint32_t column, const orc_matrix* v, uint32_t length);
int orc_get_vector(orc_matrix* v, uint32_t length, const orc_matrix* V, uint32_t column);
double orc_magnitude(const orc_matrix* m);
void orc_sum(const orc_matrix* m, double* re, double* im);

/* Givens QR of a square matrix, reference rotation order and skip rule. */
int orc_qrgr(orc_matrix* Q, orc_matrix* R, const orc_matrix* A);
int orc_is_upper_triangular(const orc_matrix* m);
/* Unshifted QR iteration H <- R Q until is_upper_triangular or max_steps.
 * Q receives the accumulated orthogonal factor (H_in = Q H_out Q^H). */
int orc_calc_triangular_h(orc_matrix* H, orc_matrix* Q, uint32_t max_steps);

/* ---- complex shifted Hessenberg QR (the fixed Ritz extraction) ----
 * H is m x m complex upper Hessenberg, row-major planes (ld = m).
 * On return wr/wi hold the eigenvalues, Z (if not NULL) the Schur vectors
 * (H = Z T Z^H), T overwrites H.  Returns 0 on convergence, -1 otherwise. */
int orc_hess_schur(int m, double* hre, double* him, double* zre, double* zim, double* wr,
                   double* wi);
/* Eigenvectors of the upper triangular T (m x m, row-major) by back
 * substitution; column i of X (row-major, ld = m) belongs to T[i][i]. */
void orc_trevc(int m, const double* tre, const double* tim, double* xre, double* xim);

/* ---- implicitly restarted Arnoldi (restated sequence with fixes) ---- */
typedef void (*orc_matvec_fn)(void* user, const double* xre, const double* xim, double* yre,
                              double* yim);

typedef struct orc_iram_opts {
  int n, m, k;         /* dimension, Krylov size, wanted count */
  int which;           /* ArnUtils sort: 0 LargestValues 1 LargestRe 2 LargestIm 3 SmallestValues 4 SmallestRe 5 SmallestIm */
  double tol;          /* residual tolerance relative to anorm */
  double anorm;        /* norm the tolerance is relative to (Frobenius) */
  int max_restarts;
  uint64_t seed;       /* start vector seed (SplitMix64 stream) */
  int real_start;      /* nonzero: start vector has zero imaginary part */
  const double* v0re;  /* optional start vector (NULL: seeded random) */
  const double* v0im;
} orc_iram_opts;

typedef struct orc_iram_result {
  int nconv;
  int restarts;
  int matvecs;
} orc_iram_result;

/* Computes k eigenpairs; evals (k complex), evecs k columns of n (row-major
 * planes: evec j element i at j*n + i); resid[j] explicit ||A x - l x||_2. */
int orc_iram(const orc_iram_opts* o, orc_matvec_fn mv, void* user, double* ev_re,
             double* ev_im, double* x_re, double* x_im, double* resid, orc_iram_result* res);

/* dense complex matvec y = A x on row-major planes (A n x n); threads > 1
 * splits rows over POSIX threads. */
void orc_dense_matvec(int n, const double* are, const double* aim, const double* xre,
                      const double* xim, double* yre, double* yim, int threads);

/* ---- seeded generators (SplitMix64, counter based) ---- */
uint64_t orc_splitmix64(uint64_t x);
double orc_uniform(uint64_t seed, uint64_t idx);         /* [0,1), 53-bit */
double orc_normal(uint64_t seed, uint64_t idx);          /* N(0,1), Box-Muller */
/* C0: A = diag(1/(1+i)) + 0.05*G/sqrt(n), real. */
void orc_gen_c0(int n, uint64_t seed, double* are);
/* C1: A = diag(d) + eps*(G_re + i G_im)/sqrt(n), d_k = exp(-k/decay) e^{2 pi i u_k}. */
void orc_gen_c1(int n, uint64_t seed, double eps, double decay, double* are, double* aim);

/* C3 (batched covariance problems): problem i of a batch with base seed
 * `seed`: n_i = 64 + floor(961 u) in [64, 1024]; X_i (s x n_i, row-major,
 * FP32) = round(65535 u)/65535 + 5 sum_{c<3} a_c b_c^T with unit-norm
 * Gaussian a_c (length s) and b_c (length n_i); the operator is
 * C_i = X_i^T X_i / s.  Returns n_i; writes X when X != NULL. */
int orc_gen_c3_problem(int i, uint64_t seed, int s, float* X);
uint64_t orc_c3_stream(uint64_t seed, int i);

/* Feed-forward network, oap training semantics in FP64 (oracle.c) */
typedef struct orc_net orc_net;
orc_net* orc_net_create(int nlayers, const int* neurons, const int* bias, const int* act);
void orc_net_free(orc_net* N);
void orc_net_set_weights(orc_net* N, int layer, const double* w);
void orc_net_get_weights(const orc_net* N, int layer, double* w);
void orc_net_forward(orc_net* N, const double* x, double* y);
void orc_net_accumulate(orc_net* N, const double* t);
void orc_net_backward(orc_net* N);
void orc_net_update(orc_net* N, double lr);
void orc_net_post_step(orc_net* N);
double orc_net_error(const orc_net* N);
void orc_net_gradient(const orc_net* N, int layer, double* g);

/* C4 convection-diffusion-reaction operator, dense row-major n x n (n = N^2) */
void orc_gen_cdr(int N, double pe, double c, double alpha, double* A);

/* Learned-start MLP: batch loss and flat gradients (oracle.c); round_bf16
 * emulates the device network's bfloat16 operand rounding */
double orc_bf16(double v);
double orc_mlp_grad(const uint32_t* dims, int ndims, const float* params, const float* X,
                    const float* T, const float* w, int B, double* grads, int round_bf16);

#ifdef __cplusplus
}
#endif

#endif
