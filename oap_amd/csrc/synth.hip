// synth.hip -- seeded synthetic operators generated directly in HBM, and
// HIP-event timing of the hot kernels (measurement support).
//
// Generators are counter based (SplitMix64 output function applied to
// seed + (idx+1) * 0x9E3779B97F4A7C15, 53-bit uniforms, Box-Muller cosine
// branch), so element idx is computed independently by any thread and the
// values equal those of a sequential SplitMix64 stream.
#include "iram.h"

namespace oaph {
namespace {

__device__ __forceinline__ double u01(uint64_t seed, uint64_t idx) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (idx + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double gauss(uint64_t seed, uint64_t idx) {
  double u1 = u01(seed, 2 * idx), u2 = u01(seed, 2 * idx + 1);
  return sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
}

// kind 0 (C0): A = diag(1/(1+i)) + p1 * G / sqrt(n), real
// kind 1 (C1): A = diag(exp(-k/p2) e^{2 pi i u_k}) + p1 (G_re + i G_im) / sqrt(n)
__global__ void k_synth(int kind, uint32_t n, uint64_t seed, double p1, double p2, double* re,
                        double* im) {
  const size_t total = (size_t)n * n;
  const double sc = p1 / sqrt((double)n);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    uint32_t r = (uint32_t)(i / n), c = (uint32_t)(i - (size_t)r * n);
    if (kind == 0) {
      double v = sc * gauss(seed, i);
      if (r == c) v += 1.0 / (1.0 + r);
      re[i] = v;
      if (im) im[i] = 0.0;
    } else {
      double vr = sc * gauss(seed, 2 * i), vi = sc * gauss(seed, 2 * i + 1);
      if (r == c) {
        double u = u01(seed ^ 0xD1A60000ULL, r);
        double mag = exp(-(double)r / p2);
        vr += mag * cos(6.283185307179586 * u);
        vi += mag * sin(6.283185307179586 * u);
      }
      re[i] = vr;
      im[i] = vi;
    }
  }
}

// C4 operator: -(alpha d_xx + d_yy) u + pe (d_x + d_y) u + c u on the N x N
// interior grid of the unit square (h = 1/(N+1), Dirichlet), second-order
// central differences, unknown (ix, iy) at ix + N iy, assembled dense with
// every coefficient rounded to FP32 (the operator is the FP32 matrix).
__global__ void k_cdr(uint32_t N, double pe, double c, double alpha, double* re, double* im) {
#pragma clang fp contract(off)
  const uint32_t n = N * N;
  const double h = 1.0 / (double)(N + 1), ih2 = 1.0 / (h * h), adv = pe / (2.0 * h);
  const double diag = (double)(float)(2.0 * alpha * ih2 + 2.0 * ih2 + c);
  const double xm = (double)(float)(-alpha * ih2 - adv), xp = (double)(float)(-alpha * ih2 + adv);
  const double ym = (double)(float)(-ih2 - adv), yp = (double)(float)(-ih2 + adv);
  const size_t total = (size_t)n * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t r = (uint32_t)(i / n), col = (uint32_t)(i - (size_t)r * n);
    const uint32_t rx = r % N;
    double v = 0.0;
    if (col == r) v = diag;
    else if (col + 1 == r && rx > 0) v = xm;
    else if (col == r + 1 && rx + 1 < N) v = xp;
    else if (col + N == r) v = ym;
    else if (col == r + N) v = yp;
    re[i] = v;
    if (im) im[i] = 0.0;
  }
}

}  // namespace
}  // namespace oaph

using namespace oaph;

extern "C" {

oaph_status oaph_synth_matrix(oaph_matrix* A, int kind, uint64_t seed, double p1, double p2) {
  OAPH_TRY(ensure_init());
  OAPH_CHECK(A && A->columns == A->rows && A->reValues, OAPH_ERR_ARG);
  OAPH_CHECK(kind == 0 || kind == 1, OAPH_ERR_ARG);
  if (kind == 1) OAPH_CHECK(A->imValues != nullptr, OAPH_ERR_PLANE);
  size_t total = (size_t)A->columns * A->rows;
  size_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_synth, dim3((unsigned)blocks), dim3(256), 0, 0, kind, A->columns, seed, p1,
                     p2, A->reValues, A->imValues);
  OAPH_HIP(hipGetLastError());
  OAPH_HIP(hipDeviceSynchronize());
  return OAPH_OK;
}

oaph_status oaph_synth_cdr(oaph_matrix* A, uint32_t grid, double pe, double c, double alpha) {
  OAPH_TRY(ensure_init());
  OAPH_CHECK(A && A->reValues && grid >= 2 && alpha > 0, OAPH_ERR_ARG);
  OAPH_CHECK(A->columns == grid * grid && A->rows == grid * grid, OAPH_ERR_SHAPE);
  const size_t total = (size_t)A->columns * A->rows;
  size_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_cdr, dim3((unsigned)blocks), dim3(256), 0, 0, grid, pe, c, alpha,
                     A->reValues, A->imValues);
  OAPH_HIP(hipGetLastError());
  OAPH_HIP(hipDeviceSynchronize());
  return OAPH_OK;
}

oaph_status oaph_time_dot_product(oaph_api* api, oaph_matrix* out, const oaph_matrix* a,
                                  const oaph_matrix* b, uint32_t iters, double* avg_ms) {
  OAPH_CHECK(api && avg_ms && iters > 0, OAPH_ERR_ARG);
  OAPH_TRY(oaph_dot_product(api, out, a, b));  // warm-up and validation
  hipEvent_t e0, e1;
  OAPH_HIP(hipEventCreate(&e0));
  OAPH_HIP(hipEventCreate(&e1));
  OAPH_HIP(hipEventRecord(e0, api->stream));
  for (uint32_t i = 0; i < iters; ++i) OAPH_TRY(oaph_dot_product(api, out, a, b));
  OAPH_HIP(hipEventRecord(e1, api->stream));
  OAPH_HIP(hipEventSynchronize(e1));
  float ms = 0;
  OAPH_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = ms / iters;
  return OAPH_OK;
}

oaph_status oaph_time_reduction(oaph_api* api, int sumsq, const oaph_matrix* m, uint32_t iters,
                                double* avg_ms) {
  OAPH_CHECK(api && m && avg_ms && iters > 0, OAPH_ERR_ARG);
  OAPH_TRY(enqueue_reduction(api, m, sumsq != 0));  // warm-up
  hipEvent_t e0, e1;
  OAPH_HIP(hipEventCreate(&e0));
  OAPH_HIP(hipEventCreate(&e1));
  OAPH_HIP(hipEventRecord(e0, api->stream));
  for (uint32_t i = 0; i < iters; ++i) OAPH_TRY(enqueue_reduction(api, m, sumsq != 0));
  OAPH_HIP(hipEventRecord(e1, api->stream));
  OAPH_HIP(hipEventSynchronize(e1));
  float ms = 0;
  OAPH_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = ms / iters;
  return OAPH_OK;
}

}  // extern "C"
