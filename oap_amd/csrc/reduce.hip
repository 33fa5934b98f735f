// reduce.hip -- one-pass grid reductions for magnitude and sum.
//
// The reference reduces per block with a shared-memory halving tree, copies
// the block partials to the host and finishes there
// (oapMatrixCuda/CuProceduresApi.cpp:765-803, GenericProceduresApi.h:207-286).
// Here each wave reduces in registers with cross-lane shuffles, the waves of
// a block combine through LDS, and the last block to finish (a device
// counter) sums the block partials in block order and writes the result to
// device memory -- one launch, no host round trip, deterministic for a given
// grid.  The counter is returned to zero by that last block, so no memset is
// needed between launches (safe under graph capture).
#include "internal.h"

namespace oaph {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 2048;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Reduce NV values per thread across the block; results valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*lds)[NV]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[wave][q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0;
      for (int w = 0; w < kThreads / 64; ++w) s += lds[w][q];
      v[q] = s;
    }
  }
}

// Last-block finish: block partials [gridDim.x][NV] summed in block order.
template <int NV>
__device__ __forceinline__ void grid_finish(double (&v)[NV], double* partials, unsigned* counter,
                                            double* result) {
  __shared__ bool is_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) partials[(size_t)blockIdx.x * NV + q] = v[q];
    __threadfence();
    unsigned done = atomicAdd(counter, 1u);
    is_last = done == gridDim.x - 1;
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  // 64 lanes of the first wave sum strided partials, then a fixed-order tree
  if (threadIdx.x < 64) {
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += 64) {
#pragma unroll
      for (int q = 0; q < NV; ++q)
        acc[q] += __builtin_nontemporal_load(&partials[(size_t)b * NV + q]);
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = wave_sum(acc[q]);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < NV; ++q) result[q] = acc[q];
      *counter = 0u;
    }
  }
}

typedef double v2d __attribute__((ext_vector_type(2)));

// sum of re^2 + im^2 (either plane may be absent)
__global__ __launch_bounds__(256) void k_sumsq(const double* __restrict__ re,
                                               const double* __restrict__ im, size_t n,
                                               double* partials, unsigned* counter,
                                               double* result) {
  __shared__ double lds[kThreads / 64][1];
  double v[1] = {0.0};
  const size_t tid = blockIdx.x * (size_t)kThreads + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * kThreads;
  const bool aligned = ((((uintptr_t)re) | ((uintptr_t)im)) & 15) == 0;
  size_t npair = aligned ? n / 2 : 0;
  double a0 = 0, a1 = 0;
  for (size_t p = tid; p < npair; p += stride) {
    if (re) {
      v2d x = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(re) + p);
      a0 += x.x * x.x;
      a1 += x.y * x.y;
    }
    if (im) {
      v2d y = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(im) + p);
      a0 += y.x * y.x;
      a1 += y.y * y.y;
    }
  }
  for (size_t i = 2 * npair + tid; i < n; i += stride) {
    if (re) a0 += re[i] * re[i];
    if (im) a0 += im[i] * im[i];
  }
  v[0] = a0 + a1;
  block_sum<1>(v, lds);
  grid_finish<1>(v, partials, counter, result);
}

// (sum re, sum im)
__global__ __launch_bounds__(256) void k_sum2(const double* __restrict__ re,
                                              const double* __restrict__ im, size_t n,
                                              double* partials, unsigned* counter,
                                              double* result) {
  __shared__ double lds[kThreads / 64][2];
  double v[2] = {0.0, 0.0};
  const size_t tid = blockIdx.x * (size_t)kThreads + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * kThreads;
  const bool aligned = ((((uintptr_t)re) | ((uintptr_t)im)) & 15) == 0;
  size_t npair = aligned ? n / 2 : 0;
  for (size_t p = tid; p < npair; p += stride) {
    if (re) {
      v2d x = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(re) + p);
      v[0] += x.x + x.y;
    }
    if (im) {
      v2d y = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(im) + p);
      v[1] += y.x + y.y;
    }
  }
  for (size_t i = 2 * npair + tid; i < n; i += stride) {
    if (re) v[0] += re[i];
    if (im) v[1] += im[i];
  }
  block_sum<2>(v, lds);
  grid_finish<2>(v, partials, counter, result);
}

}  // namespace

int reduce_blocks_for(size_t n) {
  // ~8 elements per thread minimum; at most 2048 blocks (8 per CU)
  size_t b = (n + (size_t)kThreads * 16 - 1) / ((size_t)kThreads * 16);
  if (b < 1) b = 1;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return (int)b;
}

oaph_status launch_sumsq(hipStream_t s, const double* re, const double* im, size_t n,
                         double* partials, unsigned* counter, double* result) {
  int blocks = reduce_blocks_for(n);
  hipLaunchKernelGGL(k_sumsq, dim3(blocks), dim3(kThreads), 0, s, re, im, n, partials, counter,
                     result);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status launch_sum2(hipStream_t s, const double* re, const double* im, size_t n,
                        double* partials, unsigned* counter, double* result) {
  int blocks = reduce_blocks_for(n);
  hipLaunchKernelGGL(k_sum2, dim3(blocks), dim3(kThreads), 0, s, re, im, n, partials, counter,
                     result);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

}  // namespace oaph
