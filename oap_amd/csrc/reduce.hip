// reduce.hip -- grid reductions for magnitude and sum.
//
// The reference reduces per block with a shared-memory halving tree, copies
// the block partials to the host and finishes there
// (oapMatrixCuda/CuProceduresApi.cpp:765-803, GenericProceduresApi.h:207-286).
// Here one wave per SIMD streams a contiguous chunk of the planes with eight
// 16-byte loads in flight per plane, reduces in registers and across lanes,
// the waves of a block combine through LDS and each block writes its partial
// sum; a one-wave finish kernel adds the partials in block order and writes
// the result to device memory.  No host round trip, deterministic for a given
// grid, and no grid-wide fence: a last-block finish inside the streaming
// kernel needs a device-scope release per block (the L2 of each XCD is not
// coherent with the others), which measured 171 us against 162 us for the two
// launches over 2^26 complex elements on MI355X.
#include <mutex>

#include "internal.h"

namespace oaph {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 2048;  // partial-sum capacity the callers reserve

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

typedef double v2d __attribute__((ext_vector_type(2)));

// Loads in flight per lane and plane: each wave streams a contiguous chunk
// of 16-byte pairs, issuing kU loads of each plane before it consumes them,
// with one accumulator per load slot (no dependent FMA chain across slots).
constexpr int kU = 8;

// KIND 0: sum of re^2 + im^2 (v[0]); KIND 1: (sum re, sum im).  Either plane
// may be absent.  Wave w of the grid takes pairs [w * per, (w + 1) * per).
template <int KIND>
__device__ __forceinline__ void stream_planes(const double* __restrict__ re,
                                              const double* __restrict__ im, size_t n,
                                              double (&v)[KIND == 0 ? 1 : 2]) {
  const int lane = threadIdx.x & 63;
  const size_t nw = (size_t)gridDim.x * (kThreads / 64);
  const size_t w = (size_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const bool aligned = ((((uintptr_t)re) | ((uintptr_t)im)) & 15) == 0;
  const size_t npair = aligned ? n / 2 : 0;
  const size_t group = 64 * kU;
  const size_t per = ((npair + nw - 1) / nw + group - 1) / group * group;
  const size_t beg = w * per;
  const size_t end = beg + per < npair ? beg + per : npair;
  const v2d* r2 = reinterpret_cast<const v2d*>(re);
  const v2d* i2 = reinterpret_cast<const v2d*>(im);
  double ar[kU], ai[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) ar[u] = ai[u] = 0.0;
  size_t p = beg + lane;
  for (; p + (kU - 1) * 64 < end; p += group) {
    v2d x[kU], y[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      x[u] = re ? __builtin_nontemporal_load(r2 + p + u * 64) : v2d{0.0, 0.0};
      y[u] = im ? __builtin_nontemporal_load(i2 + p + u * 64) : v2d{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (KIND == 0) {
        ar[u] = fma(x[u].x, x[u].x, fma(x[u].y, x[u].y, fma(y[u].x, y[u].x,
                    fma(y[u].y, y[u].y, ar[u]))));
      } else {
        ar[u] += x[u].x + x[u].y;
        ai[u] += y[u].x + y[u].y;
      }
    }
  }
  for (; p < end; p += 64) {
    const v2d x = re ? r2[p] : v2d{0.0, 0.0};
    const v2d y = im ? i2[p] : v2d{0.0, 0.0};
    if (KIND == 0) {
      ar[0] = fma(x.x, x.x, fma(x.y, x.y, fma(y.x, y.x, fma(y.y, y.y, ar[0]))));
    } else {
      ar[0] += x.x + x.y;
      ai[0] += y.x + y.y;
    }
  }
  // the elements the pairs leave (odd n, or unaligned planes)
  const size_t tid = (size_t)blockIdx.x * kThreads + threadIdx.x;
  for (size_t i = 2 * npair + tid; i < n; i += (size_t)gridDim.x * kThreads) {
    const double x = re ? re[i] : 0.0, y = im ? im[i] : 0.0;
    if (KIND == 0) {
      ar[0] = fma(x, x, fma(y, y, ar[0]));
    } else {
      ar[0] += x;
      ai[0] += y;
    }
  }
  double sr = 0.0, si = 0.0;
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    sr += ar[u];
    si += ai[u];
  }
  v[0] = sr;
  if (KIND == 1) v[KIND] = si;
}

// block partial sums: partials[block * NV + q]
__global__ __launch_bounds__(256) void k_sumsq(const double* __restrict__ re,
                                               const double* __restrict__ im, size_t n,
                                               double* partials) {
  __shared__ double lds[kThreads / 64][1];
  double v[1];
  stream_planes<0>(re, im, n, v);
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(256) void k_sum2(const double* __restrict__ re,
                                              const double* __restrict__ im, size_t n,
                                              double* partials) {
  __shared__ double lds[kThreads / 64][2];
  double v[2];
  stream_planes<1>(re, im, n, v);
  block_sum<2>(v, lds);
  if (threadIdx.x == 0) {
    partials[2 * (size_t)blockIdx.x] = v[0];
    partials[2 * (size_t)blockIdx.x + 1] = v[1];
  }
}

// One wave: result[q] = sum over blocks of partials[b * NV + q], lanes over
// strided blocks then a fixed-order cross-lane sum (the same bits for the
// same grid); root: the square root of the sum (NV = 1, the magnitude).
template <int NV>
__global__ __launch_bounds__(64) void k_finish(const double* __restrict__ partials, int nparts,
                                               double* result, int root) {
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 64) {
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += partials[(size_t)b * NV + q];
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = wave_sum(acc[q]);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) result[q] = (root && NV == 1) ? sqrt(acc[q]) : acc[q];
  }
}

oaph_status launch_sumsq_root(hipStream_t s, const double* re, const double* im, size_t n,
                              double* partials, double* result, int root) {
  const int blocks = reduce_blocks_for(n);
  hipLaunchKernelGGL(k_sumsq, dim3(blocks), dim3(kThreads), 0, s, re, im, n, partials);
  hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(64), 0, s, partials, blocks, result, root);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

}  // namespace

// Blocks per grid: at least one load group per lane, at most g_reduce_blocks
// (default 256: one wave per SIMD of the 256 CUs, each streaming a contiguous
// chunk with kU loads in flight per plane; 128, 384, 512 and 1024 measured
// slower over 2^26 complex elements).
int g_reduce_blocks = 256;
int reduce_blocks_for(size_t n) {
  size_t b = (n / 2 + (size_t)kThreads * kU - 1) / ((size_t)kThreads * kU);
  if (b < 1) b = 1;
  if (b > (size_t)g_reduce_blocks) b = (size_t)g_reduce_blocks;
  return (int)b;
}

oaph_status launch_sumsq(hipStream_t s, const double* re, const double* im, size_t n,
                         double* partials, double* result) {
  return launch_sumsq_root(s, re, im, n, partials, result, 0);
}

oaph_status launch_sum2(hipStream_t s, const double* re, const double* im, size_t n,
                        double* partials, double* result) {
  const int blocks = reduce_blocks_for(n);
  hipLaunchKernelGGL(k_sum2, dim3(blocks), dim3(kThreads), 0, s, re, im, n, partials);
  hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(64), 0, s, partials, blocks, result, 0);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

}  // namespace oaph

extern "C" oaph_status oaph_set_reduce_blocks(int blocks) {
  OAPH_CHECK(blocks >= 1 && blocks <= oaph::kMaxBlocks, OAPH_ERR_ARG);
  oaph::g_reduce_blocks = blocks;
  return OAPH_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free magnitude.  The block partial sums live in `scratch`, a
// file-scope device array (every GPU the library runs on holds its own copy
// until the process ends).  Host code never passes `scratch` itself (its
// host-side name is not a device address); the device address comes from
// hipGetSymbolAddress, once per device with that device current, and is kept
// per device.
namespace oaph {
namespace {
constexpr int kScratchDoubles = kMaxBlocks;  // one partial sum per block
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

oaph_status scratch_for_current_device(double** out) {
  int dev = 0;
  OAPH_HIP(hipGetDevice(&dev));
  OAPH_CHECK(dev >= 0 && dev < kMaxDevices, OAPH_ERR_UNSUPPORTED);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    OAPH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return OAPH_OK;
}
}  // namespace
}  // namespace oaph

extern "C" oaph_status oaph_magnitude_noalloc(oaph_api* api, double* out, const oaph_matrix* m) {
  using namespace oaph;
  OAPH_CHECK(api && out && m, OAPH_ERR_ARG);
  OAPH_TRY(ensure_init());
  const size_t n = numel(m);
  if (reduce_blocks_for(n) > kScratchDoubles) {
    set_last_message("oaph_magnitude_noalloc: more block partials than the static scratch holds");
    return OAPH_ERR_UNSUPPORTED;
  }
  double* partials = nullptr;
  OAPH_TRY(scratch_for_current_device(&partials));
  return launch_sumsq_root(api->stream, m->reValues, m->imValues, n, partials, out, 1);
}
