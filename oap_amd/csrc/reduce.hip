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

// Last-block finish: block partials [gridDim.x][NV] summed in block order.
template <int NV>
__device__ __forceinline__ void grid_finish(double (&v)[NV], double* partials, unsigned* counter,
                                            double* result, bool root = false) {
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
      for (int q = 0; q < NV; ++q) result[q] = (root && NV == 1) ? sqrt(acc[q]) : acc[q];
      *counter = 0u;
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
        ar[u] = fma(x[u].x, x[u].x, ar[u]);
        ar[u] = fma(x[u].y, x[u].y, ar[u]);
        ai[u] = fma(y[u].x, y[u].x, ai[u]);
        ai[u] = fma(y[u].y, y[u].y, ai[u]);
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
      ar[0] = fma(x.x, x.x, fma(x.y, x.y, ar[0]));
      ai[0] = fma(y.x, y.x, fma(y.y, y.y, ai[0]));
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
      ar[0] = fma(x, x, ar[0]);
      ai[0] = fma(y, y, ai[0]);
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
  if (KIND == 0) {
    v[0] = sr + si;
  } else {
    v[0] = sr;
    v[KIND == 0 ? 0 : 1] = si;
  }
}

// root: the result is the square root of the sum (the magnitude)
__global__ __launch_bounds__(256) void k_sumsq(const double* __restrict__ re,
                                               const double* __restrict__ im, size_t n,
                                               double* partials, unsigned* counter,
                                               double* result, int root) {
  __shared__ double lds[kThreads / 64][1];
  double v[1];
  stream_planes<0>(re, im, n, v);
  block_sum<1>(v, lds);
  grid_finish<1>(v, partials, counter, result, root != 0);
}

__global__ void k_clear_counter(unsigned* counter) { *counter = 0u; }

__global__ __launch_bounds__(256) void k_sum2(const double* __restrict__ re,
                                              const double* __restrict__ im, size_t n,
                                              double* partials, unsigned* counter,
                                              double* result) {
  __shared__ double lds[kThreads / 64][2];
  double v[2];
  stream_planes<1>(re, im, n, v);
  block_sum<2>(v, lds);
  grid_finish<2>(v, partials, counter, result);
}

}  // namespace

// Blocks per grid: at least one 16-byte load group per lane, at most
// g_reduce_blocks (default 256: one wave per SIMD of the 256 CUs, each
// streaming a contiguous chunk with kU loads in flight per plane).
int g_reduce_blocks = 256;
int reduce_blocks_for(size_t n) {
  size_t b = (n / 2 + (size_t)kThreads * kU - 1) / ((size_t)kThreads * kU);
  if (b < 1) b = 1;
  if (b > (size_t)g_reduce_blocks) b = (size_t)g_reduce_blocks;
  return (int)b;
}

oaph_status launch_sumsq(hipStream_t s, const double* re, const double* im, size_t n,
                         double* partials, unsigned* counter, double* result) {
  int blocks = reduce_blocks_for(n);
  hipLaunchKernelGGL(k_sumsq, dim3(blocks), dim3(kThreads), 0, s, re, im, n, partials, counter,
                     result, 0);
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

extern "C" oaph_status oaph_set_reduce_blocks(int blocks) {
  OAPH_CHECK(blocks >= 1 && blocks <= oaph::kMaxBlocks, OAPH_ERR_ARG);
  oaph::g_reduce_blocks = blocks;
  return OAPH_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free magnitude.  The block partials and the last-block counter
// live in `scratch`, a file-scope device array (every GPU the library runs on
// holds its own copy until the process ends): kMaxBlocks partial sums, then
// the counter.  Host code never passes `scratch` itself (its host-side name is
// not a device address); the device address comes from hipGetSymbolAddress,
// once per device with that device current, and is kept per device.
namespace oaph {
namespace {
constexpr int kScratchDoubles = kMaxBlocks + 32;  // partials, then the counter word
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
  const int blocks = reduce_blocks_for(n);
  if (blocks > kMaxBlocks) {
    set_last_message("oaph_magnitude_noalloc: more block partials than the static scratch holds");
    return OAPH_ERR_UNSUPPORTED;
  }
  double* base = nullptr;
  OAPH_TRY(scratch_for_current_device(&base));
  double* partials = base;
  unsigned* counter = reinterpret_cast<unsigned*>(base + kMaxBlocks);
  // the counter must start at zero (a fill kernel, not a memset: the call may
  // be recorded into a graph)
  hipLaunchKernelGGL(k_clear_counter, dim3(1), dim3(1), 0, api->stream, counter);
  hipLaunchKernelGGL(k_sumsq, dim3(blocks), dim3(kThreads), 0, api->stream, m->reValues,
                     m->imValues, n, partials, counter, out, 1);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}
