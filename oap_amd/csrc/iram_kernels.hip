// iram_kernels.hip -- device side of the implicitly restarted Arnoldi solver.
//
// Krylov basis layout (solver-owned, not a math::Matrix): column c of V is
// contiguous, V[c*n + i], in separate re/im planes, so that the projection
// V^H w and the update w - V h stream V row-blocks with coalesced loads.
// (The reference keeps V as an n-row x m-column math::Matrix and rebuilds a
// full transpose every Arnoldi step: ArnoldiProcedures.cpp:441-443.)
//
// Kernels:
//   k_orth<MODE>     classical Gram-Schmidt pass pieces with a one-launch grid
//                    reduction (last block finishes; counter self-resets)
//   k_normalize      v_{j+1} = f / beta
//   k_restart        one workgroup: complex Schur form of H by shifted QR, Ritz
//                    vectors of H, residual estimates, wanted ordering,
//                    exact-shift sweeps and the compressed H, Q
//   k_vq             V Q (tall-skinny x small), optionally f <- (V Q)_kk beta + f sigma
//   k_pairs<MODE>    Rayleigh quotients and residual norms of Ritz pairs
#include "iram.h"
#include "small_dense.h"

namespace oaph {
namespace iram {

namespace {

constexpr int kOrthThreads = 256;
constexpr int kRows = 64;  // rows per block in k_orth
constexpr int kCPW = kMaxM / 4 + 1;  // basis columns per wave (ncols <= m + 1)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Last-block-done protocol: returns true in every thread of the block that
// finishes last; partial results written before the call are visible to it.
__device__ __forceinline__ bool last_block(unsigned* counter) {
  __shared__ bool is_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned done = atomicAdd(counter, 1u);
    is_last = done == gridDim.x - 1;
  }
  __syncthreads();
  if (is_last) __threadfence();
  return is_last;
}

// MODE 0: s = V^H x.
// MODE 1: f = x - V sub, s = V^H f (f written to fout).
// MODE 2: f = x - V sub (written to fout), beta = ||f||; the last block writes
//         beta to the status and, when hcol >= 0, H[c][hcol] = add[c] + sub[c]
//         for c < ncols and beta to H[hrow][hcol] when hrow < m; breakdown is
//         flagged (hcol >= 0 only) when beta is below 1e-13 ||(h, beta)||.
template <int MODE>
__global__ __launch_bounds__(kOrthThreads) void k_orth(
    uint32_t n, uint32_t ncols, const double* __restrict__ Vre, const double* __restrict__ Vim,
    const double* xre, const double* xim, double* fre, double* fim, const double2* sub,
    const double2* add, double2* result, double* partials, unsigned* counter, double* Hre,
    double* Him, uint32_t m, int hcol, int hrow, Status* st) {
  __shared__ double2 lds_part[4][kRows];
  __shared__ double2 lds_fin[4][kMaxM + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * kRows + lane;
  const bool in = i < n;
  double xr = in ? xre[i] : 0.0, xi = in ? xim[i] : 0.0;
  // this wave's basis columns c = wave + 4q, all loads issued before use
  double vr[kCPW], vi[kCPW];
#pragma unroll
  for (int q = 0; q < kCPW; ++q) {
    const uint32_t c = wave + 4 * q;
    const bool ok = in && c < ncols;
    vr[q] = ok ? Vre[(size_t)c * n + i] : 0.0;
    vi[q] = ok ? Vim[(size_t)c * n + i] : 0.0;
  }
  if (MODE >= 1) {
    // this wave's share of (V sub)[i]
    double pr = 0, pi = 0;
#pragma unroll
    for (int q = 0; q < kCPW; ++q) {
      const uint32_t c = wave + 4 * q;
      if (c < ncols) {
        const double2 s = sub[c];
        pr += vr[q] * s.x - vi[q] * s.y;
        pi += vr[q] * s.y + vi[q] * s.x;
      }
    }
    lds_part[wave][lane] = make_double2(pr, pi);
    __syncthreads();
    double2 a = lds_part[0][lane], b = lds_part[1][lane], c2 = lds_part[2][lane],
            d = lds_part[3][lane];
    xr -= (a.x + b.x) + (c2.x + d.x);
    xi -= (a.y + b.y) + (c2.y + d.y);
    if (wave == 0 && in) {
      fre[i] = xr;
      fim[i] = xi;
    }
  }
  const size_t pstride = (size_t)(ncols > 0 ? ncols : 1);
  if (MODE <= 1) {
#pragma unroll
    for (int q = 0; q < kCPW; ++q) {
      const uint32_t c = wave + 4 * q;
      if (c < ncols) {
        // conj(v) * f
        double sr = wave_sum(vr[q] * xr + vi[q] * xi);
        double si = wave_sum(vr[q] * xi - vi[q] * xr);
        if (lane == 0) {
          partials[2 * ((size_t)blockIdx.x * pstride + c)] = sr;
          partials[2 * ((size_t)blockIdx.x * pstride + c) + 1] = si;
        }
      }
    }
  } else {
    if (wave == 0) {
      double s = wave_sum(xr * xr + xi * xi);
      if (lane == 0) partials[blockIdx.x] = s;
    }
  }
  if (!last_block(counter)) return;
  if (MODE <= 1) {
    // lane = column, wave w sums blocks w, w+4, ...; waves combined in order
    for (uint32_t c = lane; c < ncols; c += 64) {
      double sr = 0, si = 0;
#pragma unroll 8
      for (uint32_t b = wave; b < gridDim.x; b += 4) {
        sr += __builtin_nontemporal_load(&partials[2 * ((size_t)b * pstride + c)]);
        si += __builtin_nontemporal_load(&partials[2 * ((size_t)b * pstride + c) + 1]);
      }
      lds_fin[wave][c] = make_double2(sr, si);
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < ncols; c += blockDim.x) {
      double2 a = lds_fin[0][c], b = lds_fin[1][c], c2 = lds_fin[2][c], d = lds_fin[3][c];
      result[c] = make_double2((a.x + b.x) + (c2.x + d.x), (a.y + b.y) + (c2.y + d.y));
    }
  } else {
    if (threadIdx.x < 64) {
      double s = 0;
      for (uint32_t b = threadIdx.x; b < gridDim.x; b += 64) s += __builtin_nontemporal_load(&partials[b]);
      s = wave_sum(s);
      double hs = 0;
      if (hcol >= 0) {
        for (uint32_t c = threadIdx.x; c < ncols; c += 64) {
          double2 h = make_double2(add[c].x + sub[c].x, add[c].y + sub[c].y);
          Hre[(size_t)c * m + hcol] = h.x;
          Him[(size_t)c * m + hcol] = h.y;
          hs += h.x * h.x + h.y * h.y;
        }
      }
      hs = wave_sum(hs);
      if (threadIdx.x == 0) {
        double beta = sqrt(s);
        bool broke = hcol >= 0 && ncols > 0 && beta <= 1e-13 * sqrt(hs + s);
        if (broke) {
          beta = 0.0;
          if (st->breakdown == 0) st->breakdown = hrow;
        }
        st->beta = beta;
        if (hcol >= 0 && hrow >= 0 && (uint32_t)hrow < m) {
          Hre[(size_t)hrow * m + hcol] = beta;
          Him[(size_t)hrow * m + hcol] = 0.0;
        }
      }
    }
  }
  if (threadIdx.x == 0) *counter = 0u;
}

__global__ void k_normalize(uint32_t n, const double* fre, const double* fim, double* vre,
                            double* vim, const Status* st) {
  const double beta = st->beta;
  const double inv = beta > 0.0 ? 1.0 / beta : 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    vre[i] = fre[i] * inv;
    vim[i] = fim[i] * inv;
  }
}

// The small dense machinery (Schur form, eigenvectors, exact shifts) lives in
// small_dense.h, shared with the batched solver.
__global__ __launch_bounds__(512) void k_restart(RestartArgs a) {
  extern __shared__ double smem[];
  restart_body(a, smem);
}

// out[:, c] = sum_l V[:, l] M[l][c], c < ncols (out may not alias V).
// When fcol >= 0, column fcol is not stored; instead
// f <- (V M)[:, fcol] * beta_k + f * sigma (coefficients from the status).
constexpr int kVqChunk = 8;
__global__ __launch_bounds__(256) void k_vq(uint32_t n, uint32_t mrows, uint32_t ncols,
                                            const double* __restrict__ Vre,
                                            const double* __restrict__ Vim,
                                            const double* __restrict__ Mre,
                                            const double* __restrict__ Mim, uint32_t ldm,
                                            double* __restrict__ Ore, double* __restrict__ Oim,
                                            int fcol, double* fre, double* fim, const Status* st) {
  extern __shared__ double qs[];  // mrows x ncols complex: re block then im block
  const uint32_t total = mrows * ncols;
  for (uint32_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
    uint32_t l = idx / ncols, c = idx - l * ncols;
    qs[idx] = Mre[(size_t)l * ldm + c];
    qs[total + idx] = Mim[(size_t)l * ldm + c];
  }
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (uint32_t c0 = 0; c0 < ncols; c0 += kVqChunk) {
    double ar[kVqChunk], ai[kVqChunk];
#pragma unroll
    for (int q = 0; q < kVqChunk; ++q) { ar[q] = 0; ai[q] = 0; }
    for (uint32_t l = 0; l < mrows; ++l) {
      double vr = Vre[(size_t)l * n + i], vi = Vim[(size_t)l * n + i];
#pragma unroll
      for (int q = 0; q < kVqChunk; ++q) {
        uint32_t c = c0 + q;
        if (c < ncols) {
          double mr = qs[l * ncols + c], mi = qs[total + l * ncols + c];
          ar[q] += vr * mr - vi * mi;
          ai[q] += vr * mi + vi * mr;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kVqChunk; ++q) {
      uint32_t c = c0 + q;
      if (c >= ncols) continue;
      if ((int)c == fcol) {
        double br = st->beta_k_re, bi = st->beta_k_im, sr = st->sigma_re, si = st->sigma_im;
        double f_r = fre[i], f_i = fim[i];
        fre[i] = (ar[q] * br - ai[q] * bi) + (f_r * sr - f_i * si);
        fim[i] = (ar[q] * bi + ai[q] * br) + (f_r * si + f_i * sr);
      } else {
        Ore[(size_t)c * n + i] = ar[q];
        Oim[(size_t)c * n + i] = ai[q];
      }
    }
  }
}

// MODE 0: for each pair t < k: out[t] = (x^H y, x^H x) and lambda[t] = x^H y / x^H x.
// MODE 1: out[t].x = ||y - lambda_t x||^2 (lambda from MODE 0).
// X, AX: k columns of n (contiguous columns).  One launch each, last block finishes.
template <int MODE>
__global__ __launch_bounds__(256) void k_pairs(uint32_t n, uint32_t k, const double* Xre,
                                               const double* Xim, const double* Yre,
                                               const double* Yim, double2* lambda, double2* out,
                                               double* partials, unsigned* counter) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // each block covers 256 rows; wave w handles pairs t = w, w+4, ...
  for (uint32_t t = wave; t < k; t += 4) {
    double a0 = 0, a1 = 0, a2 = 0;
    for (uint32_t r = lane; r < 256; r += 64) {
      uint32_t i = blockIdx.x * 256 + r;
      if (i >= n) break;
      double xr = Xre[(size_t)t * n + i], xi = Xim[(size_t)t * n + i];
      double yr = Yre[(size_t)t * n + i], yi = Yim[(size_t)t * n + i];
      if (MODE == 0) {
        a0 += xr * yr + xi * yi;  // Re x^H y
        a1 += xr * yi - xi * yr;  // Im x^H y
        a2 += xr * xr + xi * xi;
      } else {
        double2 l = lambda[t];
        double rr = yr - (l.x * xr - l.y * xi), ri = yi - (l.x * xi + l.y * xr);
        a0 += rr * rr + ri * ri;
        a2 += xr * xr + xi * xi;
      }
    }
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    if (lane == 0) {
      partials[3 * ((size_t)blockIdx.x * k + t)] = a0;
      partials[3 * ((size_t)blockIdx.x * k + t) + 1] = a1;
      partials[3 * ((size_t)blockIdx.x * k + t) + 2] = a2;
    }
  }
  if (!last_block(counter)) return;
  for (uint32_t t = threadIdx.x; t < k; t += blockDim.x) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (uint32_t b = 0; b < gridDim.x; ++b) {
      s0 += __builtin_nontemporal_load(&partials[3 * ((size_t)b * k + t)]);
      s1 += __builtin_nontemporal_load(&partials[3 * ((size_t)b * k + t) + 1]);
      s2 += __builtin_nontemporal_load(&partials[3 * ((size_t)b * k + t) + 2]);
    }
    if (MODE == 0) {
      out[t] = make_double2(s0, s1);
      lambda[t] = make_double2(s0 / s2, s1 / s2);
    } else {
      out[t] = make_double2(s0, s2);
    }
  }
  if (threadIdx.x == 0) *counter = 0u;
}

// FP64 -> FP32 copy of the operator planes (mixed-precision matvec)
// dst = (float)src; *inexact is set to 1 (plain stores, all alike) when some
// element is not representable in FP32
__global__ void k_to_f32(size_t n, const double* __restrict__ src, float* __restrict__ dst,
                         unsigned* inexact) {
  bool bad = false;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double v = src[i];
    const float f = (float)v;
    dst[i] = f;
    bad |= (double)f != v;
  }
  if (bad && inexact) *inexact = 1u;
}

// x = sum_t X[:, t] (start vector of the polishing phase), written to V[:, 0]
__global__ void k_colsum(uint32_t n, uint32_t k, const double* Xre, const double* Xim,
                         double* vre, double* vim) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double sr = 0, si = 0;
    for (uint32_t t = 0; t < k; ++t) {
      sr += Xre[(size_t)t * n + i];
      si += Xim[(size_t)t * n + i];
    }
    vre[i] = sr;
    vim[i] = si;
  }
}

// seeded start vector: re = 2u-1 (SplitMix64 stream `seed`), im = 0 or 2u'-1
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t idx) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (idx + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
__global__ void k_random(uint32_t n, uint64_t seed, int cplx, double* re, double* im) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    re[i] = 2.0 * uniform01(seed, i) - 1.0;
    im[i] = cplx ? 2.0 * uniform01(seed ^ 0xA5A5ULL, i) - 1.0 : 0.0;
  }
}

}  // namespace

// ------------------------------------------------------------------ launchers
static unsigned orth_blocks(uint32_t n) { return (n + kRows - 1) / kRows; }

oaph_status orth(hipStream_t s, int mode, const Work& w, uint32_t ncols, const double* xre,
                 const double* xim, double* fre, double* fim, const double2* sub,
                 const double2* add, double2* result, int hcol, int hrow) {
  dim3 g(orth_blocks(w.n)), t(kOrthThreads);
  switch (mode) {
    case 0:
      hipLaunchKernelGGL(k_orth<0>, g, t, 0, s, w.n, ncols, w.Vre, w.Vim, xre, xim, fre, fim, sub,
                         add, result, w.partials, w.counter, w.Hre, w.Him, w.m, hcol, hrow, w.st);
      break;
    case 1:
      hipLaunchKernelGGL(k_orth<1>, g, t, 0, s, w.n, ncols, w.Vre, w.Vim, xre, xim, fre, fim, sub,
                         add, result, w.partials, w.counter, w.Hre, w.Him, w.m, hcol, hrow, w.st);
      break;
    default:
      hipLaunchKernelGGL(k_orth<2>, g, t, 0, s, w.n, ncols, w.Vre, w.Vim, xre, xim, fre, fim, sub,
                         add, result, w.partials, w.counter, w.Hre, w.Him, w.m, hcol, hrow, w.st);
      break;
  }
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

size_t orth_partials_bytes(uint32_t n, uint32_t m) {
  return (size_t)orth_blocks(n) * (m + 1) * 2 * sizeof(double);
}

oaph_status normalize(hipStream_t s, const Work& w, const double* fre, const double* fim,
                      double* vre, double* vim) {
  unsigned blocks = (w.n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_normalize, dim3(blocks), dim3(256), 0, s, w.n, fre, fim, vre, vim, w.st);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status restart(hipStream_t s, const RestartArgs& a) {
  OAPH_CHECK(a.mact >= 2 && a.mact <= kMaxM, OAPH_ERR_UNSUPPORTED);
  size_t lds = (size_t)4 * a.mact * lds_ld(a.mact) * sizeof(double);
  static bool attr_set = false;
  if (!attr_set) {
    // T and Z of a 64 x 64 problem need 130 KiB of the 160 KiB LDS
    OAPH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_restart),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 4 * kMaxM * lds_ld(kMaxM) * (int)sizeof(double)));
    attr_set = true;
  }
  hipLaunchKernelGGL(k_restart, dim3(1), dim3(512), lds, s, a);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status vq(hipStream_t s, uint32_t n, uint32_t mrows, uint32_t ncols, const double* Vre,
               const double* Vim, const double* Mre, const double* Mim, uint32_t ldm,
               double* Ore, double* Oim, int fcol, double* fre, double* fim, const Status* st) {
  size_t lds = (size_t)2 * mrows * ncols * sizeof(double);
  OAPH_CHECK(lds <= 64 * 1024, OAPH_ERR_UNSUPPORTED);
  hipLaunchKernelGGL(k_vq, dim3((n + 255) / 256), dim3(256), lds, s, n, mrows, ncols, Vre, Vim,
                     Mre, Mim, ldm, Ore, Oim, fcol, fre, fim, st);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status pairs(hipStream_t s, int mode, uint32_t n, uint32_t k, const double* Xre,
                  const double* Xim, const double* Yre, const double* Yim, double2* lambda,
                  double2* out, double* partials, unsigned* counter) {
  dim3 g((n + 255) / 256), t(256);
  if (mode == 0)
    hipLaunchKernelGGL(k_pairs<0>, g, t, 0, s, n, k, Xre, Xim, Yre, Yim, lambda, out, partials,
                       counter);
  else
    hipLaunchKernelGGL(k_pairs<1>, g, t, 0, s, n, k, Xre, Xim, Yre, Yim, lambda, out, partials,
                       counter);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

size_t pairs_partials_bytes(uint32_t n, uint32_t k) {
  return (size_t)((n + 255) / 256) * k * 3 * sizeof(double);
}

oaph_status to_f32(hipStream_t s, size_t n, const double* src, float* dst, unsigned* inexact) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)blocks), dim3(256), 0, s, n, src, dst, inexact);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status colsum(hipStream_t s, uint32_t n, uint32_t k, const double* Xre, const double* Xim,
                   double* vre, double* vim) {
  unsigned blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_colsum, dim3(blocks), dim3(256), 0, s, n, k, Xre, Xim, vre, vim);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status random_vector(hipStream_t s, uint32_t n, uint64_t seed, bool cplx, double* re,
                          double* im) {
  unsigned blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_random, dim3(blocks), dim3(256), 0, s, n, seed, cplx ? 1 : 0, re, im);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

}  // namespace iram
}  // namespace oaph
