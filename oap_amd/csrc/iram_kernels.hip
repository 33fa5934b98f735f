// iram_kernels.hip -- device side of the implicitly restarted Arnoldi solver.
//
// Krylov basis layout (solver-owned, not a math::Matrix): column c of V is
// contiguous, V[c*n + i], in separate re/im planes, so that the projection
// V^H w and the update w - V h stream V row-blocks with coalesced loads.
// (The reference keeps V as an n-row x m-column math::Matrix and rebuilds a
// full transpose every Arnoldi step: ArnoldiProcedures.cpp:441-443.)
//
// Kernels:
//   k_cgs<MODE>      one Arnoldi step's CGS2 in three launches (partial sums
//                    handed to the next launch; v_{j+1} written by the last)
//   k_residual       f - V s and ||f|| with a one-launch grid reduction (last
//                    block finishes; counter self-resets)
//   k_normalize      v = f / beta
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
constexpr int kRows = 64;  // rows per block in k_residual
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

// f = x - V sub over the first ncols basis columns (ncols may be 0), beta =
// ||f||: the last block writes beta to the status and, when hcol >= 0,
// beta to H[hrow][hcol] (hrow < m).  The start vector and the restart's
// column kk take this path; Arnoldi steps use the k_cgs chain.
__global__ __launch_bounds__(kOrthThreads) void k_residual(
    uint32_t n, uint32_t ncols, const double* __restrict__ Vre, const double* __restrict__ Vim,
    double* fre, double* fim, const double2* sub, double* partials, unsigned* counter,
    double* Hre, double* Him, uint32_t m, int hcol, int hrow, Status* st) {
  __shared__ double2 lds_part[4][kRows];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * kRows + lane;
  const bool in = i < n;
  double xr = in ? fre[i] : 0.0, xi = in ? fim[i] : 0.0;
  double pr = 0, pi = 0;
#pragma unroll
  for (int q = 0; q < kCPW; ++q) {
    const uint32_t c = wave + 4 * q;
    if (in && c < ncols) {
      const double vr = Vre[(size_t)c * n + i], vi = Vim[(size_t)c * n + i];
      const double2 s = sub[c];
      pr += vr * s.x - vi * s.y;
      pi += vr * s.y + vi * s.x;
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
  if (wave == 0) {
    double s = wave_sum(xr * xr + xi * xi);
    if (lane == 0) partials[blockIdx.x] = s;
  }
  if (!last_block(counter)) return;
  if (threadIdx.x < 64) {
    double s = 0;
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += 64) s += __builtin_nontemporal_load(&partials[b]);
    s = wave_sum(s);
    if (threadIdx.x == 0) {
      const double beta = sqrt(s);
      st->beta = beta;
      if (hcol >= 0 && hrow >= 0 && (uint32_t)hrow < m) {
        Hre[(size_t)hrow * m + hcol] = beta;
        Him[(size_t)hrow * m + hcol] = 0.0;
      }
    }
  }
  if (threadIdx.x == 0) *counter = 0u;
}

// ---------------------------------------------------------------- CGS2 chain
// One Arnoldi step's orthogonalisation in three launches with no grid-wide
// synchronisation inside a kernel: each launch leaves per-block partial sums
// and the next one starts by summing them (every block in the same order, so
// every block holds the same bits).  Block b owns rows [b rpb, (b+1) rpb),
// walked in chunks of 64 (lane = row); wave w owns basis columns w + 4q.
//   k_cgs<0>: P1[b] = V^H x over the block's rows.
//   k_cgs<1>: h = sum P1 (block 0 stores h); f1 = x - V h (written);
//             P2[b] = V^H f1, N[b] = ||f1||^2 over the block's rows.
//   k_cgs<2>: s = sum P2 (block 0 stores s); beta^2 = sum N - ||s||^2;
//             f = f1 - V s (written) and v = f / beta (when v is given);
//             block 0 writes H[c][hcol] = h_c + s_c, H[hrow][hcol] = beta, the
//             status beta, and flags breakdown (hcol >= 0) when beta <=
//             1e-13 ||(h + s, beta)||.
// beta from Pythagoras: f1 - V s is orthogonal to the (orthonormal) basis,
// so ||f||^2 = ||f1||^2 - ||s||^2; after the first pass ||s|| is of the
// order of the rounding of that pass, so the difference loses no digits
// unless f1 itself is rounding noise -- the breakdown case, where beta is
// flagged or tiny either way.  Saves the norm launch that would otherwise
// separate forming f from scaling it.
constexpr int kCgsMaxBlocks = 256;
constexpr int kCgsStride = kMaxMBig + 1;  // double2 entries per block in P1, P2

// coef[c] = sum_b part[b][c] for c < ncols (<= NC), into LDS; every thread calls.
template <int NC>
__device__ void sum_partials(const double2* __restrict__ part, uint32_t nparts, uint32_t ncols,
                             double2* coef) {
  __shared__ double2 acc[4][NC];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t c = lane; c < ncols; c += 64) {
    double sr = 0.0, si = 0.0;
#pragma unroll 16
    for (uint32_t b = wave; b < nparts; b += 4) {
      const double2 p = part[(size_t)b * kCgsStride + c];
      sr += p.x;
      si += p.y;
    }
    acc[wave][c] = make_double2(sr, si);
  }
  __syncthreads();
  for (uint32_t c = threadIdx.x; c < ncols; c += blockDim.x) {
    const double2 a = acc[0][c], b = acc[1][c], e = acc[2][c], d = acc[3][c];
    coef[c] = make_double2((a.x + b.x) + (e.x + d.x), (a.y + b.y) + (e.y + d.y));
  }
  __syncthreads();
}

struct CgsArgs {
  uint32_t n, ncols, rpb, nparts;
  uint32_t nin;  // partial sums in part_in (MODE 1: P1 may come from the matvec's blocks)
  const double* Vre;
  const double* Vim;
  const double* xre;
  const double* xim;
  double* fre;
  double* fim;
  const double2* part_in;  // P1 (MODE 1) or P2 (MODE 2)
  double2* part_out;       // P1 (MODE 0) or P2 (MODE 1)
  double* nrm;             // N: written by MODE 1, read by MODE 2
  double2* coef_out;       // h (MODE 1) or s (MODE 2), stored by block 0
  const double2* h;        // MODE 2: the first pass's coefficients
  double* vre;             // MODE 2: v = f / beta (nullptr: not written)
  double* vim;
  double* Hre;
  double* Him;
  uint32_t m;
  int hcol, hrow;
  Status* st;
};

template <int MODE>
__global__ __launch_bounds__(kOrthThreads) void k_cgs(CgsArgs a) {
  __shared__ double2 coef[kMaxM + 1];
  __shared__ double2 lds_part[4][64];
  __shared__ double inv_s;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t ncols = a.ncols, n = a.n;
  if (MODE >= 1) {
    sum_partials<kMaxM + 1>(a.part_in, a.nin, ncols, coef);
    if (blockIdx.x == 0)
      for (uint32_t c = threadIdx.x; c < ncols; c += blockDim.x) a.coef_out[c] = coef[c];
  }
  if (MODE == 2) {
    if (threadIdx.x < 64) {
      double f1n = 0, ss = 0, hs = 0;
      for (uint32_t b = lane; b < a.nparts; b += 64) f1n += a.nrm[b];
      for (uint32_t c = lane; c < ncols; c += 64) {
        const double2 s = coef[c];
        ss += s.x * s.x + s.y * s.y;
        if (a.hcol >= 0) {
          const double2 h = make_double2(a.h[c].x + s.x, a.h[c].y + s.y);
          hs += h.x * h.x + h.y * h.y;
          if (blockIdx.x == 0) {
            a.Hre[(size_t)c * a.m + a.hcol] = h.x;
            a.Him[(size_t)c * a.m + a.hcol] = h.y;
          }
        }
      }
      f1n = wave_sum(f1n);
      ss = wave_sum(ss);
      hs = wave_sum(hs);
      if (lane == 0) {
        const double b2 = f1n - ss;
        double beta = b2 > 0.0 ? sqrt(b2) : 0.0;
        const bool broke = a.hcol >= 0 && ncols > 0 && beta <= 1e-13 * sqrt(hs + beta * beta);
        if (broke) beta = 0.0;
        inv_s = beta > 0.0 ? 1.0 / beta : 0.0;
        if (blockIdx.x == 0) {
          if (broke && a.st->breakdown == 0) a.st->breakdown = a.hrow;
          a.st->beta = beta;
          if (a.hcol >= 0 && a.hrow >= 0 && (uint32_t)a.hrow < a.m) {
            a.Hre[(size_t)a.hrow * a.m + a.hcol] = beta;
            a.Him[(size_t)a.hrow * a.m + a.hcol] = 0.0;
          }
        }
      }
    }
    __syncthreads();
  }
  const double inv = MODE == 2 ? inv_s : 0.0;
  double sr[kCPW], si[kCPW];
#pragma unroll
  for (int q = 0; q < kCPW; ++q) { sr[q] = 0.0; si[q] = 0.0; }
  double nrm = 0.0;
  const uint32_t r0 = blockIdx.x * a.rpb;
  const uint32_t r1 = r0 + a.rpb < n ? r0 + a.rpb : n;
  for (uint32_t base = r0; base < r1; base += 64) {  // uniform over the block
    const uint32_t i = base + lane;
    const bool in = i < r1;
    double xr = in ? a.xre[i] : 0.0, xi = in ? a.xim[i] : 0.0;
    double vr[kCPW], vi[kCPW];
#pragma unroll
    for (int q = 0; q < kCPW; ++q) {
      const uint32_t c = wave + 4 * q;
      const bool ok = in && c < ncols;
      vr[q] = ok ? a.Vre[(size_t)c * n + i] : 0.0;
      vi[q] = ok ? a.Vim[(size_t)c * n + i] : 0.0;
    }
    if (MODE >= 1) {
      // this wave's share of (V coef)[i], combined over the waves in LDS
      double pr = 0, pi = 0;
#pragma unroll
      for (int q = 0; q < kCPW; ++q) {
        const uint32_t c = wave + 4 * q;
        if (c < ncols) {
          const double2 s = coef[c];
          pr += vr[q] * s.x - vi[q] * s.y;
          pi += vr[q] * s.y + vi[q] * s.x;
        }
      }
      if (base != r0) __syncthreads();  // the previous chunk's reads of lds_part are done
      lds_part[wave][lane] = make_double2(pr, pi);
      __syncthreads();
      const double2 p0 = lds_part[0][lane], p1 = lds_part[1][lane], p2 = lds_part[2][lane],
                    p3 = lds_part[3][lane];
      xr -= (p0.x + p1.x) + (p2.x + p3.x);
      xi -= (p0.y + p1.y) + (p2.y + p3.y);
      if (wave == 0 && in) {
        a.fre[i] = xr;
        a.fim[i] = xi;
        if (MODE == 2 && a.vre) {
          a.vre[i] = xr * inv;
          a.vim[i] = xi * inv;
        }
      }
    }
    if (MODE <= 1) {
#pragma unroll
      for (int q = 0; q < kCPW; ++q) {
        // conj(v) * x
        sr[q] += vr[q] * xr + vi[q] * xi;
        si[q] += vr[q] * xi - vi[q] * xr;
      }
    }
    if (MODE == 1) nrm += xr * xr + xi * xi;
  }
  if (MODE <= 1) {
    // the per-lane sums of every column reduced over the lanes through LDS
    // (one plane at a time): thread t sums lanes 16 (t % 4) .. +15 of
    // column t / 4 in order, and the four quarters are added as
    // (q0 + q1) + (q2 + q3)
    __shared__ double red[kMaxM + 1][64];
    double out[2];
#pragma unroll
    for (int plane = 0; plane < 2; ++plane) {
      if (plane) __syncthreads();  // the first plane's reads are done
#pragma unroll
      for (int q = 0; q < kCPW; ++q) {
        const uint32_t c = wave + 4 * q;
        if (c < ncols) red[c][lane] = plane ? si[q] : sr[q];
      }
      __syncthreads();
      double v = 0.0;
      const uint32_t c = threadIdx.x >> 2, l0 = (threadIdx.x & 3) * 16;
      if (c < ncols) {
#pragma unroll
        for (int l = 0; l < 16; ++l) v += red[c][l0 + l];
      }
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      out[plane] = v;
    }
    const uint32_t c = threadIdx.x >> 2;
    if ((threadIdx.x & 3) == 0 && c < ncols)
      a.part_out[(size_t)blockIdx.x * kCgsStride + c] = make_double2(out[0], out[1]);
    if (ncols > 64) {  // column 64 (ncols = kMaxM + 1)
      __syncthreads();
      const double pr = wave_sum(wave == 0 ? sr[16] : 0.0), pi = wave_sum(wave == 0 ? si[16] : 0.0);
      if (threadIdx.x == 0) a.part_out[(size_t)blockIdx.x * kCgsStride + 64] = make_double2(pr, pi);
    }
  }
  if (MODE == 1 && wave == 0) {
    nrm = wave_sum(nrm);
    if (lane == 0) a.nrm[blockIdx.x] = nrm;
  }
}

// The same three launches for more than kMaxM + 1 basis columns (Krylov
// dimension above kMaxM): the columns are walked in loops instead of
// per-column registers.  MODE >= 1: thread per row of the block forms
// f = x - V coef; MODE <= 1: wave w forms the block's partial sums of
// columns w, w + 4, ... (V^H x, or V^H f read back after the block barrier).
template <int MODE>
__global__ __launch_bounds__(kOrthThreads) void k_cgs_general(CgsArgs a) {
  __shared__ double2 coef[kMaxMBig + 1];
  __shared__ double inv_s;
  __shared__ double nrm_s[kOrthThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t ncols = a.ncols, n = a.n;
  const uint32_t r0 = blockIdx.x * a.rpb;
  const uint32_t r1 = r0 + a.rpb < n ? r0 + a.rpb : n;
  if (MODE >= 1) {
    sum_partials<kMaxMBig + 1>(a.part_in, a.nin, ncols, coef);
    if (blockIdx.x == 0)
      for (uint32_t c = threadIdx.x; c < ncols; c += blockDim.x) a.coef_out[c] = coef[c];
  }
  if (MODE == 2) {
    if (threadIdx.x < 64) {
      double f1n = 0, ss = 0, hs = 0;
      for (uint32_t b = lane; b < a.nparts; b += 64) f1n += a.nrm[b];
      for (uint32_t c = lane; c < ncols; c += 64) {
        const double2 s = coef[c];
        ss += s.x * s.x + s.y * s.y;
        if (a.hcol >= 0) {
          const double2 h = make_double2(a.h[c].x + s.x, a.h[c].y + s.y);
          hs += h.x * h.x + h.y * h.y;
          if (blockIdx.x == 0) {
            a.Hre[(size_t)c * a.m + a.hcol] = h.x;
            a.Him[(size_t)c * a.m + a.hcol] = h.y;
          }
        }
      }
      f1n = wave_sum(f1n);
      ss = wave_sum(ss);
      hs = wave_sum(hs);
      if (lane == 0) {
        const double b2 = f1n - ss;
        double beta = b2 > 0.0 ? sqrt(b2) : 0.0;
        const bool broke = a.hcol >= 0 && ncols > 0 && beta <= 1e-13 * sqrt(hs + beta * beta);
        if (broke) beta = 0.0;
        inv_s = beta > 0.0 ? 1.0 / beta : 0.0;
        if (blockIdx.x == 0) {
          if (broke && a.st->breakdown == 0) a.st->breakdown = a.hrow;
          a.st->beta = beta;
          if (a.hcol >= 0 && a.hrow >= 0 && (uint32_t)a.hrow < a.m) {
            a.Hre[(size_t)a.hrow * a.m + a.hcol] = beta;
            a.Him[(size_t)a.hrow * a.m + a.hcol] = 0.0;
          }
        }
      }
    }
    __syncthreads();
  }
  double nrm = 0.0;
  if (MODE >= 1) {
    const double inv = MODE == 2 ? inv_s : 0.0;
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
      double pr = 0.0, pi = 0.0;
      for (uint32_t c = 0; c < ncols; ++c) {
        const double vr = a.Vre[(size_t)c * n + i], vi = a.Vim[(size_t)c * n + i];
        const double2 s = coef[c];
        pr += vr * s.x - vi * s.y;
        pi += vr * s.y + vi * s.x;
      }
      const double xr = a.xre[i] - pr, xi = a.xim[i] - pi;
      a.fre[i] = xr;
      a.fim[i] = xi;
      if (MODE == 2 && a.vre) {
        a.vre[i] = xr * inv;
        a.vim[i] = xi * inv;
      }
      nrm += xr * xr + xi * xi;
    }
  }
  if (MODE <= 1) {
    if (MODE == 1) __syncthreads();  // the block's rows of f are written
    const double* xre = MODE == 0 ? a.xre : a.fre;
    const double* xim = MODE == 0 ? a.xim : a.fim;
    for (uint32_t c = wave; c < ncols; c += kOrthThreads / 64) {
      double sr = 0.0, si = 0.0;
      for (uint32_t i = r0 + lane; i < r1; i += 64) {
        const double vr = a.Vre[(size_t)c * n + i], vi = a.Vim[(size_t)c * n + i];
        const double xr = xre[i], xi = xim[i];
        sr += vr * xr + vi * xi;  // conj(v) x
        si += vr * xi - vi * xr;
      }
      sr = wave_sum(sr);
      si = wave_sum(si);
      if (lane == 0) a.part_out[(size_t)blockIdx.x * kCgsStride + c] = make_double2(sr, si);
    }
  }
  if (MODE == 1) {
    nrm = wave_sum(nrm);
    if (lane == 0) nrm_s[wave] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kOrthThreads / 64; ++w) t += nrm_s[w];
      a.nrm[blockIdx.x] = t;
    }
  }
}

// ---------------------------------------------------------------------------
// BF16-split projection (first Gram-Schmidt pass on the matrix cores).
// Each value v of the basis is kept as two BF16 numbers, hi = bf16(v) and
// lo = bf16(v - hi) (v to about 16 significant bits); V^H w is formed as
// hi.hi + hi.lo + lo.hi on v_mfma_f32_16x16x16_bf16 with FP32 accumulation
// per block and the per-block partial sums in FP64, the second pass in FP64.

__device__ __forceinline__ uint16_t bf16_rne_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ void bf16_split(double v, uint16_t& hi, uint16_t& lo) {
  hi = bf16_rne_bits((float)v);
  lo = bf16_rne_bits((float)(v - (double)bf16_value(hi)));
}

// Vb planes (each (m + 1) x n, column c at c * n): re hi, re lo, im hi, im lo
__global__ void k_v_to_bf16x2(uint32_t n, uint32_t c0, uint32_t c1, size_t plane,
                              const double* __restrict__ Vre, const double* __restrict__ Vim,
                              uint16_t* __restrict__ Vb) {
  const size_t total = (size_t)(c1 - c0) * n;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t i = (size_t)c0 * n + t;
    uint16_t h, l;
    bf16_split(Vre[i], h, l);
    Vb[i] = h;
    Vb[plane + i] = l;
    bf16_split(Vim[i], h, l);
    Vb[2 * plane + i] = h;
    Vb[3 * plane + i] = l;
  }
}

typedef short bs4 __attribute__((ext_vector_type(4)));
typedef float bf4 __attribute__((ext_vector_type(4)));

// part[b][c] = sum over rows r of block b of conj(V[c][r]) w[r], c < ncols <= 64,
// as one real product over the rows of [Vre | Vim] with the two right-hand
// columns [wre; wim] (re) and [wim; -wre] (im).  Wave q owns basis columns
// 16 q .. 16 q + 15 (the A rows of its 16 x 16 tiles); MFMA operand layout:
// lane l holds A(l % 16, 4 (l / 16) .. + 3) and B(4 (l / 16) .. + 3, l % 16),
// and C(4 (l / 16) + i, l % 16) for i < 4.  n and rpb are multiples of 64.
__global__ __launch_bounds__(256) void k_proj_bf16x3(uint32_t n, uint32_t ncols, uint32_t rpb,
                                                     size_t plane, const uint16_t* __restrict__ Vb,
                                                     const double* __restrict__ wre,
                                                     const double* __restrict__ wim,
                                                     double2* __restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t c = 16 * wave + (lane & 15);  // basis column of this lane's A row
  const int j = lane & 15;                     // B column of this lane
  const uint32_t kq = 4 * (lane >> 4);
  const uint32_t r0 = blockIdx.x * rpb;
  const uint32_t r1 = r0 + rpb < n ? r0 + rpb : n;
  bf4 acc = {0.f, 0.f, 0.f, 0.f};
  for (uint32_t k = r0; k < r1; k += 16) {
    const uint32_t r = k + kq;
    // w rows r .. r + 3, split
    double xr[4], xi[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      xr[e] = wre[r + e];
      xi[e] = wim[r + e];
    }
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      // A = Vre rows (half 0) or Vim rows (half 1) of column c
      bs4 ahi = {0, 0, 0, 0}, alo = {0, 0, 0, 0};
      if (c < ncols) {
        const size_t base = (size_t)(2 * half) * plane + (size_t)c * n + r;
        const uint2 h = *reinterpret_cast<const uint2*>(Vb + base);
        const uint2 l = *reinterpret_cast<const uint2*>(Vb + plane + base);
        ahi = bs4{(short)(h.x & 0xffff), (short)(h.x >> 16), (short)(h.y & 0xffff), (short)(h.y >> 16)};
        alo = bs4{(short)(l.x & 0xffff), (short)(l.x >> 16), (short)(l.y & 0xffff), (short)(l.y >> 16)};
      }
      // B column 0: re part of conj(v) w, column 1: im part
      bs4 bhi = {0, 0, 0, 0}, blo = {0, 0, 0, 0};
      if (j < 2) {
        uint16_t hh[4], ll[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double v = half == 0 ? (j == 0 ? xr[e] : xi[e]) : (j == 0 ? xi[e] : -xr[e]);
          bf16_split(v, hh[e], ll[e]);
        }
        bhi = bs4{(short)hh[0], (short)hh[1], (short)hh[2], (short)hh[3]};
        blo = bs4{(short)ll[0], (short)ll[1], (short)ll[2], (short)ll[3]};
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bhi, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, blo, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, bhi, acc, 0, 0, 0);
    }
  }
  // lane l: rows 4 (l / 16) + i (basis columns 16 wave + ...), B column l % 16
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float re = acc[i];
    const float im = __shfl_down(acc[i], 1, 64);  // the neighbour holds B column 1
    const uint32_t cc = 16 * wave + 4 * (lane >> 4) + i;
    if (j == 0 && cc < ncols) part[(size_t)blockIdx.x * kCgsStride + cc] = make_double2(re, im);
  }
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

// Krylov dimension above kMaxM: T in LDS while it fits, Z (and T beyond
// that) in the TZ work; single-shift sweeps by all threads.
__global__ __launch_bounds__(kGeneralThreads) void k_restart_general(RestartArgs a) {
  extern __shared__ double smem[];
  restart_body_t<kMaxMBig, true>(a, smem);
}

// out[:, c] = sum_l V[:, l] M[l][c], c < ncols (out may not alias V).
// When fcol >= 0, column fcol is not stored; instead
// f <- (V M)[:, fcol] * beta_k + f * sigma (coefficients from the status).
constexpr int kVqChunk = 4;
constexpr size_t kVqLdsMax = 64 * 1024;  // M staged in LDS up to this size, else read from L2
template <bool LDSM>
__global__ __launch_bounds__(256) void k_vq(uint32_t n, uint32_t mrows, uint32_t ncols,
                                            const double* __restrict__ Vre,
                                            const double* __restrict__ Vim,
                                            const double* __restrict__ Mre,
                                            const double* __restrict__ Mim, uint32_t ldm,
                                            double* __restrict__ Ore, double* __restrict__ Oim,
                                            int fcol, double* fre, double* fim, const Status* st) {
  extern __shared__ double qs[];  // mrows x ncols complex: re block then im block
  const uint32_t total = mrows * ncols;
  if (LDSM) {
    for (uint32_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
      uint32_t l = idx / ncols, c = idx - l * ncols;
      qs[idx] = Mre[(size_t)l * ldm + c];
      qs[total + idx] = Mim[(size_t)l * ldm + c];
    }
    __syncthreads();
  }
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  {
    const uint32_t c0 = blockIdx.y * kVqChunk;  // one chunk of columns per block row
    double ar[kVqChunk], ai[kVqChunk];
#pragma unroll
    for (int q = 0; q < kVqChunk; ++q) { ar[q] = 0; ai[q] = 0; }
#pragma unroll 8
    for (uint32_t l = 0; l < mrows; ++l) {
      double vr = Vre[(size_t)l * n + i], vi = Vim[(size_t)l * n + i];
#pragma unroll
      for (int q = 0; q < kVqChunk; ++q) {
        uint32_t c = c0 + q;
        if (c < ncols) {
          double mr = LDSM ? qs[l * ncols + c] : Mre[(size_t)l * ldm + c];
          double mi = LDSM ? qs[total + l * ncols + c] : Mim[(size_t)l * ldm + c];
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
// flags[0] <- 1 when an element rounds, flags[1] <- 1 when a finite element
// is beyond the FP32 range (its copy is infinite)
__global__ void k_to_f32(size_t n, const double* __restrict__ src, float* __restrict__ dst,
                         unsigned* flags) {
  bool bad = false, over = false;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double v = src[i];
    const float f = (float)v;
    dst[i] = f;
    bad |= (double)f != v;
    over |= isinf(f) && !isinf(v);
  }
  if (bad && flags) flags[0] = 1u;
  if (over && flags) flags[1] = 1u;
}

// largest |value| of an operator plane: per-thread maxima, a wave maximum,
// one atomic max per wave on the bit pattern
__global__ void k_abs_max(size_t n, const double* __restrict__ src, unsigned long long* amax) {
  unsigned long long m = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(src[i]));
    m = b > m ? b : m;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m) atomicMax(amax, m);
}

// FP64 -> FP16 copy of an operator plane scaled by a power of two: the FP16
// rounding (nearest even, one rounding) of src * scale (the product is exact)
__global__ void k_to_f16(size_t n, const double* __restrict__ src, double scale,
                         _Float16* __restrict__ dst) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (_Float16)(src[i] * scale);
}

// x = sum_{t < k} X[:, t] + w sum_{k <= t < count} X[:, t] (start vector of
// a Ritz-sum restart: the wanted Ritz vectors and the weighted guards),
// written to V[:, 0]
__global__ void k_colsum(uint32_t n, uint32_t k, uint32_t count, double w, const double* Xre,
                         const double* Xim, double* vre, double* vim) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double sr = 0, si = 0, gr = 0, gi = 0;
    for (uint32_t t = 0; t < count; ++t) {
      const double xr = Xre[(size_t)t * n + i], xi = Xim[(size_t)t * n + i];
      if (t < k) {
        sr += xr;
        si += xi;
      } else {
        gr += xr;
        gi += xi;
      }
    }
    vre[i] = sr + w * gr;
    vim[i] = si + w * gi;
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

oaph_status residual(hipStream_t s, const Work& w, uint32_t ncols, double* fre, double* fim,
                     const double2* sub, int hcol, int hrow) {
  OAPH_CHECK(ncols <= kMaxM + 1, OAPH_ERR_UNSUPPORTED);
  hipLaunchKernelGGL(k_residual, dim3(orth_blocks(w.n)), dim3(kOrthThreads), 0, s, w.n, ncols,
                     w.Vre, w.Vim, fre, fim, sub, w.partials, w.counter, w.Hre, w.Him, w.m, hcol,
                     hrow, w.st);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

namespace {

// blocks of the CGS2 chain and the rows each owns (a multiple of 64)
unsigned cgs_blocks(uint32_t n) {
  unsigned b = (n + 63) / 64;
  return b < kCgsMaxBlocks ? (b ? b : 1) : kCgsMaxBlocks;
}

CgsArgs cgs_args(const Work& w, uint32_t ncols) {
  CgsArgs a{};
  const unsigned nb = cgs_blocks(w.n);
  a.n = w.n;
  a.ncols = ncols;
  a.nparts = nb;
  a.nin = nb;
  a.rpb = ((w.n + nb - 1) / nb + 63) / 64 * 64;
  a.Vre = w.Vre;
  a.Vim = w.Vim;
  a.Hre = w.Hre;
  a.Him = w.Him;
  a.m = w.m;
  a.st = w.st;
  a.hcol = -1;
  a.hrow = -1;
  return a;
}

double2* cgs_p1(const Work& w) { return reinterpret_cast<double2*>(w.partials); }
double2* cgs_p2(const Work& w) { return cgs_p1(w) + (size_t)kCgsMaxBlocks * kCgsStride; }
double* cgs_nrm(const Work& w) {
  return reinterpret_cast<double*>(cgs_p2(w) + (size_t)kCgsMaxBlocks * kCgsStride);
}

}  // namespace

size_t orth_partials_bytes(uint32_t n) {
  const size_t cgs = (size_t)2 * kCgsMaxBlocks * kCgsStride * sizeof(double2) +
                     (size_t)kCgsMaxBlocks * sizeof(double);
  const size_t res = (size_t)orth_blocks(n) * sizeof(double);
  return cgs > res ? cgs : res;
}

oaph_status cgs_project(hipStream_t s, const Work& w, uint32_t ncols, const double* xre,
                        const double* xim) {
  OAPH_CHECK(ncols >= 1 && ncols <= kMaxMBig + 1, OAPH_ERR_UNSUPPORTED);
  CgsArgs a = cgs_args(w, ncols);
  a.xre = xre;
  a.xim = xim;
  a.part_out = cgs_p1(w);
  if (ncols > kMaxM + 1)
    hipLaunchKernelGGL(k_cgs_general<0>, dim3(a.nparts), dim3(kOrthThreads), 0, s, a);
  else
    hipLaunchKernelGGL(k_cgs<0>, dim3(a.nparts), dim3(kOrthThreads), 0, s, a);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

MvProj cgs_matvec_projection(const Work& w, uint32_t ncols) {
  MvProj p;
  p.Vre = w.Vre;
  p.Vim = w.Vim;
  p.ldv = w.n;
  p.ncols = ncols;
  p.ldp = kCgsStride;
  p.maxparts = kCgsMaxBlocks;
  p.part = cgs_p1(w);
  return p;
}

oaph_status cgs_update_project(hipStream_t s, const Work& w, uint32_t ncols, const double* xre,
                               const double* xim, double* fre, double* fim, double2* h,
                               uint32_t p1_parts) {
  OAPH_CHECK(ncols >= 1 && ncols <= kMaxMBig + 1, OAPH_ERR_UNSUPPORTED);
  OAPH_CHECK(p1_parts <= kCgsMaxBlocks, OAPH_ERR_ARG);
  CgsArgs a = cgs_args(w, ncols);
  if (p1_parts) a.nin = p1_parts;
  a.xre = xre;
  a.xim = xim;
  a.fre = fre;
  a.fim = fim;
  a.part_in = cgs_p1(w);
  a.part_out = cgs_p2(w);
  a.nrm = cgs_nrm(w);
  a.coef_out = h;
  if (ncols > kMaxM + 1)
    hipLaunchKernelGGL(k_cgs_general<1>, dim3(a.nparts), dim3(kOrthThreads), 0, s, a);
  else
    hipLaunchKernelGGL(k_cgs<1>, dim3(a.nparts), dim3(kOrthThreads), 0, s, a);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status cgs_finish(hipStream_t s, const Work& w, uint32_t ncols, double* fre, double* fim,
                       const double2* h, double2* sc, int hcol, int hrow, double* vre,
                       double* vim) {
  OAPH_CHECK(ncols >= 1 && ncols <= kMaxMBig + 1, OAPH_ERR_UNSUPPORTED);
  CgsArgs a = cgs_args(w, ncols);
  a.xre = fre;
  a.xim = fim;
  a.fre = fre;
  a.fim = fim;
  a.part_in = cgs_p2(w);
  a.nrm = cgs_nrm(w);
  a.coef_out = sc;
  a.h = h;
  a.vre = vre;
  a.vim = vim;
  a.hcol = hcol;
  a.hrow = hrow;
  if (ncols > kMaxM + 1)
    hipLaunchKernelGGL(k_cgs_general<2>, dim3(a.nparts), dim3(kOrthThreads), 0, s, a);
  else
    hipLaunchKernelGGL(k_cgs<2>, dim3(a.nparts), dim3(kOrthThreads), 0, s, a);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status v_to_bf16x2(hipStream_t s, uint32_t n, uint32_t m, uint32_t c0, uint32_t c1,
                       const double* Vre, const double* Vim, uint16_t* Vb) {
  if (c1 <= c0) return OAPH_OK;
  const size_t total = (size_t)(c1 - c0) * n;
  size_t blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_v_to_bf16x2, dim3((unsigned)blocks), dim3(256), 0, s, n, c0, c1,
                     (size_t)(m + 1) * n, Vre, Vim, Vb);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

bool proj_bf16x3_supported(uint32_t n, uint32_t ncols) {
  return n % 64 == 0 && ncols >= 1 && ncols <= 64;
}

oaph_status cgs_project_bf16x3(hipStream_t s, const Work& w, uint32_t ncols, const double* xre,
                               const double* xim, const uint16_t* Vb, uint32_t* parts) {
  OAPH_CHECK(proj_bf16x3_supported(w.n, ncols), OAPH_ERR_UNSUPPORTED);
  const CgsArgs a = cgs_args(w, ncols);
  hipLaunchKernelGGL(k_proj_bf16x3, dim3(a.nparts), dim3(256), 0, s, w.n, ncols, a.rpb,
                     (size_t)(w.m + 1) * w.n, Vb, xre, xim, cgs_p1(w));
  OAPH_HIP(hipGetLastError());
  *parts = a.nparts;
  return OAPH_OK;
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
  OAPH_CHECK(a.mact >= 2 && a.mact <= kMaxMBig, OAPH_ERR_UNSUPPORTED);
  if (a.mact > kMaxM) {
    OAPH_CHECK(a.TZ != nullptr, OAPH_ERR_ARG);
    static bool general_attr_set = false;
    if (!general_attr_set) {
      OAPH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_restart_general),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kGeneralLdsMax));
      general_attr_set = true;
    }
    hipLaunchKernelGGL(k_restart_general, dim3(1), dim3(kGeneralThreads),
                       (size_t)general_lds_bytes(a.mact), s, a);
    OAPH_HIP(hipGetLastError());
    return OAPH_OK;
  }
  const size_t lds = (size_t)restart_lds_bytes(a.mact);
  static bool attr_set = false;
  if (!attr_set) {
    // T and Z of a 64 x 64 problem need 130 KiB of the 160 KiB LDS; up to
    // m = 54 X joins them (at most 144 KiB)
    OAPH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_restart),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, kRestartLdsMax));
    attr_set = true;
  }
  hipLaunchKernelGGL(k_restart, dim3(1), dim3(512), lds, s, a);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status vq(hipStream_t s, uint32_t n, uint32_t mrows, uint32_t ncols, const double* Vre,
               const double* Vim, const double* Mre, const double* Mim, uint32_t ldm,
               double* Ore, double* Oim, int fcol, double* fre, double* fim, const Status* st) {
  const size_t lds = (size_t)2 * mrows * ncols * sizeof(double);
  // rows x column chunks: enough workgroups to spread over the CUs
  const dim3 grid((n + 63) / 64, (ncols + kVqChunk - 1) / kVqChunk);
  if (lds <= kVqLdsMax)
    hipLaunchKernelGGL(k_vq<true>, grid, dim3(64), lds, s, n, mrows, ncols, Vre, Vim,
                       Mre, Mim, ldm, Ore, Oim, fcol, fre, fim, st);
  else
    hipLaunchKernelGGL(k_vq<false>, grid, dim3(64), 0, s, n, mrows, ncols, Vre, Vim,
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

oaph_status to_f32(hipStream_t s, size_t n, const double* src, float* dst, unsigned* flags) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)blocks), dim3(256), 0, s, n, src, dst, flags);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status abs_max(hipStream_t s, size_t n, const double* src, unsigned long long* amax) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_abs_max, dim3((unsigned)blocks), dim3(256), 0, s, n, src, amax);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status to_f16(hipStream_t s, size_t n, const double* src, double scale, _Float16* dst) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_to_f16, dim3((unsigned)blocks), dim3(256), 0, s, n, src, scale, dst);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

oaph_status colsum(hipStream_t s, uint32_t n, uint32_t k, uint32_t count, double w,
                   const double* Xre, const double* Xim, double* vre, double* vim) {
  unsigned blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_colsum, dim3(blocks), dim3(256), 0, s, n, k, count, w, Xre, Xim, vre, vim);
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
