// batched.hip -- many independent real symmetric covariance eigenproblems
// solved in ONE persistent launch.
//
// Problem i: C_i = X_i^T X_i / s with X_i an s x n_i FP32 row-major matrix
// (rows padded to ldx_i = round_up(n_i, 4) with zeros).  This is the operator
// of the reference's RecToSquareApi / SquareMatrixApi (M M^T sub-blocks,
// oapMatrixCuda/RecToSquareApi.cpp:57-141) applied implicitly: C v is
// computed as X^T (X v) / s in a single pass over X -- each wave holds one
// row of X in registers, forms u_r = x_r . v by a cross-lane sum, and
// accumulates u_r x_r into its share of w -- so a matvec reads X once
// (s ldx 4 bytes) and never forms C.
//
// One workgroup per problem runs the whole implicitly restarted Lanczos
// solve (the symmetric case of the Arnoldi iteration: full CGS2
// reorthogonalisation, a real symmetric tridiagonal projection T), restarts
// with exact shifts by implicit QR sweeps on T's diagonal and off-diagonal
// (sym_restart), and computes the Ritz vectors and the explicit FP64
// residual ||C x - lambda x|| of every wanted pair.  Problems are dispatched
// largest n first so that the long ones start early.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "internal.h"

namespace oaph {
namespace batched {


constexpr int kMaxMB = 24;   // Krylov dimension limit of the batched solver
constexpr int kMaxN = 1024;  // largest problem dimension
constexpr int kMaxS = 1024;  // largest sample count

struct Problem {
  uint32_t n, ldx;
  uint64_t x_off;  // floats
  uint64_t v_off;  // doubles: (m+1) columns of n
  uint64_t out_off;  // doubles: Ritz vectors k x n
};

struct Result {
  double ev[kMaxMB];
  double resid[kMaxMB];
  int nconv;
  int restarts;
  int matvecs;
  int pad;
  long long cycles[5];  // shader clock per phase: matvec, orthogonalisation, restart, V Q, Ritz vectors
};

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double u01(uint64_t seed, uint64_t idx) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (idx + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

template <int NT>
struct Ctx {
  static constexpr int NW = NT / 64;
  static constexpr int CPT = (kMaxN + NT - 1) / NT;  // columns per thread
  static constexpr int Q4 = kMaxN / 4 / 64;          // float4 chunks per lane per row
};

// Block-wide sums of nv <= NV values per thread; out[q] receives sum q.
// Ends with a barrier, so out is readable by every thread.
template <int NT, int NV>
__device__ void block_sums(const double (&v)[NV], int nv, double* red, double* out) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    if (q < nv) {
      const double t = wave_sum(v[q]);
      if (lane == 0) red[wave * (kMaxMB + 2) + q] = t;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    double t = 0;
    for (int w = 0; w < NW; ++w) t += red[w * (kMaxMB + 2) + threadIdx.x];
    out[threadIdx.x] = t;
  }
  __syncthreads();
}

// w = X^T (X v) / s in one pass over X.  vl (LDS) holds v; the result goes to
// wl (LDS).  A row's float4 chunks are spread over SEG lanes (SEG = 16, 32 or
// 64, the smallest that covers the row), so a wave loads 64/SEG rows per
// instruction and ROWS such groups per step, the waves interleaved;
// u_r = x_r . v is a segmented cross-lane sum and u_r x_r accumulates in the
// lanes' registers.  Lanes holding the same columns are summed across the
// segments, then the waves' partial sums are added into wl in wave order
// (deterministic, no LDS partial buffer).  The rows stream through a ring of
// D register buffers: D - 1 steps of loads in flight while one is computed.
template <int NT, int ROWS, int SEG, int D>
__device__ void covariance_matvec_seg(const float* __restrict__ X, uint32_t n, uint32_t ldx,
                                      uint32_t s, const double* vl, double* wl) {
  constexpr int NW = NT / 64, RPL = 64 / SEG;
  constexpr int Q4 = SEG == 64 ? kMaxN / 4 / 64 : 1;  // float4 chunks per lane per row
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane / SEG, q0 = lane % SEG;
  const uint32_t nq = ldx / 4;  // float4 chunks per row (padding is zero)
  double vreg[Q4][4], acc[Q4][4];
#pragma unroll
  for (int t = 0; t < Q4; ++t) {
    const uint32_t q = q0 + SEG * t;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint32_t c = 4 * q + e;
      vreg[t][e] = c < n ? vl[c] : 0.0;
      acc[t][e] = 0.0;
    }
  }
  const f4* X4 = reinterpret_cast<const f4*>(X);
  // rows r0 + rr*RPL + sub of one step; the next step's rows are loaded
  // before this step's arithmetic (two steps of loads in flight)
  auto load = [&](f4 (&x)[ROWS][Q4], uint32_t r0) {
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
      const uint32_t r = r0 + rr * RPL + sub;
#pragma unroll
      for (int t = 0; t < Q4; ++t) {
        const uint32_t q = q0 + SEG * t;
        const f4 z = {0.f, 0.f, 0.f, 0.f};
        x[rr][t] = (q < nq && r < s) ? X4[(size_t)r * nq + q] : z;
      }
    }
  };
  auto step = [&](const f4 (&x)[ROWS][Q4]) {
    double u[ROWS];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
      double d = 0;
#pragma unroll
      for (int t = 0; t < Q4; ++t) {
#pragma unroll
        for (int e = 0; e < 4; ++e) d += (double)x[rr][t][e] * vreg[t][e];
      }
      u[rr] = d;
    }
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
#pragma unroll
      for (int off = SEG / 2; off > 0; off >>= 1) u[rr] += __shfl_xor(u[rr], off, 64);
    }
#pragma unroll
    for (int t = 0; t < Q4; ++t) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) acc[t][e] += u[rr] * (double)x[rr][t][e];
      }
    }
  };
  const uint32_t stride = NW * RPL * ROWS;
  f4 xr[D][ROWS][Q4];
  const uint32_t r0 = wave * RPL * ROWS;
#pragma unroll
  for (int b = 0; b + 1 < D; ++b)
    if (r0 + b * stride < s) load(xr[b], r0 + b * stride);
  for (uint32_t base = r0; base < s; base += D * stride) {
#pragma unroll
    for (int b = 0; b < D; ++b) {
      const uint32_t rr = base + b * stride;
      if (rr < s) {
        const uint32_t rn = rr + (D - 1) * stride;
        if (rn < s) load(xr[(b + D - 1) % D], rn);
        step(xr[b]);
      }
    }
  }
  // lanes with the same columns: sum across the row segments
  if (SEG < 64) {
#pragma unroll
    for (int t = 0; t < Q4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int off = SEG; off < 64; off <<= 1) acc[t][e] += __shfl_xor(acc[t][e], off, 64);
  }
  const double inv_s = 1.0 / (double)s;
  for (int w = 0; w < NW; ++w) {
    if (wave == w && sub == 0) {
#pragma unroll
      for (int t = 0; t < Q4; ++t) {
        const uint32_t q = q0 + SEG * t;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t c = 4 * q + e;
          if (c < n) {
            double v = w == 0 ? acc[t][e] : wl[c] + acc[t][e];
            wl[c] = w == NW - 1 ? v * inv_s : v;
          }
        }
      }
    }
    __syncthreads();
  }
}

template <int NT, int ROWS, int D>
__device__ void covariance_matvec(const float* __restrict__ X, uint32_t n, uint32_t ldx,
                                  uint32_t s, const double* vl, double* wl) {
  const uint32_t nq = ldx / 4;
  if (nq <= 16) covariance_matvec_seg<NT, ROWS, 16, D>(X, n, ldx, s, vl, wl);
  else if (nq <= 32) covariance_matvec_seg<NT, ROWS, 32, D>(X, n, ldx, s, vl, wl);
  else covariance_matvec_seg<NT, ROWS, 64, D>(X, n, ldx, s, vl, wl);
}

struct SolveArgs {
  const Problem* probs;
  const uint32_t* order;  // dispatch order (largest n first)
  Result* results;
  const float* X;
  double* V;
  double* out;
  uint32_t s, m, k;
  int which;
  double tol;
  uint32_t max_restarts;
  uint64_t seed;
};

// ---------------------------------------------------------------------------
// Restart of the symmetric problem: the Lanczos matrix T is real symmetric
// tridiagonal (diagonal d, off-diagonal e, e[j] = T[j+1][j]), so its Schur
// form is an eigendecomposition computed by implicit QR sweeps with
// Wilkinson shifts on (d, e) alone, and the exact shifts of the restart are
// the same sweeps with the unwanted Ritz values (ARPACK's dseigt / dsapps
// steps).  The scalar chain runs in wave 0 (every lane computes the same
// values; lane 0 stores them); the eigenvector / shift accumulation is
// row-parallel, lane r owning row r.
// ---------------------------------------------------------------------------
constexpr double kEps = 2.220446049250313e-16;

__device__ __forceinline__ bool negligible(double e, double a, double b) {
  return fabs(e) <= kEps * (fabs(a) + fabs(b)) || fabs(e) < 1e-300;
}

// One implicit symmetric QR sweep with shift mu on rows/columns [l, h] of
// (d, e), accumulating M <- M R^T on rows 0 .. nrows-1 (ld ldm).  Rotation k
// R = [[c, s], [-s, c]] on rows/columns k, k+1 zeroes the bulge (from
// (d[l] - mu, e[l]) at k = l); T <- R T R^T.
__device__ void sym_sweep(double* d, double* e, int l, int h, double mu, double* M, int ldm,
                          int nrows) {
  const int lane = threadIdx.x & 63;
  double a = d[l], b = e[l], dd = d[l + 1];
  double en = l + 1 < h ? e[l + 1] : 0.0;
  double x = a - mu, z = b;
  double mk = lane < nrows ? M[lane * ldm + l] : 0.0;
  double mk1 = lane < nrows ? M[lane * ldm + l + 1] : 0.0;
  for (int k = l; k < h; ++k) {
    // the next step's operands, loaded ahead
    const double d2 = k + 2 <= h ? d[k + 2] : 0.0;
    const double e2 = k + 2 < h ? e[k + 2] : 0.0;
    const double mk2 = (lane < nrows && k + 2 <= h) ? M[lane * ldm + k + 2] : 0.0;
    double c = 1.0, s = 0.0, r = x;
    const double q2 = x * x + z * z;
    if (z != 0.0) {
      const double inv = rsqrt(q2);
      c = x * inv;
      s = z * inv;
      r = q2 * inv;
    }
    const double c2 = c * c, s2 = s * s, cs = c * s;
    const double na = c2 * a + 2.0 * cs * b + s2 * dd;
    const double ndd = s2 * a - 2.0 * cs * b + c2 * dd;
    const double nb = cs * (dd - a) + (c2 - s2) * b;
    const double bulge = s * en, ne = c * en;
    if (lane == 0) {
      if (k > l) e[k - 1] = r;
      d[k] = na;
    }
    if (lane < nrows) M[lane * ldm + k] = c * mk + s * mk1;
    mk = -s * mk + c * mk1;
    mk1 = mk2;
    if (k + 1 < h) {
      a = ndd;
      b = ne;
      dd = d2;
      en = e2;
      x = nb;
      z = bulge;
    } else if (lane == 0) {
      d[k + 1] = ndd;
      e[k] = nb;
    }
  }
  if (lane < nrows) M[lane * ldm + h] = mk;
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ void wave_fence() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ bool sym_before(double a, double b, int which) {
  switch (which) {
    case 0: return fabs(a) > fabs(b);   // largest |lambda|
    case 1: return a > b;               // largest real part
    case 3: return fabs(a) < fabs(b);   // smallest |lambda|
    case 4: return a < b;               // smallest real part
    default: return false;              // imaginary parts are all zero
  }
}

struct SymLds {
  double alpha[kMaxMB], beta[kMaxMB];  // Lanczos T (beta[j] = T[j+1][j])
  double d[kMaxMB], e[kMaxMB];         // work copies
  double ev[kMaxMB], est[kMaxMB];
  int order[kMaxMB];
  double Z[kMaxMB * (kMaxMB + 1)];     // eigenvectors of T, then Q of the shifts
  double Y[kMaxMB * kMaxMB];           // wanted Ritz vectors of T (m x k)
  int ctl[4];                          // nconv, kk, fail
  double ctld[4];                      // beta_k, sigma, maxabs
};

// wave 0 only.  T = (alpha, beta) of size m, bm = ||f||.  Leaves in L:
// ev/est/order, Y, ctl = {nconv, kk}, and when `shift`: the kept T in
// (alpha, beta)[0..kk), Z = Q of the shifts, ctld = {beta_k, sigma}.
__device__ void sym_restart(SymLds& L, int m, int k, int which, double tol, double bm, bool shift) {
  const int lane = threadIdx.x & 63;
  const int ld = kMaxMB + 1;
  // eigendecomposition of T
  if (lane < m) {
    for (int c = 0; c < m; ++c) L.Z[lane * ld + c] = lane == c ? 1.0 : 0.0;
  }
  if (lane == 0) {
    for (int i = 0; i < m; ++i) {
      L.d[i] = L.alpha[i];
      L.e[i] = i + 1 < m ? L.beta[i] : 0.0;
    }
  }
  wave_fence();
  int h = m - 1, iters = 0, fail = 0;
  while (h > 0) {
    if (negligible(L.e[h - 1], L.d[h - 1], L.d[h])) {
      if (lane == 0) L.e[h - 1] = 0.0;
      wave_fence();
      --h;
      continue;
    }
    int l = h - 1;
    while (l > 0 && !negligible(L.e[l - 1], L.d[l - 1], L.d[l])) --l;
    if (l > 0 && lane == 0) L.e[l - 1] = 0.0;
    if (++iters > 30 * m) {
      fail = 1;
      break;
    }
    // Wilkinson shift from the trailing 2 x 2 of the block
    const double a = L.d[h - 1], b = L.e[h - 1], cc = L.d[h];
    const double del = 0.5 * (a - cc);
    const double den = del + copysign(sqrt(del * del + b * b), del);
    const double mu = den != 0.0 ? cc - b * b / den : cc - fabs(b);
    wave_fence();
    sym_sweep(L.d, L.e, l, h, mu, L.Z, ld, m);
  }
  // Ritz values, estimates bm |Z[m-1][i]|, wanted order
  double maxabs = 0.0;
  for (int i = 0; i < m; ++i) maxabs = fmax(maxabs, fabs(L.d[i]));
  if (lane < m) {
    L.ev[lane] = L.d[lane];
    L.est[lane] = bm * fabs(L.Z[(m - 1) * ld + lane]);
    int rank = 0;
    const double me = L.d[lane];
    for (int j = 0; j < m; ++j) {
      const double oj = L.d[j];
      if (sym_before(oj, me, which) || (!sym_before(me, oj, which) && j < lane)) ++rank;
    }
    L.order[rank] = lane;
  }
  wave_fence();
  int nconv = 0;
  const double tolabs = tol * fmax(maxabs, 1e-300);
  for (int t = 0; t < k && t < m; ++t)
    if (L.est[L.order[t]] <= tolabs) ++nconv;
  const int p2 = (m - k) / 2;
  int kk = k + (nconv < p2 ? nconv : p2);
  if (kk == 1 && m >= 6) kk = m / 2;
  else if (kk == 1 && m > 2) kk = 2;
  if (kk >= m) kk = m - 1;
  // wanted Ritz vectors of T
  if (lane < m)
    for (int t = 0; t < k; ++t) L.Y[lane * k + t] = L.Z[lane * ld + L.order[t]];
  if (lane == 0) {
    L.ctl[0] = nconv;
    L.ctl[1] = kk;
    L.ctl[2] = fail;
    L.ctld[2] = maxabs;
  }
  wave_fence();
  if (!shift) return;
  // exact shifts: sweeps with the unwanted Ritz values on the original T
  if (lane < m)
    for (int c = 0; c < m; ++c) L.Z[lane * ld + c] = lane == c ? 1.0 : 0.0;
  if (lane == 0)
    for (int i = 0; i < m; ++i) {
      L.d[i] = L.alpha[i];
      L.e[i] = i + 1 < m ? L.beta[i] : 0.0;
    }
  wave_fence();
  for (int t = kk; t < m; ++t) {
    const double mu = L.ev[L.order[t]];
    int l = 0;
    while (l < m - 1) {
      int hh = l;
      while (hh < m - 1 && !negligible(L.e[hh], L.d[hh], L.d[hh + 1])) ++hh;
      if (hh < m - 1 && lane == 0) L.e[hh] = 0.0;
      wave_fence();
      if (hh > l) sym_sweep(L.d, L.e, l, hh, mu, L.Z, ld, m);
      l = hh + 1;
    }
  }
  if (lane == 0) {
    for (int i = 0; i < kk; ++i) {
      L.alpha[i] = L.d[i];
      L.beta[i] = L.e[i];  // beta[kk-1] is replaced by the new ||f||
    }
    L.ctld[0] = L.e[kk - 1];                // T+[kk][kk-1]
    L.ctld[1] = L.Z[(m - 1) * ld + kk - 1];  // Q[m-1][kk-1]
  }
  wave_fence();
}

// One workgroup per problem.  The second launch bound (two waves per SIMD)
// caps the registers at 256, so two 256-thread workgroups share a CU: the
// two-row variant needed 280 without it and ran one workgroup per CU.
template <int NT, int ROWS, int D = 2>
__global__ __launch_bounds__(NT, 2) void k_batched_solve(SolveArgs a) {
  constexpr int CPT = Ctx<NT>::CPT;
  __shared__ double vl[kMaxN], wl[kMaxN];
  __shared__ double red[(NT / 64) * (kMaxMB + 2)];
  __shared__ double coef[kMaxMB + 2];
  __shared__ SymLds L;
  const uint32_t pid = a.order[blockIdx.x];
  const Problem P = a.probs[pid];
  const uint32_t n = P.n, m = a.m, k = a.k;
  const float* X = a.X + P.x_off;
  double* V = a.V + P.v_off;  // column c at V + c*n; column m holds f
  const int tid = threadIdx.x;
  constexpr int ld = kMaxMB + 1;

  // start vector: uniform in [-1, 1) from the problem's SplitMix64 stream
  {
    double ss[1] = {0};
    for (uint32_t c = tid; c < n; c += NT) {
      const double v = 2.0 * u01(a.seed + pid, c) - 1.0;
      V[c] = v;
      ss[0] += v * v;
    }
    block_sums<NT, 1>(ss, 1, red, coef);
    const double inv = 1.0 / sqrt(coef[0]);
    for (uint32_t c = tid; c < n; c += NT) V[c] *= inv;
  }
  __syncthreads();

  double* F = V + (size_t)m * n;
  uint32_t start = 0, restarts = 0, matvecs = 0;
  long long cyc[5] = {0, 0, 0, 0, 0};
  double bm = 0.0;  // ||f|| after the last Lanczos step
  for (;;) {
    for (uint32_t j = start; j < m; ++j) {
      const long long t0 = clock64();
      for (uint32_t c = tid; c < n; c += NT) vl[c] = V[(size_t)j * n + c];
      __syncthreads();
      covariance_matvec<NT, ROWS, D>(X, n, P.ldx, a.s, vl, wl);
      ++matvecs;
      const long long t1 = clock64();
      cyc[0] += t1 - t0;
      // two classical Gram-Schmidt passes against V[:, 0..j]
      double w[CPT];
#pragma unroll
      for (int q = 0; q < CPT; ++q) {
        const uint32_t c = tid + NT * q;
        w[q] = c < n ? wl[c] : 0.0;
      }
      // classical Gram-Schmidt against V[:, 0..j] with the DGKS correction
      // (ARPACK dsaitr): h = V^T w (||w||^2 in the same reduction),
      // w -= V h; when ||w|| < 0.717 ||w_before||, once more h2 = V^T w,
      // w -= V h2, alpha_j += h2[j]
      // columns are taken four at a time so that their loads are in flight
      // together
      auto project = [&](bool with_norm) {
        for (uint32_t l0 = 0; l0 <= j; l0 += 4) {
          double vv[4][CPT];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t l = l0 + u;
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
              const uint32_t c = tid + NT * q;
              vv[u][q] = (l <= j && c < n) ? V[(size_t)l * n + c] : 0.0;
            }
          }
          double p[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            p[u] = 0;
#pragma unroll
            for (int q = 0; q < CPT; ++q) p[u] += vv[u][q] * w[q];
          }
          // the four sums over the lanes with halves exchanged: partners
          // 32 apart swap two of the four partials, partners 16 apart one of
          // the remaining two, then one value per lane (7 shuffles, not 24);
          // lane 16 u ends with column l0 + u
          const int lane = tid & 63;
          const bool h32 = (lane & 32) != 0, h16 = (lane & 16) != 0;
          double a0 = h32 ? p[2] : p[0], a1 = h32 ? p[3] : p[1];
          a0 += __shfl_xor(h32 ? p[0] : p[2], 32, 64);
          a1 += __shfl_xor(h32 ? p[1] : p[3], 32, 64);
          double k = h16 ? a1 : a0;
          k += __shfl_xor(h16 ? a0 : a1, 16, 64);
#pragma unroll
          for (int off = 8; off > 0; off >>= 1) k += __shfl_xor(k, off, 64);
          if ((lane & 15) == 0 && l0 + (lane >> 4) <= j)
            red[(tid >> 6) * (kMaxMB + 2) + l0 + (lane >> 4)] = k;
        }
        if (with_norm) {
          double p = 0;
#pragma unroll
          for (int q = 0; q < CPT; ++q) p += w[q] * w[q];
          p = wave_sum(p);
          if ((tid & 63) == 0) red[(tid >> 6) * (kMaxMB + 2) + j + 1] = p;
        }
        __syncthreads();
        if ((uint32_t)tid <= j + (with_norm ? 1u : 0u)) {
          double t = 0;
          for (int wv = 0; wv < NT / 64; ++wv) t += red[wv * (kMaxMB + 2) + tid];
          coef[tid] = t;
        }
        __syncthreads();
      };
      auto subtract = [&]() {
        for (uint32_t l0 = 0; l0 <= j; l0 += 4) {
          double vv[4][CPT];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t l = l0 + u;
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
              const uint32_t c = tid + NT * q;
              vv[u][q] = (l <= j && c < n) ? V[(size_t)l * n + c] : 0.0;
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const double cl = l0 + u <= j ? coef[l0 + u] : 0.0;
#pragma unroll
            for (int q = 0; q < CPT; ++q) w[q] -= vv[u][q] * cl;
          }
        }
      };
      project(true);
      double alpha = coef[j];
      const double wnorm = sqrt(coef[j + 1]);
      subtract();
      double ss[1] = {0};
#pragma unroll
      for (int q = 0; q < CPT; ++q) ss[0] += w[q] * w[q];
      block_sums<NT, 1>(ss, 1, red, coef);
      double beta = sqrt(coef[0]);
      if (beta < 0.717 * wnorm) {
        project(false);
        alpha += coef[j];
        subtract();
        ss[0] = 0;
#pragma unroll
        for (int q = 0; q < CPT; ++q) ss[0] += w[q] * w[q];
        block_sums<NT, 1>(ss, 1, red, coef);
        beta = sqrt(coef[0]);
      }
      if (tid == 0) {
        L.alpha[j] = alpha;
        if (j + 1 < m) L.beta[j] = beta;
      }
      bm = beta;
      cyc[1] += clock64() - t1;
      const double inv = beta > 0 ? 1.0 / beta : 0.0;
      double* dst = j + 1 < m ? V + (size_t)(j + 1) * n : F;
      const double sc = j + 1 < m ? inv : 1.0;
#pragma unroll
      for (int q = 0; q < CPT; ++q) {
        const uint32_t c = tid + NT * q;
        if (c < n) dst[c] = w[q] * sc;
      }
      __syncthreads();
    }
    ++restarts;
    const bool last = restarts >= a.max_restarts;
    const long long t2 = clock64();
    if (tid < 64) sym_restart(L, (int)m, (int)k, a.which, a.tol, bm, !last);
    __syncthreads();
    const long long t3 = clock64();
    cyc[2] += t3 - t2;
    if (L.ctl[0] >= (int)k || last) break;
    // V <- V Q[:, 0..kk), f <- (V Q)[:, kk] beta_k + f sigma, v_kk = f / ||f||
    const uint32_t kk = (uint32_t)L.ctl[1];
    const double bk = L.ctld[0], sg = L.ctld[1];
    if (kk == 0 || kk >= m) break;
    double ss[1] = {0};
    for (uint32_t c = tid; c < n; c += NT) {
      double vrow[kMaxMB];
#pragma unroll
      for (int l = 0; l < kMaxMB; ++l) vrow[l] = (uint32_t)l < m ? V[(size_t)l * n + c] : 0.0;
      double fk = 0;
      for (uint32_t t = 0; t <= kk; ++t) {
        double acc = 0;
#pragma unroll
        for (int l = 0; l < kMaxMB; ++l)
          if ((uint32_t)l < m) acc += vrow[l] * L.Z[l * ld + t];
        if (t < kk) V[(size_t)t * n + c] = acc;
        else fk = acc;
      }
      const double f = fk * bk + F[c] * sg;
      F[c] = f;
      ss[0] += f * f;
    }
    block_sums<NT, 1>(ss, 1, red, coef);
    const double beta = sqrt(coef[0]);
    const double inv = beta > 0 ? 1.0 / beta : 0.0;
    for (uint32_t c = tid; c < n; c += NT) V[(size_t)kk * n + c] = F[c] * inv;
    if (tid == 0) L.beta[kk - 1] = beta;
    __syncthreads();
    cyc[3] += clock64() - t3;
    start = kk;
  }
  const long long t4 = clock64();
  // Ritz vectors x_t = V Y[:, t] (normalised), Rayleigh quotients and
  // explicit residuals ||C x - lambda x||_2 in FP64
  double* OUT = a.out + P.out_off;
  Result* R = a.results + pid;
  for (uint32_t t = 0; t < k; ++t) {
    double ss[1] = {0};
    for (uint32_t c = tid; c < n; c += NT) {
      double acc = 0;
      for (uint32_t l = 0; l < m; ++l) acc += V[(size_t)l * n + c] * L.Y[l * k + t];
      vl[c] = acc;
      ss[0] += acc * acc;
    }
    block_sums<NT, 1>(ss, 1, red, coef);
    const double inv = 1.0 / sqrt(coef[0]);
    for (uint32_t c = tid; c < n; c += NT) vl[c] *= inv;
    __syncthreads();
    covariance_matvec<NT, ROWS, D>(X, n, P.ldx, a.s, vl, wl);
    ++matvecs;
    double pr[1] = {0};
    for (uint32_t c = tid; c < n; c += NT) pr[0] += vl[c] * wl[c];
    block_sums<NT, 1>(pr, 1, red, coef);
    const double lam = coef[0];
    double rr[1] = {0};
    for (uint32_t c = tid; c < n; c += NT) {
      const double d = wl[c] - lam * vl[c];
      rr[0] += d * d;
      OUT[(size_t)t * n + c] = vl[c];
    }
    block_sums<NT, 1>(rr, 1, red, coef);
    if (tid == 0) {
      R->ev[t] = lam;
      R->resid[t] = sqrt(coef[0]);
    }
    __syncthreads();
  }
  cyc[4] += clock64() - t4;
  if (tid == 0) {
    R->nconv = L.ctl[0];
    R->restarts = (int)restarts;
    R->matvecs = (int)matvecs;
    for (int i = 0; i < 5; ++i) R->cycles[i] = cyc[i];
  }
}

// C3 synthetic problems, generated in HBM: the same recipe as
// orc_gen_c3_problem (oracle/oracle.c).  One block per problem.
__device__ __forceinline__ double gauss(uint64_t seed, uint64_t idx) {
  double u1 = u01(seed, 2 * idx), u2 = u01(seed, 2 * idx + 1);
  return sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
}
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_synth_c3(const Problem* probs, float* X, uint32_t s,
                                                  uint64_t seed) {
#pragma clang fp contract(off)
  __shared__ double a[3 * kMaxS];
  __shared__ double b[3 * kMaxN];
  __shared__ double norms[6];
  const uint32_t pid = blockIdx.x;
  const Problem P = probs[pid];
  const uint32_t n = P.n;
  const uint64_t si = mix64(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(pid + 1));
  for (uint32_t i = threadIdx.x; i < 3 * s; i += blockDim.x) a[i] = gauss(si ^ 0x2ULL, i);
  for (uint32_t i = threadIdx.x; i < 3 * n; i += blockDim.x) {
    uint32_t c = i / n, j = i - c * n;
    b[i] = gauss(si ^ 0x3ULL, (uint64_t)c * n + j);
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int c = threadIdx.x % 3;
    const bool isa = threadIdx.x < 3;
    const uint32_t len = isa ? s : n;
    const double* p = isa ? a + c * s : b + c * n;
    double t = 0;
    for (uint32_t i = 0; i < len; ++i) t += p[i] * p[i];
    norms[threadIdx.x] = 1.0 / sqrt(t);
  }
  __syncthreads();
  float* Xp = X + P.x_off;
  for (uint32_t idx = threadIdx.x; idx < s * P.ldx; idx += blockDim.x) {
    const uint32_t r = idx / P.ldx, j = idx - r * P.ldx;
    if (j >= n) {
      Xp[idx] = 0.0f;
      continue;
    }
    double v = floor(65535.0 * u01(si ^ 0x1ULL, (uint64_t)r * n + j) + 0.5) / 65535.0;
    v += 5.0 * ((a[r] * norms[0]) * (b[j] * norms[3]) + (a[s + r] * norms[1]) * (b[n + j] * norms[4]) +
                (a[2 * s + r] * norms[2]) * (b[2 * n + j] * norms[5]));
    Xp[idx] = (float)v;
  }
}

}  // namespace
}  // namespace batched
}  // namespace oaph

using namespace oaph;
using namespace oaph::batched;

namespace {

// Launch shapes the autotuner chooses between: threads per workgroup, rows
// of X each wave streams per step and the depth of the row ring (bytes in
// flight vs registers).
struct Variant {
  int threads, rows, depth;
  void (*kernel)(SolveArgs);
};
const Variant kVariants[] = {
    {256, 1, 2, k_batched_solve<256, 1>},
    {128, 1, 2, k_batched_solve<128, 1>},
    {512, 1, 2, k_batched_solve<512, 1>},
    {256, 2, 2, k_batched_solve<256, 2>},
    {512, 2, 2, k_batched_solve<512, 2>},
    {1024, 1, 2, k_batched_solve<1024, 1>},
    {128, 2, 2, k_batched_solve<128, 2>},
    {256, 1, 3, k_batched_solve<256, 1, 3>},
    {256, 1, 4, k_batched_solve<256, 1, 4>},
    {256, 2, 3, k_batched_solve<256, 2, 3>},
    {512, 1, 4, k_batched_solve<512, 1, 4>},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

}  // namespace

struct oaph_batch {
  uint32_t count = 0, s = 0;
  std::vector<uint32_t> n;
  std::vector<Problem> probs;
  std::vector<uint32_t> order;
  DevBuf dprobs, dorder, X, V, out, results;
  uint32_t m = 0, k = 0;
  int variant = 1;
  hipStream_t stream = nullptr;
  std::vector<Result> host_results;
  float last_ms = 0;
  // the solve launch captured once and replayed while its arguments hold
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  SolveArgs graph_args{};
  int graph_variant = -1;
  bool use_graph = true;
};

static void drop_graph(oaph_batch* b) {
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  b->graph_variant = -1;
}

static bool same_args(const SolveArgs& x, const SolveArgs& y) {
  return x.probs == y.probs && x.order == y.order && x.results == y.results && x.X == y.X &&
         x.V == y.V && x.out == y.out &&
         x.s == y.s && x.m == y.m && x.k == y.k && x.which == y.which && x.tol == y.tol &&
         x.max_restarts == y.max_restarts && x.seed == y.seed;
}

static oaph_status upload_problems(oaph_batch* b) {
  OAPH_TRY(b->dprobs.reserve(b->count * sizeof(Problem)));
  OAPH_HIP(hipMemcpy(b->dprobs.ptr, b->probs.data(), b->count * sizeof(Problem),
                     hipMemcpyHostToDevice));
  return OAPH_OK;
}

// Lays out V, H and outputs for (m, k) and fills the launch arguments.
static oaph_status prepare(oaph_batch* b, uint32_t hdim, uint32_t wanted, int sort_type,
                           double tol, uint32_t max_restarts, uint64_t seed, SolveArgs* a) {
  OAPH_CHECK(wanted >= 1 && hdim > wanted && tol > 0 && max_restarts >= 1, OAPH_ERR_ARG);
  OAPH_CHECK(hdim <= (uint32_t)kMaxMB, OAPH_ERR_UNSUPPORTED);
  OAPH_CHECK(sort_type >= 0 && sort_type <= 5, OAPH_ERR_ARG);
  for (uint32_t i = 0; i < b->count; ++i) OAPH_CHECK(b->n[i] >= hdim, OAPH_ERR_SHAPE);
  const uint32_t m = hdim, k = wanted;
  if (m != b->m || k != b->k) {
    uint64_t vo = 0, oo = 0;
    for (uint32_t i = 0; i < b->count; ++i) {
      Problem& p = b->probs[i];
      p.v_off = vo;
      vo += (uint64_t)(m + 1) * p.n;
      p.out_off = oo;
      oo += (uint64_t)k * p.n;
    }
    drop_graph(b);
    OAPH_TRY(upload_problems(b));
    OAPH_TRY(b->V.reserve(vo * sizeof(double)));
    OAPH_TRY(b->out.reserve(oo * sizeof(double)));
    OAPH_TRY(b->results.reserve(b->count * sizeof(Result)));
    b->m = m;
    b->k = k;
  }
  a->probs = b->dprobs.as<Problem>();
  a->order = b->dorder.as<uint32_t>();
  a->results = b->results.as<Result>();
  a->X = b->X.as<float>();
  a->V = b->V.as<double>();
  a->out = b->out.as<double>();
  a->s = b->s;
  a->m = m;
  a->k = k;
  a->which = sort_type;
  a->tol = tol;
  a->max_restarts = max_restarts;
  a->seed = seed;
  return OAPH_OK;
}

static oaph_status launch(oaph_batch* b, const SolveArgs& a, int variant) {
  const Variant& v = kVariants[variant];
  if (!b->use_graph) {
    hipLaunchKernelGGL(v.kernel, dim3(b->count), dim3(v.threads), 0, b->stream, a);
    OAPH_HIP(hipGetLastError());
    return OAPH_OK;
  }
  if (!b->graph_exec || b->graph_variant != variant || !same_args(a, b->graph_args)) {
    drop_graph(b);
    OAPH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(v.kernel, dim3(b->count), dim3(v.threads), 0, b->stream, a);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(b->stream, &b->graph);
    OAPH_HIP(le);
    OAPH_HIP(ce);
    OAPH_HIP(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    b->graph_args = a;
    b->graph_variant = variant;
  }
  OAPH_HIP(hipGraphLaunch(b->graph_exec, b->stream));
  return OAPH_OK;
}

// One timed solve with the given variant; results copied to the host.
static oaph_status timed_solve(oaph_batch* b, const SolveArgs& a, int variant) {
  hipEvent_t e0, e1;
  OAPH_HIP(hipEventCreate(&e0));
  OAPH_HIP(hipEventCreate(&e1));
  oaph_status st = OAPH_OK;
  if (hipEventRecord(e0, b->stream) != hipSuccess) st = OAPH_ERR_HIP;
  if (st == OAPH_OK) st = launch(b, a, variant);
  if (st == OAPH_OK && hipEventRecord(e1, b->stream) != hipSuccess) st = OAPH_ERR_HIP;
  if (st == OAPH_OK && hipEventSynchronize(e1) != hipSuccess) st = OAPH_ERR_HIP;
  if (st == OAPH_OK && hipEventElapsedTime(&b->last_ms, e0, e1) != hipSuccess) st = OAPH_ERR_HIP;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  OAPH_TRY(st);
  b->host_results.resize(b->count);
  OAPH_HIP(hipMemcpy(b->host_results.data(), b->results.ptr, b->count * sizeof(Result),
                     hipMemcpyDeviceToHost));
  return OAPH_OK;
}

extern "C" {

oaph_status oaph_batch_c3_sizes(uint32_t count, uint64_t seed, uint32_t* n_out) {
  OAPH_CHECK(n_out, OAPH_ERR_ARG);
  for (uint32_t i = 0; i < count; ++i) {
    // n = 64 + floor(961 u), u the first uniform of problem i's stream
    const uint64_t si = mix64(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(i + 1));
    const uint64_t w = mix64(si + 0x9E3779B97F4A7C15ULL);
    const double u = (double)(w >> 11) * (1.0 / 9007199254740992.0);
    n_out[i] = 64 + (uint32_t)(961.0 * u);
  }
  return OAPH_OK;
}

oaph_status oaph_batch_create(oaph_batch** out, uint32_t count, const uint32_t* n, uint32_t s) {
  OAPH_CHECK(out && n && count > 0, OAPH_ERR_ARG);
  OAPH_CHECK(s > 0 && s <= (uint32_t)kMaxS, OAPH_ERR_UNSUPPORTED);
  for (uint32_t i = 0; i < count; ++i)
    OAPH_CHECK(n[i] >= 2 && n[i] <= (uint32_t)kMaxN, OAPH_ERR_UNSUPPORTED);
  OAPH_TRY(ensure_init());
  oaph_batch* b = new oaph_batch();
  b->count = count;
  b->s = s;
  b->n.assign(n, n + count);
  b->probs.resize(count);
  uint64_t xo = 0;
  for (uint32_t i = 0; i < count; ++i) {
    Problem& p = b->probs[i];
    p = Problem{};
    p.n = n[i];
    p.ldx = (n[i] + 3) / 4 * 4;
    p.x_off = xo;
    xo += (uint64_t)s * p.ldx;
  }
  b->order.resize(count);
  std::iota(b->order.begin(), b->order.end(), 0u);
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](uint32_t x, uint32_t y) { return n[x] > n[y]; });
  oaph_status st = b->X.reserve(xo * sizeof(float));
  if (st == OAPH_OK) st = b->dorder.reserve(count * sizeof(uint32_t));
  if (st == OAPH_OK && hipMemset(b->X.ptr, 0, xo * sizeof(float)) != hipSuccess) st = OAPH_ERR_HIP;
  if (st == OAPH_OK &&
      hipMemcpy(b->dorder.ptr, b->order.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice) !=
          hipSuccess)
    st = OAPH_ERR_HIP;
  if (st == OAPH_OK) st = upload_problems(b);
  if (st == OAPH_OK && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    st = OAPH_ERR_HIP;
  if (st != OAPH_OK) {
    oaph_batch_destroy(b);
    return st;
  }
  *out = b;
  return OAPH_OK;
}

oaph_status oaph_batch_destroy(oaph_batch* b) {
  if (!b) return OAPH_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (DevBuf* d : {&b->dprobs, &b->dorder, &b->X, &b->V, &b->out, &b->results})
    d->release();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return OAPH_OK;
}

oaph_status oaph_batch_upload_x(oaph_batch* b, uint32_t i, const float* X) {
  OAPH_CHECK(b && X && i < b->count, OAPH_ERR_ARG);
  const Problem& p = b->probs[i];
  OAPH_HIP(hipMemcpy2D(b->X.as<float>() + p.x_off, p.ldx * sizeof(float), X, p.n * sizeof(float),
                       p.n * sizeof(float), b->s, hipMemcpyHostToDevice));
  return OAPH_OK;
}

oaph_status oaph_batch_download_x(oaph_batch* b, uint32_t i, float* X) {
  OAPH_CHECK(b && X && i < b->count, OAPH_ERR_ARG);
  const Problem& p = b->probs[i];
  OAPH_HIP(hipMemcpy2D(X, p.n * sizeof(float), b->X.as<float>() + p.x_off, p.ldx * sizeof(float),
                       p.n * sizeof(float), b->s, hipMemcpyDeviceToHost));
  return OAPH_OK;
}

oaph_status oaph_batch_synth_c3(oaph_batch* b, uint64_t seed) {
  OAPH_CHECK(b, OAPH_ERR_ARG);
  hipLaunchKernelGGL(k_synth_c3, dim3(b->count), dim3(256), 0, b->stream,
                     b->dprobs.as<Problem>(), b->X.as<float>(), b->s, seed);
  OAPH_HIP(hipGetLastError());
  OAPH_HIP(hipStreamSynchronize(b->stream));
  return OAPH_OK;
}

oaph_status oaph_batch_set_variant(oaph_batch* b, int variant) {
  OAPH_CHECK(b && variant >= 0 && variant < kNumVariants, OAPH_ERR_ARG);
  b->variant = variant;
  return OAPH_OK;
}

oaph_status oaph_batch_variant_info(int variant, int* threads, int* rows) {
  return oaph_batch_variant_shape(variant, threads, rows, nullptr);
}

oaph_status oaph_batch_variant_shape(int variant, int* threads, int* rows, int* depth) {
  OAPH_CHECK(variant >= 0 && variant < kNumVariants, OAPH_ERR_ARG);
  if (threads) *threads = kVariants[variant].threads;
  if (rows) *rows = kVariants[variant].rows;
  if (depth) *depth = kVariants[variant].depth;
  return OAPH_OK;
}

oaph_status oaph_batch_set_graph(oaph_batch* b, int enable) {
  OAPH_CHECK(b, OAPH_ERR_ARG);
  b->use_graph = enable != 0;
  if (!b->use_graph) drop_graph(b);
  return OAPH_OK;
}

oaph_status oaph_batch_solve(oaph_batch* b, uint32_t hdim, uint32_t wanted, int sort_type,
                             double tol, uint32_t max_restarts, uint64_t seed) {
  OAPH_CHECK(b, OAPH_ERR_ARG);
  SolveArgs a;
  OAPH_TRY(prepare(b, hdim, wanted, sort_type, tol, max_restarts, seed, &a));
  return timed_solve(b, a, b->variant);
}

// Times every launch variant on this batch (one warm-up and `reps` timed
// solves each, minimum kept), selects the fastest, and records it in a
// plain-text cache keyed by the batch's shape so that a later process skips
// the search.  cache_path may be NULL (no cache).
oaph_status oaph_batch_autotune(oaph_batch* b, uint32_t hdim, uint32_t wanted, int sort_type,
                                double tol, uint32_t max_restarts, uint64_t seed, uint32_t reps,
                                const char* cache_path, int* chosen, double* times_ms) {
  OAPH_CHECK(b && reps >= 1, OAPH_ERR_ARG);
  SolveArgs a;
  OAPH_TRY(prepare(b, hdim, wanted, sort_type, tol, max_restarts, seed, &a));
  hipDeviceProp_t prop;
  int dev = 0;
  OAPH_HIP(hipGetDevice(&dev));
  OAPH_HIP(hipGetDeviceProperties(&prop, dev));
  uint64_t nsum = 0, nsq = 0;
  for (uint32_t x : b->n) {
    nsum += x;
    nsq += (uint64_t)x * x;
  }
  char key[256];
  snprintf(key, sizeof key, "%s:count=%u:s=%u:nsum=%llu:nsq=%llu:m=%u:k=%u", prop.gcnArchName,
           b->count, b->s, (unsigned long long)nsum, (unsigned long long)nsq, hdim, wanted);
  for (char* c = key; *c; ++c)
    if (*c == ' ') *c = '_';
  if (cache_path && !times_ms) {
    if (FILE* f = fopen(cache_path, "r")) {
      char k2[256];
      int v = -1;
      double ms = 0;
      int found = -1;
      while (fscanf(f, "%255s %d %lf", k2, &v, &ms) == 3)
        if (strcmp(k2, key) == 0 && v >= 0 && v < kNumVariants) found = v;
      fclose(f);
      if (found >= 0) {
        b->variant = found;
        if (chosen) *chosen = found;
        return OAPH_OK;
      }
    }
  }
  int best = -1;
  double best_ms = 0;
  for (int v = 0; v < kNumVariants; ++v) {
    OAPH_TRY(timed_solve(b, a, v));  // warm-up (and graph capture)
    double t = 1e30;
    for (uint32_t r = 0; r < reps; ++r) {
      OAPH_TRY(timed_solve(b, a, v));
      t = std::min(t, (double)b->last_ms);
    }
    if (times_ms) times_ms[v] = t;
    if (best < 0 || t < best_ms) {
      best = v;
      best_ms = t;
    }
  }
  b->variant = best;
  if (chosen) *chosen = best;
  if (cache_path) {
    if (FILE* f = fopen(cache_path, "a")) {
      fprintf(f, "%s %d %.6f\n", key, best, best_ms);
      fclose(f);
    }
  }
  return OAPH_OK;
}

oaph_status oaph_batch_results(oaph_batch* b, double* evals, double* resid, uint32_t* nconv,
                               uint32_t* restarts, uint32_t* matvecs, double* kernel_ms) {
  OAPH_CHECK(b, OAPH_ERR_ARG);
  OAPH_CHECK(b->host_results.size() == b->count, OAPH_ERR_STATE);
  for (uint32_t i = 0; i < b->count; ++i) {
    const Result& r = b->host_results[i];
    for (uint32_t t = 0; t < b->k; ++t) {
      if (evals) evals[(size_t)i * b->k + t] = r.ev[t];
      if (resid) resid[(size_t)i * b->k + t] = r.resid[t];
    }
    if (nconv) nconv[i] = (uint32_t)r.nconv;
    if (restarts) restarts[i] = (uint32_t)r.restarts;
    if (matvecs) matvecs[i] = (uint32_t)r.matvecs;
  }
  if (kernel_ms) *kernel_ms = b->last_ms;
  return OAPH_OK;
}

oaph_status oaph_batch_phase_cycles(oaph_batch* b, long long* cycles) {
  OAPH_CHECK(b && cycles, OAPH_ERR_ARG);
  OAPH_CHECK(b->host_results.size() == b->count, OAPH_ERR_STATE);
  for (uint32_t i = 0; i < b->count; ++i)
    for (int p = 0; p < 5; ++p) cycles[(size_t)i * 5 + p] = b->host_results[i].cycles[p];
  return OAPH_OK;
}

oaph_status oaph_batch_eigenvectors(oaph_batch* b, uint32_t i, double* x) {
  OAPH_CHECK(b && x && i < b->count, OAPH_ERR_ARG);
  OAPH_CHECK(b->host_results.size() == b->count, OAPH_ERR_STATE);
  const Problem& p = b->probs[i];
  OAPH_HIP(hipMemcpy(x, b->out.as<double>() + p.out_off, (size_t)b->k * p.n * sizeof(double),
                     hipMemcpyDeviceToHost));
  return OAPH_OK;
}

}  // extern "C"
