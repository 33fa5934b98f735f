// small_dense.h -- device code for the small dense part of the Arnoldi
// restart (complex Schur form by shifted QR, eigenvectors, wanted ordering,
// exact-shift sweeps) of the single-problem restart kernel (k_restart; the
// batched solver's symmetric restart has its own code).  Same iteration as
// orc_hess_schur / orc_iram in oracle/oracle.c.  Every function is called
// by all threads of the block:
// rotations split their row/column updates over the threads, barriers
// separate the row and column halves of each rotation.
#pragma once

#include "iram.h"

namespace oaph {
namespace iram {
namespace {

struct cd {
  double re, im;
};
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ double cabs2(cd a) { return a.re * a.re + a.im * a.im; }

// Block-wide ordering point.  A one-wave block needs no hardware barrier:
// the LDS executes a wave's instructions in issue order, and the compiler
// keeps possibly-aliasing LDS accesses in program order (the fence stops it
// from moving memory operations across this point).
__device__ __forceinline__ void bsync() {
  if (blockDim.x > 64) {
    __syncthreads();
  } else {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  }
}

// G = [[c, s], [-conj s, c]] with real c, G [f; g] = [r; 0] (LAPACK zlartg
// convention).  Reciprocal square roots on the common path; hypot-based
// scaling when a modulus is near the under/overflow thresholds.
__device__ void cgivens_safe(cd f, cd g, double& c, cd& s);
// One reciprocal square root on the common path: with t = 1 / sqrt(|f|^2 r^2),
// r^2 = |f|^2 + |g|^2, c = |f|^2 t = |f| / r and s = f conj(g) t, which keeps
// the dependent chain short (FP64 results feed the next instruction only
// after tens of cycles on one wave).
__device__ __forceinline__ void cgivens(cd f, cd g, double& c, cd& s) {
  const double af2 = f.re * f.re + f.im * f.im, ag2 = g.re * g.re + g.im * g.im;
  const double pr = f.re * g.re + f.im * g.im, pi = f.im * g.re - f.re * g.im;  // f conj(g)
  if (ag2 == 0.0) { c = 1; s = {0, 0}; return; }
  if (!(af2 > 1e-150 && ag2 > 1e-150 && af2 < 1e150 && ag2 < 1e150)) {
    cgivens_safe(f, g, c, s);
    return;
  }
  const double t = rsqrt(af2 * (af2 + ag2));
  c = af2 * t;
  s = {pr * t, pi * t};
}
__device__ void cgivens_safe(cd f, cd g, double& c, cd& s) {
  double af = hypot(f.re, f.im), ag = hypot(g.re, g.im);
  if (ag == 0) { c = 1; s = {0, 0}; return; }
  if (af == 0) { c = 0; s = {g.re / ag, -g.im / ag}; return; }
  double rho = hypot(af, ag);
  c = af / rho;
  double ur = f.re / af, ui = f.im / af;
  s = {(ur * g.re + ui * g.im) / rho, (ui * g.re - ur * g.im) / rho};
}

// LDS-resident m x m complex matrix, row-major
struct Mat {
  double* re;
  double* im;
  int ld;
  __device__ cd get(int r, int c) const { return {re[r * ld + c], im[r * ld + c]}; }
  __device__ void set(int r, int c, cd v) { re[r * ld + c] = v.re; im[r * ld + c] = v.im; }
};

// rows i, i+1 <- G rows, columns j0..j1 (lanes split the columns)
__device__ void rot_rows(Mat M, int i, int j0, int j1, double c, cd s) {
  for (int j = j0 + (int)threadIdx.x; j <= j1; j += blockDim.x) {
    cd a = M.get(i, j), b = M.get(i + 1, j);
    cd na = {c * a.re + (s.re * b.re - s.im * b.im), c * a.im + (s.re * b.im + s.im * b.re)};
    cd nb = {-(s.re * a.re + s.im * a.im) + c * b.re, -(s.re * a.im - s.im * a.re) + c * b.im};
    M.set(i, j, na);
    M.set(i + 1, j, nb);
  }
}
// columns j, j+1 <- columns G^H, rows r0..r1
__device__ void rot_cols(Mat M, int j, int r0, int r1, double c, cd s) {
  for (int r = r0 + (int)threadIdx.x; r <= r1; r += blockDim.x) {
    cd a = M.get(r, j), b = M.get(r, j + 1);
    cd na = {a.re * c + (b.re * s.re + b.im * s.im), a.im * c + (b.im * s.re - b.re * s.im)};
    cd nb = {-(a.re * s.re - a.im * s.im) + b.re * c, -(a.re * s.im + a.im * s.re) + b.im * c};
    M.set(r, j, na);
    M.set(r, j + 1, nb);
  }
}

// Rotation log of a sweep (rotation kk acts on rows/columns kk, kk+1).
struct RotLog {
  double* c;
  cd* s;
};

// Single-wave ordering point: LDS operations of one wave complete in issue
// order; the fences keep the compiler from moving them across this point.
__device__ __forceinline__ void wsync() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ double lane_read(double v, int l) {
  const long long bits = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ cd lane_read(cd v, int l) { return {lane_read(v.re, l), lane_read(v.im, l)}; }

// row pair (a above b) <- G (a; b), G = [[c, s], [-conj s, c]]  (rot_rows)
__device__ __forceinline__ cd grow_top(cd a, cd b, double c, cd s) {
  return {c * a.re + (s.re * b.re - s.im * b.im), c * a.im + (s.re * b.im + s.im * b.re)};
}
__device__ __forceinline__ cd grow_bot(cd a, cd b, double c, cd s) {
  return {-(s.re * a.re + s.im * a.im) + c * b.re, -(s.re * a.im - s.im * a.re) + c * b.im};
}
// column pair (a left of b) <- (a, b) G^H  (rot_cols)
__device__ __forceinline__ cd gcol_left(cd a, cd b, double c, cd s) {
  return {a.re * c + (b.re * s.re + b.im * s.im), a.im * c + (b.im * s.re - b.re * s.im)};
}
__device__ __forceinline__ cd gcol_right(cd a, cd b, double c, cd s) {
  return {-(a.re * s.re - a.im * s.im) + b.re * c, -(a.re * s.im + a.im * s.re) + b.im * c};
}

// One bulge step's row phase: the rotation from (f, g) (cgivens) and the
// rotated row pair G (a; b).  With c = |f|^2 t and s = f conj(g) t, the pair
// is t (|f|^2 a + f conj(g) b, -conj(f) g a + |f|^2 b): the products that
// need only f, g, a and b are formed while the reciprocal square root is in
// flight, and t scales them last (one dependent multiply after t instead of
// the rotation's three).  The range test that picks the hypot-based
// rotation runs beside it; the fallback recomputes everything.
// givens_prep forms the parts that need only (f, g), so that a caller can
// place them before the first use of row data still being loaded.
struct GivensPrep {
  double af2, ag2, pr, pi, x;
};
__device__ __forceinline__ GivensPrep givens_prep(cd f, cd g) {
  GivensPrep q;
  q.af2 = f.re * f.re + f.im * f.im;
  q.ag2 = g.re * g.re + g.im * g.im;
  q.pr = f.re * g.re + f.im * g.im;  // f conj(g)
  q.pi = f.im * g.re - f.re * g.im;
  q.x = q.af2 * (q.af2 + q.ag2);
  return q;
}
__device__ __forceinline__ void givens_rows(const GivensPrep& q, cd f, cd g, cd a, cd b,
                                            double& c, cd& s, cd& na, cd& nb) {
  const double af2 = q.af2, ag2 = q.ag2, pr = q.pr, pi = q.pi;
  const cd ua = {af2 * a.re + (pr * b.re - pi * b.im), af2 * a.im + (pr * b.im + pi * b.re)};
  const cd ub = {af2 * b.re - (pr * a.re + pi * a.im), af2 * b.im - (pr * a.im - pi * a.re)};
  const double t = rsqrt(q.x);
  c = af2 * t;
  s = {pr * t, pi * t};
  na = {ua.re * t, ua.im * t};
  nb = {ub.re * t, ub.im * t};
  if (!(af2 > 1e-150 && af2 < 1e150 && ag2 < 1e150)) {
    if (ag2 == 0.0) {
      c = 1.0;
      s = {0.0, 0.0};
    } else {
      cgivens_safe(f, g, c, s);
    }
    na = grow_top(a, b, c, s);
    nb = grow_bot(a, b, c, s);
  }
}

constexpr int kDefer = 8;  // rotations per deferred batch

// The column halves of rotations [k1, k2) that no later rotation of the
// sweep reads: rows r < kk of T (row r receives rotations max(k1, r+1) ..
// k2-1 in order) and every row of Q.  One row per worker, carried in
// registers across the batch.
__device__ void chase_deferred(Mat T, Mat Q, int kdim, int k1, int k2, const double* lc,
                               const cd* ls, int worker, int nworkers) {
  const int nt = k2 - 1 > 0 ? k2 - 1 : 0;  // T rows 0 .. k2-2
  for (int task = worker; task < nt + kdim; task += nworkers) {
    const bool isq = task >= nt;
    Mat M = isq ? Q : T;
    const int r = isq ? task - nt : task;
    const int kb = isq ? k1 : (r + 1 > k1 ? r + 1 : k1);
    if (kb >= k2) continue;
    cd a = M.get(r, kb);
    for (int kk = kb; kk < k2; ++kk) {
      const cd b = M.get(r, kk + 1);
      const double c = lc[kk];
      const cd s = ls[kk];
      M.set(r, kk, gcol_left(a, b, c, s));
      a = gcol_right(a, b, c, s);
    }
    M.set(r, k2, a);
  }
}

// One implicit single-shift QR step with shift mu on rows/columns [lo, hi] of
// T (kdim <= 64), accumulating Q on the right: rotation kk (lo <= kk < hi) is
// computed from (T[kk][kk-1], T[kk+1][kk-1]) (from (T[lo][lo] - mu,
// T[lo+1][lo]) for kk = lo), applied to rows kk, kk+1 over columns >= kk-1
// (>= kk at lo), to columns kk, kk+1 over rows <= min(kk+2, hi) of T and all
// rows of Q, and the bulge entry T[kk+1][kk-1] is zeroed.  Same steps as
// shifted_sweep in oracle/oracle.c.
//
// Critical path in registers: wave 0 holds rows kk and kk+1 with lane j =
// column j, prefetches row kk+2 from LDS, rotates the row pair in place,
// applies the column pair to rows kk .. kk+2 with lane reads of the four or
// six entries involved, hands the next (f, g) on in registers and writes row
// kk back once it is final.  The rest of each column rotation (rows < kk of
// T, all of Q) is applied by the other waves in batches of kDefer rotations,
// one batch behind (disjoint rows); a one-wave block applies each batch
// itself.  The block is split at negligible subdiagonals by the caller
// (shifted_sweep): a sweep started at the top of a reducible T stays the
// identity.
__device__ void chase(Mat T, Mat Q, int kdim, int lo, int hi, cd mu, RotLog log) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool solo = blockDim.x <= 64;
  const int j = lane;
  cd a = {0.0, 0.0}, b = {0.0, 0.0}, f = {0.0, 0.0}, g = {0.0, 0.0};
  if (wave == 0) {
    if (j < kdim) {
      a = T.get(lo, j);
      b = T.get(lo + 1, j);
    }
    f = csub(lane_read(a, lo), mu);
    g = lane_read(b, lo);
  }
  for (int kb = lo;; kb += kDefer) {
    const bool crit = kb < hi;
    const int ke = crit ? (kb + kDefer < hi ? kb + kDefer : hi) : hi;
    if (wave == 0 && crit) {
      int j0 = lo;
      for (int kk = kb; kk < ke; ++kk) {
        const bool have2 = kk + 2 <= hi;
        cd nxt = {0.0, 0.0};
        if (have2 && j < kdim && j >= kk) nxt = T.get(kk + 2, j);
        double c;
        cd s;
        cgivens(f, g, c, s);
        if (lane == 0) {
          log.c[kk] = c;
          log.s[kk] = s;
        }
        j0 = kk > lo ? kk - 1 : kk;
        cd na = a, nb = b;
        if (j >= j0 && j < kdim) {
          na = grow_top(a, b, c, s);
          nb = grow_bot(a, b, c, s);
        }
        if (kk > lo && j == kk - 1) nb = cd{0.0, 0.0};
        const cd t00 = lane_read(na, kk), t01 = lane_read(na, kk + 1);
        const cd t10 = lane_read(nb, kk), t11 = lane_read(nb, kk + 1);
        cd u20 = {0.0, 0.0}, u21 = {0.0, 0.0};
        if (have2) {
          const cd t21 = lane_read(nxt, kk + 1);
          u20 = gcol_left(cd{0.0, 0.0}, t21, c, s);
          u21 = gcol_right(cd{0.0, 0.0}, t21, c, s);
        }
        if (j == kk) {
          na = gcol_left(t00, t01, c, s);
          nb = gcol_left(t10, t11, c, s);
          nxt = u20;
        } else if (j == kk + 1) {
          na = gcol_right(t00, t01, c, s);
          nb = gcol_right(t10, t11, c, s);
          nxt = u21;
        }
        // row kk is final for the critical path
        if (j >= j0 && j < kdim) T.set(kk, j, na);
        if (kk > lo && j == kk - 1) T.set(kk + 1, kk - 1, cd{0.0, 0.0});
        f = gcol_left(t10, t11, c, s);
        g = u20;
        a = nb;
        b = nxt;
      }
      if (ke == hi && j >= j0 && j < kdim) T.set(hi, j, a);  // the last row of the block
    }
    if (solo) {
      if (crit) {
        wsync();
        chase_deferred(T, Q, kdim, kb, ke, log.c, log.s, lane, 64);
        wsync();
      }
    } else {
      if (wave >= 1 && kb > lo)
        chase_deferred(T, Q, kdim, kb - kDefer, kb < hi ? kb : hi, log.c, log.s, tid - 64,
                       (int)blockDim.x - 64);
      __syncthreads();
    }
    if (!crit) break;
  }
}

// A train of ns (<= waves of the block) implicit single-shift QR sweeps over
// rows/columns [lo, hi] of T (kdim <= 64), accumulating Q: bulge w, with
// shift mu (each thread passes the shift of its own wave), is chased by wave
// w, three positions behind bulge w - 1.  Every step has a row phase (each
// active wave computes its rotation from the bulge column and rotates its row
// pair, lanes over the columns) and a column phase (its column pair of T over
// rows <= min(k + 2, hi) and of Q, lanes over the rows), separated by block
// barriers.  Row pairs of different bulges are disjoint, and so are column
// pairs; against ns consecutive chase() sweeps the only reordering is a row
// rotation of bulge w before a column rotation of bulge w - 1 on the same
// entries, and left and right products commute.  Same result as the ns
// sweeps up to rounding, with ns rotations per step instead of one.
__device__ void chase_train(Mat T, Mat Q, int kdim, int lo, int hi, int ns, cd mu) {
  const int lane = threadIdx.x & 63;
  // the wave index in a scalar register: the per-step conditions below are
  // then scalar branches
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int L = hi - lo;
  const int steps = L + 3 * (ns - 1);
  const int jc = lane < kdim ? lane : kdim - 1;  // clamped lane index: loads stay in the matrix
  // (f, g) of this bulge's next step: column k of rows k + 1 and k + 2 after
  // its column phase, kept in registers (no other bulge writes them before
  // that step), so the next rotation starts without waiting for LDS
  cd fn = {0.0, 0.0}, gn = {0.0, 0.0};
  for (int t = 0; t < steps; ++t) {
    const int p = t - 3 * wave;
    const bool act = wave < ns && p >= 0 && p < L;
    const int k = lo + p;
    double c = 1.0;
    cd s = {0.0, 0.0}, qx = {0.0, 0.0}, qy = {0.0, 0.0};
    if (act) {
      const int j = lane, j0 = k > lo ? k - 1 : k;
      const bool mine = j >= j0 && j < kdim;
      // (f, g) = column j0 of the row pair: the bulge column, or the shifted
      // leading column at the top of the block
      if (k == lo) {
        fn = csub(T.get(lo, lo), mu);
        gn = T.get(lo + 1, lo);
      }
      const cd f = fn, g = gn;
      // every lane loads (clamped), so the LDS wait count after these loads
      // does not depend on which lanes are active
      const cd a = T.get(k, jc), b = T.get(k + 1, jc);
      // Q's column pair is written only by this wave's column phases
      qx = Q.get(jc, k);
      qy = Q.get(jc, k + 1);
      // the rotation's first products while the row loads are in flight
      const GivensPrep gp = givens_prep(f, g);
      __builtin_amdgcn_sched_barrier(0);
      cd na, nb;
      givens_rows(gp, f, g, a, b, c, s, na, nb);
      if (mine) {
        T.set(k, j, na);
        T.set(k + 1, j, (k > lo && j == k - 1) ? cd{0.0, 0.0} : nb);
      }
    }
    __syncthreads();
    if (act) {
      const int r = lane, rmax = k + 2 < hi ? k + 2 : hi;
      const int rc = r <= rmax ? r : rmax;
      const cd x = T.get(rc, k), y = T.get(rc, k + 1);
      const cd nx = gcol_left(x, y, c, s);
      if (r <= rmax) {
        T.set(r, k, nx);
        T.set(r, k + 1, gcol_right(x, y, c, s));
      }
      if (k + 1 < hi) {
        fn = lane_read(nx, k + 1);
        gn = lane_read(nx, k + 2);
      }
      if (r < kdim) {
        Q.set(r, k, gcol_left(qx, qy, c, s));
        Q.set(r, k + 1, gcol_right(qx, qy, c, s));
      }
    }
    __syncthreads();
  }
}

// Bit i (1 <= i < kdim) set when the subdiagonal entry (i, i-1) is
// negligible: |t_{i,i-1}| <= eps (|t_{i-1,i-1}| + |t_{i,i}|).  Lane i tests
// entry i; every wave computes the same mask.
__device__ __forceinline__ unsigned long long negligible_mask(Mat T, int kdim) {
  const double eps = 2.220446049250313e-16;
  const int i = threadIdx.x & 63;
  bool neg = false;
  if (i >= 1 && i < kdim) {
    cd a = T.get(i - 1, i - 1), b = T.get(i, i), sd = T.get(i, i - 1);
    double d = fabs(a.re) + fabs(a.im) + fabs(b.re) + fabs(b.im);
    double sub = fabs(sd.re) + fabs(sd.im);
    neg = sub <= eps * d || sub < 1e-300;
  }
  return __ballot(neg);
}

__device__ void shifted_sweep(Mat T, Mat Q, int kdim, cd mu, RotLog log) {
  const unsigned long long mask = negligible_mask(T, kdim);
  bsync();
  {
    const int i = threadIdx.x;
    if (i < 64 && ((mask >> i) & 1ull)) T.set(i, i - 1, {0, 0});
  }
  bsync();
  int lo = 0;
  while (lo < kdim) {
    // the block [lo, hi] ends before the next negligible subdiagonal above lo
    const unsigned long long above = lo >= 63 ? 0ull : mask & ~((2ull << lo) - 1ull);
    const int hi = above ? __builtin_ctzll(above) - 1 : kdim - 1;
    if (hi > lo) chase(T, Q, kdim, lo, hi, mu, log);
    lo = hi + 1;
  }
}

// ns (<= waves) exact-shift sweeps as one train per unreduced block (blocks
// split at the negligible subdiagonals found before the train).
__device__ void shifted_train(Mat T, Mat Q, int kdim, int ns, const cd* mus) {
  const unsigned long long mask = negligible_mask(T, kdim);
  __syncthreads();
  {
    const int i = threadIdx.x;
    if (i < 64 && ((mask >> i) & 1ull)) T.set(i, i - 1, {0, 0});
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const cd mu = mus[wave < ns ? wave : 0];
  int lo = 0;
  while (lo < kdim) {
    const unsigned long long above = lo >= 63 ? 0ull : mask & ~((2ull << lo) - 1ull);
    const int hi = above ? __builtin_ctzll(above) - 1 : kdim - 1;
    if (hi > lo) chase_train(T, Q, kdim, lo, hi, ns, mu);
    lo = hi + 1;
  }
}

// The Schur iteration on the window [lo, ihi] of T (the rest of T already
// reduced around it): deflation at the bottom, Wilkinson shifts.
__device__ int schur_window(Mat T, Mat Z, int mm, int lo, int ihi, RotLog log) {
  int iter = 0, total = 0;
  while (ihi >= lo) {
    // largest l in (lo, ihi] with a negligible subdiagonal (lo if none)
    const unsigned long long upto = ihi >= 63 ? ~0ull : ((2ull << ihi) - 1ull);
    const unsigned long long above = lo >= 63 ? 0ull : ~((2ull << lo) - 1ull);
    const unsigned long long neg = negligible_mask(T, mm) & upto & above;
    const int l = neg ? 63 - __builtin_clzll(neg) : lo;
    bsync();
    if (l > lo && threadIdx.x == 0) T.set(l, l - 1, {0, 0});
    bsync();
    if (l == ihi) {
      --ihi;
      iter = 0;
      continue;
    }
    if (++total > 100 * mm) return -1;
    cd mu, mu2 = {0.0, 0.0};
    bool two = false;
    if (iter > 0 && iter % 10 == 0) {
      cd d = T.get(ihi, ihi), sd = T.get(ihi, ihi - 1);
      mu = {d.re + 0.75 * fabs(sd.re), d.im};
    } else {
      // both eigenvalues of the trailing 2 x 2: the one nearer t[ihi][ihi]
      // (Wilkinson) first; a window of 4 or more takes both as a train of two
      cd a = T.get(ihi - 1, ihi - 1), b = T.get(ihi - 1, ihi), c = T.get(ihi, ihi - 1),
         d = T.get(ihi, ihi);
      cd h = {0.5 * (a.re - d.re), 0.5 * (a.im - d.im)};
      cd q = cadd(cmul(h, h), cmul(b, c));
      double mod = hypot(q.re, q.im);
      double sr = sqrt(0.5 * (mod + q.re)), si = sqrt(0.5 * (mod - q.re));
      if (q.im < 0) si = -si;
      cd t1 = {h.re + sr, h.im + si}, t2 = {h.re - sr, h.im - si};
      const bool near1 = cabs2(t1) < cabs2(t2);
      mu = near1 ? cadd(d, t1) : cadd(d, t2);
      mu2 = near1 ? cadd(d, t2) : cadd(d, t1);
      two = ihi - l >= 3 && blockDim.x >= 128;
    }
    ++iter;
    if (two)
      chase_train(T, Z, mm, l, ihi, 2, (threadIdx.x >> 6) == 0 ? mu : mu2);
    else
      chase(T, Z, mm, l, ihi, mu, log);
  }
  return 0;
}

__device__ __forceinline__ int schur(Mat T, Mat Z, int mm, RotLog log) {
  return schur_window(T, Z, mm, 0, mm - 1, log);
}

// ---------------------------------------------------------------------------
// Perfect-shift Schur form (m <= kAberthMaxM): the eigenvalues of the
// Hessenberg H first, by the Ehrlich-Aberth simultaneous iteration on
// det(H - zI) (Hyman's recurrence: with x_{m-1} = 1, row r of (H - zI) x = 0
// gives x_{r-1}, and row 0 leaves q(z), det(H - zI) = c q(z); the
// z-derivative runs beside it, and q'/q is all the iteration needs).  Then
// the Schur iteration uses them as perfect shifts, smallest modulus first: a
// train of eight of them over the active window splits its trailing 8 x 8
// block off (|T[b][b-1]| at the rounding level of ||H||), whose eight
// eigenvalues then deflate one per single perfect-shift sweep at the
// bottom.  Where a split or a deflation does not happen (forward
// instability) the Wilkinson iteration takes over that window; the last
// window of up to twelve goes by single perfect shifts.  CPU emulation on
// Arnoldi H of order 48 (C1 construction, first and later cycles): 430-655
// train steps against 2,250-2,690 with Wilkinson pairs, eigenvalues within
// 1e-14 of LAPACK; 11-15 Aberth iterations.
constexpr int kAberthMaxM = 64;   // eight register slots of x, x' per wave
constexpr int kAberthMaxIter = 64;

// a / b with b scaled by its larger component (Smith's form: no overflow of
// |b|^2 for the large values Hyman's recurrence can reach)
__device__ __forceinline__ cd cdiv(cd a, cd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = 1.0 / (b.re + b.im * r);
    return {(a.re + a.im * r) * d, (a.im - a.re * r) * d};
  }
  const double r = b.re / b.im, d = 1.0 / (b.re * r + b.im);
  return {(a.re * r + a.im) * d, (a.im * r - a.re) * d};
}

// |t_{i,i-1}| <= 8 eps ||H||_F (a backward-stable split) or the local test
__device__ __forceinline__ bool split_ok(Mat T, int i, double hnorm) {
  const double eps = 2.220446049250313e-16;
  const cd sd = T.get(i, i - 1);
  if (fabs(sd.re) + fabs(sd.im) <= 8.0 * eps * hnorm) return true;
  const cd a = T.get(i - 1, i - 1), b = T.get(i, i);
  const double d = fabs(a.re) + fabs(a.im) + fabs(b.re) + fabs(b.im);
  const double sub = fabs(sd.re) + fabs(sd.im);
  return sub <= eps * d || sub < 1e-300;
}

// Block of BS rows of Hyman's recurrence (bottom row bot, lane = root), its
// running sums S, D complete for every column below the block: solves it
// from the bottom row up (x_{r-1} = -S_r / h_{r,r-1}, x'_{r-1} = -D_r /
// h_{r,r-1}), adding each new column to the block's rows above at once, and
// publishes the BS new columns x, x' to pb (zeros past row 0); row 0 leaves
// q, q' in qbuf.
template <int BS>
__device__ __forceinline__ void ab_solve(double (&sre)[BS], double (&sim)[BS], double (&dre)[BS],
                                         double (&dim)[BS], const double (&hre)[BS],
                                         const double (&him)[BS], int bot, cd zi,
                                         const cd* inv_sub, double* pb, double* qbuf, int lane) {
#pragma unroll
  for (int i = 0; i < BS; ++i) {
    const int r = bot - i;
    if (r >= 1) {
      const cd ih = inv_sub[r];
      const cd xn = cmul({-sre[i], -sim[i]}, ih);
      const cd dn = cmul({-dre[i], -dim[i]}, ih);
      pb[(4 * i) * 64 + lane] = xn.re;
      pb[(4 * i + 1) * 64 + lane] = xn.im;
      pb[(4 * i + 2) * 64 + lane] = dn.re;
      pb[(4 * i + 3) * 64 + lane] = dn.im;
#pragma unroll
      for (int i2 = i + 1; i2 < BS; ++i2) {
        cd h = {lane_read(hre[i2], r - 1), lane_read(him[i2], r - 1)};
        if (i2 == i + 1) h = csub(h, zi);  // row r - 1, column r - 1
        sre[i2] = fma(h.re, xn.re, fma(-h.im, xn.im, sre[i2]));
        sim[i2] = fma(h.re, xn.im, fma(h.im, xn.re, sim[i2]));
        dre[i2] = fma(h.re, dn.re, fma(-h.im, dn.im, dre[i2]));
        dim[i2] = fma(h.re, dn.im, fma(h.im, dn.re, dim[i2]));
        if (i2 == i + 1) {
          dre[i2] -= xn.re;
          dim[i2] -= xn.im;
        }
      }
    } else {
      if (r == 0) {
        qbuf[lane] = sre[i];
        qbuf[64 + lane] = sim[i];
        qbuf[128 + lane] = dre[i];
        qbuf[192 + lane] = dim[i];
      }
      pb[(4 * i) * 64 + lane] = 0.0;
      pb[(4 * i + 1) * 64 + lane] = 0.0;
      pb[(4 * i + 2) * 64 + lane] = 0.0;
      pb[(4 * i + 3) * 64 + lane] = 0.0;
    }
  }
}

// The BS columns j = bot - 1 - i (x, x' in xs, ds) into the running sums
// of BS rows (H rows hre, him, lane = column); no diagonal terms (rows
// above the columns, or the caller adds them).
template <int BS>
__device__ __forceinline__ void ab_update(double (&sre)[BS], double (&sim)[BS], double (&dre)[BS],
                                          double (&dim)[BS], const double (&hre)[BS],
                                          const double (&him)[BS], int bot, const cd (&xs)[BS],
                                          const cd (&ds)[BS]) {
#pragma unroll
  for (int i = 0; i < BS; ++i) {
    const int j = bot - 1 - i >= 0 ? bot - 1 - i : 0;
#pragma unroll
    for (int i2 = 0; i2 < BS; ++i2) {
      const cd h = {lane_read(hre[i2], j), lane_read(him[i2], j)};
      sre[i2] = fma(h.re, xs[i].re, fma(-h.im, xs[i].im, sre[i2]));
      sim[i2] = fma(h.re, xs[i].im, fma(h.im, xs[i].re, sim[i2]));
      dre[i2] = fma(h.re, ds[i].re, fma(-h.im, ds[i].im, dre[i2]));
      dim[i2] = fma(h.re, ds[i].im, fma(h.im, ds[i].re, dim[i2]));
    }
  }
}

// The Aberth iteration of aberth_roots (below) for BS rows per block.
constexpr int kAbBlocks = 16;  // row blocks of the recurrence (two per wave)
template <int BS>  // rows per block: ceil(mm / 16)
__device__ bool aberth_roots_t(Mat T, int mm, double hnorm, cd* z, double* part, int* flags,
                               cd* inv_sub, const double* warm_re, const double* warm_im) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const bool act = lane < mm;
  // starting points: a circle around trace / mm with the RMS radius of the
  // spectrum bound ||H - cI||_F / sqrt(mm) (wave 0)
  cd tr = {0.0, 0.0};
  for (int i = 0; i < mm; ++i) tr = cadd(tr, T.get(i, i));
  const cd c0 = {tr.re / mm, tr.im / mm};
  if (wave == 0) {
    double f2 = 0.0;
    for (int i = lane; i < mm * mm; i += 64) {
      const int r = i / mm, cc = i - r * mm;
      if (r <= cc + 1) {
        cd t = T.get(r, cc);
        if (r == cc) t = csub(t, c0);
        f2 += cabs2(t);
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) f2 += __shfl_xor(f2, off, 64);
    const double rad = sqrt(f2 / mm) + 1e-300;
    if (act) {
      const double ang = 6.283185307179586 * lane / mm + 0.4;
      if (warm_re) {
        // the previous restart's Ritz values (6-8 iterations instead of
        // 9-13 in emulation), kept apart by a tiny turn
        z[lane] = {warm_re[lane] + 1e-8 * rad * cos(ang), warm_im[lane] + 1e-8 * rad * sin(ang)};
      } else {
        z[lane] = {c0.re + rad * cos(ang), c0.im + rad * sin(ang)};
      }
    }
    if (lane == 0) {
      flags[0] = 0;  // 1: every root converged
      flags[2] = 0;  // iterations taken
    }
    // the reciprocal subdiagonals 1 / t_{r, r-1} (the recurrence divides by
    // them every row)
    for (int r = 1 + lane; r < mm; r += 64) inv_sub[r] = cdiv({1.0, 0.0}, T.get(r, r - 1));
  }
  __syncthreads();
  bool done = !act;  // wave 0's view
  double prev = 1e300;
  const int nb = (mm + BS - 1) / BS;
  // this wave's rows of H, lane = column: block K = wave + 8 s, row mm - 1 - K BS - i
  double hre[2][BS], him[2][BS];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < BS; ++i) {
      const int K = wave + 8 * s, r = mm - 1 - K * BS - i;
      const bool in = K < nb && r >= 0 && lane < mm && lane >= r - 1;
      const cd h = in ? T.get(r, lane) : cd{0.0, 0.0};
      hre[s][i] = h.re;
      him[s][i] = h.im;
    }
  double* qbuf = part + 2 * BS * 4 * 64;  // q, q' of row 0 (4 x 64 doubles)
  double* sump = part + 2 * 4 * 4 * 64 + 4 * 64;  // the waves' Aberth sums (8 x 2 x 64)
  for (int it = 0; it < kAberthMaxIter; ++it) {
    // no early exit from the loop: the iterations left once every root has
    // converged skip their work (an early exit here miscompiled for gfx950)
    const bool all_done = flags[0] != 0;
    const cd zi = act ? z[lane] : cd{0.0, 0.0};
    // running sums of this wave's rows: S = sum_j (H - zI)_{rj} x_j and
    // D = sum_j (H - zI)_{rj} x'_j - x_r; first column mm - 1 (x = 1, x' = 0)
    double sre[2][BS], sim[2][BS], dre[2][BS], dim[2][BS];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < BS; ++i) {
        const bool diag = (wave + 8 * s) == 0 && i == 0;  // row mm - 1
        cd h = {lane_read(hre[s][i], mm - 1), lane_read(him[s][i], mm - 1)};
        if (diag) h = csub(h, zi);
        sre[s][i] = h.re;
        sim[s][i] = h.im;
        dre[s][i] = diag ? -1.0 : 0.0;
        dim[s][i] = 0.0;
      }
    if (!all_done) {
      if (wave == 0) ab_solve<BS>(sre[0], sim[0], dre[0], dim[0], hre[0], him[0], mm - 1, zi, inv_sub,
                                  part, qbuf, lane);
      for (int K = 0; K + 1 < nb; ++K) {
        __syncthreads();
        // block K's columns j = bot - 1 - i into the rows still to be solved
        // (blocks after K); the owner of block K + 1 takes its own rows, the
        // diagonal term and the solve first, at raised priority (the chain
        // every other wave waits for), then its rows of a later block
        const int bot = mm - 1 - K * BS;
        const double* pb = part + (size_t)(K & 1) * (BS * 4 * 64);
        double* pn = part + (size_t)((K + 1) & 1) * (BS * 4 * 64);
        if (wave + 8 <= K || (wave <= K && wave + 8 >= nb)) continue;  // no rows left here
        cd xs[BS], ds[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) {
          xs[i] = {pb[(4 * i) * 64 + lane], pb[(4 * i + 1) * 64 + lane]};
          ds[i] = {pb[(4 * i + 2) * 64 + lane], pb[(4 * i + 3) * 64 + lane]};
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          if (wave + 8 * s != K + 1) continue;  // the same for the whole wave
          __builtin_amdgcn_s_setprio(3);
          ab_update<BS>(sre[s], sim[s], dre[s], dim[s], hre[s], him[s], bot, xs, ds);
          // the diagonal term of row bot - BS and column bot - BS: -z x and
          // -(z x' + x)
          const cd xj = xs[BS - 1], dj = ds[BS - 1];
          const cd zx = cmul(zi, xj), zd = cmul(zi, dj);
          sre[s][0] -= zx.re;
          sim[s][0] -= zx.im;
          dre[s][0] -= zd.re + xj.re;
          dim[s][0] -= zd.im + xj.im;
          ab_solve<BS>(sre[s], sim[s], dre[s], dim[s], hre[s], him[s], bot - BS, zi, inv_sub, pn,
                       qbuf, lane);
          __builtin_amdgcn_s_setprio(0);
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int K2 = wave + 8 * s;
          if (K2 <= K + 1 || K2 >= nb) continue;  // the same for the whole wave
          ab_update<BS>(sre[s], sim[s], dre[s], dim[s], hre[s], him[s], bot, xs, ds);
        }
      }
      __syncthreads();  // q, q' of row 0 from the last block's solve
      // the Aberth sums sum_{j != i} 1 / (z_i - z_j), j = wave + 8 q on
      // this wave: conj(d) / |d|^2 with the hardware reciprocal and one
      // Newton step (the differences stay far from the overflow range)
      cd ps = {0.0, 0.0};
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int j = wave + 8 * q;
        if (j < mm) {
          const cd d = csub(zi, z[j]);
          const double a2 = cabs2(d);
          double inv = __builtin_amdgcn_rcp(a2);
          inv = fma(fma(-a2, inv, 1.0), inv, inv);
          if (j != lane) ps = cadd(ps, {d.re * inv, -d.im * inv});
        }
      }
      sump[wave * 128 + lane] = ps.re;
      sump[wave * 128 + 64 + lane] = ps.im;
      __syncthreads();
    }
    // wave 0: the Aberth step w = 1 / (q'/q - sum_{j != i} 1 / (z_i - z_j))
    if (wave == 0 && !all_done) {
      cd w = {0.0, 0.0};
      bool hit = false;
      if (!done) {
        const cd q0 = {qbuf[lane], qbuf[64 + lane]}, qd0 = {qbuf[128 + lane], qbuf[192 + lane]};
        if (q0.re == 0.0 && q0.im == 0.0) {
          hit = true;  // z is a root
        } else {
          const cd ratio = cdiv(qd0, q0);
          cd sum = {0.0, 0.0};
#pragma unroll
          for (int w2 = 0; w2 < 8; ++w2) sum = cadd(sum, {sump[w2 * 128 + lane], sump[w2 * 128 + 64 + lane]});
          w = cdiv({1.0, 0.0}, csub(ratio, sum));
        }
      }
      wsync();  // every lane has read z before any writes
      if (!done) {
        if (hit) {
          done = true;
        } else {
          const cd zn = csub(zi, w);
          z[lane] = zn;
          const double rel = sqrt(cabs2(w)) / fmax(sqrt(cabs2(zn)), 1e-300 + 1e-16 * hnorm);
          // converged, or stalled at the rounding level
          if (rel <= 16.0 * 2.220446049250313e-16 || (rel < 1e-12 && rel > 0.5 * prev)) done = true;
          prev = rel;
        }
      }
      const bool left = __ballot(!done) != 0ull;
      if (lane == 0 && !left) flags[0] = 1;
      if (lane == 0) flags[2] = it + 1;
    }
    __syncthreads();
  }
  // every root converged, finite, and the roots sum to the trace (wave 0)
  if (wave == 0) {
    const cd zi = act ? z[lane] : cd{0.0, 0.0};
    double sre = act ? zi.re : 0.0, sim = act ? zi.im : 0.0, sab = act ? sqrt(cabs2(zi)) : 0.0;
    bool bad = !done || (act && !(fabs(zi.re) < 1e300 && fabs(zi.im) < 1e300));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      sre += __shfl_xor(sre, off, 64);
      sim += __shfl_xor(sim, off, 64);
      sab += __shfl_xor(sab, off, 64);
    }
    const bool trace_ok = hypot(sre - tr.re, sim - tr.im) <= 1e-12 * (sab + hnorm);
    const bool ok = __ballot(bad) == 0ull && trace_ok;
    if (lane == 0) flags[1] = ok ? 1 : 0;
  }
  __syncthreads();
  return flags[1] != 0;
}


// All mm eigenvalues of H (in T, unreduced: every subdiagonal nonzero) into
// z (LDS, mm entries), by all 8 waves of the block; lane i iterates root i.
// Hyman's recurrence for det(H - zI) (x_{mm-1} = 1, x_{r-1} = -(sum_{j>=r}
// (H - zI)_{rj} x_j) / h_{r,r-1}, and the same for x' = dx/dz) is a back
// substitution by rows: its rows form nb <= 16 blocks of BS = ceil(mm / 16)
// rows from the bottom, block K owned by wave K % 8, which keeps the
// block's rows of H in registers (lane = column, read back by readlane) and
// the running sums of those rows (lane = root).  The owner solves its block
// row by row once every column below it is in the sums and publishes the
// block's new x, x' in LDS (part: 2 x BS x 4 x 64 doubles, double-buffered
// by block parity); after one barrier every wave adds them to the sums of
// its rows still to be solved, with fused multiply-adds and no branches
// inside a block.  Row 0 leaves q(z) = det(H - zI) / (the product of the
// subdiagonals, up to sign) and q'(z) in LDS, and wave 0 takes the Aberth
// step for every root.  Returns (the same in every thread) whether all
// roots converged, are finite, and sum to the trace; flags[2] = the
// iterations taken.
__device__ bool aberth_roots(Mat T, int mm, double hnorm, cd* z, double* part, int* flags,
                             cd* inv_sub, const double* warm_re = nullptr,
                             const double* warm_im = nullptr) {
  switch ((mm + kAbBlocks - 1) / kAbBlocks) {
    case 1: return aberth_roots_t<1>(T, mm, hnorm, z, part, flags, inv_sub, warm_re, warm_im);
    case 2: return aberth_roots_t<2>(T, mm, hnorm, z, part, flags, inv_sub, warm_re, warm_im);
    case 3: return aberth_roots_t<3>(T, mm, hnorm, z, part, flags, inv_sub, warm_re, warm_im);
    default: return aberth_roots_t<4>(T, mm, hnorm, z, part, flags, inv_sub, warm_re, warm_im);
  }
}

// Right eigenvectors of H (in T, upper Hessenberg, mm <= 64) for the
// eigenvalues ev[order[t]], t < cnt, by inverse iteration (the method of
// LAPACK's zlaein, on the leading rows and columns 0..bend[i] of H, bend[i]
// the last row of the unreduced block that holds eigenvalue i, as zhsein
// does; the vector is zero below): Gaussian elimination of H - mu I with partial pivoting
// between the carried row and the next row of H (a zero pivot becomes
// eps3 = eps ||H||_F), then U x = v by back substitution with v = eps3 (1,
// ..., 1) and, where x did not grow by zlaein's factor (||x||_1 >= 0.1 /
// sqrt(mm)), two more of zlaein's starting vectors.  Wave w < nwe takes t =
// w, w + nwe, ...: the lanes hold the columns of the carried row, the pivot
// rows go to the wave's packed U in LDS (ubuf: mm (mm + 1) / 2 entries per
// wave, row j from column j on) and the back substitution runs by columns
// (lane r accumulates row r).  A vector is accepted when it grew and its
// residual ||(H - mu I) y||_2 <= 64 mm eps ||H||_F.  Writes y / ||y|| to
// column t of Y (leading dimension ky, rows < mm) and the Ritz estimate
// beta |y_{mm-1}| to est[order[t]].  Returns (the same in every thread)
// whether every vector was accepted.
__device__ bool inverse_vectors(Mat T, int mm, double hnorm, const cd* ev, const int* order,
                                const int* bend, int t0, int cnt, cd* ubuf, int nwe, double* Yre,
                                double* Yim, int ky, double beta, double* est, int* flag) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const double eps = 2.220446049250313e-16;
  const double eps3 = fmax(eps * hnorm, 1e-300);
  if (threadIdx.x == 0) *flag = 0;
  __syncthreads();
  const bool act = lane < mm;
  cd* U = ubuf + (size_t)wave * (mm * (mm + 1) / 2);
  bool bad = false;
  for (int t = t0 + wave; wave < nwe && t < cnt; t += nwe) {
    const cd mu = ev[order[t]];
    const int ms = __builtin_amdgcn_readfirstlane(bend[order[t]] + 1);  // leading block
    const bool in = lane < ms;
    const double rootn = sqrt((double)ms);
    const int roff = lane * ms - lane * (lane - 1) / 2;  // packed start of row `lane`
    // carried row: row 0 of H - mu I
    cd w = in ? T.get(0, lane) : cd{0.0, 0.0};
    if (lane == 0) w = csub(w, mu);
    cd invd = {0.0, 0.0};  // lane j: 1 / U[j][j]
    int off = 0;           // packed start of row j
    for (int j = 0; j < ms - 1; ++j) {
      cd h = (in && lane >= j) ? T.get(j + 1, lane) : cd{0.0, 0.0};
      if (lane == j + 1) h = csub(h, mu);
      const cd p = lane_read(w, j);
      const cd q = T.get(j + 1, j);
      const bool swap = cabs2(q) > cabs2(p);  // the same in every lane
      cd piv = swap ? q : p;
      if (piv.re == 0.0 && piv.im == 0.0) piv = {eps3, 0.0};
      const cd ip = cdiv({1.0, 0.0}, piv);
      const cd l = cmul(swap ? p : q, ip);
      const cd prow = swap ? h : w;  // pivot row
      const cd orow = swap ? w : h;
      if (in && lane >= j) U[off + lane - j] = lane == j ? piv : prow;
      if (lane == j) invd = ip;
      w = csub(orow, cmul(l, prow));
      off += ms - j;
    }
    {
      cd piv = lane_read(w, ms - 1);
      if (piv.re == 0.0 && piv.im == 0.0) piv = {eps3, 0.0};
      if (lane == ms - 1) {
        U[off] = piv;
        invd = cdiv({1.0, 0.0}, piv);
      }
    }
    wsync();
    cd x = {0.0, 0.0};
    bool grew = false;
    for (int its = 0; its < 3 && !grew; ++its) {
      // zlaein's starting vectors: eps3 (1, ..., 1), then eps3 / (sqrt(ms) + 1)
      // everywhere but eps3 first, minus eps3 sqrt(ms) at position ms - its
      double v = eps3;
      if (its > 0) {
        v = lane == 0 ? eps3 : eps3 / (rootn + 1.0);
        if (lane == ms - its) v -= eps3 * rootn;
      }
      cd acc = {in ? v : 0.0, 0.0};
      for (int i = ms - 1; i >= 0; --i) {
        const cd xi = lane_read(cmul(acc, invd), i);
        if (lane == i) x = xi;
        if (lane < i) acc = csub(acc, cmul(U[roff + i - lane], xi));
      }
      double n1 = in ? fabs(x.re) + fabs(x.im) : 0.0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) n1 += __shfl_xor(n1, o, 64);
      n1 = lane_read(n1, 0);  // lanes may differ in the last bit: lane 0 decides
      grew = n1 >= 0.1 / rootn && n1 < 1e300;
    }
    double n2 = in ? cabs2(x) : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) n2 += __shfl_xor(n2, o, 64);
    n2 = lane_read(n2, 0);
    const double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
    const cd y = {x.re * inv, x.im * inv};
    // residual (H - mu I) y, y broadcast through the wave's U area
    wsync();
    if (act) U[lane] = y;
    wsync();
    cd rr = {0.0, 0.0};
    if (act) {
      for (int c = lane > 0 ? lane - 1 : 0; c < mm; ++c) rr = cadd(rr, cmul(T.get(lane, c), U[c]));
      rr = csub(rr, cmul(mu, y));
    }
    double r2 = act ? cabs2(rr) : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r2 += __shfl_xor(r2, o, 64);
    r2 = lane_read(r2, 0);
    const double rtol = 64.0 * mm * eps * hnorm;
    if (!grew || !(r2 <= rtol * rtol) || !(n2 > 0.0)) bad = true;
    if (act) {
      Yre[(size_t)lane * ky + t] = y.re;
      Yim[(size_t)lane * ky + t] = y.im;
    }
    if (lane == mm - 1) est[order[t]] = beta * sqrt(cabs2(y));
    wsync();  // the residual's reads of U before the next vector's stores
  }
  if (bad && lane == 0) *flag = 1;
  __syncthreads();
  return *flag == 0;
}

// Window [lo, hi] whose eigenvalues are the cnt (<= 16) values seeds: one
// single perfect-shift sweep per eigenvalue (the unused seed nearest the
// trailing diagonal entry), each deflating at the bottom; the Wilkinson
// iteration finishes the window when a sweep does not deflate.
__device__ int perfect_block(Mat T, Mat Z, int mm, int lo, int hi, const cd* seeds, int cnt,
                             double hnorm, RotLog log) {
  unsigned used = 0u;
  while (hi > lo) {
    bsync();
    if (split_ok(T, hi, hnorm)) {
      bsync();
      if (threadIdx.x == 0) T.set(hi, hi - 1, {0.0, 0.0});
      --hi;
      continue;
    }
    const cd t = T.get(hi, hi);
    int best = -1;
    double bd = 0.0;
    for (int i = 0; i < cnt; ++i) {
      if ((used >> i) & 1u) continue;
      const double d = cabs2(csub(seeds[i], t));
      if (best < 0 || d < bd) { best = i; bd = d; }
    }
    if (best < 0) return schur_window(T, Z, mm, lo, hi, log);
    bsync();
    chase(T, Z, mm, lo, hi, seeds[best], log);
    bsync();
    if (split_ok(T, hi, hnorm)) {
      used |= 1u << best;
      bsync();
      if (threadIdx.x == 0) T.set(hi, hi - 1, {0.0, 0.0});
      --hi;
    } else {
      return schur_window(T, Z, mm, lo, hi, log);
    }
  }
  bsync();
  return 0;
}

// The Schur form from the perfect shifts seeds (all mm eigenvalues, ascending
// modulus); block of >= 8 waves.
__device__ int schur_peel(Mat T, Mat Z, int mm, const cd* seeds, double hnorm, RotLog log,
                          int* peels = nullptr) {
  const int wave = threadIdx.x >> 6;
  int ihi = mm - 1, next = 0;
  while (ihi > 0) {
    bsync();
    if (split_ok(T, ihi, hnorm)) {
      bsync();
      if (threadIdx.x == 0) T.set(ihi, ihi - 1, {0.0, 0.0});
      --ihi;
      continue;
    }
    // the unused seeds are the window's eigenvalues only while every
    // deflation so far consumed one
    const int remaining = mm - next;
    if (remaining != ihi + 1) return schur_window(T, Z, mm, 0, ihi, log);
    if (remaining <= 12) return perfect_block(T, Z, mm, 0, ihi, seeds + next, remaining, hnorm, log);
    bsync();
    chase_train(T, Z, mm, 0, ihi, 8, seeds[next + (wave < 8 ? wave : 0)]);
    const int b = ihi - 7;
    bsync();
    if (!split_ok(T, b, hnorm)) return schur_window(T, Z, mm, 0, ihi, log);
    bsync();
    if (threadIdx.x == 0) T.set(b, b - 1, {0.0, 0.0});
    const int f = perfect_block(T, Z, mm, b, ihi, seeds + next, 8, hnorm, log);
    if (peels && threadIdx.x == 0) ++*peels;
    if (f) return f;
    next += 8;
    ihi = b - 1;
  }
  bsync();
  return 0;
}


__host__ __device__ constexpr int lds_ld(int m) { return m | 1; }

// ---------------------------------------------------------------------------
// General Krylov dimension (kMaxM < m <= kMaxMBig).  T lives in LDS while it
// fits (general_t_in_lds), else in global memory like Z (the TZ work, L2
// resident); the sweeps run as bulge trains (chase_train_general) of two
// Wilkinson shifts in the Schur phase and up to one exact shift per wave,
// with the lanes walking 64-wide chunks of the rows and columns and Z
// updated after each train from the rotation logs, one row of Z per thread
// (the logs live in the eigenvector work a.X, idle in both phases).  Same
// iteration as shifted_sweep in oracle/oracle.c.
constexpr int kGeneralThreads = 256;
constexpr int kGeneralLdsMax = 128 * 1024;
__host__ __device__ constexpr bool general_t_in_lds(int m) {
  return 2 * m * lds_ld(m) * (int)sizeof(double) <= kGeneralLdsMax;
}
__host__ __device__ constexpr int general_lds_bytes(int m) {
  return general_t_in_lds(m) ? 2 * m * lds_ld(m) * (int)sizeof(double) : 0;
}

// |t_{i,i-1}| <= eps (|t_{i-1,i-1}| + |t_{i,i}|) (the test of negligible_mask)
__device__ __forceinline__ bool sub_negligible(Mat T, int i) {
  const double eps = 2.220446049250313e-16;
  const cd a = T.get(i - 1, i - 1), b = T.get(i, i), sd = T.get(i, i - 1);
  const double d = fabs(a.re) + fabs(a.im) + fabs(b.re) + fabs(b.im);
  const double sub = fabs(sd.re) + fabs(sd.im);
  return sub <= eps * d || sub < 1e-300;
}

// Z[:, lo..hi] <- Z[:, lo..hi] G_lo^H ... G_{hi-1}^H from the rotation log
__device__ void apply_log_to_z(Mat Z, int kdim, int lo, int hi, RotLog log) {
  for (int r = threadIdx.x; r < kdim; r += blockDim.x) {
    cd a = Z.get(r, lo);
#pragma unroll 4
    for (int kk = lo; kk < hi; ++kk) {
      const cd b = Z.get(r, kk + 1);
      const double c = log.c[kk];
      const cd s = log.s[kk];
      Z.set(r, kk, gcol_left(a, b, c, s));
      a = gcol_right(a, b, c, s);
    }
    Z.set(r, hi, a);
  }
}

// A train of ns (<= waves) implicit single-shift sweeps over rows/columns
// [lo, hi] of T for the general restart (kdim <= kMaxMBig): the steps of
// chase_train (bulge w chased by wave w three rows behind bulge w - 1, a row
// phase and a column phase between block barriers, each bulge's next (f, g)
// kept in registers) with the lanes walking the columns and rows in chunks
// of 64.  Z takes the column rotations after the train, sweep by sweep from
// the rotation logs (lgc, lgs: ns logs of kMaxMBig entries in global
// memory): the train applies Z's column pairs in an order that differs from
// sweep by sweep only by swapping rotations of disjoint column pairs.
__device__ void chase_train_general(Mat T, Mat Z, int kdim, int lo, int hi, int ns, cd mu,
                                    double* lgc, cd* lgs) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int L = hi - lo;
  const int steps = L + 3 * (ns - 1);
  cd fn = {0.0, 0.0}, gn = {0.0, 0.0};
  for (int t = 0; t < steps; ++t) {
    const int p = t - 3 * wave;
    const bool act = wave < ns && p >= 0 && p < L;
    const int k = lo + p;
    double c = 1.0;
    cd s = {0.0, 0.0};
    if (act) {
      if (k == lo) {
        fn = csub(T.get(lo, lo), mu);
        gn = T.get(lo + 1, lo);
      }
      cgivens(fn, gn, c, s);
      if (lane == 0) {
        lgc[wave * kMaxMBig + k] = c;
        lgs[wave * kMaxMBig + k] = s;
      }
      const int j0 = k > lo ? k - 1 : k;
      for (int j = j0 + lane; j < kdim; j += 64) {
        const cd a = T.get(k, j), b = T.get(k + 1, j);
        T.set(k, j, grow_top(a, b, c, s));
        T.set(k + 1, j, (k > lo && j == k - 1) ? cd{0.0, 0.0} : grow_bot(a, b, c, s));
      }
    }
    __syncthreads();
    if (act) {
      const int rmax = k + 2 < hi ? k + 2 : hi;
      cd f1 = {0.0, 0.0}, g1 = {0.0, 0.0};
      for (int r = lane; r <= rmax; r += 64) {
        const cd x = T.get(r, k), y = T.get(r, k + 1);
        const cd nx = gcol_left(x, y, c, s);
        T.set(r, k, nx);
        T.set(r, k + 1, gcol_right(x, y, c, s));
        if (r == k + 1) f1 = nx;
        if (r == k + 2) g1 = nx;
      }
      if (k + 1 < hi) {
        fn = lane_read(f1, (k + 1) & 63);
        gn = lane_read(g1, (k + 2) & 63);
      }
    }
    __syncthreads();
  }
  for (int w = 0; w < ns; ++w)
    apply_log_to_z(Z, kdim, lo, hi, RotLog{lgc + w * kMaxMBig, lgs + w * kMaxMBig});
  __syncthreads();
}

// Complex Schur form of T (mm x mm), accumulating Z: deflation at the
// bottom, both eigenvalues of the trailing 2 x 2 as a train of two (one
// shift when the window is below 4; exceptional shifts every tenth
// iteration), as schur().  scr: one LDS int; lgc, lgs: the train logs.
__device__ int schur_general(Mat T, Mat Z, int mm, int* scr, double* lgc, cd* lgs) {
  int ihi = mm - 1, iter = 0, total = 0;
  while (ihi >= 0) {
    if (threadIdx.x == 0) scr[0] = 0;
    __syncthreads();
    // largest l <= ihi with a negligible subdiagonal (0 if none)
    for (int i = 1 + (int)threadIdx.x; i <= ihi; i += blockDim.x)
      if (sub_negligible(T, i)) atomicMax(scr, i);
    __syncthreads();
    const int l = scr[0];
    if (l > 0 && threadIdx.x == 0) T.set(l, l - 1, {0, 0});
    __syncthreads();
    if (l == ihi) {
      --ihi;
      iter = 0;
      continue;
    }
    if (++total > 100 * mm) return -1;
    cd mu, mu2 = {0.0, 0.0};
    bool two = false;
    if (iter > 0 && iter % 10 == 0) {
      const cd d = T.get(ihi, ihi), sd = T.get(ihi, ihi - 1);
      mu = {d.re + 0.75 * fabs(sd.re), d.im};
    } else {
      const cd a = T.get(ihi - 1, ihi - 1), b = T.get(ihi - 1, ihi), c = T.get(ihi, ihi - 1),
               d = T.get(ihi, ihi);
      const cd h = {0.5 * (a.re - d.re), 0.5 * (a.im - d.im)};
      const cd q = cadd(cmul(h, h), cmul(b, c));
      const double mod = hypot(q.re, q.im);
      double sr = sqrt(0.5 * (mod + q.re)), si = sqrt(0.5 * (mod - q.re));
      if (q.im < 0) si = -si;
      const cd t1 = {h.re + sr, h.im + si}, t2 = {h.re - sr, h.im - si};
      const bool near1 = cabs2(t1) < cabs2(t2);
      mu = near1 ? cadd(d, t1) : cadd(d, t2);
      mu2 = near1 ? cadd(d, t2) : cadd(d, t1);
      two = ihi - l >= 3 && blockDim.x >= 128;
    }
    ++iter;
    __syncthreads();  // every wave has read the window before the train writes it
    chase_train_general(T, Z, mm, l, ihi, two ? 2 : 1, (threadIdx.x >> 6) == 1 ? mu2 : mu, lgc,
                        lgs);
  }
  return 0;
}

// The exact-shift sweeps of the general restart, in trains of up to one
// shift per wave: before each train T is split at its negligible
// subdiagonals (set to zero) and every unreduced block [lo, hi] gets the
// train.  negl: kMaxMBig LDS flags.
__device__ void shifts_general(Mat T, Mat Z, int mm, int nshift, const cd* mus,
                               unsigned char* negl, double* lgc, cd* lgs) {
  const int nw = blockDim.x / 64;
  for (int t0 = 0; t0 < nshift; t0 += nw) {
    const int ns = nshift - t0 < nw ? nshift - t0 : nw;
    for (int i = 1 + (int)threadIdx.x; i < mm; i += blockDim.x) {
      const bool ng = sub_negligible(T, i);
      negl[i] = ng ? 1 : 0;
      if (ng) T.set(i, i - 1, {0, 0});
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const cd mu = mus[t0 + (wave < ns ? wave : 0)];
    int lo = 0;
    while (lo < mm) {
      int hi = lo;
      while (hi + 1 < mm && !negl[hi + 1]) ++hi;
      if (hi > lo) chase_train_general(T, Z, mm, lo, hi, ns, mu, lgc, lgs);
      lo = hi + 1;
    }
    __syncthreads();  // the flags are read before the next train writes them
  }
}

// Dynamic LDS of the restart kernel: T and Z, and the eigenvector work X
// beside them when the three fit (m <= 54); larger m keep X in global memory.
constexpr int kRestartLdsMax = 144 * 1024;
__host__ __device__ constexpr bool x_in_lds(int m) {
  return 6 * m * lds_ld(m) * (int)sizeof(double) <= kRestartLdsMax;
}
// The Aberth iteration's published columns (2 x 4 x 4 x 64 doubles), q, q'
// (4 x 64) and the waves' sums (8 x 2 x 64) follow T (in the Z area, free until the eigenvectors); for
// small m the launch allocates beyond the restart's own need to hold them.
constexpr int kAberthPartDoubles = 2 * 4 * 4 * 64 + 4 * 64 + 8 * 2 * 64;
__host__ __device__ constexpr int restart_lds_bytes(int m) {
  return ((x_in_lds(m) ? 6 : 4) * m * lds_ld(m) > 2 * m * lds_ld(m) + kAberthPartDoubles
              ? (x_in_lds(m) ? 6 : 4) * m * lds_ld(m)
              : 2 * m * lds_ld(m) + kAberthPartDoubles) *
         (int)sizeof(double);
}

__device__ bool ritz_before(cd a, cd b, int which) {
  double ma = cabs2(a), mb = cabs2(b);
  switch (which) {
    case 0: return ma > mb;
    case 1: return a.re > b.re;
    case 2: return a.im > b.im;
    case 3: return ma < mb;
    case 4: return a.re < b.re;
    default: return a.im < b.im;
  }
}

// The restart of one problem by one workgroup.  GENERAL = false (k_restart,
// 8 waves, mact <= kMaxM): T and Z in LDS, mact x mact complex each with the
// odd leading dimension lds_ld(mact), so that a column walk (one row per
// lane) spreads over the LDS banks (4 mact lds_ld(mact) doubles), then X,
// the eigenvectors of T (column i for T[i][i]; 2 mact lds_ld(mact) doubles)
// when x_in_lds(mact), else in the global work a.X (m x m complex); the
// Schur form and the exact shifts run as bulge trains.  GENERAL = true
// (k_restart_general, mact <= kMaxMBig): Z in the global work a.TZ, T in LDS
// when general_t_in_lds(mact) else in a.TZ, X in a.X, single-shift sweeps
// (chase_train_general).  NM: capacity of the per-value LDS arrays.  Outputs:
// status (nconv, kk, sorted Ritz values, estimates), Y (the first ky Ritz
// vectors of H in wanted order), and with a.do_shifts the compressed H and
// the accumulated Q of the exact shifts.
template <int NM, bool GENERAL>
__device__ void restart_body_t(const RestartArgs& a, double* smem) {
  // A breakdown flagged during this factorisation (the same value for every
  // thread: the orthogonalisation kernels have finished) limits the active
  // block to the invariant leading block, with no shifts; one below k or 2
  // leaves the restart to the host (a fresh vector), so nothing is done.
  int mm = a.mact;
  bool do_shifts = a.do_shifts != 0;
  {
    const int bd = a.st->breakdown;
    if (bd > 0) {
      if (bd < a.k || bd < 2) return;  // uniform: every thread returns
      if (bd < mm) mm = bd;
      do_shifts = false;
    }
  }
  const size_t mm2 = (size_t)a.m * a.m;
  const int ld = lds_ld(mm);
  // where T and Z live: LDS, or the TZ planes (leading dimension a.m); the
  // launch sized the LDS for a.mact (>= mm)
  const bool t_lds = !GENERAL || general_t_in_lds(a.mact);
  const bool z_lds = !GENERAL;
  Mat T = t_lds ? Mat{smem, smem + mm * ld, ld} : Mat{a.TZ, a.TZ + mm2, a.m};
  Mat Z = z_lds ? Mat{smem + 2 * mm * ld, smem + 3 * mm * ld, ld}
                : Mat{a.TZ + 2 * mm2, a.TZ + 3 * mm2, a.m};
  __shared__ cd ev[NM];
  __shared__ double est[NM];
  __shared__ int order[NM];
  __shared__ double tnorm_s;
  __shared__ int kk_s;  // kept size, broadcast through LDS (uniform loop bound)
  __shared__ int scr_s[1];
  __shared__ unsigned char negl_s[GENERAL ? NM : 1];
  Status* st = a.st;
  const int lane = threadIdx.x, nthr = blockDim.x;
  __shared__ double tmax_s[1024];

  if (lane == 0) st->clk[0] = clock64();
  __shared__ double log_c[NM];
  __shared__ cd log_s[NM];
  const RotLog log{log_c, log_s};
  int fail = 0;
  // method 2: the Aberth eigenvalues and inverse-iteration eigenvectors took
  // the place of the Schur form (the same in every thread)
  bool inv = false;
  if (a.phase == 2) {
    // the Schur form of the first launch (what of it lives in LDS)
    for (int idx = lane; idx < mm * mm; idx += nthr) {
      int r = idx / mm, c = idx - r * mm;
      const size_t g = (size_t)r * a.m + c;
      if (t_lds) T.set(r, c, {a.TZ[g], a.TZ[mm2 + g]});
      if (z_lds) Z.set(r, c, {a.TZ[2 * mm2 + g], a.TZ[3 * mm2 + g]});
    }
    fail = st->schur_fail;
    __syncthreads();
  } else {
    for (int idx = lane; idx < mm * mm; idx += nthr) {
      int r = idx / mm, c = idx - r * mm;
      double hr = r <= c + 1 ? a.Hre[(size_t)r * a.m + c] : 0.0;
      double hi = r <= c + 1 ? a.Him[(size_t)r * a.m + c] : 0.0;
      T.set(r, c, {hr, hi});
      if (GENERAL) Z.set(r, c, {r == c ? 1.0 : 0.0, 0.0});
    }
    __syncthreads();
    if (GENERAL) {
      fail = schur_general(T, Z, mm, scr_s, a.Xre, reinterpret_cast<cd*>(a.Xre + 4 * kMaxMBig));
      if (lane == 0) st->schur_path = 0;
    } else {
      // perfect shifts where possible: the eigenvalues of the unreduced H by
      // the Aberth iteration on wave 0, its scratch in the Z and X areas
      // (Z is set to the identity afterwards)
      __shared__ cd seeds_s[NM];
      __shared__ int bend_s[NM];  // last row of the unreduced block of each eigenvalue
      int peel = 0;
      bool split_vals = false;  // eigenvalues of a split H, block by block
      double hnorm = 0.0;
      const bool try_peel = a.schur_method >= 1 && mm <= kAberthMaxM && mm >= 3 && nthr >= 512;
      if (try_peel) {
        double f2 = 0.0;
        for (int idx = lane; idx < mm * mm; idx += nthr) {
          const int r = idx / mm, c = idx - r * mm;
          if (r <= c + 1) f2 += cabs2(T.get(r, c));
        }
        tmax_s[lane] = f2;
        __syncthreads();
        if (lane == 0) {
          double v = 0.0;
          for (int t = 0; t < nthr; ++t) v += tmax_s[t];
          tnorm_s = sqrt(v);
        }
        __syncthreads();
        hnorm = tnorm_s;
        __shared__ int aberth_flags[3];
        const unsigned long long negl = negligible_mask(T, mm);  // the same in every wave
        if (negl == 0ull) {
          // the previous restart of this factorisation left its Ritz values
          // in the status: warm starting points
          const bool warm = a.warm && st->ev_count == mm;
          peel = aberth_roots(T, mm, hnorm, ev, smem + 2 * mm * ld, aberth_flags, seeds_s,
                              warm ? st->ev_re : nullptr, warm ? st->ev_im : nullptr)
                     ? 1
                     : 0;
          for (int i = lane; i < mm; i += nthr) bend_s[i] = mm - 1;
        } else if (a.schur_method == 2 && a.phase == 0) {
          // H splits at its negligible subdiagonals (set to zero, as the
          // Schur iteration does): the eigenvalues block by block, Aberth
          // for blocks of three or more, closed forms below
          if (lane < mm && lane >= 1 && ((negl >> lane) & 1ull)) T.set(lane, lane - 1, {0.0, 0.0});
          __syncthreads();
          bool ok = true;
          int lo = 0;
          while (lo < mm) {  // the same blocks in every thread
            int hi = lo;
            while (hi + 1 < mm && !((negl >> (hi + 1)) & 1ull)) ++hi;
            const int sz = hi - lo + 1;
            if (sz >= 3) {
              const Mat W{T.re + lo * ld + lo, T.im + lo * ld + lo, ld};
              ok = aberth_roots(W, sz, hnorm, ev + lo, smem + 2 * mm * ld, aberth_flags,
                                seeds_s + lo) && ok;
            } else if (lane == 0) {
              if (sz == 1) {
                ev[lo] = T.get(lo, lo);
              } else {
                // (a + d) / 2 +- sqrt(((a - d) / 2)^2 + b c)
                const cd aa = T.get(lo, lo), b = T.get(lo, hi), c = T.get(hi, lo),
                         d = T.get(hi, hi);
                const cd h = {0.5 * (aa.re - d.re), 0.5 * (aa.im - d.im)};
                const cd q = cadd(cmul(h, h), cmul(b, c));
                const double mod = hypot(q.re, q.im);
                double sr = sqrt(0.5 * (mod + q.re)), si = sqrt(0.5 * (mod - q.re));
                if (q.im < 0) si = -si;
                const cd mid = {0.5 * (aa.re + d.re), 0.5 * (aa.im + d.im)};
                ev[lo] = {mid.re + sr, mid.im + si};
                ev[hi] = {mid.re - sr, mid.im - si};
              }
            }
            for (int i = lo + lane; i <= hi; i += nthr) bend_s[i] = hi;
            __syncthreads();
            lo = hi + 1;
          }
          split_vals = ok;
        }
        if (peel) {
          // smallest modulus first (stable rank sort)
          for (int i = lane; i < mm; i += nthr) {
            const double mi = cabs2(ev[i]);
            int rank = 0;
            for (int j = 0; j < mm; ++j) {
              const double mj = cabs2(ev[j]);
              if (mj < mi || (mj == mi && j < i)) ++rank;
            }
            seeds_s[rank] = ev[i];
          }
        }
        __syncthreads();
      }
      if ((peel || split_vals) && a.schur_method == 2 && a.phase == 0) {
        // the Aberth values in wanted order, then the wanted eigenvectors of
        // H by inverse iteration, U factors in the Z and X areas; the Schur
        // form is formed only where a vector is not accepted
        for (int i = lane; i < mm; i += nthr) {
          int rank = 0;
          for (int j = 0; j < mm; ++j) {
            const bool bj = ritz_before(ev[j], ev[i], a.which);
            const bool bi = ritz_before(ev[i], ev[j], a.which);
            if (bj || (!bi && j < i)) ++rank;
          }
          order[rank] = i;
          est[i] = INFINITY;  // estimates are formed for the ky wanted ones only
        }
        __syncthreads();
        const int avail = restart_lds_bytes(a.mact) / (int)sizeof(double) - 2 * mm * ld;
        int nwe = avail / (mm * (mm + 1));
        if (nwe > nthr / 64) nwe = nthr / 64;
        const int cnt = a.ky < mm ? a.ky : mm;
        const int cntw = a.k < cnt ? a.k : cnt;
        // the wanted vectors first; the guards (a Ritz-vector-sum restart's
        // extra terms) only when the wanted estimates pass, i.e. when this
        // restart can end the cycles -- otherwise their columns are zeroed
        if (nwe >= 1)
          inv = inverse_vectors(T, mm, hnorm, ev, order, bend_s, 0, cntw,
                                reinterpret_cast<cd*>(smem + 2 * mm * ld), nwe, a.Yre, a.Yim,
                                a.ky, st->beta, est, scr_s);
        if (inv && cnt > cntw) {
          double mx = 0.0;
          for (int i = 0; i < mm; ++i) mx = fmax(mx, sqrt(cabs2(ev[i])));
          const double tabs = a.tol * (a.anorm > 0 ? a.anorm : fmax(mx, 1e-300));
          int nc = 0;
          for (int t = 0; t < cntw; ++t)
            if (est[order[t]] <= tabs) ++nc;
          if (nc >= a.k) {
            inv = inverse_vectors(T, mm, hnorm, ev, order, bend_s, cntw, cnt,
                                  reinterpret_cast<cd*>(smem + 2 * mm * ld), nwe, a.Yre, a.Yim,
                                  a.ky, st->beta, est, scr_s);
          } else {
            for (int idx = lane; idx < mm * (cnt - cntw); idx += nthr) {
              const int r = idx / (cnt - cntw), t = cntw + idx % (cnt - cntw);
              a.Yre[(size_t)r * a.ky + t] = 0.0;
              a.Yim[(size_t)r * a.ky + t] = 0.0;
            }
            __syncthreads();
          }
        }
      }
      if (!inv) {
        for (int idx = lane; idx < mm * mm; idx += nthr) {
          int r = idx / mm, c = idx - r * mm;
          Z.set(r, c, {r == c ? 1.0 : 0.0, 0.0});
        }
        __syncthreads();
        fail = peel ? schur_peel(T, Z, mm, seeds_s, hnorm, log) : schur(T, Z, mm, log);
      }
      if (lane == 0) st->schur_path = inv ? 2 : peel;
    }
    __syncthreads();
  }
  if (lane == 0) st->clk[1] = clock64();
  // eigenvectors of T by back substitution: lane i owns column i of X
  // X in LDS when it fits in what the launch allocated for a.mact (each
  // lane's dependent back substitution then reads LDS, not L2), else the
  // global work a.X
  const bool xl = !GENERAL && 6 * mm * ld * (int)sizeof(double) <= restart_lds_bytes(a.mact);
  double* Xre = xl ? smem + 4 * mm * ld : a.Xre;
  double* Xim = xl ? smem + 5 * mm * ld : a.Xim;
  const size_t xld = xl ? (size_t)ld : (size_t)a.m;
  if (!inv) {
    for (int i = lane; i < mm; i += nthr) ev[i] = T.get(i, i);
    // ||T|| (max modulus of the upper triangle) for the trevc safeguard
    double tm = 0;
    for (int idx = lane; idx < mm * mm; idx += nthr) {
      int r = idx / mm, c = idx - r * mm;
      if (c >= r) {
        const cd t = T.get(r, c);
        tm = fmax(tm, hypot(t.re, t.im));
      }
    }
    tmax_s[lane] = tm;
    __syncthreads();
    if (lane == 0) {
      double v = 0;
      for (int t = 0; t < nthr; ++t) v = fmax(v, tmax_s[t]);
      tnorm_s = v;
    }
    __syncthreads();
    const double small = fmax(tnorm_s * 2.220446049250313e-16, 1e-300);
    const double beta = st->beta;
    for (int i = lane; i < mm; i += nthr) {
      cd li = ev[i];
      for (int j = 0; j < mm; ++j) {
        Xre[j * xld + i] = j == i ? 1.0 : 0.0;
        Xim[j * xld + i] = 0.0;
      }
      for (int j = i - 1; j >= 0; --j) {
        cd s = {0, 0};
        for (int l = j + 1; l <= i; ++l) {
          cd x = {Xre[l * xld + i], Xim[l * xld + i]};
          s = cadd(s, cmul(T.get(j, l), x));
        }
        cd d = csub(T.get(j, j), li);
        if (hypot(d.re, d.im) < small) d = {small, 0.0};
        double den = cabs2(d);
        Xre[j * xld + i] = -(s.re * d.re + s.im * d.im) / den;
        Xim[j * xld + i] = -(s.im * d.re - s.re * d.im) / den;
      }
      // y = Z x: ||y|| = ||x|| (Z unitary); residual estimate beta |y[mm-1]| / ||y||
      double nx = 0;
      cd ylast = {0, 0};
      for (int l = 0; l <= i; ++l) {
        cd x = {Xre[l * xld + i], Xim[l * xld + i]};
        nx += cabs2(x);
        ylast = cadd(ylast, cmul(Z.get(mm - 1, l), x));
      }
      est[i] = beta * hypot(ylast.re, ylast.im) / sqrt(nx);
    }
  }
  __syncthreads();
  if (lane == 0) st->clk[2] = clock64();
  if (a.phase == 1) {
    // values, estimates and the Schur form for the host's comparator
    if (lane == 0) {
      double maxabs = 0;
      for (int i = 0; i < mm; ++i) maxabs = fmax(maxabs, sqrt(cabs2(ev[i])));
      st->maxabs = maxabs;
      st->schur_fail = fail;
      st->ev_count = mm;
    }
    for (int i = lane; i < mm; i += nthr) {
      st->ev_re[i] = ev[i].re;
      st->ev_im[i] = ev[i].im;
      st->est[i] = est[i];
      st->order[i] = i;
    }
    for (int idx = lane; idx < mm * mm; idx += nthr) {
      int r = idx / mm, c = idx - r * mm;
      const size_t g = (size_t)r * a.m + c;
      if (t_lds) {
        const cd t = T.get(r, c);
        a.TZ[g] = t.re;
        a.TZ[mm2 + g] = t.im;
      }
      if (z_lds) {
        const cd z = Z.get(r, c);
        a.TZ[2 * mm2 + g] = z.re;
        a.TZ[3 * mm2 + g] = z.im;
      }
    }
    return;  // uniform: every thread returns
  }
  if (a.phase == 2) {
    for (int i = lane; i < mm; i += nthr) order[i] = st->order[i];
  } else if (!inv) {
    // stable rank sort by the wanted order
    for (int i = lane; i < mm; i += nthr) {
      int rank = 0;
      for (int j = 0; j < mm; ++j) {
        bool bj = ritz_before(ev[j], ev[i], a.which);
        bool bi = ritz_before(ev[i], ev[j], a.which);
        if (bj || (!bi && j < i)) ++rank;
      }
      order[rank] = i;
    }
  }
  __syncthreads();
  if (lane == 0) {
    double maxabs = 0;
    for (int i = 0; i < mm; ++i) maxabs = fmax(maxabs, sqrt(cabs2(ev[i])));
    double tolabs = a.tol * (a.anorm > 0 ? a.anorm : fmax(maxabs, 1e-300));
    int nconv = 0;
    for (int t = 0; t < a.k && t < mm; ++t)
      if (est[order[t]] <= tolabs) ++nconv;
    // kept size k + min(nconv, (m - k) / 2) (ARPACK's nev adjustment):
    // keeping only k lets an unconverged wanted eigenvalue be pushed out by
    // the shifts, and the solve then converges to a neighbour outside the
    // wanted set on clustered spectra
    const int p2 = (mm - a.k) / 2;
    int kk = a.k + (nconv < p2 ? nconv : p2);
    if (kk == 1 && mm >= 6) kk = mm / 2;
    else if (kk == 1 && mm > 2) kk = 2;
    if (kk < a.keep) kk = a.keep;
    if (kk >= mm) kk = mm - 1;
    st->nconv = nconv;
    st->kk = kk;
    kk_s = kk;
    st->schur_fail = fail;
    st->maxabs = maxabs;
    st->tolabs = tolabs;
    for (int t = 0; t < mm; ++t) {
      st->ev_re[t] = ev[order[t]].re;
      st->ev_im[t] = ev[order[t]].im;
      st->est[t] = est[order[t]];
      st->order[t] = order[t];
    }
    st->ev_count = mm;
  }
  __syncthreads();
  if (lane == 0) st->clk[3] = clock64();
  // Ritz vectors of H in wanted order: Y[:, t] = Z x_{order[t]} / ||x||, t < ky
  for (int t = 0; !inv && t < a.ky && t < mm; ++t) {
    int i = order[t];
    double nx = 0;
    for (int l = 0; l <= i; ++l)
      nx += Xre[l * xld + i] * Xre[l * xld + i] +
            Xim[l * xld + i] * Xim[l * xld + i];
    double inv = 1.0 / sqrt(nx);
    for (int r = lane; r < mm; r += nthr) {
      cd y = {0, 0};
      for (int l = 0; l <= i; ++l)
        y = cadd(y, cmul(Z.get(r, l), {Xre[l * xld + i], Xim[l * xld + i]}));
      a.Yre[(size_t)r * a.ky + t] = y.re * inv;
      a.Yim[(size_t)r * a.ky + t] = y.im * inv;
    }
  }
  __syncthreads();
  if (lane == 0) st->clk[4] = clock64();
  if (!do_shifts) return;  // uniform: every thread returns
  // exact shifts: T <- H, Q <- I (Q reuses Z), sweeps with the unwanted values
  const int kk = kk_s;
  for (int idx = lane; idx < mm * mm; idx += nthr) {
    int r = idx / mm, c = idx - r * mm;
    double hr = r <= c + 1 ? a.Hre[(size_t)r * a.m + c] : 0.0;
    double hi = r <= c + 1 ? a.Him[(size_t)r * a.m + c] : 0.0;
    T.set(r, c, {hr, hi});
    Z.set(r, c, {r == c ? 1.0 : 0.0, 0.0});
  }
  __syncthreads();
  // the p = mm - kk unwanted values, in trains of up to one per wave
  const int nw = nthr / 64;
  __shared__ cd mus_s[NM];
  for (int t = kk + lane; t < mm; t += nthr) mus_s[t - kk] = ev[order[t]];
  __syncthreads();
  if (GENERAL) {
    shifts_general(T, Z, mm, mm - kk, mus_s, negl_s, a.Xre,
                   reinterpret_cast<cd*>(a.Xre + 4 * kMaxMBig));
  } else if (nw >= 2) {
    for (int t0 = 0; t0 < mm - kk; t0 += nw) {
      const int ns = mm - kk - t0 < nw ? mm - kk - t0 : nw;
      shifted_train(T, Z, mm, ns, mus_s + t0);
    }
  } else {
    for (int t = kk; t < mm; ++t) shifted_sweep(T, Z, mm, ev[order[t]], log);
  }
  __syncthreads();
  if (lane == 0) st->clk[5] = clock64();
  for (int idx = lane; idx < mm * mm; idx += nthr) {
    int r = idx / mm, c = idx - r * mm;
    const cd z = Z.get(r, c), t = T.get(r, c);
    a.Qre[(size_t)r * a.m + c] = z.re;
    a.Qim[(size_t)r * a.m + c] = z.im;
    bool keep = r < kk && c < kk;
    a.Hre[(size_t)r * a.m + c] = keep ? t.re : 0.0;
    a.Him[(size_t)r * a.m + c] = keep ? t.im : 0.0;
  }
  if (lane == 0) {
    const cd bk = T.get(kk, kk - 1), sg = Z.get(mm - 1, kk - 1);
    st->beta_k_re = bk.re;
    st->beta_k_im = bk.im;
    st->sigma_re = sg.re;
    st->sigma_im = sg.im;
  }
  __syncthreads();
}

__device__ __forceinline__ void restart_body(const RestartArgs& a, double* smem) {
  restart_body_t<kMaxM, false>(a, smem);
}

}  // namespace
}  // namespace iram
}  // namespace oaph
