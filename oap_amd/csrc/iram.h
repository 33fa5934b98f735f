// iram.h -- shared declarations of the Arnoldi solver (host driver + kernels).
#pragma once

#include "internal.h"

namespace oaph {
namespace iram {

constexpr int kMaxM = 64;      // Krylov dimension limit of the LDS-resident restart kernel
constexpr int kMaxMBig = 256;  // Krylov dimension limit of the general restart kernel

// Device-resident solver status; the host reads it once per restart cycle.
struct Status {
  int nconv;
  int kk;
  int breakdown;   // first index j+1 whose Arnoldi residual vanished (0 = none)
  int schur_fail;
  double beta;     // norm of the latest residual vector f
  double beta_k_re, beta_k_im, sigma_re, sigma_im;
  double maxabs;   // max |Ritz value|
  double tolabs;   // absolute residual tolerance used for nconv
  double ev_re[kMaxMBig], ev_im[kMaxMBig], est[kMaxMBig];  // Ritz values in wanted order, estimates
  int order[kMaxMBig];
  long long clk[8];  // shader-clock stamps of the restart phases (profiling)
  int schur_path;    // last restart: 2 Aberth values + inverse iteration, 1 perfect-shift Schur form, 0 Wilkinson
  int ev_count;      // Ritz values the last restart left in ev_re / ev_im
};

// Device buffers the orthogonalisation kernels use.
struct Work {
  uint32_t n = 0, m = 0;
  double* Vre = nullptr;  // column c at Vre + c*n
  double* Vim = nullptr;
  double* Hre = nullptr;  // m x m, row-major, ld m
  double* Him = nullptr;
  double* partials = nullptr;
  unsigned* counter = nullptr;
  Status* st = nullptr;
};

struct RestartArgs {
  double* Hre;
  double* Him;
  double* Qre;
  double* Qim;
  double* Yre;  // m x ky Ritz vectors of H in wanted order (leading dimension ky)
  double* Yim;
  double* Xre;  // m x m work
  double* Xim;
  Status* st;
  int m;        // leading dimension of H, Q, X
  int mact;     // active size
  int k;
  int ky;       // Ritz vectors written to Y (k wanted + guards)
  int which;
  int do_shifts;
  double tol;
  double anorm;  // <= 0: relative to max |Ritz value|
  // 0: the whole restart with the built-in order `which`; 1: Schur form,
  // Ritz values and estimates only (status in natural order, T and Z stored
  // to TZ); 2: T and Z loaded from TZ, the order taken from the status (set
  // by the host's comparator), then the rest of the restart
  int phase = 0;
  int schur_method = 2;  // 2: Aberth eigenvalues + inverse iteration; 1: perfect-shift Schur form
  int warm = 0;          // 1: an earlier restart of this factorisation left its Ritz values
  int keep = 0;          // smallest kept size of an exact-shift restart (0: k + min(nconv, p / 2))
  double* TZ = nullptr;  // 4 m x m planes (T re, im, Z re, im), leading dimension m
                         // (always given when mact > kMaxM: the general kernel's T and Z)
};

// One Arnoldi step's CGS2 against V[:, 0..ncols): cgs_project (V^H x),
// cgs_update_project (f = x - V h, h stored), cgs_finish (f -= V s, s stored,
// H column hcol and beta into row hrow when hcol >= 0, v = f / beta when vre
// is given).  The three launches hand partial sums over w.partials.  The
// matvec that forms x can write cgs_project's partial sums itself
// (cgs_matvec_projection); cgs_update_project then sums p1_parts of them
// (0: the cgs_project launch's).
oaph_status cgs_project(hipStream_t s, const Work& w, uint32_t ncols, const double* xre,
                        const double* xim);
MvProj cgs_matvec_projection(const Work& w, uint32_t ncols);
oaph_status cgs_update_project(hipStream_t s, const Work& w, uint32_t ncols, const double* xre,
                               const double* xim, double* fre, double* fim, double2* h,
                               uint32_t p1_parts = 0);
oaph_status cgs_finish(hipStream_t s, const Work& w, uint32_t ncols, double* fre, double* fim,
                       const double2* h, double2* sc, int hcol, int hrow, double* vre,
                       double* vim);
// f <- f - V[:, 0..ncols) sub, status beta = ||f||, H[hrow][hcol] = beta when
// hcol >= 0 and hrow < m.
oaph_status residual(hipStream_t s, const Work& w, uint32_t ncols, double* fre, double* fim,
                     const double2* sub, int hcol, int hrow);
size_t orth_partials_bytes(uint32_t n);
oaph_status normalize(hipStream_t s, const Work& w, const double* fre, const double* fim,
                      double* vre, double* vim);
oaph_status restart(hipStream_t s, const RestartArgs& a);
// The BF16-split basis copy Vb (4 planes of (m + 1) x n: re hi, re lo, im hi,
// im lo) of columns [c0, c1), and the first Gram-Schmidt pass on the matrix
// cores from it (partial sums into the P1 area that cgs_update_project sums;
// *parts = their count).  Supported for n % 64 == 0 and ncols <= 64.
oaph_status v_to_bf16x2(hipStream_t s, uint32_t n, uint32_t m, uint32_t c0, uint32_t c1,
                       const double* Vre, const double* Vim, uint16_t* Vb);
bool proj_bf16x3_supported(uint32_t n, uint32_t ncols);
oaph_status cgs_project_bf16x3(hipStream_t s, const Work& w, uint32_t ncols, const double* xre,
                               const double* xim, const uint16_t* Vb, uint32_t* parts);
oaph_status vq(hipStream_t s, uint32_t n, uint32_t mrows, uint32_t ncols, const double* Vre,
               const double* Vim, const double* Mre, const double* Mim, uint32_t ldm,
               double* Ore, double* Oim, int fcol, double* fre, double* fim, const Status* st);
oaph_status pairs(hipStream_t s, int mode, uint32_t n, uint32_t k, const double* Xre,
                  const double* Xim, const double* Yre, const double* Yim, double2* lambda,
                  double2* out, double* partials, unsigned* counter);
size_t pairs_partials_bytes(uint32_t n, uint32_t k);
// dst = (float)src; flags (may be NULL): flags[0] set to 1 when an element
// rounds, flags[1] when a finite element overflows the FP32 range
// precision of the operator a matvec reads (the solver's FP64 operator or
// its reduced-precision copies)
enum : int { kOpFP64 = 0, kOpFP32 = 1, kOpF16 = 2 };
oaph_status to_f32(hipStream_t s, size_t n, const double* src, float* dst, unsigned* flags);
// *amax <- max(*amax, max |src|) over n values (bit patterns of non-negative
// doubles order as unsigned integers; a NaN sorts above every number)
oaph_status abs_max(hipStream_t s, size_t n, const double* src, unsigned long long* amax);
// dst = FP16 of src * scale (round to nearest even, once)
oaph_status to_f16(hipStream_t s, size_t n, const double* src, double scale, _Float16* dst);
// V[:, 0] = sum of the first k columns of X plus w times the next count - k
oaph_status colsum(hipStream_t s, uint32_t n, uint32_t k, uint32_t count, double w,
                   const double* Xre, const double* Xim, double* vre, double* vim);
oaph_status random_vector(hipStream_t s, uint32_t n, uint64_t seed, bool cplx, double* re,
                          double* im);

}  // namespace iram
}  // namespace oaph
