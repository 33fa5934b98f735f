// facade_example.cpp -- reference-style C++ caller of the drop-in facade
// (include/oap/oap.h).  It is written the way oap's own callers use
// CuHArnoldiCallback (oap2dt3dDevice/MainAPExecutor.cpp:129-174,
// oapArnoldiDeviceTests/oapArnoldiPackageMatricesTests.cpp:61-78): a device
// matrix, a MultiplyFunc that calls CuProceduresApi::dotProduct, setters,
// execute, testOutcome.  Input: a real n x n matrix in row-major text on
// stdin ("n" then n*n values); output: one JSON line.
#include <cstdio>
#include <vector>

#include "oap/oap.h"

static void multiply(math::Matrix* w, math::Matrix* v, oap::CuProceduresApi& api, void* user,
                     oap::VecMultiplicationType) {
  api.dotProduct(w, static_cast<math::Matrix*>(user), v);
}

int main() {
  unsigned n = 0;
  if (std::scanf("%u", &n) != 1 || n == 0) return 2;
  std::vector<double> a((size_t)n * n);
  for (auto& x : a)
    if (std::scanf("%lf", &x) != 1) return 2;
  try {
    math::Matrix* A = oap::cuda::NewDeviceReMatrix(n, n);
    oap::cuda::CopyHostArraysToDeviceMatrix(A, a.data(), nullptr);
    const unsigned k = 4, m = 16;
    std::vector<double> re(k), im(k);
    CuHArnoldiCallback arnoldi;
    arnoldi.setCallback(multiply, A);
    arnoldi.setQRType(oap::QRType::QRGR);
    arnoldi.setSortType(ArnUtils::SortLargestValues);
    arnoldi.setCheckTolerance(1e-10);
    arnoldi.setOutputsEigenvalues(re.data(), im.data());
    arnoldi.setOutputType(ArnUtils::DEVICE);
    math::MatrixInfo info(true, false, n, n);
    arnoldi.execute(m, k, info, ArnUtils::DEVICE);
    std::printf("{\"re\": [");
    for (unsigned t = 0; t < k; ++t) std::printf("%s%.17g", t ? ", " : "", re[t]);
    std::printf("], \"im\": [");
    for (unsigned t = 0; t < k; ++t) std::printf("%s%.17g", t ? ", " : "", im[t]);
    std::printf("], \"resid\": [");
    for (unsigned t = 0; t < k; ++t) std::printf("%s%.17g", t ? ", " : "", arnoldi.testOutcome(t));
    std::printf("]}\n");
    oap::cuda::DeleteDeviceMatrix(A);
  } catch (const oap::Error& e) {
    std::fprintf(stderr, "oap error: %s\n", e.what());
    return 1;
  }
  return 0;
}
