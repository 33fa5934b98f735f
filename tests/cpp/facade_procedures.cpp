// facade_procedures.cpp -- reference-style C++ caller of the procedures and
// matrix utilities of the drop-in facade (include/oap/oap.h) beyond the
// Arnoldi path: MatrixEx with transposeEx (as oapDeviceTests/
// oapMatrixCudaTest.cpp:280-305 calls it), the periodic product, the
// (columns, rows) forms, addSubstract, magnitude2, calcTriangularHStep,
// SetMatrix, the dims copies and the value accessors.  Fixed small inputs;
// the results are printed as one JSON line for the test to check.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "oap/oap.h"

static std::string arr(const std::vector<double>& v) {
  std::string s = "[";
  char buf[64];
  for (size_t i = 0; i < v.size(); ++i) {
    std::snprintf(buf, sizeof(buf), "%s%.17g", i ? "," : "", v[i]);
    s += buf;
  }
  return s + "]";
}

static std::vector<double> re_of(const math::Matrix* m) {
  std::vector<double> v((size_t)m->columns * m->rows);
  oap::cuda::CopyDeviceMatrixToHostArrays(v.data(), nullptr, m);
  return v;
}

// a comparator with no built-in counterpart: largest real part of
// lambda e^{-i pi/4}
static bool byRotatedRe(const EigenPair& i, const EigenPair& j) {
  return (i.re() + i.im()) > (j.re() + j.im());
}

int main() {
  try {
    oap::CuProceduresApi api;
    // transposeEx: 4 x 10 input (columns 4, rows 10), region {0, 10, 0, 2}
    std::vector<double> in(40);
    for (size_t i = 0; i < in.size(); ++i) in[i] = (double)(i % 7) - 2.0;
    math::Matrix* din = oap::cuda::NewDeviceReMatrix(4, 10);
    oap::cuda::CopyHostArraysToDeviceMatrix(din, in.data(), nullptr);
    math::Matrix* dout = oap::cuda::NewDeviceReMatrix(10, 4);
    api.setZeroMatrix(dout);
    MatrixEx* ex = oap::cuda::NewDeviceMatrixEx();
    MatrixEx hex = {0, 10, 0, 2};
    oap::cuda::SetMatrixEx(ex, &hex);
    api.transposeEx(dout, din, ex);
    oap::cuda::DeleteDeviceMatrixEx(ex);
    const std::vector<double> tex = re_of(dout);

    // dotProductPeriodic: A 2 x 2, B 2 x 6 (columns 2, rows 6) -> C 2 x 6
    std::vector<double> a = {1, 2, 3, 4}, b = {1, 0, 0, 1, 2, 1, 1, 3, -1, 2, 0.5, 4};
    math::Matrix* dA = oap::cuda::NewDeviceReMatrix(2, 2);
    math::Matrix* dB = oap::cuda::NewDeviceReMatrix(2, 6);
    math::Matrix* dC = oap::cuda::NewDeviceReMatrix(2, 6);
    oap::cuda::CopyHostArraysToDeviceMatrix(dA, a.data(), nullptr);
    oap::cuda::CopyHostArraysToDeviceMatrix(dB, b.data(), nullptr);
    api.dotProductPeriodic(dC, dA, dB);
    const std::vector<double> per = re_of(dC);

    // addSubstract over 2 x 1 of 2 x 2 matrices: out += a - b
    math::Matrix* dO = oap::cuda::NewDeviceReMatrix(2, 2);
    std::vector<double> o = {10, 20, 30, 40};
    oap::cuda::CopyHostArraysToDeviceMatrix(dO, o.data(), nullptr);
    api.addSubstract(dO, dA, dA, 2, 1);
    api.add(dO, dO, dA, 1, 2);
    const std::vector<double> addsub = re_of(dO);

    floatt m2 = 0, m1 = 0;
    api.magnitude2(m2, dA);
    api.magnitudeOpt(m1, dA);

    // calcTriangularHStep on a 3 x 3 Hessenberg matrix
    std::vector<double> h = {4, 1, 2, 3, 5, 1, 0, 2, 6};
    math::Matrix* dH = oap::cuda::NewDeviceReMatrix(3, 3);
    math::Matrix* dQ = oap::cuda::NewDeviceReMatrix(3, 3);
    oap::cuda::CopyHostArraysToDeviceMatrix(dH, h.data(), nullptr);
    api.calcTriangularHStep(dH, dQ, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    const std::vector<double> h1 = re_of(dH), q1 = re_of(dQ);

    // SetMatrix, the dims copy and the value accessors
    math::Matrix* dM = oap::cuda::NewDeviceReMatrix(4, 3);
    api.setZeroMatrix(dM);
    oap::cuda::SetReMatrix(dM, dA, 1, 1);
    oap::cuda::SetReValue(dM, 9.5, 0, 2);
    const double v21 = oap::cuda::GetReValue(dM, 2, 1);
    std::vector<double> host(4 * 3, -1.0);
    math::Matrix hm = {4, 3, 4, 3, host.data(), nullptr};
    uintt dims[2][2][2] = {{{0, 2}, {0, 2}}, {{1, 2}, {1, 2}}};
    oap::cuda::CopyDeviceMatrixToHostMatrixDims(&hm, dM, dims);
    oap::cuda::SetZeroRow(dM, 1);
    const std::vector<double> mset = re_of(dM);

    // CuHArnoldiDefault with the caller's comparator on a 64 x 64 complex
    // operator built from a formula the test repeats
    const unsigned n = 64;
    std::vector<double> hre(n * n), him(n * n);
    for (unsigned r = 0; r < n; ++r)
      for (unsigned c = 0; c < n; ++c) {
        const double off = 0.01 * std::sin(7.0 * r + 3.0 * c);
        hre[r * n + c] = r == c ? std::cos(0.1 * r) / (1.0 + 0.05 * r) : off;
        him[r * n + c] = r == c ? std::sin(0.1 * r) / (1.0 + 0.05 * r) : 0.5 * off;
      }
    math::Matrix* dOp = oap::cuda::NewDeviceMatrix(true, true, n, n);
    oap::cuda::CopyHostArraysToDeviceMatrix(dOp, hre.data(), him.data());
    std::vector<double> ere(3), eim(3);
    CuHArnoldiDefault arnoldi;
    arnoldi.setMatrix(dOp);
    arnoldi.setSortType(byRotatedRe);
    arnoldi.setCheckTolerance(1e-10);
    arnoldi.setOutputType(ArnUtils::DEVICE);
    arnoldi.setOutputsEigenvalues(ere.data(), eim.data());
    arnoldi.execute(24, 3, math::MatrixInfo(true, true, n, n));
    std::vector<double> eig = {ere[0], eim[0], ere[1], eim[1], ere[2], eim[2]};
    oap::cuda::DeleteDeviceMatrix(dOp);

    std::printf("{\"eig_rotated\": %s, \"transpose_ex\": %s, \"periodic\": %s, \"addsub\": %s, \"magnitude2\": %.17g, "
                "\"magnitude\": %.17g, \"h1\": %s, \"q1\": %s, \"value21\": %.17g, "
                "\"dims_host\": %s, \"set_matrix\": %s}\n",
                arr(eig).c_str(), arr(tex).c_str(), arr(per).c_str(), arr(addsub).c_str(), m2, m1, arr(h1).c_str(),
                arr(q1).c_str(), v21, arr(host).c_str(), arr(mset).c_str());
    for (math::Matrix* m : {din, dout, dA, dB, dC, dO, dH, dQ, dM}) oap::cuda::DeleteDeviceMatrix(m);
  } catch (const oap::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
