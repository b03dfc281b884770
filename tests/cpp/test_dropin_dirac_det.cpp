// The reference's DiracDeterminant unit tests (src/QMCWaveFunctions/tests/test_dirac_det.cpp,
// test_dirac_matrix.cpp), run against qmcx::DelayedUpdateHIP plugged in as DU_TYPE.
// MiniDeterminant below follows DiracDeterminant<DU_TYPE>'s calls into its engine
// (DiracDeterminant.cpp:49-155): invert_transpose on recompute, getInvRow + dot for a
// ratio, acceptRow on accept, updateInvMat on completeUpdates.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "qmcx_dropin.hpp"

template <typename T>
struct Matrix {
  int r = 0, c = 0;
  std::vector<T> v;
  Matrix() = default;
  Matrix(int rows, int cols) : r(rows), c(cols), v((size_t)rows * cols, T(0)) {}
  int rows() const { return r; }
  int cols() const { return c; }
  T* data() { return v.data(); }
  const T* data() const { return v.data(); }
  T& operator()(int i, int j) { return v[(size_t)i * c + j]; }
  const T& operator()(int i, int j) const { return v[(size_t)i * c + j]; }
};
template <typename T>
struct Vector {
  std::vector<T> v;
  explicit Vector(int n = 0) : v(n, T(0)) {}
  int size() const { return (int)v.size(); }
  T* data() { return v.data(); }
  const T* data() const { return v.data(); }
  T& operator[](int i) { return v[i]; }
};

static int failures = 0, checks = 0;
// Catch2 Approx as configured in the reference: |a - b| <= 100 FLT_EPSILON (1 + |b|)
static void require_approx(double a, double b, const char* what) {
  ++checks;
  if (std::fabs(a - b) > 100 * 1.1920928955078125e-07 * (1 + std::fabs(b))) {
    ++failures;
    std::printf("FAILED %s: %.12g vs %.12g\n", what, a, b);
  }
}

template <typename DU>
struct MiniDeterminant {
  int n;
  Matrix<double> psiM, psiM_temp;
  Vector<double> invRow;
  DU updateEng;
  double LogValue = 0, PhaseValue = 0, curRatio = 1;
  MiniDeterminant(int norb, int delay) : n(norb), psiM(norb, norb), psiM_temp(norb, norb), invRow(norb) {
    updateEng.resize(norb, delay);
  }
  void recompute(const Matrix<double>& logdet) {   // DiracDeterminant::recompute (:264-280)
    psiM_temp = logdet;
    updateEng.invert_transpose(psiM_temp, psiM, LogValue, PhaseValue);
  }
  double ratio(int row, const Vector<double>& psiV) {   // ratioGrad_compute (:103-124)
    updateEng.getInvRow(psiM, row, invRow);
    double s = 0;
    for (int j = 0; j < n; ++j) s += invRow.v[j] * psiV.v[j];
    curRatio = s;
    return s;
  }
  void acceptMove(int row, const Vector<double>& psiV) {   // acceptMove (:129-145)
    PhaseValue += curRatio > 0 ? 0.0 : M_PI;
    LogValue += std::log(std::fabs(curRatio));
    updateEng.acceptRow(psiM, row, psiV);
  }
  void completeUpdates() { updateEng.updateInvMat(psiM); }   // (:148-155)
};

// plain host reference: log|det| of a row-major matrix by Gaussian elimination with pivoting
static double host_logdet(Matrix<double> a) {
  const int n = a.rows();
  double ld = 0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int r = k + 1; r < n; ++r)
      if (std::fabs(a(r, k)) > std::fabs(a(p, k))) p = r;
    for (int c = 0; c < n; ++c) std::swap(a(k, c), a(p, c));
    ld += std::log(std::fabs(a(k, k)));
    for (int r = k + 1; r < n; ++r) {
      const double l = a(r, k) / a(k, k);
      for (int c = k; c < n; ++c) a(r, c) -= l * a(k, c);
    }
  }
  return ld;
}

using DU = qmcx::DelayedUpdateHIP<double, double>;

int main() {
  // FakeSPO (test_dirac_det.cpp:74-127)
  const double a3[3][3] = {{2.3, 4.5, 2.6}, {0.5, 8.5, 3.3}, {1.8, 4.4, 4.9}};
  const double v3[3] = {1.9, 2.0, 3.1};
  const double a4[4][4] = {{2.3, 4.5, 2.6, 1.2}, {0.5, 8.5, 3.3, 0.3}, {1.8, 4.4, 4.9, 2.8}, {0.8, 4.1, 3.2, 1.1}};
  const double v4[3][4] = {{3.2, 0.5, 5.9, 3.7}, {0.3, 1.4, 3.9, 8.2}, {3.3, 5.4, 4.9, 2.2}};

  {  // DiracDeterminant_first (test_dirac_det.cpp:196-248)
    MiniDeterminant<DU> ddb(3, 1);
    Matrix<double> logdet(3, 3);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) logdet(j, i) = a3[i][j];
    ddb.recompute(logdet);
    const double b[3][3] = {{0.6159749342, -0.2408954682, -0.1646081192},
                            {0.07923894288, 0.1496231042, -0.1428117337},
                            {-0.2974298429, -0.04586322768, 0.3927890292}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) require_approx(ddb.psiM(i, j), b[i][j], "first: psiM");
    Vector<double> psiV(3);
    for (int j = 0; j < 3; ++j) psiV[j] = v3[j];
    require_approx(ddb.ratio(0, psiV), 0.178276269185, "first: ratio");
    ddb.acceptMove(0, psiV);
    ddb.completeUpdates();
    const double b2[3][3] = {{3.455170657, -1.35124809, -0.9233316353},
                             {0.05476311768, 0.1591951095, -0.1362710138},
                             {-2.235099338, 0.7119205298, 0.9105960265}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) require_approx(ddb.psiM(i, j), b2[i][j], "first: updated psiM");
  }

  for (int delay : {1, 2}) {  // DiracDeterminant_second / _delayed_update (test_dirac_det.cpp:252-515)
    MiniDeterminant<DU> ddc(4, delay);
    Matrix<double> logdet(4, 4);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) logdet(j, i) = a4[i][j];
    ddc.recompute(logdet);
    Matrix<double> cur = logdet;
    double logcur = host_logdet(cur);
    for (int r = 0; r < 3; ++r) {
      Vector<double> psiV(4);
      for (int j = 0; j < 4; ++j) psiV[j] = v4[r][j];
      const double ratio = ddc.ratio(r, psiV);
      for (int j = 0; j < 4; ++j) cur(r, j) = v4[r][j];
      const double lognew = host_logdet(cur);
      require_approx(std::fabs(ratio), std::exp(lognew - logcur), "sequence: |ratio|");
      logcur = lognew;
      ddc.acceptMove(r, psiV);
      if (delay == 2 && r == 0) ddc.completeUpdates();   // the reference test forces the SM path here
    }
    ddc.completeUpdates();
    // fresh inverse of the final matrix through a second engine
    MiniDeterminant<DU> fresh(4, 1);
    fresh.recompute(cur);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) require_approx(ddc.psiM(i, j), fresh.psiM(i, j), "sequence: final psiM");
    require_approx(ddc.LogValue, fresh.LogValue, "sequence: log value");
  }

  std::printf("%s: %d checks, %d failures\n", failures ? "FAILED" : "All tests passed", checks, failures);
  return failures ? 1 : 0;
}
