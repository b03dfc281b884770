// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A from-scratch, all-FP64 CPU restatement of miniqmc's VMC particle-move
// path.  It is the checker for the HIP product in miniqmc_amd/: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only to
// compare against or to time a CPU baseline.  Nothing in the product links it.
//
// Reference semantics restated here (file:line under the miniqmc tree):
//   * Philox4x64-10 counter RNG + Box-Muller: the project's own walker stream
//     (the reference uses std::mt19937 + BoxMuller2, Utilities/StdRandom.h:19-119,
//     Utilities/RandomGenerator.h:46-79; its libstdc++ stream is parity unpinned).
//   * CrystalLattice::toUnit_floor            Particle/Lattice/CrystalLattice.h:185-196
//   * find_reduced_basis / found_shorter_base Particle/Lattice/LatticeAnalyzer.h:134-184
//   * calcWignerSeitzRadius                   Particle/Lattice/LatticeAnalyzer.h:86-116
//   * DTD_BConds<T,3,PPPG+SOA>::computeDistances Particle/Lattice/ParticleBConds.h:99-217
//   * DistanceTableAA/BA row semantics        Particle/DistanceTableAA.h:85-113, DistanceTableBA.h:83-105
//   * getSplineBound / prefactors / evaluate_vgh Numerics/Spline2/MultiBsplineEvalHelper.hpp:43-113,
//                                             MultiBsplineData.hpp:52-80, MultiBspline.hpp:163-280
//   * einspline_spo::evaluate (d2psi = h_xx)  QMCWaveFunctions/einspline_spo.hpp:204-223
//   * BsplineFunctor                          QMCWaveFunctions/Jastrow/BsplineFunctor.h:93-339
//   * OneBodyJastrow                          QMCWaveFunctions/Jastrow/OneBodyJastrow.h:110-283
//   * TwoBodyJastrow                          QMCWaveFunctions/Jastrow/TwoBodyJastrow.h:251-430
//   * ThreeBodyJastrow (eeI) with PolynomialFunctor3D
//                                             QMCWaveFunctions/Jastrow/ThreeBodyJastrow.h:30-677,
//                                             Jastrow/PolynomialFunctor3D.h:34-560
//   * DiracMatrix::invert_transpose, computeLogDet QMCWaveFunctions/DiracMatrix.h:26-121
//   * DelayedUpdate getInvRow/acceptRow/updateInvMat QMCWaveFunctions/DelayedUpdate.h:96-231
//   * DiracDeterminant                        QMCWaveFunctions/DiracDeterminant.cpp:80-280
//   * WaveFunction                            QMCWaveFunctions/WaveFunction.cpp:178-279
//   * VMC sweep                               Drivers/miniqmc.cpp:422-467
//   * WaveFunction::evaluateRatios (virtual moves of NonLocalPP)
//                                             QMCWaveFunctions/WaveFunction.cpp:281-297,
//                                             DiracDeterminant.cpp:216-224, SPOSet.h:72-83,
//                                             TwoBodyJastrow.h:103-107, OneBodyJastrow.h:136-140
//
// Every function here is pinned by tests/test_oracle.py against the
// reference's own KATs (test_dirac_matrix.cpp, test_dirac_det.cpp,
// test_particle.cpp) or against closed forms / brute force; see the header of
// each test for which check pins which function.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// Hot loops get an AVX2 clone beside the baseline x86-64 one; the loader picks
// the clone by the host CPU's feature bits.  No FMA contraction and each output
// element keeps its own sequential summation order (vectorised across
// elements), so both clones give bit-identical results.
#define ORC_HOT __attribute__((target_clones("avx2", "default")))

// ---------------------------------------------------------------- RNG
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11), published constants.
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
  unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
  hi = (uint64_t)(p >> 64);
  lo = (uint64_t)p;
}

static void philox4x64_10(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
  uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint64_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += W0; k1 += W1; }
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(M0, c0, hi0, lo0);
    mulhilo64(M1, c2, hi1, lo1);
    const uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

// The walker move stream: for (seed, walker, sweep, electron) four uniforms in
// [0,1) from one Philox block (counter = {electron, sweep, walker, 0},
// key = {seed, 0x51A7}).  u[0..1] -> Box-Muller pair (n0,n1), u[2..3] -> (n2, n3);
// the acceptance uniform comes from a second block with counter word 3 = 1.
static void move_randoms(uint64_t seed, uint64_t walker, uint64_t sweep, uint64_t iel,
                         double normal[3], double& uacc) {
  const uint64_t key[2] = {seed, 0x51A7ULL};
  uint64_t ctr[4] = {iel, sweep, walker, 0};
  uint64_t o[4];
  philox4x64_10(ctr, key, o);
  const double two_pi = 6.283185307179586476925286766559;
  double r0 = std::sqrt(-2.0 * std::log(1.0 - u01(o[0])));
  double r1 = std::sqrt(-2.0 * std::log(1.0 - u01(o[2])));
  normal[0] = r0 * std::cos(two_pi * u01(o[1]));
  normal[1] = r0 * std::sin(two_pi * u01(o[1]));
  normal[2] = r1 * std::cos(two_pi * u01(o[3]));
  ctr[3] = 1;
  philox4x64_10(ctr, key, o);
  uacc = u01(o[0]);
}

// ---------------------------------------------------------------- 3x3 helpers
static void inverse3(const double a[3][3], double inv[3][3]) {
  const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  const double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
  const double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
  const double id = 1.0 / det;
  inv[0][0] = c00 * id;
  inv[1][0] = c01 * id;
  inv[2][0] = c02 * id;
  inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
  inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
}
static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// ---------------------------------------------------------------- lattice
struct Lattice {
  double R[3][3];        // rows are the cell vectors a_i
  double G[3][3];        // R^-1: u = r . G
  double rb[3][3];       // reduced basis (rows)
  double g[3][3];        // inverse of rb
  double corner[8][3];
  double ws_radius;
};

// LatticeAnalyzer.h:134-158: try to replace the longest vector by a shorter combination.
static bool shorter_base(double rb[3][3]) {
  const double eps = 10.0 * std::numeric_limits<float>::epsilon();
  int imax = 0;
  double r2max = dot3(rb[0], rb[0]);
  for (int i = 1; i < 3; ++i) {
    double r2 = dot3(rb[i], rb[i]);
    if (r2 - r2max > eps) { r2max = r2; imax = i; }
  }
  double cand[4][3];
  for (int d = 0; d < 3; ++d) {
    cand[0][d] = rb[0][d] + rb[1][d] - rb[2][d];
    cand[1][d] = rb[0][d] + rb[2][d] - rb[1][d];
    cand[2][d] = rb[1][d] + rb[2][d] - rb[0][d];
    cand[3][d] = rb[0][d] + rb[1][d] + rb[2][d];
  }
  for (int i = 0; i < 4; ++i)
    if (dot3(cand[i], cand[i]) - r2max < -eps) {
      for (int d = 0; d < 3; ++d) rb[imax][d] = cand[i][d];
      return true;
    }
  return false;
}

// LatticeAnalyzer.h:160-184
static void reduce_basis(double rb[3][3]) {
  for (;;) {
    double saved[3][3];
    std::memcpy(saved, rb, sizeof(saved));
    bool changed = false;
    for (int i = 0; i < 3; ++i) {
      rb[i][0] = rb[i][1] = rb[i][2] = 0.0;
      changed = shorter_base(rb);
      for (int d = 0; d < 3; ++d) rb[i][d] = saved[i][d];
      if (changed) break;
    }
    if (!changed && !shorter_base(rb)) return;
  }
}

static void lattice_set(Lattice& L, const double* R9) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) L.R[i][j] = R9[3 * i + j];
  inverse3(L.R, L.G);
  // Wigner-Seitz radius: half the shortest non-zero lattice translation over {-1,0,1}^3.
  double rmin = 0.5 * std::numeric_limits<double>::max();
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j)
      for (int k = -1; k <= 1; ++k)
        if (i || j || k) {
          double v[3];
          for (int d = 0; d < 3; ++d) v[d] = i * L.R[0][d] + j * L.R[1][d] + k * L.R[2][d];
          rmin = std::min(rmin, dot3(v, v));
        }
  L.ws_radius = 0.5 * std::sqrt(rmin);
  std::memcpy(L.rb, L.R, sizeof(L.rb));
  reduce_basis(L.rb);
  inverse3(L.rb, L.g);
  const double (*b)[3] = L.rb;
  for (int d = 0; d < 3; ++d) {
    L.corner[0][d] = 0.0;
    L.corner[1][d] = -b[0][d];
    L.corner[2][d] = -b[1][d];
    L.corner[3][d] = -b[2][d];
    L.corner[4][d] = -(b[0][d] + b[1][d]);
    L.corner[5][d] = -(b[0][d] + b[2][d]);
    L.corner[6][d] = -(b[1][d] + b[2][d]);
    L.corner[7][d] = -(b[0][d] + b[1][d] + b[2][d]);
  }
}

// CrystalLattice.h:185-196
static void to_unit_floor(const double G[3][3], const double r[3], double u[3]) {
  const double eps = std::numeric_limits<double>::epsilon();
  for (int j = 0; j < 3; ++j) {
    double v = r[0] * G[0][j] + r[1] * G[1][j] + r[2] * G[2][j];
    if (-eps < v && v < 0.0) v = 0.0;
    else v -= std::floor(v);
    u[j] = v;
  }
}

// One target range [t0, t1) with a fixed flip sign; branch-free so that the
// corner search vectorises across targets.
ORC_HOT static void min_image_range(const Lattice& L, const double pos[3], const double* __restrict px,
                                    const double* __restrict py, const double* __restrict pz, int t0, int t1,
                                    double flip, double* __restrict r, double* __restrict dx,
                                    double* __restrict dy, double* __restrict dz) {
  const double g00 = L.g[0][0], g01 = L.g[0][1], g02 = L.g[0][2];
  const double g10 = L.g[1][0], g11 = L.g[1][1], g12 = L.g[1][2];
  const double g20 = L.g[2][0], g21 = L.g[2][1], g22 = L.g[2][2];
  const double r00 = L.rb[0][0], r01 = L.rb[0][1], r02 = L.rb[0][2];
  const double r10 = L.rb[1][0], r11 = L.rb[1][1], r12 = L.rb[1][2];
  const double r20 = L.rb[2][0], r21 = L.rb[2][1], r22 = L.rb[2][2];
  double cx[8], cy[8], cz[8];
  for (int c = 0; c < 8; ++c) { cx[c] = L.corner[c][0]; cy[c] = L.corner[c][1]; cz[c] = L.corner[c][2]; }
  const double p0 = pos[0], p1 = pos[1], p2 = pos[2];
  for (int t = t0; t < t1; ++t) {
    const double d0 = (px[t] - p0) * flip;
    const double d1 = (py[t] - p1) * flip;
    const double d2 = (pz[t] - p2) * flip;
    const double s0 = -std::floor(d0 * g00 + d1 * g10 + d2 * g20);
    const double s1 = -std::floor(d0 * g01 + d1 * g11 + d2 * g21);
    const double s2 = -std::floor(d0 * g02 + d1 * g12 + d2 * g22);
    const double ex = d0 + s0 * r00 + s1 * r10 + s2 * r20;
    const double ey = d1 + s0 * r01 + s1 * r11 + s2 * r21;
    const double ez = d2 + s0 * r02 + s1 * r12 + s2 * r22;
    double best = ex * ex + ey * ey + ez * ez, bx = ex, by = ey, bz = ez;
#pragma GCC unroll 8
    for (int c = 1; c < 8; ++c) {   // the first strictly smaller corner wins
      const double x = ex + cx[c], y = ey + cy[c], z = ez + cz[c];
      const double q = x * x + y * y + z * z;
      const bool lt = q < best;
      best = lt ? q : best;
      bx = lt ? x : bx;
      by = lt ? y : by;
      bz = lt ? z : bz;
    }
    r[t] = std::sqrt(best);
    dx[t] = flip * bx;
    dy[t] = flip * by;
    dz[t] = flip * bz;
  }
}

// Minimum-image distances from `pos` to targets [0,n) stored SoA (x[], y[], z[]).
// flip_ind: targets with index < flip_ind use (target - pos), the others
// use -(pos - target) so that row/column pairs are exact negatives.
// ParticleBConds.h:154-217.
static void min_image_row(const Lattice& L, const double pos[3], const double* px, const double* py,
                          const double* pz, int n, int flip_ind, double* r, double* dx, double* dy, double* dz) {
  const int f = std::min(std::max(flip_ind, 0), n);
  min_image_range(L, pos, px, py, pz, 0, f, 1.0, r, dx, dy, dz);
  min_image_range(L, pos, px, py, pz, f, n, -1.0, r, dx, dy, dz);
}

// ---------------------------------------------------------------- radial B-spline functor
// Uniform cubic B-spline basis, in power form t^3, t^2, t, 1 for the four
// coefficients of a knot interval (BsplineFunctor.h:64-86 uses the same basis).
static const double BS_V[4][4] = {{-1.0 / 6, 3.0 / 6, -3.0 / 6, 1.0 / 6},
                                  {3.0 / 6, -6.0 / 6, 0.0, 4.0 / 6},
                                  {-3.0 / 6, 3.0 / 6, 3.0 / 6, 1.0 / 6},
                                  {1.0 / 6, 0.0, 0.0, 0.0}};
// derivative rows: coefficients of t^2, t, 1
static const double BS_D[4][3] = {{-0.5, 1.0, -0.5}, {1.5, -2.0, 0.0}, {-1.5, 1.0, 0.5}, {0.5, 0.0, 0.0}};
// second-derivative rows: coefficients of t, 1
static const double BS_DD[4][2] = {{-1.0, 1.0}, {3.0, -2.0}, {-3.0, 1.0}, {1.0, 0.0}};

struct Functor {
  int nparams = 0;
  double rcut = 0.0, cusp = 0.0, dr = 0.0, drinv = 0.0;
  std::vector<double> coef;  // nparams + 4
  bool present = false;
};

// BsplineFunctor.h:93-130 (resize/reset/setupParameters)
static void functor_setup(Functor& f, int n, double rcut, double cusp, const double* p) {
  f.nparams = n;
  f.rcut = rcut;
  f.cusp = cusp;
  f.dr = rcut / (double)(n + 1);
  f.drinv = 1.0 / f.dr;
  f.coef.assign(n + 4, 0.0);
  f.coef[1] = p[0];
  f.coef[2] = p[1];
  f.coef[0] = p[1] - 2.0 * f.dr * cusp;
  for (int i = 2; i < n; ++i) f.coef[i + 1] = p[i];
  f.present = true;
}

// u, (du/dr)/r, d2u/dr2 at one distance (BsplineFunctor.h:265-339 per element).
static inline void functor_vgl(const Functor& f, double r, double& u, double& du, double& d2u) {
  const double rinv = 1.0 / r;
  const double x = r * f.drinv;
  const int i = (int)x;
  const double t = x - (double)i;
  const double t2 = t * t, t3 = t2 * t;
  const double* c = &f.coef[i];
  double v = 0, d = 0, dd = 0;
  for (int m = 0; m < 4; ++m) {
    v += c[m] * (BS_V[m][0] * t3 + BS_V[m][1] * t2 + BS_V[m][2] * t + BS_V[m][3]);
    d += c[m] * (BS_D[m][0] * t2 + BS_D[m][1] * t + BS_D[m][2]);
    dd += c[m] * (BS_DD[m][0] * t + BS_DD[m][1]);
  }
  u = v;
  du = f.drinv * rinv * d;
  d2u = f.drinv * f.drinv * dd;
}

static inline double functor_v(const Functor& f, double r) {
  const double x = r * f.drinv;
  const int i = (int)x;
  const double t = x - (double)i;
  const double t2 = t * t, t3 = t2 * t;
  const double* c = &f.coef[i];
  double v = 0;
  for (int m = 0; m < 4; ++m) v += c[m] * (BS_V[m][0] * t3 + BS_V[m][1] * t2 + BS_V[m][2] * t + BS_V[m][3]);
  return v;
}

// ---------------------------------------------------------------- three-body (eeI) polynomial functor
// PolynomialFunctor3D (Jastrow/PolynomialFunctor3D.h:34-560): with L = rcut/2 and C = 3,
//   F(r12, r1I, r2I) = sum_{l,m <= neI, n <= nee} gamma(l,m,n) r1I^l r2I^m r12^n
//                      * ((r1I - L)(r2I - L))^C               for r1I < L and r2I < L, else 0,
// gamma symmetric in (l, m).  The free parameters fill gamma after the linear constraints are
// eliminated (resize :65-172, reset_gamma :180-260):
//   e-e no cusp, k = 0..2 neI:       sum_{l+m=k} gamma(l,m,1) = 0
//   e-I no cusp, k = 0..neI+nee:     sum_{m+n=k} C gamma(0,m,n) - L gamma(1,m,n) = 0
// by row reduction of the constraint matrix over the unknowns (m <= l), scanning columns in
// order with the largest pivot of each column; the columns left without a pivot take the
// parameters in order, and each pivot column is solved from its row.
struct Poly3D {
  bool present = false;
  int neI = 0, nee = 0;
  double L = 0.0;
  std::vector<double> g;   // gamma[l][m][n], (neI+1) x (neI+1) x (nee+1)
  inline double gamma(int l, int m, int n) const { return g[((size_t)l * (neI + 1) + m) * (nee + 1) + n]; }
};
constexpr int kPolyC = 3;

// returns the number of free parameters the shape takes, or -1 when nparams differs
static int poly3d_setup(Poly3D& f, int neI, int nee, double rcut, const double* params, int nparams) {
  const double L = 0.5 * rcut;
  // unknowns: (l, m, n) with m <= l, numbered for m, then l >= m, then n
  std::vector<int> idx((size_t)(neI + 1) * (neI + 1) * (nee + 1), -1);
  auto at = [&](int l, int m, int n) -> int& { return idx[((size_t)l * (neI + 1) + m) * (nee + 1) + n]; };
  int nu = 0;
  for (int m = 0; m <= neI; ++m)
    for (int l = m; l <= neI; ++l)
      for (int n = 0; n <= nee; ++n) at(l, m, n) = at(m, l, n) = nu++;
  const int nc = (2 * neI + 1) + (neI + nee + 1);
  if (nparams != nu - nc) return -1;
  std::vector<double> A((size_t)nc * nu, 0.0);
  auto a = [&](int r, int c) -> double& { return A[(size_t)r * nu + c]; };
  for (int k = 0; k <= 2 * neI; ++k)
    for (int m = 0; m <= k; ++m) {
      const int l = k - m;
      if (l > neI || m > neI || l < m) continue;
      a(k, at(l, m, 1)) = l > m ? 2.0 : 1.0;   // gamma(l,m,1) and gamma(m,l,1) are one unknown
    }
  const int r0 = 2 * neI + 1;
  for (int k = 0; k <= neI + nee; ++k) {
    // (m, n) with m + n = k: C gamma(0,m,n) - L gamma(1,m,n)
    for (int m = 0; m <= std::min(k, neI); ++m) {
      const int n = k - m;
      if (n > nee) continue;
      a(r0 + k, at(0, m, n)) = (double)kPolyC;
      a(r0 + k, at(1, m, n)) = -L;
    }
  }
  // reduced row echelon form, pivot columns in scan order
  std::vector<int> pivot_col(nc, -1);
  std::vector<char> free_col(nu, 1);
  int col = -1;
  for (int r = 0; r < nc; ++r) {
    int best = r;
    double bmax = 0.0;
    do {
      ++col;
      best = r;
      bmax = std::fabs(a(r, col));
      for (int q = r + 1; q < nc; ++q)
        if (std::fabs(a(q, col)) > bmax) { bmax = std::fabs(a(q, col)); best = q; }
    } while (bmax < 1e-6);
    for (int c = 0; c < nu; ++c) std::swap(a(r, c), a(best, c));
    const double inv = 1.0 / a(r, col);
    for (int c = 0; c < nu; ++c) a(r, c) *= inv;
    for (int q = 0; q < nc; ++q) {
      if (q == r) continue;
      const double v = a(q, col);
      if (v != 0.0)
        for (int c = 0; c < nu; ++c) a(q, c) -= v * a(r, c);
    }
    pivot_col[r] = col;
    free_col[col] = 0;
  }
  std::vector<double> gv(nu, 0.0);
  int p = 0;
  for (int c = 0; c < nu; ++c)
    if (free_col[c]) gv[c] = params[p++];
  for (int r = 0; r < nc; ++r) {
    const int c0 = pivot_col[r];
    double v = 0.0;
    for (int c = 0; c < nu; ++c)
      if (c != c0) v -= a(r, c) * gv[c];
    gv[c0] = v;
  }
  f.present = true;
  f.neI = neI;
  f.nee = nee;
  f.L = L;
  f.g.assign((size_t)(neI + 1) * (neI + 1) * (nee + 1), 0.0);
  for (int l = 0; l <= neI; ++l)
    for (int m = 0; m <= neI; ++m)
      for (int n = 0; n <= nee; ++n) f.g[((size_t)l * (neI + 1) + m) * (nee + 1) + n] = gv[at(l, m, n)];
  return nu - nc;
}

// Value and derivatives of F at (r12, r1, r2) (r1 < L, r2 < L), in the layout of
// evaluateVGL (PolynomialFunctor3D.h:445-556): out = {F, F_12/r12, F_1/r1, F_2/r2, F_12,12,
// F_1,1, F_2,2, F_12,1/(r12 r1), F_12,2/(r12 r2)}.  The polynomial P and its partials are summed
// directly, then the cutoff factor B^C, B = (r1 - L)(r2 - L), enters by the product rule.
static void poly3d_vgl(const Poly3D& f, double r12, double r1, double r2, double out[9]) {
  const int NE = f.neI, NN = f.nee;
  double pw1[8], pw2[8], pw0[8];
  pw1[0] = pw2[0] = pw0[0] = 1.0;
  for (int i = 1; i <= std::max(NE, NN); ++i) {
    pw1[i] = pw1[i - 1] * r1;
    pw2[i] = pw2[i - 1] * r2;
    pw0[i] = pw0[i - 1] * r12;
  }
  auto pw = [](const double* p, int e) { return e < 0 ? 0.0 : p[e]; };
  double P = 0, P0 = 0, P1 = 0, P2 = 0, P00 = 0, P11 = 0, P22 = 0, P01 = 0, P02 = 0;
  for (int l = 0; l <= NE; ++l)
    for (int m = 0; m <= NE; ++m)
      for (int n = 0; n <= NN; ++n) {
        const double g = f.gamma(l, m, n);
        if (g == 0.0) continue;
        P += g * pw1[l] * pw2[m] * pw0[n];
        P0 += g * pw1[l] * pw2[m] * n * pw(pw0, n - 1);
        P1 += g * l * pw(pw1, l - 1) * pw2[m] * pw0[n];
        P2 += g * pw1[l] * m * pw(pw2, m - 1) * pw0[n];
        P00 += g * pw1[l] * pw2[m] * n * (n - 1) * pw(pw0, n - 2);
        P11 += g * l * (l - 1) * pw(pw1, l - 2) * pw2[m] * pw0[n];
        P22 += g * pw1[l] * m * (m - 1) * pw(pw2, m - 2) * pw0[n];
        P01 += g * l * pw(pw1, l - 1) * pw2[m] * n * pw(pw0, n - 1);
        P02 += g * pw1[l] * m * pw(pw2, m - 1) * n * pw(pw0, n - 1);
      }
  const double a1 = r1 - f.L, a2 = r2 - f.L, B = a1 * a2;
  const double BC = std::pow(B, kPolyC), BC1 = std::pow(B, kPolyC - 1), BC2 = std::pow(B, kPolyC - 2);
  const double C = (double)kPolyC;
  const double F = P * BC;
  const double F0 = P0 * BC, F00 = P00 * BC;
  const double F1 = P1 * BC + P * C * BC1 * a2;
  const double F2 = P2 * BC + P * C * BC1 * a1;
  const double F11 = P11 * BC + 2.0 * P1 * C * BC1 * a2 + P * C * (C - 1.0) * BC2 * a2 * a2;
  const double F22 = P22 * BC + 2.0 * P2 * C * BC1 * a1 + P * C * (C - 1.0) * BC2 * a1 * a1;
  const double F01 = P01 * BC + P0 * C * BC1 * a2;
  const double F02 = P02 * BC + P0 * C * BC1 * a1;
  out[0] = F;
  out[1] = F0 / r12;
  out[2] = F1 / r1;
  out[3] = F2 / r2;
  out[4] = F00;
  out[5] = F11;
  out[6] = F22;
  out[7] = F01 / (r12 * r1);
  out[8] = F02 / (r12 * r2);
}

static double poly3d_v(const Poly3D& f, double r12, double r1, double r2) {
  double P = 0;
  double p1 = 1.0;
  for (int l = 0; l <= f.neI; ++l, p1 *= r1) {
    double p2 = 1.0;
    for (int m = 0; m <= f.neI; ++m, p2 *= r2) {
      double p0 = 1.0;
      for (int n = 0; n <= f.nee; ++n, p0 *= r12) P += f.gamma(l, m, n) * p1 * p2 * p0;
    }
  }
  return P * std::pow((r1 - f.L) * (r2 - f.L), kPolyC);
}

// ---------------------------------------------------------------- 3D multi B-spline
struct Spline {
  int ng[3] = {0, 0, 0};
  int norb = 0, npad = 0;
  int prec = 4;                 // 4 = float table, 8 = double table
  const float* cf = nullptr;    // borrowed
  const double* cd = nullptr;   // borrowed
  Lattice lat;                  // lattice used to map r -> u (SPOSet Lattice)
  inline double coef(size_t idx) const { return prec == 4 ? (double)cf[idx] : cd[idx]; }
};

// MultiBsplineEvalHelper.hpp:43-64
static inline void spline_bound(double x, double& t, int& ind, int nmax) {
  if (x < 0) {
    ind = 0;
    t = 0.0;
  } else {
    double ip;
    t = std::modf(x, &ip);
    ind = (int)ip;
    if (ind > nmax) {
      ind = nmax;
      t = 1.0 - std::numeric_limits<double>::epsilon();
    }
  }
}

// cubic B-spline weights, first and second derivatives at fraction t
// (MultiBsplineData.hpp:52-80)
static inline void weights(double t, double a[4], double da[4], double d2a[4]) {
  for (int m = 0; m < 4; ++m) {
    a[m] = ((BS_V[m][0] * t + BS_V[m][1]) * t + BS_V[m][2]) * t + BS_V[m][3];
    da[m] = (BS_D[m][0] * t + BS_D[m][1]) * t + BS_D[m][2];
    d2a[m] = BS_DD[m][0] * t + BS_DD[m][1];
  }
}

// 64-term stencil sum over every orbital for one (i, j) column of the stencil;
// each orbital's ten accumulators keep the same summation order whatever the
// vector width.  One body, instantiated for float and double tables.
#define ORC_SPLINE_COLUMN(NAME, C)                                                                              \
  ORC_HOT static void NAME(const C* __restrict q, size_t zs, int n, const double* c, const double* dc,          \
                           const double* d2c, const double* w, double* __restrict acc) {                        \
    double* __restrict a0 = acc;                                                                                 \
    double* __restrict a1 = acc + n;                                                                             \
    double* __restrict a2 = acc + 2 * n;                                                                         \
    double* __restrict a3 = acc + 3 * n;                                                                         \
    double* __restrict a4 = acc + 4 * n;                                                                         \
    double* __restrict a5 = acc + 5 * n;                                                                         \
    double* __restrict a6 = acc + 6 * n;                                                                         \
    double* __restrict a7 = acc + 7 * n;                                                                         \
    double* __restrict a8 = acc + 8 * n;                                                                         \
    double* __restrict a9 = acc + 9 * n;                                                                         \
    const C* __restrict q0p = q;                                                                                 \
    const C* __restrict q1p = q + zs;                                                                            \
    const C* __restrict q2p = q + 2 * zs;                                                                        \
    const C* __restrict q3p = q + 3 * zs;                                                                        \
    const double c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];                                                     \
    const double e0 = dc[0], e1 = dc[1], e2 = dc[2], e3 = dc[3];                                                 \
    const double f0 = d2c[0], f1 = d2c[1], f2 = d2c[2], f3 = d2c[3];                                             \
    const double w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4], w5 = w[5];                               \
    for (int o = 0; o < n; ++o) {                                                                                \
      const double q0 = (double)q0p[o], q1 = (double)q1p[o], q2 = (double)q2p[o], q3 = (double)q3p[o];          \
      const double s0 = ((c0 * q0 + c1 * q1) + c2 * q2) + c3 * q3;                                              \
      const double s1 = ((e0 * q0 + e1 * q1) + e2 * q2) + e3 * q3;                                              \
      const double s2 = ((f0 * q0 + f1 * q1) + f2 * q2) + f3 * q3;                                              \
      a0[o] += w0 * s0;                                                                                          \
      a1[o] += w1 * s0;                                                                                          \
      a2[o] += w2 * s0;                                                                                          \
      a3[o] += w0 * s1;                                                                                          \
      a4[o] += w3 * s0;                                                                                          \
      a5[o] += w4 * s0;                                                                                          \
      a6[o] += w1 * s1;                                                                                          \
      a7[o] += w5 * s0;                                                                                          \
      a8[o] += w2 * s1;                                                                                          \
      a9[o] += w0 * s2;                                                                                          \
    }                                                                                                            \
  }
ORC_SPLINE_COLUMN(spline_column_f, float)
ORC_SPLINE_COLUMN(spline_column_d, double)

// value, gradient (3), hessian (xx,xy,xz,yy,yz,zz) for every orbital at the
// fractional point u; out has 10 planes of `norb` (MultiBspline.hpp:163-280).
static void spline_vgh_unit(const Spline& s, const double u[3], double* out) {
  int ix, iy, iz;
  double tx, ty, tz;
  spline_bound(u[0] * s.ng[0], tx, ix, s.ng[0] - 1);
  spline_bound(u[1] * s.ng[1], ty, iy, s.ng[1] - 1);
  spline_bound(u[2] * s.ng[2], tz, iz, s.ng[2] - 1);
  double a[4], b[4], c[4], da[4], db[4], dc[4], d2a[4], d2b[4], d2c[4];
  weights(tx, a, da, d2a);
  weights(ty, b, db, d2b);
  weights(tz, c, dc, d2c);
  const size_t zs = s.npad, ys = (size_t)(s.ng[2] + 3) * zs, xs = (size_t)(s.ng[1] + 3) * ys;
  const int n = s.norb;
  double* acc = out;
  std::fill(acc, acc + 10 * (size_t)n, 0.0);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      // weights of (v, d/dx, d/dy, d2/dx2, d2/dxdy, d2/dy2) for this column
      const double w[6] = {a[i] * b[j], da[i] * b[j], a[i] * db[j], d2a[i] * b[j], da[i] * db[j], a[i] * d2b[j]};
      const size_t base = (ix + i) * xs + (iy + j) * ys + iz * zs;
      if (s.prec == 4) spline_column_f(s.cf + base, zs, n, c, dc, d2c, w, acc);
      else spline_column_d(s.cd + base, zs, n, c, dc, d2c, w, acc);
    }
  const double gx = s.ng[0], gy = s.ng[1], gz = s.ng[2];  // delta_inv of a unit-cube grid
  const double sc[10] = {1.0, gx, gy, gz, gx * gx, gx * gy, gx * gz, gy * gy, gy * gz, gz * gz};
  for (int p = 0; p < 10; ++p)
    for (int o = 0; o < n; ++o) out[p * n + o] = acc[p * n + o] * sc[p];
}

// ---------------------------------------------------------------- determinant
// LU with partial pivoting of a row-major n x n matrix (rows = electrons).
// Returns log|det| and sets phase (0 or pi); inv receives A^-1 (row-major).
// DiracMatrix.h:26-38 (log det from the U diagonal and pivot parity), :93-121.
static double lu_log_inverse(int n, const double* A, double* inv, double& phase) {
  std::vector<double> lu(A, A + (size_t)n * n);
  std::vector<int> piv(n);
  int sign = 1;
  double logdet = 0.0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(lu[(size_t)k * n + k]);
    for (int r = k + 1; r < n; ++r) {
      double v = std::fabs(lu[(size_t)r * n + k]);
      if (v > best) { best = v; p = r; }
    }
    piv[k] = p;
    if (p != k) {
      for (int c = 0; c < n; ++c) std::swap(lu[(size_t)k * n + c], lu[(size_t)p * n + c]);
      sign = -sign;
    }
    const double d = lu[(size_t)k * n + k];
    if (d < 0) sign = -sign;
    logdet += std::log(std::fabs(d));
    const double dinv = 1.0 / d;
    for (int r = k + 1; r < n; ++r) {
      double l = lu[(size_t)r * n + k] * dinv;
      lu[(size_t)r * n + k] = l;
      if (l != 0.0)
        for (int c = k + 1; c < n; ++c) lu[(size_t)r * n + c] -= l * lu[(size_t)k * n + c];
    }
  }
  phase = sign > 0 ? 0.0 : M_PI;
  // inverse: solve A X = I column by column; P A = L U
  std::vector<double> col(n);
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) col[i] = (i == j) ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k)
      if (piv[k] != k) std::swap(col[k], col[piv[k]]);
    for (int i = 0; i < n; ++i) {
      double s = col[i];
      for (int k = 0; k < i; ++k) s -= lu[(size_t)i * n + k] * col[k];
      col[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = col[i];
      for (int k = i + 1; k < n; ++k) s -= lu[(size_t)i * n + k] * col[k];
      col[i] = s / lu[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + j] = col[i];
  }
  return logdet;
}

// psiM = (A^-1)^T with A[i][j] = phi_j(r_i); DiracMatrix.h:93-121 and
// DiracDeterminant.cpp:264-280.
static double invert_transpose(int n, const double* A, double* psiM, double& phase) {
  std::vector<double> inv((size_t)n * n);
  double ld = lu_log_inverse(n, A, inv.data(), phase);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) psiM[(size_t)i * n + j] = inv[(size_t)j * n + i];
  return ld;
}

// Fahy's Sherman-Morrison in one pass over A (DelayedUpdate.h:138-147, k = 1):
// t_m = A[m].u - [m == row], A[m] -= b t_m v.  Row m's update needs only t_m,
// so eight rows are dotted together (each summed in j order) and updated at once.
ORC_HOT static void sm_update(int n, double* __restrict A, const double* __restrict u, int row, double b,
                              const double* __restrict v) {
  int m = 0;
  for (; m + 8 <= n; m += 8) {
    double* a0 = A + (size_t)m * n;
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < n; ++j) {
      const double uj = u[j];
#pragma GCC unroll 8
      for (int i = 0; i < 8; ++i) t[i] += a0[(size_t)i * n + j] * uj;
    }
    for (int i = 0; i < 8; ++i) {
      const double bt = b * (t[i] - (m + i == row ? 1.0 : 0.0));
      double* __restrict ar = a0 + (size_t)i * n;
      for (int j = 0; j < n; ++j) ar[j] -= bt * v[j];
    }
  }
  for (; m < n; ++m) {
    double* __restrict ar = A + (size_t)m * n;
    double t = 0;
    for (int j = 0; j < n; ++j) t += ar[j] * u[j];
    const double bt = b * (t - (m == row ? 1.0 : 0.0));
    for (int j = 0; j < n; ++j) ar[j] -= bt * v[j];
  }
}
static double dot_n(int n, const double* a, const double* b) {
  double s = 0;
  for (int j = 0; j < n; ++j) s += a[j] * b[j];
  return s;
}
// A[m][j] -= sum_e T[m][e] W[e][j], vectorised over j, e summed in order
ORC_HOT static void woodbury_apply(int n, int k, int c, double* __restrict A, const double* __restrict T,
                                   const double* __restrict W) {
  std::vector<double> row(n);
  for (int m = 0; m < n; ++m) {
    std::fill(row.begin(), row.end(), 0.0);
    for (int e = 0; e < c; ++e) {
      const double t = T[(size_t)m * k + e];
      const double* w = W + (size_t)e * n;
      for (int j = 0; j < n; ++j) row[j] += t * w[j];
    }
    double* a = A + (size_t)m * n;
    for (int j = 0; j < n; ++j) a[j] -= row[j];
  }
}

// The rank-k delayed (Woodbury) inverse update; DelayedUpdate.h:29-231.
struct DelayedUpdate {
  int n = 0, k = 1, c = 0;
  std::vector<double> U, V, Binv, p, tmat, tmp;
  std::vector<int> rows;
  void resize(int norb, int delay) {
    n = norb;
    k = delay;
    c = 0;
    U.assign((size_t)k * n, 0.0);
    V.assign((size_t)k * n, 0.0);
    Binv.assign((size_t)k * k, 0.0);
    p.assign(k, 0.0);
    tmat.assign((size_t)n * k, 0.0);
    tmp.assign(n, 0.0);
    rows.assign(k, 0);
  }
  // invRow = Ainv[row] - V^T Binv^T U Ainv[row]; the intermediate
  // Binv^T U Ainv[row] is parked in row c of Binv, where acceptRow reads it.
  void get_inv_row(const double* A, int row, double* inv_row) {
    const double* a = A + (size_t)row * n;
    std::copy(a, a + n, inv_row);
    if (c == 0) return;
    for (int d = 0; d < c; ++d) {
      double s = 0;
      for (int j = 0; j < n; ++j) s += U[(size_t)d * n + j] * a[j];
      p[d] = s;
    }
    double* y = &Binv[(size_t)c * k];
    for (int r = 0; r < c; ++r) {
      double s = 0;
      for (int q = 0; q < c; ++q) s += Binv[(size_t)q * k + r] * p[q];
      y[r] = s;
    }
    for (int d = 0; d < c; ++d)
      for (int j = 0; j < n; ++j) inv_row[j] -= V[(size_t)d * n + j] * y[d];
  }
  void accept_row(double* A, int row, const double* psiv) {
    std::copy(A + (size_t)row * n, A + (size_t)row * n + n, &V[(size_t)c * n]);
    std::copy(psiv, psiv + n, &U[(size_t)c * n]);
    rows[c] = row;
    for (int d = 0; d <= c; ++d) {
      double s = 0;
      for (int j = 0; j < n; ++j) s += V[(size_t)d * n + j] * psiv[j];
      p[d] = -s;
    }
    double y = -p[c];
    for (int i = 0; i < c; ++i) y += Binv[(size_t)c * k + i] * p[i];
    y = 1.0 / y;
    Binv[(size_t)c * k + c] = y;
    for (int s = 0; s < c; ++s) {
      double t = 0;
      for (int r = 0; r < c; ++r) t += Binv[(size_t)s * k + r] * p[r];
      Binv[(size_t)s * k + c] = y * t;
    }
    for (int s = 0; s < c; ++s)
      for (int r = 0; r < c; ++r) Binv[(size_t)s * k + r] -= Binv[(size_t)c * k + r] * Binv[(size_t)s * k + c];
    for (int i = 0; i < c; ++i) Binv[(size_t)c * k + i] *= -y;
    ++c;
    if (c == k) update_inv(A);
  }
  void update_inv(double* A) {
    if (c == 0) return;
    if (c == 1) {
      // Fahy's variant of Sherman-Morrison
      sm_update(n, A, U.data(), rows[0], Binv[0], V.data());
    } else {
      // T = A U^T - E ; W = Binv V ; A -= T W
      for (int m = 0; m < n; ++m)
        for (int d = 0; d < c; ++d) tmat[(size_t)m * k + d] = dot_n(n, &U[(size_t)d * n], A + (size_t)m * n);
      for (int d = 0; d < c; ++d) tmat[(size_t)rows[d] * k + d] -= 1.0;
      std::vector<double> W((size_t)c * n, 0.0);
      for (int e = 0; e < c; ++e)
        for (int j = 0; j < n; ++j) {
          double s = 0;
          for (int d = 0; d < c; ++d) s += Binv[(size_t)e * k + d] * V[(size_t)d * n + j];
          W[(size_t)e * n + j] = s;
        }
      woodbury_apply(n, k, c, A, tmat.data(), W.data());
    }
    c = 0;
  }
};

struct Determinant {
  int first = 0, n = 0;
  std::vector<double> psiM, dpsiM, d2psiM;  // psiM n x n ; dpsiM n x n x 3 ; d2psiM n x n
  std::vector<double> row_norm2;            // |A row i|^2 of the Slater matrix (sensitivity only)
  std::vector<double> psiV, dpsiV, d2psiV, inv_row;
  double log_value = 0.0, phase = 0.0, cur_ratio = 1.0;
  int inv_row_id = -1;
  DelayedUpdate eng;
};

// ---------------------------------------------------------------- system and walkers
struct System {
  Lattice lat;
  int nions = 0, nspecies = 0, nup = 0, ndn = 0, nel = 0;
  std::vector<double> ion_x, ion_y, ion_z;
  std::vector<int> ion_species;
  std::vector<int> species_first, species_last;
  std::vector<Functor> j1;   // per ion species
  Functor j2[2];             // [0] same spin, [1] opposite spin
  std::vector<Poly3D> j3;    // eeI functor per ion species and spin pair: [species * 2 + (same ? 0 : 1)]
  bool has_j3 = false;
  Spline spo;
};

struct Walker {
  const System* sys = nullptr;
  int N = 0;
  std::vector<double> Rx, Ry, Rz;
  // AA row of the active electron at its current position, and the proposed row
  std::vector<double> dist_old, dxo, dyo, dzo, dist_new, dxn, dyn, dzn;
  std::vector<double> ei_r, ei_dx, ei_dy, ei_dz, ei_rn, ei_dxn, ei_dyn, ei_dzn;
  double active_pos[3];
  int active = -1;
  // J2
  std::vector<double> Uat, dUx, dUy, dUz, d2Uat;
  std::vector<double> cur_u, cur_du, cur_d2u, old_u, old_du, old_d2u;
  double cur_Uat = 0.0, j2_log = 0.0;
  // J1
  std::vector<double> Vat, Gx, Gy, Gz, Lap;
  std::vector<double> U1, dU1, d2U1;
  double curAt = 0.0, curLap = 0.0, curGrad[3] = {0, 0, 0}, j1_log = 0.0;
  Determinant det[2];
  // J3 (eeI): Uat, dUat, d2Uat of ThreeBodyJastrow.h:50-58, the electron-ion rows of every
  // electron (the eI table, [N][M]) and the moving electron's per-electron terms
  std::vector<double> U3, dU3x, dU3y, dU3z, d2U3;
  std::vector<double> eiTr, eiTx, eiTy, eiTz;
  std::vector<double> nUk, ndUx, ndUy, ndUz, nd2Uk, oUk, odUx, odUy, odUz, od2Uk;
  double cur_U3 = 0.0, cur_dU3[3] = {0, 0, 0}, cur_d2U3 = 0.0, j3_log = 0.0;
  double log_value = 0.0;
  std::vector<double> G, L;  // per electron (3N, N)
};

static inline const Functor& j2f(const System& s, int i, int j) {
  return s.j2[((i < s.nup) == (j < s.nup)) ? 0 : 1];
}

// AA row for electron i at position pos; flip index i (DistanceTableAA.h:85-101)
static void aa_row(const Walker& w, int i, const double pos[3], double* r, double* dx, double* dy, double* dz) {
  min_image_row(w.sys->lat, pos, w.Rx.data(), w.Ry.data(), w.Rz.data(), w.N, i, r, dx, dy, dz);
}
static void ei_row(const Walker& w, const double pos[3], double* r, double* dx, double* dy, double* dz) {
  const System& s = *w.sys;
  min_image_row(s.lat, pos, s.ion_x.data(), s.ion_y.data(), s.ion_z.data(), s.nions, 0, r, dx, dy, dz);
}

// TwoBodyJastrow::computeU3 (TwoBodyJastrow.h:251-276): u, du/r, d2u over a row,
// zero for r >= rcut and for the self index.
static void j2_u3(const Walker& w, int iat, const double* dist, int jmax, double* u, double* du, double* d2u) {
  const System& s = *w.sys;
  for (int j = 0; j < jmax; ++j) {
    u[j] = du[j] = d2u[j] = 0.0;
    const Functor& f = j2f(s, iat, j);
    if (j != iat && dist[j] < f.rcut) functor_vgl(f, dist[j], u[j], du[j], d2u[j]);
  }
}

// OneBodyJastrow::computeU3 (OneBodyJastrow.h:209-245)
static void j1_u3(const Walker& w, const double* dist, double* u, double* du, double* d2u) {
  const System& s = *w.sys;
  for (int c = 0; c < s.nions; ++c) {
    u[c] = du[c] = d2u[c] = 0.0;
    const Functor& f = s.j1[s.ion_species[c]];
    if (f.present && dist[c] < f.rcut) functor_vgl(f, dist[c], u[c], du[c], d2u[c]);
  }
}

// ThreeBodyJastrow::computeU3 with computeU3_engine (ThreeBodyJastrow.h:441-625) for electron jel
// with electron-ion row (rjI, displacement ion - electron) and electron-electron row (rjk,
// displacement k - jel): over the ions I with r_jI < L_I and the electrons k != jel (k < jel
// with `triangle`) with r_kI < L_I, F(r_jk, r_jI, r_kI) of (ion species, spin pair):
//   Uj += F;  dUj += F_jI/r_jI d_jI + F_jk/r_jk d_jk  (d = displacements);
//   d2Uj -= F_jk,jk + F_jI,jI + 2 (F_jk/r_jk + F_jI/r_jI) + 2 F_jk,jI/(r_jk r_jI) d_jk.d_jI;
//   Uk[k] += F;  dUk[k] += F_kI/r_kI d_kI - F_jk/r_jk d_jk;
//   d2Uk[k] -= F_jk,jk + F_kI,kI + 2 (F_jk/r_jk + F_kI/r_kI) - 2 F_jk,kI/(r_jk r_kI) d_kI.d_jk.
// The other electrons' ion rows come from the walker's eI table.
static void j3_u3(const Walker& w, int jel, const double* rjI, const double* xjI, const double* yjI,
                  const double* zjI, const double* rjk, const double* xjk, const double* yjk, const double* zjk,
                  bool triangle, double& Uj, double dUj[3], double& d2Uj, double* Uk, double* dUx, double* dUy,
                  double* dUz, double* d2Uk) {
  const System& s = *w.sys;
  const int N = w.N, M = s.nions;
  const int kmax = triangle ? jel : N;
  Uj = d2Uj = 0.0;
  dUj[0] = dUj[1] = dUj[2] = 0.0;
  for (int k = 0; k < kmax; ++k) Uk[k] = dUx[k] = dUy[k] = dUz[k] = d2Uk[k] = 0.0;
  const bool jup = jel < s.nup;
  for (int I = 0; I < M; ++I) {
    const int sp = s.ion_species[I];
    const Poly3D& f0 = s.j3[2 * sp];
    if (!f0.present || !(rjI[I] < f0.L)) continue;
    for (int k = 0; k < kmax; ++k) {
      const size_t kI = (size_t)k * M + I;
      if (k == jel || !(w.eiTr[kI] < f0.L)) continue;
      const Poly3D& f = s.j3[2 * sp + (((k < s.nup) == jup) ? 0 : 1)];
      double o[9];
      poly3d_vgl(f, rjk[k], rjI[I], w.eiTr[kI], o);
      const double jk[3] = {xjk[k], yjk[k], zjk[k]}, jI[3] = {xjI[I], yjI[I], zjI[I]};
      const double kIv[3] = {w.eiTx[kI], w.eiTy[kI], w.eiTz[kI]};
      Uj += o[0];
      d2Uj -= o[4] + o[5] + 2.0 * (o[1] + o[2]) + 2.0 * o[7] * dot3(jk, jI);
      for (int d = 0; d < 3; ++d) dUj[d] += o[2] * jI[d] + o[1] * jk[d];
      Uk[k] += o[0];
      dUx[k] += o[3] * kIv[0] - o[1] * jk[0];
      dUy[k] += o[3] * kIv[1] - o[1] * jk[1];
      dUz[k] += o[3] * kIv[2] - o[1] * jk[2];
      d2Uk[k] -= o[4] + o[6] + 2.0 * (o[1] + o[3]) - 2.0 * o[8] * dot3(kIv, jk);
    }
  }
}

// ThreeBodyJastrow::computeU (ThreeBodyJastrow.h:403-439): value only, for electron jel's rows
static double j3_v(const Walker& w, int jel, const double* rjI, const double* rjk) {
  const System& s = *w.sys;
  const int N = w.N, M = s.nions;
  const bool jup = jel < s.nup;
  double Uj = 0.0;
  for (int I = 0; I < M; ++I) {
    const int sp = s.ion_species[I];
    const Poly3D& f0 = s.j3[2 * sp];
    if (!f0.present || !(rjI[I] < f0.L)) continue;
    for (int k = 0; k < N; ++k) {
      const size_t kI = (size_t)k * M + I;
      if (k == jel || !(w.eiTr[kI] < f0.L)) continue;
      Uj += poly3d_v(s.j3[2 * sp + (((k < s.nup) == jup) ? 0 : 1)], rjk[k], rjI[I], w.eiTr[kI]);
    }
  }
  return Uj;
}

// SPO evaluation at Cartesian r: psi, dpsi (fractional-coordinate gradient),
// d2psi = h_xx (einspline_spo.hpp:204-223)
static void spo_eval(const System& s, const double r[3], double* psi, double* dpsi, double* d2psi,
                     std::vector<double>& scratch) {
  double u[3];
  to_unit_floor(s.spo.lat.G, r, u);
  const int n = s.spo.norb;
  scratch.resize(10 * (size_t)n);
  spline_vgh_unit(s.spo, u, scratch.data());
  for (int o = 0; o < n; ++o) {
    psi[o] = scratch[o];
    dpsi[3 * o + 0] = scratch[1 * n + o];
    dpsi[3 * o + 1] = scratch[2 * n + o];
    dpsi[3 * o + 2] = scratch[3 * n + o];
    d2psi[o] = scratch[4 * n + o];
  }
}

static void walker_init(Walker& w, const System& s, const double* R, int delay) {
  w.sys = &s;
  const int N = s.nel;
  w.N = N;
  w.Rx.resize(N); w.Ry.resize(N); w.Rz.resize(N);
  for (int i = 0; i < N; ++i) { w.Rx[i] = R[3 * i]; w.Ry[i] = R[3 * i + 1]; w.Rz[i] = R[3 * i + 2]; }
  for (auto* v : {&w.dist_old, &w.dxo, &w.dyo, &w.dzo, &w.dist_new, &w.dxn, &w.dyn, &w.dzn, &w.Uat, &w.dUx,
                  &w.dUy, &w.dUz, &w.d2Uat, &w.cur_u, &w.cur_du, &w.cur_d2u, &w.old_u, &w.old_du, &w.old_d2u,
                  &w.Vat, &w.Gx, &w.Gy, &w.Gz, &w.Lap})
    v->assign(N, 0.0);
  const int M = s.nions;
  for (auto* v : {&w.ei_r, &w.ei_dx, &w.ei_dy, &w.ei_dz, &w.ei_rn, &w.ei_dxn, &w.ei_dyn, &w.ei_dzn, &w.U1,
                  &w.dU1, &w.d2U1})
    v->assign(M, 0.0);
  w.G.assign(3 * (size_t)N, 0.0);
  w.L.assign(N, 0.0);
  if (s.has_j3) {
    for (auto* v : {&w.U3, &w.dU3x, &w.dU3y, &w.dU3z, &w.d2U3, &w.nUk, &w.ndUx, &w.ndUy, &w.ndUz, &w.nd2Uk, &w.oUk,
                    &w.odUx, &w.odUy, &w.odUz, &w.od2Uk})
      v->assign(N, 0.0);
    for (auto* v : {&w.eiTr, &w.eiTx, &w.eiTy, &w.eiTz}) v->assign((size_t)N * M, 0.0);
  }
  for (int d = 0; d < 2; ++d) {
    Determinant& D = w.det[d];
    D.first = d == 0 ? 0 : s.nup;
    D.n = d == 0 ? s.nup : s.ndn;
    const size_t n = D.n;
    D.psiM.assign(n * n, 0.0);
    D.dpsiM.assign(3 * n * n, 0.0);
    D.d2psiM.assign(n * n, 0.0);
    D.psiV.assign(n, 0.0);
    D.dpsiV.assign(3 * n, 0.0);
    D.d2psiV.assign(n, 0.0);
    D.inv_row.assign(n, 0.0);
    D.eng.resize(D.n, delay);
  }
}

static inline Determinant& det_of(Walker& w, int iat) { return iat < w.sys->nup ? w.det[0] : w.det[1]; }

// WaveFunction::evaluateLog (WaveFunction.cpp:178-195): determinant recompute +
// J1/J2 from scratch; G and L are accumulated.
static double evaluate_log(Walker& w) {
  const System& s = *w.sys;
  const int N = w.N;
  std::fill(w.G.begin(), w.G.end(), 0.0);
  std::fill(w.L.begin(), w.L.end(), 0.0);
  std::vector<double> scratch;
  double logv = 0.0;
  for (int d = 0; d < 2; ++d) {
    Determinant& D = w.det[d];
    const int n = D.n;
    std::vector<double> A((size_t)n * n), grad(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      const int iat = D.first + i;
      double r[3] = {w.Rx[iat], w.Ry[iat], w.Rz[iat]};
      spo_eval(s, r, &A[(size_t)i * n], &D.dpsiM[3 * (size_t)i * n], &D.d2psiM[(size_t)i * n], scratch);
    }
    D.row_norm2.assign(n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) D.row_norm2[i] += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    D.log_value = invert_transpose(n, A.data(), D.psiM.data(), D.phase);
    D.eng.c = 0;
    D.inv_row_id = -1;
    for (int i = 0; i < n; ++i) {
      const int iat = D.first + i;
      double gx = 0, gy = 0, gz = 0, lap = 0;
      for (int j = 0; j < n; ++j) {
        const double a = D.psiM[(size_t)i * n + j];
        gx += a * D.dpsiM[3 * ((size_t)i * n + j) + 0];
        gy += a * D.dpsiM[3 * ((size_t)i * n + j) + 1];
        gz += a * D.dpsiM[3 * ((size_t)i * n + j) + 2];
        lap += a * D.d2psiM[(size_t)i * n + j];
      }
      w.G[3 * iat] += gx; w.G[3 * iat + 1] += gy; w.G[3 * iat + 2] += gz;
      w.L[iat] += lap - (gx * gx + gy * gy + gz * gz);
    }
    logv += D.log_value;
  }
  // J1 recompute (OneBodyJastrow.h:110-119)
  {
    double jl = 0;
    for (int i = 0; i < N; ++i) {
      double pos[3] = {w.Rx[i], w.Ry[i], w.Rz[i]};
      ei_row(w, pos, w.ei_r.data(), w.ei_dx.data(), w.ei_dy.data(), w.ei_dz.data());
      j1_u3(w, w.ei_r.data(), w.U1.data(), w.dU1.data(), w.d2U1.data());
      double v = 0, lap = 0, gx = 0, gy = 0, gz = 0;
      for (int c = 0; c < s.nions; ++c) {
        v += w.U1[c];
        lap += w.d2U1[c] + 2.0 * w.dU1[c];
        gx += w.dU1[c] * w.ei_dx[c];
        gy += w.dU1[c] * w.ei_dy[c];
        gz += w.dU1[c] * w.ei_dz[c];
      }
      w.Vat[i] = v; w.Gx[i] = gx; w.Gy[i] = gy; w.Gz[i] = gz; w.Lap[i] = lap;
      jl -= v;
    }
    w.j1_log = jl;
  }
  // J2 recompute over the lower triangle (TwoBodyJastrow.h:356-400)
  {
    std::fill(w.Uat.begin(), w.Uat.end(), 0.0);
    std::fill(w.dUx.begin(), w.dUx.end(), 0.0);
    std::fill(w.dUy.begin(), w.dUy.end(), 0.0);
    std::fill(w.dUz.begin(), w.dUz.end(), 0.0);
    std::fill(w.d2Uat.begin(), w.d2Uat.end(), 0.0);
    std::vector<double> r(N), dx(N), dy(N), dz(N), u(N), du(N), d2u(N);
    for (int i = 0; i < N; ++i) {
      double pos[3] = {w.Rx[i], w.Ry[i], w.Rz[i]};
      aa_row(w, i, pos, r.data(), dx.data(), dy.data(), dz.data());
      j2_u3(w, i, r.data(), i, u.data(), du.data(), d2u.data());
      double su = 0, lap = 0, gx = 0, gy = 0, gz = 0;
      for (int j = 0; j < i; ++j) {
        su += u[j];
        lap += d2u[j] + 2.0 * du[j];
        gx += du[j] * dx[j];
        gy += du[j] * dy[j];
        gz += du[j] * dz[j];
      }
      w.Uat[i] = su;
      w.dUx[i] = gx; w.dUy[i] = gy; w.dUz[i] = gz;
      w.d2Uat[i] = -lap;
      for (int j = 0; j < i; ++j) {
        w.Uat[j] += u[j];
        w.d2Uat[j] -= d2u[j] + 2.0 * du[j];
        w.dUx[j] -= du[j] * dx[j];
        w.dUy[j] -= du[j] * dy[j];
        w.dUz[j] -= du[j] * dz[j];
      }
    }
    double sum = 0;
    for (int i = 0; i < N; ++i) sum += w.Uat[i];
    w.j2_log = -0.5 * sum;
  }
  for (int i = 0; i < N; ++i) {
    w.G[3 * i] += w.Gx[i] + w.dUx[i];
    w.G[3 * i + 1] += w.Gy[i] + w.dUy[i];
    w.G[3 * i + 2] += w.Gz[i] + w.dUz[i];
    w.L[i] += -w.Lap[i] + w.d2Uat[i];
  }
  w.j3_log = 0.0;
  if (s.has_j3) {
    // ThreeBodyJastrow::recompute (ThreeBodyJastrow.h:363-401): the eI table, then for every
    // electron its terms against the electrons before it, both sides accumulated
    const int M = s.nions;
    for (int i = 0; i < N; ++i) {
      double pos[3] = {w.Rx[i], w.Ry[i], w.Rz[i]};
      const size_t o = (size_t)i * M;
      ei_row(w, pos, &w.eiTr[o], &w.eiTx[o], &w.eiTy[o], &w.eiTz[o]);
    }
    std::vector<double> r(N), dx(N), dy(N), dz(N);
    for (int i = 0; i < N; ++i) {
      double pos[3] = {w.Rx[i], w.Ry[i], w.Rz[i]};
      aa_row(w, i, pos, r.data(), dx.data(), dy.data(), dz.data());
      const size_t o = (size_t)i * M;
      double Uj, dUj[3], d2Uj;
      j3_u3(w, i, &w.eiTr[o], &w.eiTx[o], &w.eiTy[o], &w.eiTz[o], r.data(), dx.data(), dy.data(), dz.data(), true,
            Uj, dUj, d2Uj, w.nUk.data(), w.ndUx.data(), w.ndUy.data(), w.ndUz.data(), w.nd2Uk.data());
      w.U3[i] = Uj;
      w.dU3x[i] = dUj[0]; w.dU3y[i] = dUj[1]; w.dU3z[i] = dUj[2];
      w.d2U3[i] = d2Uj;
      for (int k = 0; k < i; ++k) {
        w.U3[k] += w.nUk[k];
        w.d2U3[k] += w.nd2Uk[k];
        w.dU3x[k] += w.ndUx[k]; w.dU3y[k] += w.ndUy[k]; w.dU3z[k] += w.ndUz[k];
      }
    }
    double sum = 0;
    for (int i = 0; i < N; ++i) {
      sum += w.U3[i];
      w.G[3 * i] += w.dU3x[i];
      w.G[3 * i + 1] += w.dU3y[i];
      w.G[3 * i + 2] += w.dU3z[i];
      w.L[i] += w.d2U3[i];
    }
    w.j3_log = -0.5 * sum;
  }
  w.log_value = logv + w.j1_log + w.j2_log + w.j3_log;
  return w.log_value;
}

// ParticleSet::setActive -> AA and BA rows of the current position.
static void set_active(Walker& w, int iat) {
  double pos[3] = {w.Rx[iat], w.Ry[iat], w.Rz[iat]};
  aa_row(w, iat, pos, w.dist_old.data(), w.dxo.data(), w.dyo.data(), w.dzo.data());
  w.dist_old[iat] = std::numeric_limits<double>::max();
  ei_row(w, pos, w.ei_r.data(), w.ei_dx.data(), w.ei_dy.data(), w.ei_dz.data());
}

// WaveFunction::evalGrad (WaveFunction.cpp:197-208)
static void eval_grad(Walker& w, int iat, double g[3]) {
  Determinant& D = det_of(w, iat);
  const int row = iat - D.first, n = D.n;
  D.inv_row_id = row;
  D.eng.get_inv_row(D.psiM.data(), row, D.inv_row.data());
  double gx = 0, gy = 0, gz = 0;
  for (int j = 0; j < n; ++j) {
    const double a = D.inv_row[j];
    const double* dp = &D.dpsiM[3 * ((size_t)row * n + j)];
    gx += a * dp[0]; gy += a * dp[1]; gz += a * dp[2];
  }
  g[0] = gx + w.Gx[iat] + w.dUx[iat];
  g[1] = gy + w.Gy[iat] + w.dUy[iat];
  g[2] = gz + w.Gz[iat] + w.dUz[iat];
  if (w.sys->has_j3) {   // ThreeBodyJastrow::evalGrad (ThreeBodyJastrow.h:275)
    g[0] += w.dU3x[iat];
    g[1] += w.dU3y[iat];
    g[2] += w.dU3z[iat];
  }
}

// ParticleSet::makeMove
static void make_move(Walker& w, int iat, const double displ[3]) {
  w.active = iat;
  w.active_pos[0] = w.Rx[iat] + displ[0];
  w.active_pos[1] = w.Ry[iat] + displ[1];
  w.active_pos[2] = w.Rz[iat] + displ[2];
  aa_row(w, iat, w.active_pos, w.dist_new.data(), w.dxn.data(), w.dyn.data(), w.dzn.data());
  ei_row(w, w.active_pos, w.ei_rn.data(), w.ei_dxn.data(), w.ei_dyn.data(), w.ei_dzn.data());
}

// WaveFunction::ratioGrad (WaveFunction.cpp:210-223) with the per-component
// ratioGrad of DiracDeterminant.cpp:92-124, OneBodyJastrow.h:259-268,
// TwoBodyJastrow.h:294-304.
static double ratio_grad(Walker& w, int iat, double g[3], double* det_ratio_out, std::vector<double>& scratch) {
  const System& s = *w.sys;
  Determinant& D = det_of(w, iat);
  const int row = iat - D.first, n = D.n;
  spo_eval(s, w.active_pos, D.psiV.data(), D.dpsiV.data(), D.d2psiV.data(), scratch);
  if (D.inv_row_id != row) {
    D.inv_row_id = row;
    D.eng.get_inv_row(D.psiM.data(), row, D.inv_row.data());
  }
  double ratio = 0, gx = 0, gy = 0, gz = 0;
  for (int j = 0; j < n; ++j) {
    const double a = D.inv_row[j];
    ratio += a * D.psiV[j];
    gx += a * D.dpsiV[3 * j];
    gy += a * D.dpsiV[3 * j + 1];
    gz += a * D.dpsiV[3 * j + 2];
  }
  D.cur_ratio = ratio;
  if (det_ratio_out) *det_ratio_out = ratio;
  g[0] = gx / ratio; g[1] = gy / ratio; g[2] = gz / ratio;
  // J1
  j1_u3(w, w.ei_rn.data(), w.U1.data(), w.dU1.data(), w.d2U1.data());
  {
    double v = 0, lap = 0, ax = 0, ay = 0, az = 0;
    for (int c = 0; c < s.nions; ++c) {
      v += w.U1[c];
      lap += w.d2U1[c] + 2.0 * w.dU1[c];
      ax += w.dU1[c] * w.ei_dxn[c];
      ay += w.dU1[c] * w.ei_dyn[c];
      az += w.dU1[c] * w.ei_dzn[c];
    }
    w.curAt = v; w.curLap = lap;
    w.curGrad[0] = ax; w.curGrad[1] = ay; w.curGrad[2] = az;
    g[0] += ax; g[1] += ay; g[2] += az;
    ratio *= std::exp(w.Vat[iat] - v);
  }
  // J2
  j2_u3(w, iat, w.dist_new.data(), w.N, w.cur_u.data(), w.cur_du.data(), w.cur_d2u.data());
  {
    double su = 0, ax = 0, ay = 0, az = 0;
    for (int j = 0; j < w.N; ++j) {
      su += w.cur_u[j];
      ax += w.cur_du[j] * w.dxn[j];
      ay += w.cur_du[j] * w.dyn[j];
      az += w.cur_du[j] * w.dzn[j];
    }
    w.cur_Uat = su;
    g[0] += ax; g[1] += ay; g[2] += az;
    ratio *= std::exp(w.Uat[iat] - su);
  }
  // J3 (ThreeBodyJastrow::ratioGrad, ThreeBodyJastrow.h:277-300): the proposal's rows
  if (s.has_j3) {
    j3_u3(w, iat, w.ei_rn.data(), w.ei_dxn.data(), w.ei_dyn.data(), w.ei_dzn.data(), w.dist_new.data(),
          w.dxn.data(), w.dyn.data(), w.dzn.data(), false, w.cur_U3, w.cur_dU3, w.cur_d2U3, w.nUk.data(),
          w.ndUx.data(), w.ndUy.data(), w.ndUz.data(), w.nd2Uk.data());
    g[0] += w.cur_dU3[0]; g[1] += w.cur_dU3[1]; g[2] += w.cur_dU3[2];
    ratio *= std::exp(w.U3[iat] - w.cur_U3);
  }
  return ratio;
}

// Sensitivity of the move's determinant ratio rho = psiV . c (c = psiM[row], the inverse's column
// of the row) to relative perturbations of size eps of the Slater matrix A and of psiV:
// |d rho| <= eps (|r| |A|_F + |psiV|) |c|, r = psiM psiV the ratios of putting psiV in each row.
// Returned relative to |rho|.  Needs the current inverse (delay 1).
static double move_sensitivity(Walker& w, int iat) {
  Determinant& D = det_of(w, iat);
  const int row = iat - D.first, n = D.n;
  double r2 = 0, c2 = 0, p2 = 0, a2 = 0;
  for (int i = 0; i < n; ++i) {
    const double* pm = &D.psiM[(size_t)i * n];
    double ri = 0;
    for (int j = 0; j < n; ++j) ri += pm[j] * D.psiV[j];
    r2 += ri * ri;
    a2 += D.row_norm2[i];
  }
  for (int j = 0; j < n; ++j) {
    c2 += D.psiM[(size_t)row * n + j] * D.psiM[(size_t)row * n + j];
    p2 += D.psiV[j] * D.psiV[j];
  }
  return (std::sqrt(r2 * a2) + std::sqrt(p2)) * std::sqrt(c2) / std::fabs(D.cur_ratio);
}

// WaveFunction::acceptMove + ParticleSet::acceptMove
static void accept_move(Walker& w, int iat) {
  Determinant& D = det_of(w, iat);
  const int row = iat - D.first, n = D.n;
  // DiracDeterminant::acceptMove (DiracDeterminant.cpp:129-145)
  D.phase += D.cur_ratio > 0 ? 0.0 : M_PI;
  D.log_value += std::log(std::fabs(D.cur_ratio));
  D.eng.accept_row(D.psiM.data(), row, D.psiV.data());
  D.inv_row_id = -1;
  if (!D.row_norm2.empty()) {
    double nv = 0;
    for (int j = 0; j < n; ++j) nv += D.psiV[j] * D.psiV[j];
    D.row_norm2[row] = nv;
  }
  std::copy(D.dpsiV.begin(), D.dpsiV.end(), D.dpsiM.begin() + 3 * (size_t)row * n);
  std::copy(D.d2psiV.begin(), D.d2psiV.end(), D.d2psiM.begin() + (size_t)row * n);
  D.cur_ratio = 1.0;
  // J1 (OneBodyJastrow.h:271-283)
  w.j1_log += w.Vat[iat] - w.curAt;
  w.Vat[iat] = w.curAt;
  w.Gx[iat] = w.curGrad[0]; w.Gy[iat] = w.curGrad[1]; w.Gz[iat] = w.curGrad[2];
  w.Lap[iat] = w.curLap;
  // J2 (TwoBodyJastrow.h:307-353)
  j2_u3(w, iat, w.dist_old.data(), w.N, w.old_u.data(), w.old_du.data(), w.old_d2u.data());
  double cur_d2 = 0, cgx = 0, cgy = 0, cgz = 0;
  for (int j = 0; j < w.N; ++j) {
    const double du = w.cur_u[j] - w.old_u[j];
    const double newl = w.cur_d2u[j] + 2.0 * w.cur_du[j];
    const double dl = w.old_d2u[j] + 2.0 * w.old_du[j] - newl;
    w.Uat[j] += du;
    w.d2Uat[j] += dl;
    cur_d2 -= newl;
    const double ngx = w.cur_du[j] * w.dxn[j], ngy = w.cur_du[j] * w.dyn[j], ngz = w.cur_du[j] * w.dzn[j];
    w.dUx[j] -= ngx - w.old_du[j] * w.dxo[j];
    w.dUy[j] -= ngy - w.old_du[j] * w.dyo[j];
    w.dUz[j] -= ngz - w.old_du[j] * w.dzo[j];
    cgx += ngx; cgy += ngy; cgz += ngz;
  }
  w.j2_log += w.Uat[iat] - w.cur_Uat;
  w.Uat[iat] = w.cur_Uat;
  w.dUx[iat] = cgx; w.dUy[iat] = cgy; w.dUz[iat] = cgz;
  w.d2Uat[iat] = cur_d2;
  // J3 (ThreeBodyJastrow::acceptMove, ThreeBodyJastrow.h:302-361): the terms at the old position
  // from its rows, every electron's sums moved by the difference, the eI row replaced
  if (w.sys->has_j3) {
    const int M = w.sys->nions;
    const size_t o = (size_t)iat * M;
    double Uold, dUold[3], d2old;
    j3_u3(w, iat, &w.eiTr[o], &w.eiTx[o], &w.eiTy[o], &w.eiTz[o], w.dist_old.data(), w.dxo.data(), w.dyo.data(),
          w.dzo.data(), false, Uold, dUold, d2old, w.oUk.data(), w.odUx.data(), w.odUy.data(), w.odUz.data(),
          w.od2Uk.data());
    for (int k = 0; k < w.N; ++k) {
      w.U3[k] += w.nUk[k] - w.oUk[k];
      w.d2U3[k] += w.nd2Uk[k] - w.od2Uk[k];
      w.dU3x[k] += w.ndUx[k] - w.odUx[k];
      w.dU3y[k] += w.ndUy[k] - w.odUy[k];
      w.dU3z[k] += w.ndUz[k] - w.odUz[k];
    }
    w.j3_log += Uold - w.cur_U3;
    w.U3[iat] = w.cur_U3;
    w.dU3x[iat] = w.cur_dU3[0]; w.dU3y[iat] = w.cur_dU3[1]; w.dU3z[iat] = w.cur_dU3[2];
    w.d2U3[iat] = w.cur_d2U3;
    for (int c = 0; c < M; ++c) {
      w.eiTr[o + c] = w.ei_rn[c];
      w.eiTx[o + c] = w.ei_dxn[c]; w.eiTy[o + c] = w.ei_dyn[c]; w.eiTz[o + c] = w.ei_dzn[c];
    }
  }
  // positions
  w.Rx[iat] = w.active_pos[0]; w.Ry[iat] = w.active_pos[1]; w.Rz[iat] = w.active_pos[2];
  w.active = -1;
}

static void complete_updates(Walker& w) {
  for (int d = 0; d < 2; ++d) {
    w.det[d].inv_row_id = -1;
    w.det[d].eng.update_inv(w.det[d].psiM.data());
  }
}

// WaveFunction::evaluateGL (WaveFunction.cpp:261-279), not from scratch.
static double evaluate_gl(Walker& w) {
  const System& s = *w.sys;
  (void)s;
  const int N = w.N;
  std::fill(w.G.begin(), w.G.end(), 0.0);
  std::fill(w.L.begin(), w.L.end(), 0.0);
  double logv = 0;
  for (int d = 0; d < 2; ++d) {
    Determinant& D = w.det[d];
    const int n = D.n;
    for (int i = 0; i < n; ++i) {
      const int iat = D.first + i;
      double gx = 0, gy = 0, gz = 0, lap = 0;
      for (int j = 0; j < n; ++j) {
        const double a = D.psiM[(size_t)i * n + j];
        gx += a * D.dpsiM[3 * ((size_t)i * n + j) + 0];
        gy += a * D.dpsiM[3 * ((size_t)i * n + j) + 1];
        gz += a * D.dpsiM[3 * ((size_t)i * n + j) + 2];
        lap += a * D.d2psiM[(size_t)i * n + j];
      }
      w.G[3 * iat] += gx; w.G[3 * iat + 1] += gy; w.G[3 * iat + 2] += gz;
      w.L[iat] += lap - (gx * gx + gy * gy + gz * gz);
    }
    logv += D.log_value;
  }
  double v1 = 0, u2 = 0;
  for (int i = 0; i < N; ++i) {
    v1 += w.Vat[i];
    u2 += w.Uat[i];
    w.G[3 * i] += w.Gx[i] + w.dUx[i];
    w.G[3 * i + 1] += w.Gy[i] + w.dUy[i];
    w.G[3 * i + 2] += w.Gz[i] + w.dUz[i];
    w.L[i] += -w.Lap[i] + w.d2Uat[i];
  }
  w.j1_log = -v1;
  w.j2_log = -0.5 * u2;
  w.j3_log = 0.0;
  if (w.sys->has_j3) {   // ThreeBodyJastrow::evaluateGL (ThreeBodyJastrow.h:627-645)
    double u3 = 0;
    for (int i = 0; i < N; ++i) {
      u3 += w.U3[i];
      w.G[3 * i] += w.dU3x[i];
      w.G[3 * i + 1] += w.dU3y[i];
      w.G[3 * i + 2] += w.dU3z[i];
      w.L[i] += w.d2U3[i];
    }
    w.j3_log = -0.5 * u3;
  }
  w.log_value = logv + w.j1_log + w.j2_log + w.j3_log;
  return w.log_value;
}

}  // namespace orc

using namespace orc;

// ================================================================ C ABI (test-side)
extern "C" {

void orc_philox4x64_10(const uint64_t* ctr, const uint64_t* key, uint64_t* out) { philox4x64_10(ctr, key, out); }

void orc_move_randoms(uint64_t seed, uint64_t walker, uint64_t sweep, uint64_t iel, double* normal3, double* uacc) {
  move_randoms(seed, walker, sweep, iel, normal3, *uacc);
}

// lattice: 9 doubles (rows = cell vectors).  out: G[9], rb[9], g[9], corners[24], ws_radius
void orc_lattice_info(const double* R9, double* G9, double* rb9, double* g9, double* corners24, double* ws) {
  Lattice L;
  lattice_set(L, R9);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      G9[3 * i + j] = L.G[i][j];
      rb9[3 * i + j] = L.rb[i][j];
      g9[3 * i + j] = L.g[i][j];
    }
  for (int c = 0; c < 8; ++c)
    for (int d = 0; d < 3; ++d) corners24[3 * c + d] = L.corner[c][d];
  *ws = L.ws_radius;
}

void orc_to_unit_floor(const double* R9, int npts, const double* r, double* u) {
  Lattice L;
  lattice_set(L, R9);
  for (int p = 0; p < npts; ++p) to_unit_floor(L.G, r + 3 * p, u + 3 * p);
}

// distances from pos to n targets (AoS [n][3]); out r[n], dr[n][3]
void orc_distance_row(const double* R9, const double* pos, int n, const double* targets, int flip_ind, double* r,
                      double* dr) {
  Lattice L;
  lattice_set(L, R9);
  std::vector<double> x(n), y(n), z(n), dx(n), dy(n), dz(n);
  for (int t = 0; t < n; ++t) { x[t] = targets[3 * t]; y[t] = targets[3 * t + 1]; z[t] = targets[3 * t + 2]; }
  min_image_row(L, pos, x.data(), y.data(), z.data(), n, flip_ind, r, dx.data(), dy.data(), dz.data());
  for (int t = 0; t < n; ++t) { dr[3 * t] = dx[t]; dr[3 * t + 1] = dy[t]; dr[3 * t + 2] = dz[t]; }
}

// radial functor: u, du/r, d2u at npts distances (0 for r >= rcut)
void orc_functor_vgl(int nparams, double rcut, double cusp, const double* params, int npts, const double* r,
                     double* u, double* du, double* d2u, double* coefs_out) {
  Functor f;
  functor_setup(f, nparams, rcut, cusp, params);
  for (int p = 0; p < npts; ++p) {
    if (r[p] < f.rcut) functor_vgl(f, r[p], u[p], du[p], d2u[p]);
    else u[p] = du[p] = d2u[p] = 0.0;
  }
  if (coefs_out)
    for (int i = 0; i < nparams + 4; ++i) coefs_out[i] = f.coef[i];
}

// spline VGH at fractional points; prec 4 -> float table, 8 -> double table
// out: [npts][10][norb]
void orc_spline_vgh_unit(const int* ng, int norb, int npad, int prec, const void* coefs, int npts, const double* u,
                         double* out) {
  Spline s;
  s.ng[0] = ng[0]; s.ng[1] = ng[1]; s.ng[2] = ng[2];
  s.norb = norb; s.npad = npad; s.prec = prec;
  if (prec == 4) s.cf = (const float*)coefs; else s.cd = (const double*)coefs;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < npts; ++p) spline_vgh_unit(s, u + 3 * p, out + (size_t)p * 10 * norb);
}

// psiM = (A^-1)^T, returns log|det A| and phase
double orc_invert_transpose(int n, const double* A, double* psiM, double* phase) {
  return invert_transpose(n, A, psiM, *phase);
}

// Delayed-update engine exposed for KAT/sequence tests
void* orc_du_create(int norb, int delay) {
  auto* e = new DelayedUpdate;
  e->resize(norb, delay);
  return e;
}
void orc_du_destroy(void* h) { delete (DelayedUpdate*)h; }
void orc_du_get_inv_row(void* h, const double* A, int row, double* inv_row) {
  ((DelayedUpdate*)h)->get_inv_row(A, row, inv_row);
}
void orc_du_accept_row(void* h, double* A, int row, const double* psiv) { ((DelayedUpdate*)h)->accept_row(A, row, psiv); }
void orc_du_update_inv(void* h, double* A) { ((DelayedUpdate*)h)->update_inv(A); }
int orc_du_delay_count(void* h) { return ((DelayedUpdate*)h)->c; }

// ---------------------------------------------------------------- whole-system driver
struct OrcSystem {
  System sys;
  std::vector<float> coef_f;
  std::vector<double> coef_d;
};

void* orc_system_create(const double* lattice9, int nions, const double* ion_pos, const int* ion_species,
                        int nspecies, int nup, int ndn) {
  auto* h = new OrcSystem;
  System& s = h->sys;
  lattice_set(s.lat, lattice9);
  s.nions = nions;
  s.nspecies = nspecies;
  s.nup = nup;
  s.ndn = ndn;
  s.nel = nup + ndn;
  s.ion_x.resize(nions); s.ion_y.resize(nions); s.ion_z.resize(nions);
  s.ion_species.assign(ion_species, ion_species + nions);
  for (int i = 0; i < nions; ++i) {
    s.ion_x[i] = ion_pos[3 * i]; s.ion_y[i] = ion_pos[3 * i + 1]; s.ion_z[i] = ion_pos[3 * i + 2];
  }
  s.j1.assign(nspecies, Functor());
  return h;
}
void orc_system_destroy(void* h) { delete (OrcSystem*)h; }
void orc_system_set_j1(void* h, int species, int nparams, double rcut, double cusp, const double* params) {
  functor_setup(((OrcSystem*)h)->sys.j1[species], nparams, rcut, cusp, params);
}
void orc_system_set_j2(void* h, int pair, int nparams, double rcut, double cusp, const double* params) {
  functor_setup(((OrcSystem*)h)->sys.j2[pair], nparams, rcut, cusp, params);
}
// Coefficients are copied (float table); prec selects float or double evaluation storage.
void orc_system_set_spline(void* h, const int* ng, int norb, int npad, const float* coefs, const double* spo_lattice9,
                           int prec) {
  OrcSystem* o = (OrcSystem*)h;
  Spline& sp = o->sys.spo;
  sp.ng[0] = ng[0]; sp.ng[1] = ng[1]; sp.ng[2] = ng[2];
  sp.norb = norb; sp.npad = npad;
  const size_t cnt = (size_t)(ng[0] + 3) * (ng[1] + 3) * (ng[2] + 3) * npad;
  if (prec == 8) {
    o->coef_d.assign(coefs, coefs + cnt);
    sp.prec = 8; sp.cd = o->coef_d.data();
  } else {
    o->coef_f.assign(coefs, coefs + cnt);
    sp.prec = 4; sp.cf = o->coef_f.data();
  }
  lattice_set(sp.lat, spo_lattice9);
}

double orc_system_ws_radius(void* h) { return ((OrcSystem*)h)->sys.lat.ws_radius; }

// ThreeBodyJastrow::addFunc (ThreeBodyJastrow.h:118-152) with a PolynomialFunctor3D of shape
// (neI, nee), cutoff rcut and free parameters params[nparams] (nio.hpp:170-260): the functor of
// ion species `species` and spin pair `pair` (0 = same, 1 = opposite).  Returns the number of
// free parameters of the shape, or -1 when nparams differs (the functor is then not set).
int orc_system_set_j3(void* h, int species, int pair, int neI, int nee, double rcut, const double* params,
                      int nparams) {
  System& s = ((OrcSystem*)h)->sys;
  if ((int)s.j3.size() < 2 * s.nspecies) s.j3.assign(2 * s.nspecies, Poly3D());
  Poly3D f;
  const int r = poly3d_setup(f, neI, nee, rcut, params, nparams);
  if (r < 0) return r;
  s.j3[2 * species + pair] = f;
  s.has_j3 = true;
  return r;
}

// The functor alone: gamma[(neI+1)^2 (nee+1)] and, at npts points (r12, r1, r2), out[npts][9] as
// poly3d_vgl (F zero with its derivatives where r1 >= L or r2 >= L) and the value alone v[npts].
int orc_poly3d(int neI, int nee, double rcut, const double* params, int nparams, int npts, const double* r12,
               const double* r1, const double* r2, double* gamma, double* out, double* v) {
  Poly3D f;
  const int r = poly3d_setup(f, neI, nee, rcut, params, nparams);
  if (r < 0) return r;
  std::copy(f.g.begin(), f.g.end(), gamma);
  for (int p = 0; p < npts; ++p) {
    double* o = out + 9 * (size_t)p;
    if (!(r1[p] < f.L) || !(r2[p] < f.L)) {
      std::fill(o, o + 9, 0.0);
      v[p] = 0.0;
      continue;
    }
    poly3d_vgl(f, r12[p], r1[p], r2[p], o);
    v[p] = poly3d_v(f, r12[p], r1[p], r2[p]);
  }
  return r;
}

// SPO evaluation at Cartesian points (through toUnit_floor with the SPO lattice):
// psi[npts][norb], dpsi[npts][norb][3], d2psi[npts][norb]
void orc_spo_evaluate(void* h, int npts, const double* r, double* psi, double* dpsi, double* d2psi) {
  const System& s = ((OrcSystem*)h)->sys;
  const int n = s.spo.norb;
#pragma omp parallel
  {
    std::vector<double> scratch;
#pragma omp for
    for (int p = 0; p < npts; ++p)
      spo_eval(s, r + 3 * p, psi + (size_t)p * n, dpsi + (size_t)p * 3 * n, d2psi + (size_t)p * n, scratch);
  }
}

// Run `nsweeps` VMC sweeps for `nw` walkers (OpenMP over walkers).
//   R        [nw][N][3] initial positions; updated in place
//   walker0  global index of the first walker (selects its random stream)
//   mode     0: accept if u < 0.5 (miniqmc.cpp:445); 1: Metropolis u < ratio^2
//   log_init [nw] evaluateLog before the first sweep
//   log_out  [nw][nsweeps] LogValue after each sweep's evaluateGL
//   trace    optional [nw][nsweeps][N][9]: ratio, det_ratio, accept, grad_now[3], grad_new[3]
//   G, L     optional [nw][N][3], [nw][N] after the last sweep
//   naccept  [nw]
// Returns the wall time spent in the sweeps (seconds, excluding evaluateLog).
//   sens     [nw][nsweeps][N] or null: move_sensitivity of every move (delay 1 only)
double orc_run_vmc(void* h, int nw, int walker0, double* R, int nsweeps, uint64_t seed, double tau, int mode,
                   int delay, double* log_init, double* log_out, double* trace, double* G, double* L,
                   long long* naccept, int nthreads, double* sens) {
  const System& s = ((OrcSystem*)h)->sys;
  const int N = s.nel;
  std::vector<Walker> ws(nw);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int iw = 0; iw < nw; ++iw) {
    walker_init(ws[iw], s, R + (size_t)iw * 3 * N, delay);
    log_init[iw] = evaluate_log(ws[iw]);
  }
  const double sqtau = std::sqrt(tau);
  double t0 = 0.0, t1 = 0.0;
#ifdef _OPENMP
  t0 = omp_get_wtime();
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int iw = 0; iw < nw; ++iw) {
    Walker& w = ws[iw];
    std::vector<double> scratch;
    long long acc = 0;
    for (int sw = 0; sw < nsweeps; ++sw) {
      for (int iel = 0; iel < N; ++iel) {
        double nrm[3], uacc;
        move_randoms(seed, (uint64_t)(walker0 + iw), (uint64_t)sw, (uint64_t)iel, nrm, uacc);
        set_active(w, iel);
        double gnow[3];
        eval_grad(w, iel, gnow);
        double displ[3] = {sqtau * nrm[0], sqtau * nrm[1], sqtau * nrm[2]};
        make_move(w, iel, displ);
        double gnew[3], dratio;
        const double ratio = ratio_grad(w, iel, gnew, &dratio, scratch);
        if (sens) sens[((size_t)iw * nsweeps + sw) * N + iel] = delay == 1 ? move_sensitivity(w, iel) : 0.0;
        const bool ok = (mode == 0) ? (uacc < 0.5) : (uacc < ratio * ratio);
        if (ok) {
          accept_move(w, iel);
          ++acc;
        } else {
          w.active = -1;
        }
        if (trace) {
          double* t = trace + (((size_t)iw * nsweeps + sw) * N + iel) * 9;
          t[0] = ratio; t[1] = dratio; t[2] = ok ? 1.0 : 0.0;
          t[3] = gnow[0]; t[4] = gnow[1]; t[5] = gnow[2];
          t[6] = gnew[0]; t[7] = gnew[1]; t[8] = gnew[2];
        }
      }
      complete_updates(w);
      log_out[(size_t)iw * nsweeps + sw] = evaluate_gl(w);
    }
    naccept[iw] = acc;
    for (int i = 0; i < N; ++i) {
      R[((size_t)iw * N + i) * 3 + 0] = w.Rx[i];
      R[((size_t)iw * N + i) * 3 + 1] = w.Ry[i];
      R[((size_t)iw * N + i) * 3 + 2] = w.Rz[i];
    }
    if (G) std::copy(w.G.begin(), w.G.end(), G + (size_t)iw * 3 * N);
    if (L) std::copy(w.L.begin(), w.L.end(), L + (size_t)iw * N);
  }
#ifdef _OPENMP
  t1 = omp_get_wtime();
#endif
  return t1 - t0;
}

// evaluateLog only (fresh), with G and L: for checking the incremental state
// WaveFunction::evaluateRatios (WaveFunction.cpp:281-297) for electron iel at nvp virtual
// positions vpos[nvp][3], after evaluateLog of R:
//   determinant  psi(vpos) . invRow(iel)  (DiracDeterminant.cpp:216-224, SPOSet.h:72-83)
//   J1           exp(Vat[iel] - sum_I u(|r_I - vpos|))      (OneBodyJastrow.h:136-140)
//   J2           exp(Uat[iel] - sum_{j != iel} u(|r_j - vpos|)) (TwoBodyJastrow.h:103-107)
// The virtual particle set's tables are BA rows from the point (VirtualParticleSet.cpp:35-56).
static void evaluate_ratios_at(Walker& w, int iel, int nvp, const double* vpos, double* ratios);

void orc_evaluate_ratios(void* h, const double* R, int iel, int nvp, const double* vpos, double* ratios) {
  const System& s = ((OrcSystem*)h)->sys;
  Walker w;
  walker_init(w, s, R, 1);
  evaluate_log(w);
  evaluate_ratios_at(w, iel, nvp, vpos, ratios);
}

// the same for several electrons of one configuration with one evaluateLog: electron iels[q]
// at vpos[q][nvp][3] -> ratios[q][nvp]
void orc_evaluate_ratios_many(void* h, const double* R, int nq, const int* iels, int nvp, const double* vpos,
                              double* ratios) {
  const System& s = ((OrcSystem*)h)->sys;
  Walker w;
  walker_init(w, s, R, 1);
  evaluate_log(w);
  for (int q = 0; q < nq; ++q)
    evaluate_ratios_at(w, iels[q], nvp, vpos + 3 * (size_t)q * nvp, ratios + (size_t)q * nvp);
}

}  // extern "C"

static void evaluate_ratios_at(Walker& w, int iel, int nvp, const double* vpos, double* ratios) {
  const System& s = *w.sys;
  Determinant& D = det_of(w, iel);
  const int n = D.n, row = iel - D.first, N = w.N;
  std::vector<double> psi(n), dpsi(3 * (size_t)n), d2psi(n), scratch;
  std::vector<double> r(std::max(N, s.nions)), dx(r.size()), dy(r.size()), dz(r.size());
  for (int k = 0; k < nvp; ++k) {
    const double* p = vpos + 3 * k;
    spo_eval(s, p, psi.data(), dpsi.data(), d2psi.data(), scratch);
    const double det_ratio = dot_n(n, psi.data(), &D.psiM[(size_t)row * n]);
    ei_row(w, p, r.data(), dx.data(), dy.data(), dz.data());
    double v1 = 0.0;
    for (int c = 0; c < s.nions; ++c) {
      const Functor& f = s.j1[s.ion_species[c]];
      if (f.present && r[c] < f.rcut) v1 += functor_v(f, r[c]);
    }
    min_image_row(s.lat, p, w.Rx.data(), w.Ry.data(), w.Rz.data(), N, 0, r.data(), dx.data(), dy.data(), dz.data());
    double v2 = 0.0;
    for (int j = 0; j < N; ++j) {
      const Functor& f = j2f(s, iel, j);
      if (j != iel && r[j] < f.rcut) v2 += functor_v(f, r[j]);
    }
    double lj = (w.Vat[iel] - v1) + (w.Uat[iel] - v2);
    if (s.has_j3) {   // ThreeBodyJastrow::evaluateRatios (ThreeBodyJastrow.h:262-273)
      std::vector<double> rI(s.nions), ax(s.nions), ay(s.nions), az(s.nions);
      ei_row(w, p, rI.data(), ax.data(), ay.data(), az.data());
      lj += w.U3[iel] - j3_v(w, iel, rI.data(), r.data());
    }
    ratios[k] = det_ratio * std::exp(lj);
  }
}

extern "C" {

// evaluateLog split by component (the WaveFunctionComponent list of WaveFunction.cpp:178-195):
// part 0 = up determinant, 1 = down determinant, 2 = J1, 3 = J2, 4 = J3 (zero without one).
// logs[5]; G[5][N][3]; L[5][N].
// The determinant parts carry the reference's SPO convention (gradients in the SPO lattice's
// fractional coordinates, d2psi = h_xx); the Jastrow parts are Cartesian.
void orc_evaluate_log_parts(void* h, const double* R, double* logs, double* G, double* L) {
  const System& s = ((OrcSystem*)h)->sys;
  Walker w;
  walker_init(w, s, R, 1);
  evaluate_log(w);
  const int N = w.N;
  std::fill(G, G + 5 * 3 * (size_t)N, 0.0);
  std::fill(L, L + 5 * (size_t)N, 0.0);
  for (int d = 0; d < 2; ++d) {
    const Determinant& D = w.det[d];
    const int n = D.n;
    logs[d] = D.log_value;
    for (int i = 0; i < n; ++i) {
      const int iat = D.first + i;
      double gx = 0, gy = 0, gz = 0, lap = 0;
      for (int j = 0; j < n; ++j) {
        const double a = D.psiM[(size_t)i * n + j];
        gx += a * D.dpsiM[3 * ((size_t)i * n + j) + 0];
        gy += a * D.dpsiM[3 * ((size_t)i * n + j) + 1];
        gz += a * D.dpsiM[3 * ((size_t)i * n + j) + 2];
        lap += a * D.d2psiM[(size_t)i * n + j];
      }
      double* g = G + ((size_t)d * N + iat) * 3;
      g[0] = gx; g[1] = gy; g[2] = gz;
      L[(size_t)d * N + iat] = lap - (gx * gx + gy * gy + gz * gz);
    }
  }
  logs[2] = w.j1_log;
  logs[3] = w.j2_log;
  logs[4] = w.j3_log;
  for (int i = 0; i < N; ++i) {
    double* g1 = G + ((size_t)2 * N + i) * 3;
    double* g2 = G + ((size_t)3 * N + i) * 3;
    g1[0] = w.Gx[i]; g1[1] = w.Gy[i]; g1[2] = w.Gz[i];
    g2[0] = w.dUx[i]; g2[1] = w.dUy[i]; g2[2] = w.dUz[i];
    L[(size_t)2 * N + i] = -w.Lap[i];
    L[(size_t)3 * N + i] = w.d2Uat[i];
    if (s.has_j3) {
      double* g3 = G + ((size_t)4 * N + i) * 3;
      g3[0] = w.dU3x[i]; g3[1] = w.dU3y[i]; g3[2] = w.dU3z[i];
      L[(size_t)4 * N + i] = w.d2U3[i];
    }
  }
}

double orc_evaluate_log(void* h, const double* R, double* G, double* L) {
  const System& s = ((OrcSystem*)h)->sys;
  Walker w;
  walker_init(w, s, R, 1);
  double v = evaluate_log(w);
  if (G) std::copy(w.G.begin(), w.G.end(), G);
  if (L) std::copy(w.L.begin(), w.L.end(), L);
  return v;
}

}  // extern "C"
