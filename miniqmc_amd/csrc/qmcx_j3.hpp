// Three-body electron-electron-ion Jastrow on the device:
// ThreeBodyJastrow<PolynomialFunctor3D> (QMCWaveFunctions/Jastrow/ThreeBodyJastrow.h:30-677,
// Jastrow/PolynomialFunctor3D.h:34-560).  Included by qmcx_kernels.hpp.
//
// For ions I and electron pairs (j, k) with both electrons within L_I = rcut/2 of the ion,
//   U_j = sum_{k != j, I} F(r_jk, r_jI, r_kI),   log J3 = -1/2 sum_j U_j,
// with F = P(r_jk, r_jI, r_kI) ((r_jI - L)(r_kI - L))^3 and P a polynomial of degree <= 3 in
// each argument.  Per walker the state is U, grad U (x, y, z) and lap U of every electron (the
// reference's Uat, dUat, d2Uat), plus which ions each electron sits within the cutoff of (one
// bit per ion, the reference's elecs_inside lists turned around) and, per electron position,
// the list of those ions with distance and displacement.
#pragma once

namespace qmcx {

constexpr int kJ3Deg = 3;                      // N_eI, N_ee <= 3
constexpr int kJ3G = (kJ3Deg + 1) * (kJ3Deg + 1) * (kJ3Deg + 1);
constexpr int kJ3Near = 6;                     // ions within the cutoff of one electron position
constexpr int kJ3Rec = 1 + 5 * kJ3Near;        // near-ion record: count, then (ion, r, dx, dy, dz)
constexpr int kJ3Pairs = 128;                  // (electron, ion) pairs of one move phase
constexpr int kJ3MaxIons = 256;                // ions a three-body Jastrow takes (mask words)

// gamma[l][m][n] zero-padded to degree 3; L = rcut / 2
template <typename T>
struct Poly3 {
  T L;
  T g[kJ3G];
};

// Horner steps in one variable for coefficients c[0..3]: value, first and second derivative
template <typename T>
__device__ inline void horner3(const T c[4], T x, T& v, T& d, T& dd) {
  v = ((c[3] * x + c[2]) * x + c[1]) * x + c[0];
  d = (T(3) * c[3] * x + T(2) * c[2]) * x + c[1];
  dd = T(6) * c[3] * x + T(2) * c[2];
}

// F and its derivatives at (r12, r1, r2) with r1, r2 < L, in the layout of evaluateVGL
// (PolynomialFunctor3D.h:445-556): o = {F, F_12/r12, F_1/r1, F_2/r2, F_12,12, F_1,1, F_2,2,
// F_12,1/(r12 r1), F_12,2/(r12 r2)}.  The polynomial is reduced one variable at a time (r12,
// then r2, then r1), each by Horner's rule with its derivatives; the cutoff factor B^3,
// B = (r1 - L)(r2 - L), enters by the product rule.
template <typename T>
__device__ inline void poly3_vgl(const Poly3<T>& f, T r12, T r1, T r2, T o[9]) {
  T A[4], A2[4], A22[4], Bq[4], B2[4], Cq[4];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    T q[4], q1[4], q2[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) horner3<T>(&f.g[(l * 4 + m) * 4], r12, q[m], q1[m], q2[m]);
    T dd;
    horner3<T>(q, r2, A[l], A2[l], A22[l]);
    horner3<T>(q1, r2, Bq[l], B2[l], dd);
    T d1;
    horner3<T>(q2, r2, Cq[l], d1, dd);
  }
  T P, P1, P11, P2, P22, P0, P00, P01, P02, u, uu;
  horner3<T>(A, r1, P, P1, P11);
  horner3<T>(A2, r1, P2, u, uu);
  horner3<T>(A22, r1, P22, u, uu);
  horner3<T>(Bq, r1, P0, P01, uu);
  horner3<T>(Cq, r1, P00, u, uu);
  horner3<T>(B2, r1, P02, u, uu);
  const T a1 = r1 - f.L, a2 = r2 - f.L, B = a1 * a2;
  const T B1 = B * B, B3 = B1 * B;   // B^2, B^3 (C = 3)
  o[0] = P * B3;
  o[1] = P0 * B3 / r12;
  o[2] = (P1 * B3 + T(3) * P * B1 * a2) / r1;
  o[3] = (P2 * B3 + T(3) * P * B1 * a1) / r2;
  o[4] = P00 * B3;
  o[5] = P11 * B3 + T(6) * P1 * B1 * a2 + T(6) * P * B * a2 * a2;
  o[6] = P22 * B3 + T(6) * P2 * B1 * a1 + T(6) * P * B * a1 * a1;
  o[7] = (P01 * B3 + T(3) * P0 * B1 * a2) / (r12 * r1);
  o[8] = (P02 * B3 + T(3) * P0 * B1 * a1) / (r12 * r2);
}

template <typename T>
__device__ inline T poly3_v(const Poly3<T>& f, T r12, T r1, T r2) {
  T A[4];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    T q[4], d, dd;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const T* c = &f.g[(l * 4 + m) * 4];
      q[m] = ((c[3] * r12 + c[2]) * r12 + c[1]) * r12 + c[0];
    }
    horner3<T>(q, r2, A[l], d, dd);
  }
  const T P = ((A[3] * r1 + A[2]) * r1 + A[1]) * r1 + A[0];
  const T B = (r1 - f.L) * (r2 - f.L);
  return P * B * B * B;
}

// The engine's terms of one pair (ThreeBodyJastrow.h:483-567): electron j (whose sums are
// accumulated), electron k, ion I; displacements jk = r_k - r_j, jI = r_I - r_j, kI = r_I - r_k.
//   j side: U += F; G += F_1/r1 jI + F_12/r12 jk;
//           L -= F_12,12 + F_1,1 + 2 (F_12/r12 + F_1/r1) + 2 F_12,1/(r12 r1) jk.jI
//   k side: U += F; G += F_2/r2 kI - F_12/r12 jk;
//           L -= F_12,12 + F_2,2 + 2 (F_12/r12 + F_2/r2) - 2 F_12,2/(r12 r2) kI.jk
template <typename T>
__device__ inline void j3_pair_terms(const T o[9], const T jk[3], const T jI[3], const T kI[3], T js[5], T ks[5]) {
  const T jkjI = jk[0] * jI[0] + jk[1] * jI[1] + jk[2] * jI[2];
  const T kIjk = kI[0] * jk[0] + kI[1] * jk[1] + kI[2] * jk[2];
  js[0] = o[0];
  js[1] = o[2] * jI[0] + o[1] * jk[0];
  js[2] = o[2] * jI[1] + o[1] * jk[1];
  js[3] = o[2] * jI[2] + o[1] * jk[2];
  js[4] = -(o[4] + o[5] + T(2) * (o[1] + o[2]) + T(2) * o[7] * jkjI);
  ks[0] = o[0];
  ks[1] = o[3] * kI[0] - o[1] * jk[0];
  ks[2] = o[3] * kI[1] - o[1] * jk[1];
  ks[3] = o[3] * kI[2] - o[1] * jk[2];
  ks[4] = -(o[4] + o[6] + T(2) * (o[1] + o[3]) - T(2) * o[8] * kIjk);
}

}  // namespace qmcx
