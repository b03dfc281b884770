// The fused move kernel: one workgroup owns one walker and moves the electrons
// of one chunk in order (setActive, evalGrad, makeMove, ratioGrad, accept/reject;
// src/Drivers/miniqmc.cpp:426-459).  Included by qmcx_kernels.hpp.
//
// Everything a move needs from one step to the next lives in LDS for the
// whole chunk: positions, the J2 sums, and with delay rank K > 1 the stale
// inverse rows of the chunk's electrons (Ainv is not modified until the flush).
// The products U_d . Ainv[row'] of an accepted row with the chunk's later rows
// are formed at accept time, so getInvRow (DelayedUpdate.h:96-114) reads LDS
// only.  The spline stencil loads of a move are issued first and consumed after
// the first barrier; barriers inside a move wait for LDS traffic only
// (lds_sync), so those loads stay in flight across it.
#pragma once

namespace qmcx {

constexpr int kRed = 17;   // reduction slots: gnow(3) ratio(1) gnew(3) J2(5) J1(5)
constexpr int kMoveThreads = 512;

struct MoveLds {
  unsigned oR, oIon, oRng, oRed, oInv, oPsi, oBinv, oP, oPart, oRows, oSpec, oFn, oCell, oSpoG, oJ2, oAc, oQ, total;
};
__host__ __device__ inline unsigned lds_take(unsigned& o, size_t bytes) {
  unsigned r = o;
  o = (unsigned)((o + bytes + 15) & ~size_t(15));
  return r;
}
// E = electrons per chunk; the inverse-row cache is used when K > 1
template <typename T>
__host__ __device__ inline MoveLds move_lds(int N, int M, int n, int K, int S, int npad, int E) {
  MoveLds L;
  unsigned o = 0;
  const bool cache = K > 1;
  L.oR = lds_take(o, sizeof(double) * 3 * N);
  L.oIon = lds_take(o, sizeof(double) * 3 * M);
  L.oRng = lds_take(o, sizeof(double) * 4 * E);
  L.oRed = lds_take(o, sizeof(double) * kRed * 16);
  L.oInv = lds_take(o, sizeof(T) * n);
  L.oPsi = lds_take(o, sizeof(T) * 5 * n);
  L.oBinv = lds_take(o, cache ? sizeof(double) * K * K : 0);
  L.oP = lds_take(o, sizeof(double) * (K + 4));
  // the spline partials; reused by the Sherman-Morrison pass (n doubles)
  L.oPart = lds_take(o, sizeof(T) * S * 5 * npad > sizeof(double) * n ? sizeof(T) * S * 5 * npad : sizeof(double) * n);
  L.oRows = lds_take(o, sizeof(int) * K);
  L.oSpec = lds_take(o, sizeof(int) * M);
  L.oFn = lds_take(o, sizeof(Functor<T>) * (kMaxSpecies + 2));
  L.oCell = lds_take(o, sizeof(Cell<T>));
  L.oSpoG = lds_take(o, sizeof(double) * 12);
  L.oJ2 = lds_take(o, sizeof(T) * 5 * N);
  L.oAc = lds_take(o, cache ? sizeof(T) * (size_t)E * n : 0);
  L.oQ = lds_take(o, cache ? sizeof(double) * (size_t)K * E : 0);
  L.total = o;
  return L;
}

// ------------------------------------------------------------------ spline, split into issue and compute
template <typename T, int PER>
struct StencilLoads {
  typename V16<T>::type v[PER][4];
  T tx, ty, tz;
};

// lane (q, s): q = 16-byte orbital group, s = slice owning PER of the 16 (i,j) stencil columns
template <typename T, int PER>
__device__ inline void stencil_issue(const SysP<T>& S, const double* spoG, double x, double y, double z, int q, int s,
                                     StencilLoads<T, PER>& L) {
  using VT = typename V16<T>::type;
  constexpr int VW = V16<T>::n;
  int ix[3];
  double tt[3];
  spline_locate(spoG, S.ng, x, y, z, ix, tt);
  const T* base0 = S.coefs + (size_t)ix[0] * S.xs + (size_t)ix[1] * S.ys + (size_t)ix[2] * S.zs + (size_t)q * VW;
  const size_t zs = S.zs / VW;
#pragma unroll
  for (int pp = 0; pp < PER; ++pp) {
    const int p = s * PER + pp;
    const VT* src = reinterpret_cast<const VT*>(base0 + (size_t)(p >> 2) * S.xs + (size_t)(p & 3) * S.ys);
#pragma unroll
    for (int k = 0; k < 4; ++k) L.v[pp][k] = __builtin_nontemporal_load(src + k * zs);
  }
  L.tx = (T)tt[0];
  L.ty = (T)tt[1];
  L.tz = (T)tt[2];
}

// FP64 variant: loads and sums in batches of G columns inside one phase
template <typename T, int PER, int G>
__device__ inline void stencil_direct(const SysP<T>& S, const double* spoG, double x, double y, double z, int q, int s,
                                      T* part) {
  using VT = typename V16<T>::type;
  constexpr int VW = V16<T>::n;
  int ix[3];
  double tt[3];
  spline_locate(spoG, S.ng, x, y, z, ix, tt);
  const T* base0 = S.coefs + (size_t)ix[0] * S.xs + (size_t)ix[1] * S.ys + (size_t)ix[2] * S.zs + (size_t)q * VW;
  const size_t zs = S.zs / VW;
  T c[4], dc[4], d2c[4];
  bspline_weights<T>((T)tt[2], c, dc, d2c);
  const T tx = (T)tt[0], ty = (T)tt[1];
  T acc[5][VW];
#pragma unroll
  for (int o = 0; o < 5; ++o)
#pragma unroll
    for (int e = 0; e < VW; ++e) acc[o][e] = 0;
#pragma unroll
  for (int g0 = 0; g0 < PER; g0 += G) {
    VT v[G][4];
#pragma unroll
    for (int pp = 0; pp < G; ++pp) {
      const int p = s * PER + g0 + pp;
      const VT* src = reinterpret_cast<const VT*>(base0 + (size_t)(p >> 2) * S.xs + (size_t)(p & 3) * S.ys);
#pragma unroll
      for (int k = 0; k < 4; ++k) v[pp][k] = __builtin_nontemporal_load(src + k * zs);
    }
#pragma unroll
    for (int pp = 0; pp < G; ++pp) {
      const int p = s * PER + g0 + pp;
      T ai, dai, d2ai, bj, dbj, d2bj;
      knot_weight<T>(p >> 2, tx, ai, dai, d2ai);
      knot_weight<T>(p & 3, ty, bj, dbj, d2bj);
      const T w00 = ai * bj, w10 = dai * bj, w01 = ai * dbj, w20 = d2ai * bj;
#pragma unroll
      for (int e = 0; e < VW; ++e) {
        const T q0 = v[pp][0][e], q1 = v[pp][1][e], q2 = v[pp][2][e], q3 = v[pp][3][e];
        const T s0 = c[0] * q0 + c[1] * q1 + c[2] * q2 + c[3] * q3;
        const T s1 = dc[0] * q0 + dc[1] * q1 + dc[2] * q2 + dc[3] * q3;
        acc[0][e] += w00 * s0;
        acc[1][e] += w10 * s0;
        acc[2][e] += w01 * s0;
        acc[3][e] += w00 * s1;
        acc[4][e] += w20 * s0;
      }
    }
  }
#pragma unroll
  for (int o = 0; o < 5; ++o)
#pragma unroll
    for (int e = 0; e < VW; ++e) part[((size_t)s * 5 + o) * S.npad + q * VW + e] = acc[o][e];
}

// v, gx, gy, gz, hxx partial sums of this lane's columns -> part[s][5][npad]
template <typename T, int PER>
__device__ inline void stencil_compute(const SysP<T>& S, const StencilLoads<T, PER>& L, int q, int s, T* part) {
  constexpr int VW = V16<T>::n;
  T c[4], dc[4], d2c[4];
  bspline_weights<T>(L.tz, c, dc, d2c);
  T acc[5][VW];
#pragma unroll
  for (int o = 0; o < 5; ++o)
#pragma unroll
    for (int e = 0; e < VW; ++e) acc[o][e] = 0;
#pragma unroll
  for (int pp = 0; pp < PER; ++pp) {
    const int p = s * PER + pp;
    T ai, dai, d2ai, bj, dbj, d2bj;
    knot_weight<T>(p >> 2, L.tx, ai, dai, d2ai);
    knot_weight<T>(p & 3, L.ty, bj, dbj, d2bj);
    const T w00 = ai * bj, w10 = dai * bj, w01 = ai * dbj, w20 = d2ai * bj;
#pragma unroll
    for (int e = 0; e < VW; ++e) {
      const T q0 = L.v[pp][0][e], q1 = L.v[pp][1][e], q2 = L.v[pp][2][e], q3 = L.v[pp][3][e];
      const T s0 = c[0] * q0 + c[1] * q1 + c[2] * q2 + c[3] * q3;
      const T s1 = dc[0] * q0 + dc[1] * q1 + dc[2] * q2 + dc[3] * q3;
      acc[0][e] += w00 * s0;
      acc[1][e] += w10 * s0;
      acc[2][e] += w01 * s0;
      acc[3][e] += w00 * s1;
      acc[4][e] += w20 * s0;
    }
  }
#pragma unroll
  for (int o = 0; o < 5; ++o)
#pragma unroll
    for (int e = 0; e < VW; ++e) part[((size_t)s * 5 + o) * S.npad + q * VW + e] = acc[o][e];
}

// Sum the wave partials red[v*16 + w] (w < nw); every lane of every wave gets
// the same totals, in the same order.
__device__ inline void combine_partials(const double* red, int nw, double tot[kRed]) {
  const int lane = threadIdx.x & 63, q = lane & 15;
#pragma unroll
  for (int base = 0; base < kRed; base += 4) {
    const int v = base + (lane >> 4);
    double x = (v < kRed && q < nw) ? red[v * 16 + q] : 0.0;
    x += __shfl_xor(x, 8, 64);
    x += __shfl_xor(x, 4, 64);
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 1, 64);
#pragma unroll
    for (int g = 0; g < 4; ++g)
      if (base + g < kRed) tot[base + g] = __shfl(x, g * 16, 64);
  }
}

// J2 pair terms of electron j against the moving electron at (x, y, z)
template <typename T>
__device__ inline void j2_pair(const Cell<T>& C, const Functor<T>& f, double xj, double yj, double zj, double x, double y,
                               double z, T flip, T& u, T& du, T& d2u, T& dx, T& dy, T& dz) {
  const T r = min_image(C, (T)(xj - x), (T)(yj - y), (T)(zj - z), flip, dx, dy, dz);
  u = du = d2u = 0;
  if (r < f.rcut) functor_vgl(f, r, u, du, d2u);
}

__device__ inline void wave_lds_fence() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

#ifdef QMCX_PHASE_TIMING
// wall_clock64 ticks per phase, accumulated by lane 0 of workgroup 0
__device__ unsigned long long qmcx_phase_ticks[12];
#define QMCX_TICK(slot)                                                  \
  if (w == 0 && tid == 0) {                                              \
    const unsigned long long now_ = wall_clock64();                      \
    qmcx_phase_ticks[slot] += now_ - tick_;                              \
    tick_ = now_;                                                        \
  }
#else
#define QMCX_TICK(slot)
#endif

template <typename T, int JPT, int PER>
__global__ __launch_bounds__(kMoveThreads) void k_moves(SysP<T> S, BatchP<T> B, MoveArgs a) {
  constexpr int VW = V16<T>::n;
  constexpr int SL = 16 / PER;
  // FP32 keeps the stencil loads in flight across the first barrier; FP64 (twice the
  // registers per load) loads and sums inside phase 2
  constexpr bool kEarly = sizeof(T) == 4;
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  const int N = S.N, n = S.n, M = S.M, K = B.K;
  const int E = a.e_end - a.e_begin;
  const int r0 = a.e_begin - a.det * n;
  const MoveLds L = move_lds<T>(N, M, n, K, SL, S.npad, E);
  double* sR = (double*)(smem + L.oR);
  double* sIon = (double*)(smem + L.oIon);
  double* sRng = (double*)(smem + L.oRng);
  double* sRed = (double*)(smem + L.oRed);
  T* sInv = (T*)(smem + L.oInv);
  T* sPsi = (T*)(smem + L.oPsi);
  double* sBinv = (double*)(smem + L.oBinv);
  double* sP = (double*)(smem + L.oP);
  T* sPart = (T*)(smem + L.oPart);
  int* sRows = (int*)(smem + L.oRows);
  int* sSpec = (int*)(smem + L.oSpec);
  Functor<T>* sFn = (Functor<T>*)(smem + L.oFn);
  Cell<T>* sCell = (Cell<T>*)(smem + L.oCell);
  double* sSpoG = (double*)(smem + L.oSpoG);
  T* sJ2 = (T*)(smem + L.oJ2);
  T* sAc = (T*)(smem + L.oAc);
  double* sQ = (double*)(smem + L.oQ);
#ifdef QMCX_PHASE_TIMING
  unsigned long long tick_ = wall_clock64();
#endif
  const size_t moff = ((size_t)w * 2 + a.det) * n * n;
  T* Ainv = B.Ainv + moff;
  T* Amat = B.A + moff;
  T* dpsiM = B.dpsiM + ((size_t)w * 2 + a.det) * (size_t)n * 3 * n;
  T* d2psiM = B.d2psiM + moff;
  T* J2 = B.J2 + (size_t)w * 5 * N;
  T* J1 = B.J1 + (size_t)w * 5 * N;
  T* Vb = B.Vbuf + (size_t)w * K * n;
  const int sweep = B.sweep[0];

  // ---- chunk set-up: walker state into LDS
  for (int i = tid; i < 3 * N; i += nt) sR[i] = B.R[(size_t)w * 3 * N + i];
  for (int i = tid; i < 5 * N; i += nt) sJ2[i] = J2[i];
  for (int i = tid; i < 3 * M; i += nt) sIon[i] = S.ions[i];
  for (int i = tid; i < M; i += nt) sSpec[i] = S.species[i];
  {
    const int words = (int)(sizeof(Functor<T>) * (kMaxSpecies + 2) / 4);
    for (int i = tid; i < words; i += nt) ((int*)sFn)[i] = ((const int*)S.fn)[i];
  }
  if (tid == 0) *sCell = *S.cell;
  if (tid < 9) sSpoG[tid] = S.spoG[tid];
  if (K > 1)
    for (size_t i = tid; i < (size_t)E * n; i += nt) sAc[i] = Ainv[(size_t)r0 * n + i];
  for (int e = tid; e < E; e += nt) {
    double nr[3], u;
    move_randoms(a.seed, (uint64_t)(a.walker0 + w), (uint64_t)sweep, (uint64_t)(a.e_begin + e), nr, u);
    sRng[e * 4 + 0] = nr[0];
    sRng[e * 4 + 1] = nr[1];
    sRng[e * 4 + 2] = nr[2];
    sRng[e * 4 + 3] = u;
  }
  int c = 0;
  double logd = 0, ph = 0, jl1 = 0, jl2 = 0;
  unsigned long long nacc = 0;
  __syncthreads();
  const Cell<T>& C = *sCell;
  const int nq = S.npad / VW;
  const int sq = tid % nq, ss = tid / nq;
  const bool spline_lane = ss < SL;

  for (int e = 0; e < E; ++e) {
    const int iel = a.e_begin + e, row = r0 + e;
    QMCX_TICK(0)
    // ---- phase 1: issue this move's global loads, then getInvRow's small algebra
    const double x0 = sR[iel], y0 = sR[N + iel], z0 = sR[2 * N + iel];
    const bool given = a.displ != nullptr;
    const double nx = x0 + (given ? a.displ[3 * w + 0] : a.sqrt_tau * sRng[e * 4 + 0]);
    const double ny = y0 + (given ? a.displ[3 * w + 1] : a.sqrt_tau * sRng[e * 4 + 1]);
    const double nz = z0 + (given ? a.displ[3 * w + 2] : a.sqrt_tau * sRng[e * 4 + 2]);
    StencilLoads<T, kEarly ? PER : 1> st;
    if constexpr (kEarly)
      if (spline_lane) stencil_issue<T, PER>(S, sSpoG, nx, ny, nz, sq, ss, st);
    // stored state of the moving electron (J1 in HBM, J2 in LDS), read before any update
    const T Vat_old = J1[iel], g1x = J1[N + iel], g1y = J1[2 * N + iel], g1z = J1[3 * N + iel];
    const T Uat_old = sJ2[iel], g2x = sJ2[N + iel], g2y = sJ2[2 * N + iel], g2z = sJ2[3 * N + iel];
    T dp0 = 0, dp1 = 0, dp2 = 0, inv_sm = 0;
    if (tid < n) {
      const T* dp = dpsiM + (size_t)row * 3 * n;
      dp0 = dp[tid];
      dp1 = dp[n + tid];
      dp2 = dp[2 * n + tid];
      if (K == 1) inv_sm = Ainv[(size_t)row * n + tid];
    }
    if (K > 1 && c > 0) {
      // y = Binv^T p with p_d = U_d . Ainv[row] formed when row d was accepted; parked in row c of Binv
      for (int r = tid; r < c; r += nt) {
        double s = 0;
        for (int q = 0; q < c; ++q) s += sBinv[q * K + r] * sQ[q * E + e];
        sBinv[c * K + r] = s;
      }
    }
    lds_sync();
    QMCX_TICK(1)

    // ---- phase 2: invRow, evalGrad of the determinant, J1/J2 at the proposal, spline sums
    double gd0 = 0, gd1 = 0, gd2 = 0;
    T vinv = 0;
    if (tid < n) {
      double v;
      if (K > 1) {
        v = (double)sAc[(size_t)e * n + tid];
        const double* y = sBinv + c * K;
        for (int d = 0; d < c; ++d) v -= (double)sAc[(size_t)(sRows[d] - r0) * n + tid] * y[d];
      } else {
        v = (double)inv_sm;
      }
      vinv = (T)v;
      sInv[tid] = vinv;
      gd0 = (double)vinv * (double)dp0;
      gd1 = (double)vinv * (double)dp1;
      gd2 = (double)vinv * (double)dp2;
    }
    for (int j = tid + nt; j < n; j += nt) {
      // only when n exceeds the workgroup (not the case for the shipped systems)
      double v = K > 1 ? (double)sAc[(size_t)e * n + j] : (double)Ainv[(size_t)row * n + j];
      if (K > 1)
        for (int d = 0; d < c; ++d) v -= (double)sAc[(size_t)(sRows[d] - r0) * n + j] * sBinv[c * K + d];
      sInv[j] = (T)v;
      const T* dp = dpsiM + (size_t)row * 3 * n;
      gd0 += v * (double)dp[j];
      gd1 += v * (double)dp[n + j];
      gd2 += v * (double)dp[2 * n + j];
    }
    {
      T r7 = 0, r8 = 0, r9 = 0, r10 = 0, r11 = 0, r12 = 0, r13 = 0, r14 = 0, r15 = 0, r16 = 0;
      const bool up = iel < n;
#pragma unroll
      for (int q = 0; q < JPT; ++q) {
        const int j = tid + q * nt;
        if (j < N && j != iel) {
          const Functor<T>& f = sFn[kMaxSpecies + (((j < n) == up) ? 0 : 1)];
          T u, du, d2u, dx, dy, dz;
          j2_pair(C, f, sR[j], sR[N + j], sR[2 * N + j], nx, ny, nz, j < iel ? T(1) : T(-1), u, du, d2u, dx, dy, dz);
          r7 += u;
          r8 += du * dx;
          r9 += du * dy;
          r10 += du * dz;
          r11 += d2u + T(2) * du;
        }
      }
      for (int ci = tid; ci < M; ci += nt) {
        const Functor<T>& f = sFn[sSpec[ci]];
        if (f.ncoef > 0) {
          T dx, dy, dz;
          const T r = min_image(C, (T)(sIon[ci] - nx), (T)(sIon[M + ci] - ny), (T)(sIon[2 * M + ci] - nz), T(-1),
                                dx, dy, dz);
          if (r < f.rcut) {
            T u, du, d2u;
            functor_vgl(f, r, u, du, d2u);
            r12 += u;
            r13 += d2u + T(2) * du;
            r14 += du * dx;
            r15 += du * dy;
            r16 += du * dz;
          }
        }
      }
      double v[13] = {gd0, gd1, gd2, r7, r8, r9, r10, r11, r12, r13, r14, r15, r16};
#pragma unroll
      for (int i = 0; i < 13; ++i) v[i] = wave_sum(v[i]);
      if (lane == 0) {
        sRed[0 * 16 + wid] = v[0];
        sRed[1 * 16 + wid] = v[1];
        sRed[2 * 16 + wid] = v[2];
#pragma unroll
        for (int i = 3; i < 13; ++i) sRed[(i + 4) * 16 + wid] = v[i];
      }
    }
    if constexpr (kEarly) {
      if (spline_lane) stencil_compute<T, PER>(S, st, sq, ss, sPart);
    } else {
      if (spline_lane) stencil_direct<T, PER, (PER > 4 ? 4 : PER)>(S, sSpoG, nx, ny, nz, sq, ss, sPart);
    }
    lds_sync();
    QMCX_TICK(2)

    // ---- phase 3: SPO values (einspline_spo.hpp:204-223) and ratioGrad_compute (DiracDeterminant.cpp:103-124)
    {
      const T gx = (T)S.ng[0], gy = (T)S.ng[1], gz = (T)S.ng[2];
      double d0 = 0, d1 = 0, d2 = 0, d3 = 0;
      for (int j = tid; j < n; j += nt) {
        T o5[5];
#pragma unroll
        for (int o = 0; o < 5; ++o) {
          T sum = 0;
          for (int s2 = 0; s2 < SL; ++s2) sum += sPart[((size_t)s2 * 5 + o) * S.npad + j];
          o5[o] = sum;
        }
        o5[1] *= gx;
        o5[2] *= gy;
        o5[3] *= gz;
        o5[4] *= gx * gx;
#pragma unroll
        for (int o = 0; o < 5; ++o) sPsi[o * n + j] = o5[o];
        const double v = (double)(j == tid ? vinv : sInv[j]);
        d0 += v * (double)o5[0];
        d1 += v * (double)o5[1];
        d2 += v * (double)o5[2];
        d3 += v * (double)o5[3];
      }
      d0 = wave_sum(d0);
      d1 = wave_sum(d1);
      d2 = wave_sum(d2);
      d3 = wave_sum(d3);
      if (lane == 0) {
        sRed[3 * 16 + wid] = d0;
        sRed[4 * 16 + wid] = d1;
        sRed[5 * 16 + wid] = d2;
        sRed[6 * 16 + wid] = d3;
      }
    }
    lds_sync();
    double red[kRed];
    combine_partials(sRed, nw, red);
    QMCX_TICK(3)

    // ---- phase 4: acceptance and updates
    const double rdet = red[3];
    const double ratio = rdet * exp((double)(Vat_old - (T)red[12])) * exp((double)(Uat_old - (T)red[7]));
    const double uacc = sRng[e * 4 + 3];
    const bool accept = a.decide != nullptr ? (a.decide[w] != 0) : a.mode == 0 ? (uacc < 0.5) : (uacc < ratio * ratio);
    if (B.trace != nullptr && tid == 0) {
      double* t = B.trace + ((size_t)w * N + iel) * 9;
      t[0] = ratio;
      t[1] = rdet;
      t[2] = accept ? 1.0 : 0.0;
      t[3] = red[0] + (double)g1x + (double)g2x;
      t[4] = red[1] + (double)g1y + (double)g2y;
      t[5] = red[2] + (double)g1z + (double)g2z;
      t[6] = red[4] / rdet + red[14] + red[8];
      t[7] = red[5] / rdet + red[15] + red[9];
      t[8] = red[6] / rdet + red[16] + red[10];
    }
    if (accept) {
      if (tid == 0) {
        logd += log(fabs(rdet));
        ph += rdet > 0.0 ? 0.0 : M_PI;
        jl1 += (double)(Vat_old - (T)red[12]);
        jl2 += (double)(Uat_old - (T)red[7]);
        ++nacc;
      }
      // new Slater row and its derivatives (DiracDeterminant.cpp:135-141); stale inverse row for the flush
      for (int j = tid; j < n; j += nt) {
        Amat[(size_t)row * n + j] = sPsi[j];
        dpsiM[(size_t)row * 3 * n + j] = sPsi[n + j];
        dpsiM[(size_t)row * 3 * n + n + j] = sPsi[2 * n + j];
        dpsiM[(size_t)row * 3 * n + 2 * n + j] = sPsi[3 * n + j];
        d2psiM[(size_t)row * n + j] = sPsi[4 * n + j];
        if (K > 1) Vb[(size_t)c * n + j] = sAc[(size_t)e * n + j];
      }
      if (K == 1) {
        // Fahy/Sherman-Morrison (DelayedUpdate.h:166-173):
        // Ainv[m] -= ((Ainv[m].psiV - [m == row]) / ratio) * Ainv_old[row]
        double* tm = (double*)sPart;
        for (int m0 = wid * 4; m0 < n; m0 += nw * 4) {
          double sm[4] = {0, 0, 0, 0};
          const T* am = Ainv + (size_t)m0 * n;
          const int nr = min(4, n - m0);
          for (int j = lane; j < n; j += 64) {
            const double pv = (double)sPsi[j];
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (r < nr) sm[r] += (double)am[(size_t)r * n + j] * pv;
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) sm[r] = wave_sum(sm[r]);
          if (lane < nr) {
            const double mine = lane == 0 ? sm[0] : lane == 1 ? sm[1] : lane == 2 ? sm[2] : sm[3];
            tm[m0 + lane] = (mine - (m0 + lane == row ? 1.0 : 0.0)) / rdet;
          }
        }
        lds_sync();
        for (int m = wid; m < n; m += nw) {
          const T t = (T)tm[m];
          T* am = Ainv + (size_t)m * n;
          for (int j = lane; j < n; j += 64) am[j] -= t * sInv[j];
        }
      } else {
        // acceptRow (DelayedUpdate.h:124-153): p_d = -V_d . psiV (d <= c) and, for the
        // chunk's later rows r', U_c . Ainv[r'] used by their getInvRow
        const int nd = c + 1 + (E - 1 - e);
        for (int t2 = wid; t2 < nd; t2 += nw) {
          const T* v = t2 <= c ? (t2 == c ? sAc + (size_t)e * n : sAc + (size_t)(sRows[t2] - r0) * n)
                               : sAc + (size_t)(e + (t2 - c)) * n;
          double s = 0;
          for (int j = lane; j < n; j += 64) s += (double)v[j] * (double)sPsi[j];
          s = wave_sum(s);
          if (lane == 0) {
            if (t2 <= c) sP[t2] = -s;
            else sQ[c * E + e + (t2 - c)] = s;
          }
        }
        lds_sync();
        if (wid == 0) {
          // the capacitance inverse grows by one row and column, FP64, one wave
          double y = -sP[c];
          for (int i = 0; i < c; ++i) y += sBinv[c * K + i] * sP[i];
          const double yi = 1.0 / y;
          for (int s2 = lane; s2 < c; s2 += 64) {
            double t = 0;
            for (int r = 0; r < c; ++r) t += sBinv[s2 * K + r] * sP[r];
            sBinv[s2 * K + c] = yi * t;
          }
          wave_lds_fence();
          for (int s2 = lane; s2 < c; s2 += 64) {
            const double bsc = sBinv[s2 * K + c];
            for (int r = 0; r < c; ++r) sBinv[s2 * K + r] -= sBinv[c * K + r] * bsc;
          }
          wave_lds_fence();
          for (int i = lane; i < c; i += 64) sBinv[c * K + i] *= -yi;
          if (lane == 0) {
            sBinv[c * K + c] = yi;
            sRows[c] = row;
          }
        }
        ++c;
      }
      // J1 (OneBodyJastrow.h:271-283) and J2 (TwoBodyJastrow.h:307-353); the pair
      // terms at the old and the accepted position are recomputed here
      const bool up = iel < n;
#pragma unroll
      for (int q = 0; q < JPT; ++q) {
        const int j = tid + q * nt;
        if (j < N && j != iel) {
          const Functor<T>& f = sFn[kMaxSpecies + (((j < n) == up) ? 0 : 1)];
          const T flip = j < iel ? T(1) : T(-1);
          const double xj = sR[j], yj = sR[N + j], zj = sR[2 * N + j];
          T ou, odu, od2u, odx, ody, odz, cu, cdu, cd2u, cdx, cdy, cdz;
          j2_pair(C, f, xj, yj, zj, x0, y0, z0, flip, ou, odu, od2u, odx, ody, odz);
          j2_pair(C, f, xj, yj, zj, nx, ny, nz, flip, cu, cdu, cd2u, cdx, cdy, cdz);
          const T newl = cd2u + T(2) * cdu;
          sJ2[j] += cu - ou;
          sJ2[4 * N + j] += od2u + T(2) * odu - newl;
          sJ2[N + j] -= cdu * cdx - odu * odx;
          sJ2[2 * N + j] -= cdu * cdy - odu * ody;
          sJ2[3 * N + j] -= cdu * cdz - odu * odz;
        } else if (j == iel) {
          sJ2[iel] = (T)red[7];
          sJ2[N + iel] = (T)red[8];
          sJ2[2 * N + iel] = (T)red[9];
          sJ2[3 * N + iel] = (T)red[10];
          sJ2[4 * N + iel] = (T)(-red[11]);
          J1[iel] = (T)red[12];
          J1[N + iel] = (T)red[14];
          J1[2 * N + iel] = (T)red[15];
          J1[3 * N + iel] = (T)red[16];
          J1[4 * N + iel] = (T)red[13];
          sR[iel] = nx;
          sR[N + iel] = ny;
          sR[2 * N + iel] = nz;
        }
      }
    }
    // with K = 1 the next move reads rows of Ainv that other waves just wrote
    if (K == 1) __syncthreads();
    else lds_sync();
    QMCX_TICK(4)
  }

  for (int i = tid; i < 3 * N; i += nt) B.R[(size_t)w * 3 * N + i] = sR[i];
  for (int i = tid; i < 5 * N; i += nt) J2[i] = sJ2[i];
  if (tid == 0) {
    B.logdet[w * 2 + a.det] += logd;
    B.phase[w * 2 + a.det] += ph;
    B.jlog[w * 2 + 0] += jl1;
    B.jlog[w * 2 + 1] += jl2;
    B.nacc[w] += nacc;
    B.ndelay[w] = K > 1 ? c : 0;
  }
  if (K > 1) {
    for (int i = tid; i < c * K; i += nt) B.Binv[(size_t)w * K * K + i] = sBinv[i];
    for (int i = tid; i < c; i += nt) B.rows[(size_t)w * K + i] = sRows[i];
  }
}

}  // namespace qmcx
