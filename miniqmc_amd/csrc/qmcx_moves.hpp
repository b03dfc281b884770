// The fused move kernel: one workgroup owns one walker and moves the electrons
// of one chunk in order (setActive, evalGrad, makeMove, ratioGrad, accept/reject;
// src/Drivers/miniqmc.cpp:426-459).  Included by qmcx_kernels.hpp.
//
// Everything a move needs from one step to the next lives in LDS for the
// whole chunk: positions, the J2 sums, and with delay rank K > 1 the stale
// inverse rows of the chunk's electrons (Ainv is not modified until the flush).
// The products U_d . Ainv[row'] of an accepted row with the chunk's later rows
// are formed at accept time, so getInvRow (DelayedUpdate.h:96-114) reads LDS
// only.  The SPO values at the move's proposal come precomputed from
// k_spo_proposals (a sweep's proposals do not depend on its accept/reject
// chain); the move loads its 5 x n row first, without branches, and the
// barriers inside a move wait for LDS traffic only (lds_sync), so those loads
// stay in flight across them.
#pragma once

namespace qmcx {

constexpr int kRed = 12;   // reduction slots: gnow(3) ratio(1) gnew(3) J2(5)
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
__host__ __device__ inline MoveLds move_lds(int N, int M, int n, int K, int E) {
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
  L.oPart = lds_take(o, cache ? 0 : sizeof(double) * n);   // Sherman-Morrison scratch
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

// Sum the wave partials red[v*16 + w] (w < nw); every lane of every wave gets
// the same totals, in the same order.
__device__ inline void combine_partials(const double* red, int nw, double tot[kRed]) {
  // lane v < kRed sums slot v over the waves in order; the totals are broadcast as uniform values
  const int lane = threadIdx.x & 63;
  const int v = lane < kRed ? lane : 0;
  double x = 0.0;
  for (int q = 0; q < nw; ++q) x += red[v * 16 + q];
  const long long bits = __double_as_longlong(x);
  const int lo = (int)(bits & 0xffffffffLL), hi = (int)(bits >> 32);
#pragma unroll
  for (int i = 0; i < kRed; ++i) {
    const unsigned l = (unsigned)__builtin_amdgcn_readlane(lo, i), h = (unsigned)__builtin_amdgcn_readlane(hi, i);
    tot[i] = __longlong_as_double((long long)(((unsigned long long)h << 32) | l));
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

// KD: delayed updates (K > 1: inverse-row cache, Woodbury bordering) or Sherman-Morrison (K = 1)
template <typename T, int JPT, bool KD>
__global__ __launch_bounds__(kMoveThreads) void k_moves(SysP<T> S, BatchP<T> B, MoveArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  const int N = S.N, n = S.n, M = S.M, K = KD ? B.K : 1;
  const int E = a.e_end - a.e_begin;
  const int r0 = a.e_begin - a.det * n;
  const MoveLds L = move_lds<T>(N, M, n, K, E);
  double* sR = (double*)(smem + L.oR);
  double* sIon = (double*)(smem + L.oIon);
  double* sRng = (double*)(smem + L.oRng);
  double* sRed = (double*)(smem + L.oRed);
  T* sInv = (T*)(smem + L.oInv);
  T* sPsi = (T*)(smem + L.oPsi);
  double* sBinv = (double*)(smem + L.oBinv);
  double* sP = (double*)(smem + L.oP);
  double* sTm = (double*)(smem + L.oPart);
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
  // precomputed SPO rows of this chunk's proposals: [e][5][npad]
  const T* spo = (const T*)a.spo + (((size_t)w * 2 + a.det) * n + r0) * 5 * S.npad;

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
  if (KD)
    for (size_t i = tid; i < (size_t)E * n; i += nt) sAc[i] = Ainv[(size_t)r0 * n + i];
  for (int e = tid; e < E; e += nt) {
    double nr[3], u;
    move_randoms(a.seed, (uint64_t)(a.walker0 + w), (uint64_t)sweep, (uint64_t)(a.e_begin + e), nr, u);
    // the move's displacement and acceptance uniform; a caller's displacement and accept
    // decision (the flex stage calls) take their place, the decision as -inf (accept) / +inf
    const bool given = a.displ != nullptr;
    sRng[e * 4 + 0] = given ? a.displ[3 * w + 0] : a.sqrt_tau * nr[0];
    sRng[e * 4 + 1] = given ? a.displ[3 * w + 1] : a.sqrt_tau * nr[1];
    sRng[e * 4 + 2] = given ? a.displ[3 * w + 2] : a.sqrt_tau * nr[2];
    sRng[e * 4 + 3] = a.decide == nullptr ? u : (a.decide[w] != 0 ? -INFINITY : INFINITY);
  }
  int c = 0;
  double logd = 0, ph = 0, jl1 = 0, jl2 = 0;
  unsigned long long nacc = 0;
  __syncthreads();
  const Cell<T> C = *sCell;   // in registers for the minimum-image searches
  // Rows a move reads from HBM that no earlier move of the chunk changes: the moving electron's
  // J1 state, its dpsiM row and the SPO row at its proposal.  Move e+1's are loaded at the start
  // of move e (software pipelining; index clamped, so the loads are unconditional).
  const int jr = tid < n ? tid : n - 1;
  T nJ1[4], ndp[3], np5[5];
  double nj1p[5];
  const double* j1pw = a.j1p + (((size_t)w * 2 + a.det) * n + r0) * 8;
  auto prefetch = [&](int ee) {
    const int ie = a.e_begin + ee;
    nJ1[0] = J1[ie];
    nJ1[1] = J1[N + ie];
    nJ1[2] = J1[2 * N + ie];
    nJ1[3] = J1[3 * N + ie];
    const T* dq = dpsiM + (size_t)(r0 + ee) * 3 * n;
    ndp[0] = dq[jr];
    ndp[1] = dq[n + jr];
    ndp[2] = dq[2 * n + jr];
    const T* pq = spo + (size_t)ee * 5 * S.npad;
#pragma unroll
    for (int o = 0; o < 5; ++o) np5[o] = pq[(size_t)o * S.npad + jr];
#pragma unroll
    for (int k = 0; k < 5; ++k) nj1p[k] = j1pw[(size_t)ee * 8 + k];
  };
  prefetch(0);

  for (int e = 0; e < E; ++e) {
    const int iel = a.e_begin + e, row = r0 + e;
    QMCX_TICK(0)
    // ---- phase 1: issue this move's global loads, then getInvRow's small algebra
    const double x0 = sR[iel], y0 = sR[N + iel], z0 = sR[2 * N + iel];
    const double nx = x0 + sRng[e * 4 + 0];
    const double ny = y0 + sRng[e * 4 + 1];
    const double nz = z0 + sRng[e * 4 + 2];
    // stored state of the moving electron (J1 in HBM, J2 in LDS), its dpsiM / inverse rows and
    // the SPO row at its proposal, read before any update; unconditional loads (clamped
    // index), so that no branch merge forces an early wait on them
    const T Vat_old = nJ1[0], g1x = nJ1[1], g1y = nJ1[2], g1z = nJ1[3];
    const T Uat_old = sJ2[iel], g2x = sJ2[N + iel], g2y = sJ2[2 * N + iel], g2z = sJ2[3 * N + iel];
    const T dp0 = ndp[0], dp1 = ndp[1], dp2 = ndp[2];
    T p5[5];
#pragma unroll
    for (int o = 0; o < 5; ++o) p5[o] = np5[o];
    // J1 at the proposal: sum u, sum (d2u + 2 du/r), gradient (precomputed with the SPO row)
    double j1v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) j1v[k] = nj1p[k];
    prefetch(e + 1 < E ? e + 1 : e);
    T inv_sm = 0;
    if constexpr (!KD) inv_sm = Ainv[(size_t)row * n + jr];   // changed by the previous accept
    const T* prow = spo + (size_t)e * 5 * S.npad;
    if (KD && c > 0) {
      // y = Binv^T p with p_d = U_d . Ainv[row] formed when row d was accepted; parked in row c of Binv
      for (int r = tid; r < c; r += nt) {
        double s = 0;
        int q = 0;
        for (; q + 4 <= c; q += 4) {
          const double b0 = sBinv[q * K + r], b1 = sBinv[(q + 1) * K + r], b2 = sBinv[(q + 2) * K + r],
                       b3 = sBinv[(q + 3) * K + r];
          const double p0 = sQ[q * E + e], p1 = sQ[(q + 1) * E + e], p2 = sQ[(q + 2) * E + e], p3 = sQ[(q + 3) * E + e];
          s += b0 * p0;
          s += b1 * p1;
          s += b2 * p2;
          s += b3 * p3;
        }
        for (; q < c; ++q) s += sBinv[q * K + r] * sQ[q * E + e];
        sBinv[c * K + r] = s;
      }
    }
    lds_sync();
    QMCX_TICK(1)

    // ---- phase 2: invRow, evalGrad of the determinant and ratioGrad_compute against the
    // proposal's SPO row (einspline_spo.hpp:204-223, DiracDeterminant.cpp:103-124), J1/J2 at
    // the proposal; all partial sums reduced in one round
    double gd0 = 0, gd1 = 0, gd2 = 0, d0 = 0, d1 = 0, d2 = 0, d3 = 0;
    T vinv = 0;
    if (tid < n) {
      double v;
      if (KD) {
        v = (double)sAc[(size_t)e * n + tid];
        const double* y = sBinv + c * K;
        // four delayed rows' LDS reads in flight at a time; subtraction order unchanged
        int d = 0;
        for (; d + 4 <= c; d += 4) {
          const int e0 = sRows[d] - r0, e1 = sRows[d + 1] - r0, e2 = sRows[d + 2] - r0, e3 = sRows[d + 3] - r0;
          const double y0 = y[d], y1 = y[d + 1], y2 = y[d + 2], y3 = y[d + 3];
          const double a0 = (double)sAc[(size_t)e0 * n + tid], a1 = (double)sAc[(size_t)e1 * n + tid];
          const double a2 = (double)sAc[(size_t)e2 * n + tid], a3 = (double)sAc[(size_t)e3 * n + tid];
          v -= a0 * y0;
          v -= a1 * y1;
          v -= a2 * y2;
          v -= a3 * y3;
        }
        for (; d < c; ++d) v -= (double)sAc[(size_t)(sRows[d] - r0) * n + tid] * y[d];
      } else {
        v = (double)inv_sm;
      }
      vinv = (T)v;
      sInv[tid] = vinv;
      gd0 = (double)vinv * (double)dp0;
      gd1 = (double)vinv * (double)dp1;
      gd2 = (double)vinv * (double)dp2;
#pragma unroll
      for (int o = 0; o < 5; ++o) sPsi[o * n + tid] = p5[o];
      d0 = (double)vinv * (double)p5[0];
      d1 = (double)vinv * (double)p5[1];
      d2 = (double)vinv * (double)p5[2];
      d3 = (double)vinv * (double)p5[3];
    }
    for (int j = tid + nt; j < n; j += nt) {
      // only when n exceeds the workgroup (not the case for the shipped systems)
      double v = KD ? (double)sAc[(size_t)e * n + j] : (double)Ainv[(size_t)row * n + j];
      if (KD)
        for (int d = 0; d < c; ++d) v -= (double)sAc[(size_t)(sRows[d] - r0) * n + j] * sBinv[c * K + d];
      sInv[j] = (T)v;
      const T* dp = dpsiM + (size_t)row * 3 * n;
      gd0 += v * (double)dp[j];
      gd1 += v * (double)dp[n + j];
      gd2 += v * (double)dp[2 * n + j];
      const double vt = (double)(T)v;
      T o5[5];
#pragma unroll
      for (int o = 0; o < 5; ++o) {
        o5[o] = prow[(size_t)o * S.npad + j];
        sPsi[o * n + j] = o5[o];
      }
      d0 += vt * (double)o5[0];
      d1 += vt * (double)o5[1];
      d2 += vt * (double)o5[2];
      d3 += vt * (double)o5[3];
    }
    QMCX_TICK(5)
    {
      T r7 = 0, r8 = 0, r9 = 0, r10 = 0, r11 = 0;
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
      QMCX_TICK(6)
      QMCX_TICK(7)
      // slots: gnow 0-2, det ratio 3, its gradient 4-6, J2 7-11
      double v[kRed] = {gd0, gd1, gd2, d0, d1, d2, d3, r7, r8, r9, r10, r11};
#pragma unroll
      for (int i = 0; i < kRed; ++i) v[i] = wave_sum(v[i]);
      if (lane == 0)
#pragma unroll
        for (int i = 0; i < kRed; ++i) sRed[i * 16 + wid] = v[i];
    }
    lds_sync();
    QMCX_TICK(2)

    double red[kRed];
    combine_partials(sRed, nw, red);
    QMCX_TICK(3)

    // ---- phase 4: acceptance and updates
    const double rdet = red[3];
    const double ratio = rdet * exp((double)(Vat_old - (T)j1v[0])) * exp((double)(Uat_old - (T)red[7]));
    const double uacc = sRng[e * 4 + 3];
    const bool accept = a.mode == 0 ? (uacc < 0.5) : (uacc < ratio * ratio);
    if (B.trace != nullptr && tid == 0) {
      double* t = B.trace + ((size_t)w * N + iel) * 9;
      t[0] = ratio;
      t[1] = rdet;
      t[2] = accept ? 1.0 : 0.0;
      t[3] = red[0] + (double)g1x + (double)g2x;
      t[4] = red[1] + (double)g1y + (double)g2y;
      t[5] = red[2] + (double)g1z + (double)g2z;
      t[6] = red[4] / rdet + j1v[2] + red[8];
      t[7] = red[5] / rdet + j1v[3] + red[9];
      t[8] = red[6] / rdet + j1v[4] + red[10];
    }
    if (accept) {
      if (tid == 0) {
        logd += log(fabs(rdet));
        ph += rdet > 0.0 ? 0.0 : M_PI;
        jl1 += (double)(Vat_old - (T)j1v[0]);
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
        if (KD) Vb[(size_t)c * n + j] = sAc[(size_t)e * n + j];
      }
      if (!KD) {
        // Fahy/Sherman-Morrison (DelayedUpdate.h:166-173):
        // Ainv[m] -= ((Ainv[m].psiV - [m == row]) / ratio) * Ainv_old[row]
        double* tm = sTm;
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
          J1[iel] = (T)j1v[0];
          J1[N + iel] = (T)j1v[2];
          J1[2 * N + iel] = (T)j1v[3];
          J1[3 * N + iel] = (T)j1v[4];
          J1[4 * N + iel] = (T)j1v[1];
          sR[iel] = nx;
          sR[N + iel] = ny;
          sR[2 * N + iel] = nz;
        }
      }
    }
    // with K = 1 the next move reads rows of Ainv that other waves just wrote
    if (!KD) __syncthreads();
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
    B.ndelay[w] = KD ? c : 0;
  }
  if (KD) {
    for (int i = tid; i < c * K; i += nt) B.Binv[(size_t)w * K * K + i] = sBinv[i];
    for (int i = tid; i < c; i += nt) B.rows[(size_t)w * K + i] = sRows[i];
  }
}

}  // namespace qmcx
