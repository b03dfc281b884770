// Kernels of the VMC particle-move engine.  Included once, by qmcx_capi.hip.
//
// Data layout in HBM (per batch of W walkers, n = orbitals = electrons per spin,
// N = 2n electrons, M ions, K = delay rank):
//   R       double [W][3][N]       positions, SoA
//   A       T      [W][2][n][n]    Slater rows phi_j(r_i) (row = electron)
//   Ainv    T      [W][2][n][n]    psiM = (A^-1)^T (DiracDeterminant::psiM)
//   dpsiM   T      [W][2][n][3][n] orbital gradients of each row, SoA per row
//   d2psiM  T      [W][2][n][n]    h_xx of each row (einspline_spo.hpp:220)
//   J2, J1  T      [W][5][N]       Uat,dUx,dUy,dUz,d2Uat  /  Vat,Gx,Gy,Gz,Lap
//   Vbuf    T      [W][K][n]       stale inverse rows of the delayed electrons
//   Binv    T      [W][K][K]       inverse of the delayed capacitance matrix
//   rows    int    [W][K]          delayed rows;  ndelay int [W]
//   logdet, phase, jlog double [W][2]
#pragma once
#include "qmcx_device.hpp"

namespace qmcx {

template <typename T>
struct SysP {
  Cell<T> cell;
  double spoG[9];
  int ng[3];
  int N, n, M, nspecies;
  int npad;
  long long xs, ys, zs;
  const T* coefs;
  const double* ions;         // [3][M]
  const int* species;         // [M]
  const Functor<T>* fn;       // [kMaxSpecies] J1 then [2] J2
};

template <typename T>
struct BatchP {
  int W, K;
  double* R;
  T* A;
  T* Ainv;
  T* dpsiM;
  T* d2psiM;
  T* J2;
  T* J1;
  T* Vbuf;
  T* Binv;
  int* rows;
  int* ndelay;
  double* logdet;
  double* phase;
  double* jlog;
  unsigned long long* nacc;
  double* trace;   // [W][N][9] or null
  int* sweep;      // [0] global sweep index, [1] index of the first sweep of the current run
};

struct MoveArgs {
  int e_begin, e_end, det, mode, slices, walker0;
  double sqrt_tau;
  uint64_t seed;
};

// ------------------------------------------------------------------ LDS plan
struct MoveLds {
  size_t oR, oIon, oRng, oRedD, oInv0, oInv, oPsi, oBinv, oP, oRedT, oPart, oRows, oSpec, oFn, total;
};
__host__ __device__ inline size_t lds_take(size_t& o, size_t bytes) {
  size_t r = o;
  o = (o + bytes + 15) & ~size_t(15);
  return r;
}
template <typename T>
__host__ __device__ inline MoveLds move_lds(int N, int M, int n, int K, int S, int npad) {
  MoveLds L;
  size_t o = 0;
  L.oR = lds_take(o, sizeof(double) * 3 * N);
  L.oIon = lds_take(o, sizeof(double) * 3 * M);
  L.oRng = lds_take(o, sizeof(double) * 64 * 4);
  L.oRedD = lds_take(o, sizeof(double) * 16 * 16);
  L.oInv0 = lds_take(o, sizeof(T) * n);
  L.oInv = lds_take(o, sizeof(T) * n);
  L.oPsi = lds_take(o, sizeof(T) * 5 * n);
  L.oBinv = lds_take(o, sizeof(T) * K * K);
  L.oP = lds_take(o, sizeof(T) * (K + 4));
  L.oRedT = lds_take(o, sizeof(T) * 18 * 16);
  L.oPart = lds_take(o, sizeof(T) * S * 5 * npad);
  L.oRows = lds_take(o, sizeof(int) * K);
  L.oSpec = lds_take(o, sizeof(int) * M);
  L.oFn = lds_take(o, sizeof(Functor<T>) * (kMaxSpecies + 2));
  L.total = o;
  return L;
}

// ------------------------------------------------------------------ spline gather
// One workgroup evaluates NO outputs (5: v, gx, gy, gz, hxx; 10: full VGH) of
// every orbital at one point.  Lane (q, s): q = 16-byte orbital group, s =
// stencil slice owning 16/S of the (i,j) columns of the 4x4x4 stencil; each
// lane issues 4*16/S coalesced 16-byte loads along the orbital axis.  Slice
// partials go to LDS `part` [S][NO][npad] and are summed by `spline_finish`.
template <typename T, int NO>
__device__ inline void spline_gather(const SysP<T>& S, double x, double y, double z, int slices, T* part) {
  using VT = typename V16<T>::type;
  constexpr int VW = V16<T>::n;
  const int nq = S.npad / VW;
  const int tid = threadIdx.x;
  const int q = tid % nq, s = tid / nq;
  if (s >= slices) return;
  int ix[3];
  double tt[3];
  spline_locate(S.spoG, S.ng, x, y, z, ix, tt);
  T a[4], da[4], d2a[4], b[4], db[4], d2b[4], c[4], dc[4], d2c[4];
  bspline_weights<T>((T)tt[0], a, da, d2a);
  bspline_weights<T>((T)tt[1], b, db, d2b);
  bspline_weights<T>((T)tt[2], c, dc, d2c);
  T acc[NO][VW];
#pragma unroll
  for (int o = 0; o < NO; ++o)
#pragma unroll
    for (int e = 0; e < VW; ++e) acc[o][e] = 0;
  const int per = 16 / slices;
  const T* base0 = S.coefs + (size_t)ix[0] * S.xs + (size_t)ix[1] * S.ys + (size_t)ix[2] * S.zs + (size_t)q * VW;
  for (int pp = 0; pp < per; ++pp) {
    const int p = s * per + pp;
    const int i = p >> 2, j = p & 3;
    const VT* src = reinterpret_cast<const VT*>(base0 + (size_t)i * S.xs + (size_t)j * S.ys);
    VT v4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v4[k] = __builtin_nontemporal_load(src + (size_t)k * (S.zs / VW));
    const T w00 = a[i] * b[j], w10 = da[i] * b[j], w01 = a[i] * db[j], w20 = d2a[i] * b[j];
    const T w11 = da[i] * db[j], w02 = a[i] * d2b[j];
#pragma unroll
    for (int e = 0; e < VW; ++e) {
      T qv[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) qv[k] = v4[k][e];
      const T s0 = c[0] * qv[0] + c[1] * qv[1] + c[2] * qv[2] + c[3] * qv[3];
      const T s1 = dc[0] * qv[0] + dc[1] * qv[1] + dc[2] * qv[2] + dc[3] * qv[3];
      acc[0][e] += w00 * s0;
      acc[1][e] += w10 * s0;
      acc[2][e] += w01 * s0;
      acc[3][e] += w00 * s1;
      acc[4][e] += w20 * s0;
      if constexpr (NO == 10) {
        const T s2 = d2c[0] * qv[0] + d2c[1] * qv[1] + d2c[2] * qv[2] + d2c[3] * qv[3];
        acc[5][e] += w11 * s0;
        acc[6][e] += w10 * s1;
        acc[7][e] += w02 * s0;
        acc[8][e] += w01 * s1;
        acc[9][e] += w00 * s2;
      }
    }
  }
#pragma unroll
  for (int o = 0; o < NO; ++o)
#pragma unroll
    for (int e = 0; e < VW; ++e) part[((size_t)s * NO + o) * S.npad + q * VW + e] = acc[o][e];
}

// Sum the slice partials and scale by the grid's delta_inv powers.
// out(o, j) for j < norb is written at out[o*ostride + j].
template <typename T, int NO, typename TO>
__device__ inline void spline_finish(const SysP<T>& S, int slices, const T* part, TO* out, int ostride, int norb) {
  const T gx = (T)S.ng[0], gy = (T)S.ng[1], gz = (T)S.ng[2];
  const T scale[10] = {T(1), gx, gy, gz, gx * gx, gx * gy, gx * gz, gy * gy, gy * gz, gz * gz};
  for (int idx = threadIdx.x; idx < NO * norb; idx += blockDim.x) {
    const int o = idx / norb, j = idx - o * norb;
    T sum = 0;
    for (int s = 0; s < slices; ++s) sum += part[((size_t)s * NO + o) * S.npad + j];
    out[(size_t)o * ostride + j] = (TO)(sum * scale[o]);
  }
}

// ------------------------------------------------------------------ fused moves
// One workgroup = one walker; electrons [e_begin, e_end) of determinant `det`
// are moved in order.  With K = 1 every accepted move updates Ainv in place
// (Sherman-Morrison, DelayedUpdate.h:166-173); with K > 1 accepted rows are
// delayed (DelayedUpdate.h:124-153) and k_flush applies them afterwards.
template <typename T>
__global__ __launch_bounds__(1024) void k_moves(SysP<T> S, BatchP<T> B, MoveArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int N = S.N, n = S.n, M = S.M, K = B.K;
  const MoveLds L = move_lds<T>(N, M, n, K, a.slices, S.npad);
  double* sR = (double*)(smem + L.oR);
  double* sIon = (double*)(smem + L.oIon);
  double* sRng = (double*)(smem + L.oRng);
  T* sInv0 = (T*)(smem + L.oInv0);
  T* sInv = (T*)(smem + L.oInv);
  T* sPsi = (T*)(smem + L.oPsi);
  T* sBinv = (T*)(smem + L.oBinv);
  T* sP = (T*)(smem + L.oP);
  T* sRedT = (T*)(smem + L.oRedT);
  T* sPart = (T*)(smem + L.oPart);
  int* sRows = (int*)(smem + L.oRows);
  int* sSpec = (int*)(smem + L.oSpec);
  Functor<T>* sFn = (Functor<T>*)(smem + L.oFn);

  for (int i = tid; i < 3 * N; i += nt) sR[i] = B.R[(size_t)w * 3 * N + i];
  for (int i = tid; i < 3 * M; i += nt) sIon[i] = S.ions[i];
  for (int i = tid; i < M; i += nt) sSpec[i] = S.species[i];
  {
    const int words = (int)(sizeof(Functor<T>) * (kMaxSpecies + 2) / 4);
    const int* src = (const int*)S.fn;
    int* dst = (int*)sFn;
    for (int i = tid; i < words; i += nt) dst[i] = src[i];
  }
  const size_t moff = ((size_t)w * 2 + a.det) * n * n;
  T* Ainv = B.Ainv + moff;
  T* Amat = B.A + moff;
  T* dpsiM = B.dpsiM + ((size_t)w * 2 + a.det) * (size_t)n * 3 * n;
  T* d2psiM = B.d2psiM + moff;
  T* J2 = B.J2 + (size_t)w * 5 * N;
  T* J1 = B.J1 + (size_t)w * 5 * N;
  T* Vb = B.Vbuf + (size_t)w * K * n;
  const int sweep = B.sweep[0];
  int c = 0;
  double logd = 0, ph = 0, jl1 = 0, jl2 = 0;
  unsigned long long nacc = 0;
  const int JPT = (N + nt - 1) / nt;   // electrons owned per lane (<= 2, checked on the host)
  __syncthreads();

  for (int iel = a.e_begin; iel < a.e_end; ++iel) {
    const int row = iel - a.det * n;
    const int slot = (iel - a.e_begin) & 63;
    if (slot == 0) {
      if (tid < 64 && iel + tid < a.e_end) {
        double nr[3], u;
        move_randoms(a.seed, (uint64_t)(a.walker0 + w), (uint64_t)sweep, (uint64_t)(iel + tid), nr, u);
        sRng[tid * 4 + 0] = nr[0];
        sRng[tid * 4 + 1] = nr[1];
        sRng[tid * 4 + 2] = nr[2];
        sRng[tid * 4 + 3] = u;
      }
      __syncthreads();
    }
    // stored state of the moving electron, read before any lane can modify it
    const T Vat_old = J1[iel], Uat_old = J2[iel];
    const T g1x = J1[N + iel], g1y = J1[2 * N + iel], g1z = J1[3 * N + iel];
    const T g2x = J2[N + iel], g2y = J2[2 * N + iel], g2z = J2[3 * N + iel];

    // ---- getInvRow (DelayedUpdate.h:96-114): p = U Ainv[row]
    for (int j = tid; j < n; j += nt) sInv0[j] = Ainv[(size_t)row * n + j];
    if (c > 0) {
      const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
      const T* ar = Ainv + (size_t)row * n;
      for (int d = wid; d < c; d += nw) {
        const T* u = Amat + (size_t)sRows[d] * n;
        T s = 0;
        for (int j = lane; j < n; j += 64) s += u[j] * ar[j];
        s = wave_sum(s);
        if (lane == 0) sP[d] = s;
      }
    }
    __syncthreads();
    if (c > 0) {
      // y = Binv^T p, parked in row c of Binv (acceptRow reads it)
      for (int r = tid; r < c; r += nt) {
        T s = 0;
        for (int q = 0; q < c; ++q) s += sBinv[q * K + r] * sP[q];
        sBinv[c * K + r] = s;
      }
      __syncthreads();
    }

    // ---- proposal
    const double x0 = sR[iel], y0 = sR[N + iel], z0 = sR[2 * N + iel];
    const double nx = x0 + a.sqrt_tau * sRng[slot * 4 + 0];
    const double ny = y0 + a.sqrt_tau * sRng[slot * 4 + 1];
    const double nz = z0 + a.sqrt_tau * sRng[slot * 4 + 2];

    T red[17];
#pragma unroll
    for (int i = 0; i < 17; ++i) red[i] = 0;

    // invRow = Ainv[row] - V^T y ; evalGrad of the determinant (DiracDeterminant.cpp:80-89)
    for (int j = tid; j < n; j += nt) {
      T v = sInv0[j];
      for (int d = 0; d < c; ++d) v -= Ainv[(size_t)sRows[d] * n + j] * sBinv[c * K + d];
      sInv[j] = v;
      const T* dp = dpsiM + (size_t)row * 3 * n;
      red[0] += v * dp[j];
      red[1] += v * dp[n + j];
      red[2] += v * dp[2 * n + j];
    }

    // distance rows (old and proposed position) + J2 terms, one electron per lane slot
    T ou[2], odu[2], od2u[2], odx[2], ody[2], odz[2], cu[2], cdu[2], cd2u[2], cdx[2], cdy[2], cdz[2];
    const bool up = iel < n;
    for (int q = 0; q < JPT; ++q) {
      const int j = tid + q * nt;
      ou[q] = odu[q] = od2u[q] = odx[q] = ody[q] = odz[q] = 0;
      cu[q] = cdu[q] = cd2u[q] = cdx[q] = cdy[q] = cdz[q] = 0;
      if (j < N && j != iel) {
        const T flip = j < iel ? T(1) : T(-1);
        const Functor<T>& f = sFn[kMaxSpecies + (((j < n) == up) ? 0 : 1)];
        const double xj = sR[j], yj = sR[N + j], zj = sR[2 * N + j];
        T r = min_image(S.cell, (T)(xj - x0), (T)(yj - y0), (T)(zj - z0), flip, odx[q], ody[q], odz[q]);
        if (r < f.rcut) functor_vgl(f, r, ou[q], odu[q], od2u[q]);
        r = min_image(S.cell, (T)(xj - nx), (T)(yj - ny), (T)(zj - nz), flip, cdx[q], cdy[q], cdz[q]);
        if (r < f.rcut) functor_vgl(f, r, cu[q], cdu[q], cd2u[q]);
        red[7] += cu[q];
        red[8] += cdu[q] * cdx[q];
        red[9] += cdu[q] * cdy[q];
        red[10] += cdu[q] * cdz[q];
        red[11] += cd2u[q] + T(2) * cdu[q];
      }
    }
    // J1 against the ions at the proposed position (OneBodyJastrow.h:259-268)
    for (int ci = tid; ci < M; ci += nt) {
      const Functor<T>& f = sFn[sSpec[ci]];
      if (f.ncoef > 0) {
        T dx, dy, dz;
        const T r = min_image(S.cell, (T)(sIon[ci] - nx), (T)(sIon[M + ci] - ny), (T)(sIon[2 * M + ci] - nz), T(-1),
                              dx, dy, dz);
        if (r < f.rcut) {
          T u, du, d2u;
          functor_vgl(f, r, u, du, d2u);
          red[12] += u;
          red[13] += d2u + T(2) * du;
          red[14] += du * dx;
          red[15] += du * dy;
          red[16] += du * dz;
        }
      }
    }
    // SPO at the proposed position (einspline_spo::evaluate, einspline_spo.hpp:204-223)
    spline_gather<T, 5>(S, nx, ny, nz, a.slices, sPart);
    __syncthreads();
    spline_finish<T, 5, T>(S, a.slices, sPart, sPsi, n, n);
    __syncthreads();
    // ratioGrad_compute (DiracDeterminant.cpp:103-124)
    for (int j = tid; j < n; j += nt) {
      const T v = sInv[j];
      red[3] += v * sPsi[j];
      red[4] += v * sPsi[n + j];
      red[5] += v * sPsi[2 * n + j];
      red[6] += v * sPsi[3 * n + j];
    }
    block_sum<T, 17>(red, sRedT);

    const T rdet = red[3];
    const double ratio = (double)rdet * exp((double)(Vat_old - red[12])) * exp((double)(Uat_old - red[7]));
    const double uacc = sRng[slot * 4 + 3];
    const bool accept = a.mode == 0 ? (uacc < 0.5) : (uacc < ratio * ratio);
    if (B.trace != nullptr && tid == 0) {
      double* t = B.trace + ((size_t)w * N + iel) * 9;
      t[0] = ratio;
      t[1] = (double)rdet;
      t[2] = accept ? 1.0 : 0.0;
      t[3] = (double)(red[0] + g1x + g2x);
      t[4] = (double)(red[1] + g1y + g2y);
      t[5] = (double)(red[2] + g1z + g2z);
      t[6] = (double)(red[4] / rdet + red[14] + red[8]);
      t[7] = (double)(red[5] / rdet + red[15] + red[9]);
      t[8] = (double)(red[6] / rdet + red[16] + red[10]);
    }
    if (accept) {
      if (tid == 0) {
        logd += log(fabs((double)rdet));
        ph += rdet > T(0) ? 0.0 : M_PI;
        jl1 += (double)(Vat_old - red[12]);
        jl2 += (double)(Uat_old - red[7]);
        ++nacc;
      }
      // new Slater row and its derivatives (DiracDeterminant.cpp:135-141)
      for (int j = tid; j < n; j += nt) {
        Amat[(size_t)row * n + j] = sPsi[j];
        dpsiM[(size_t)row * 3 * n + j] = sPsi[n + j];
        dpsiM[(size_t)row * 3 * n + n + j] = sPsi[2 * n + j];
        dpsiM[(size_t)row * 3 * n + 2 * n + j] = sPsi[3 * n + j];
        d2psiM[(size_t)row * n + j] = sPsi[4 * n + j];
      }
      if (K == 1) {
        // Fahy/Sherman-Morrison: Ainv[m] -= ((Ainv[m].psiV - delta_m,row) / ratio) * Ainv_old[row]
        const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
        const T rinv = T(1) / rdet;
        for (int m = wid; m < n; m += nw) {
          T* am = Ainv + (size_t)m * n;
          T s = 0;
          for (int j = lane; j < n; j += 64) s += am[j] * sPsi[j];
          s = wave_sum(s);
          const T t = (s - (m == row ? T(1) : T(0))) * rinv;
          for (int j = lane; j < n; j += 64) am[j] -= t * sInv0[j];
        }
      } else {
        // acceptRow (DelayedUpdate.h:124-153)
        for (int j = tid; j < n; j += nt) Vb[(size_t)c * n + j] = sInv0[j];
        {
          const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
          for (int d = wid; d <= c; d += nw) {
            const T* v = (d == c) ? sInv0 : Vb + (size_t)d * n;
            T s = 0;
            for (int j = lane; j < n; j += 64) s += v[j] * sPsi[j];
            s = wave_sum(s);
            if (lane == 0) sP[d] = -s;
          }
        }
        __syncthreads();
        T y = -sP[c];
        for (int i = 0; i < c; ++i) y += sBinv[c * K + i] * sP[i];
        const T yi = T(1) / y;
        for (int s2 = tid; s2 < c; s2 += nt) {
          T t = 0;
          for (int r = 0; r < c; ++r) t += sBinv[s2 * K + r] * sP[r];
          sBinv[s2 * K + c] = yi * t;
        }
        __syncthreads();
        for (int idx = tid; idx < c * c; idx += nt) {
          const int s2 = idx / c, r = idx - s2 * c;
          sBinv[s2 * K + r] -= sBinv[c * K + r] * sBinv[s2 * K + c];
        }
        __syncthreads();
        for (int i = tid; i < c; i += nt) sBinv[c * K + i] *= -yi;
        if (tid == 0) {
          sBinv[c * K + c] = yi;
          sRows[c] = row;
        }
        ++c;
      }
      // J1 (OneBodyJastrow.h:271-283) and J2 (TwoBodyJastrow.h:307-353)
      for (int q = 0; q < JPT; ++q) {
        const int j = tid + q * nt;
        if (j < N && j != iel) {
          const T newl = cd2u[q] + T(2) * cdu[q];
          J2[j] += cu[q] - ou[q];
          J2[4 * N + j] += od2u[q] + T(2) * odu[q] - newl;
          J2[N + j] -= cdu[q] * cdx[q] - odu[q] * odx[q];
          J2[2 * N + j] -= cdu[q] * cdy[q] - odu[q] * ody[q];
          J2[3 * N + j] -= cdu[q] * cdz[q] - odu[q] * odz[q];
        } else if (j == iel) {
          J2[iel] = red[7];
          J2[N + iel] = red[8];
          J2[2 * N + iel] = red[9];
          J2[3 * N + iel] = red[10];
          J2[4 * N + iel] = -red[11];
          J1[iel] = red[12];
          J1[N + iel] = red[14];
          J1[2 * N + iel] = red[15];
          J1[3 * N + iel] = red[16];
          J1[4 * N + iel] = red[13];
          sR[iel] = nx;
          sR[N + iel] = ny;
          sR[2 * N + iel] = nz;
        }
      }
    }
    __syncthreads();
  }

  for (int i = tid; i < 3 * N; i += nt) B.R[(size_t)w * 3 * N + i] = sR[i];
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

// ------------------------------------------------------------------ delayed flush
// updateInvMat for K > 1 (DelayedUpdate.h:176-193), rows [m0, m0+FB) of one walker:
//   T  = Ainv U^T - E          (U rows = new Slater rows, E_{m,d} = [m == rows_d])
//   T' = T Binv
//   Ainv -= T' V               (V rows = stale inverse rows)
constexpr int kFlushRows = 16;
template <typename T>
__global__ __launch_bounds__(256) void k_flush_generic(SysP<T> S, BatchP<T> B, int det) {
  const int w = blockIdx.y, m0 = blockIdx.x * kFlushRows;
  const int c = B.ndelay[w];
  if (c == 0) return;
  const int n = S.n, K = B.K, tid = threadIdx.x, nt = blockDim.x;
  extern __shared__ __align__(16) unsigned char smem[];
  T* sT = (T*)smem;                     // [FB][K]
  T* sT2 = sT + kFlushRows * K;         // [FB][K]
  int* sRows = (int*)(sT2 + kFlushRows * K);
  const size_t moff = ((size_t)w * 2 + det) * n * n;
  T* Ainv = B.Ainv + moff;
  const T* Amat = B.A + moff;
  const T* Vb = B.Vbuf + (size_t)w * K * n;
  const T* Bi = B.Binv + (size_t)w * K * K;
  for (int i = tid; i < c; i += nt) sRows[i] = B.rows[(size_t)w * K + i];
  __syncthreads();
  const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  for (int idx = wid; idx < kFlushRows * c; idx += nw) {
    const int m = idx / c, d = idx - m * c;
    T s = 0;
    if (m0 + m < n) {
      const T* ar = Ainv + (size_t)(m0 + m) * n;
      const T* ur = Amat + (size_t)sRows[d] * n;
      for (int j = lane; j < n; j += 64) s += ar[j] * ur[j];
    }
    s = wave_sum(s);
    if (lane == 0) sT[m * K + d] = s - ((m0 + m) == sRows[d] ? T(1) : T(0));
  }
  __syncthreads();
  for (int idx = tid; idx < kFlushRows * c; idx += nt) {
    const int m = idx / c, d = idx - m * c;
    T s = 0;
    for (int e = 0; e < c; ++e) s += sT[m * K + e] * Bi[e * K + d];
    sT2[m * K + d] = s;
  }
  __syncthreads();
  for (int j = tid; j < n; j += nt) {
    T acc[kFlushRows];
#pragma unroll
    for (int m = 0; m < kFlushRows; ++m) acc[m] = 0;
    for (int d = 0; d < c; ++d) {
      const T v = Vb[(size_t)d * n + j];
#pragma unroll
      for (int m = 0; m < kFlushRows; ++m) acc[m] += sT2[m * K + d] * v;
    }
#pragma unroll
    for (int m = 0; m < kFlushRows; ++m)
      if (m0 + m < n) Ainv[(size_t)(m0 + m) * n + j] -= acc[m];
  }
}

// ------------------------------------------------------------------ FP64 LU + inverse
// One workgroup per matrix.  Right-looking LU with partial pivoting (largest
// |a| in the column, first index on ties, as LAPACK getrf), log|det| from the
// U diagonal, sign from pivots and diagonal signs (DiracMatrix.h:26-38);
// then column t of a^-1 is solved by lane t, which is row t of psiM = (a^-1)^T.
template <typename TI, typename TO>
__global__ __launch_bounds__(1024) void k_lu_invert(int n, const TI* A, TO* psiM, double* ws, double* logdet,
                                                    double* phase, int ld_stride) {
  const int mat = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  extern __shared__ __align__(16) unsigned char smem[];
  int* piv = (int*)smem;                           // [n]
  double* rv = (double*)(smem + ((n * 4 + 15) & ~15));   // [16]
  int* ri = (int*)(rv + 16);                       // [16]
  double* LU = ws + (size_t)mat * 2 * n * n;
  double* X = LU + (size_t)n * n;
  const TI* a = A + (size_t)mat * n * n;
  for (size_t i = tid; i < (size_t)n * n; i += nt) LU[i] = (double)a[i];
  __syncthreads();
  const int lane = tid & 63, wid = tid >> 6, nw = (nt + 63) >> 6;
  double logacc = 0.0;
  int sign = 1;
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = n;
    for (int r = k + tid; r < n; r += nt) {
      const double v = fabs(LU[(size_t)r * n + k]);
      if (v > best) { best = v; bi = r; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { rv[wid] = best; ri[wid] = bi; }
    __syncthreads();
    best = rv[0];
    bi = ri[0];
    for (int q = 1; q < nw; ++q)
      if (rv[q] > best || (rv[q] == best && ri[q] < bi)) { best = rv[q]; bi = ri[q]; }
    const int p = bi;
    if (tid == 0) piv[k] = p;
    if (p != k)
      for (int col = tid; col < n; col += nt) {
        const double t = LU[(size_t)k * n + col];
        LU[(size_t)k * n + col] = LU[(size_t)p * n + col];
        LU[(size_t)p * n + col] = t;
      }
    __syncthreads();
    const double d = LU[(size_t)k * n + k];
    if (p != k) sign = -sign;
    if (d < 0) sign = -sign;
    logacc += log(fabs(d));
    const double dinv = 1.0 / d;
    for (int r = k + 1 + tid; r < n; r += nt) LU[(size_t)r * n + k] *= dinv;
    __syncthreads();
    for (int r = k + 1 + wid; r < n; r += nw) {
      const double l = LU[(size_t)r * n + k];
      for (int col = k + 1 + lane; col < n; col += 64) LU[(size_t)r * n + col] -= l * LU[(size_t)k * n + col];
    }
    __syncthreads();
  }
  if (tid == 0) {
    logdet[(size_t)mat * ld_stride] = logacc;
    phase[(size_t)mat * ld_stride] = sign > 0 ? 0.0 : M_PI;
  }
  for (int t = tid; t < n; t += nt) {
    double* x = X + (size_t)t * n;
    for (int i = 0; i < n; ++i) x[i] = (i == t) ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k) {
      const int p = piv[k];
      if (p != k) { const double tmp = x[k]; x[k] = x[p]; x[p] = tmp; }
    }
    for (int i = 0; i < n; ++i) {
      double s = x[i];
      const double* li = LU + (size_t)i * n;
      for (int k = 0; k < i; ++k) s -= li[k] * x[k];
      x[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = x[i];
      const double* ui = LU + (size_t)i * n;
      for (int k = i + 1; k < n; ++k) s -= ui[k] * x[k];
      x[i] = s / ui[i];
    }
    TO* out = psiM + (size_t)mat * n * n + (size_t)t * n;
    for (int i = 0; i < n; ++i) out[i] = (TO)x[i];
  }
}

// ------------------------------------------------------------------ SPO rows (evaluateLog)
// SPO at every electron of every walker -> A, dpsiM, d2psiM
// (evaluate_notranspose, SPOSet.h:94; DiracDeterminant.cpp:264-280).
template <typename T>
__global__ __launch_bounds__(1024) void k_spo_rows(SysP<T> S, BatchP<T> B, int slices) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* part = (T*)smem;
  T* out = part + (size_t)slices * 5 * S.npad;  // [5][n]
  const int e = blockIdx.x, w = blockIdx.y, N = S.N, n = S.n;
  const double* R = B.R + (size_t)w * 3 * N;
  spline_gather<T, 5>(S, R[e], R[N + e], R[2 * N + e], slices, part);
  __syncthreads();
  spline_finish<T, 5, T>(S, slices, part, out, n, n);
  __syncthreads();
  const int det = e < n ? 0 : 1, row = e - det * n;
  const size_t moff = ((size_t)w * 2 + det) * n * n;
  T* Arow = B.A + moff + (size_t)row * n;
  T* dp = B.dpsiM + ((size_t)w * 2 + det) * (size_t)n * 3 * n + (size_t)row * 3 * n;
  T* d2 = B.d2psiM + moff + (size_t)row * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    Arow[j] = out[j];
    dp[j] = out[n + j];
    dp[n + j] = out[2 * n + j];
    dp[2 * n + j] = out[3 * n + j];
    d2[j] = out[4 * n + j];
  }
}

// stand-alone VGH at arbitrary points: out[p][10][norb] (FP64)
template <typename T>
__global__ __launch_bounds__(1024) void k_spline_vgh(SysP<T> S, const double* r, double* out, int norb, int slices) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* part = (T*)smem;
  const int p = blockIdx.x;
  spline_gather<T, 10>(S, r[3 * p], r[3 * p + 1], r[3 * p + 2], slices, part);
  __syncthreads();
  spline_finish<T, 10, double>(S, slices, part, out + (size_t)p * 10 * norb, norb, norb);
}

// stand-alone distance rows: lane per (source, target)
template <typename T>
__global__ void k_distance_rows(Cell<T> C, int npts, const double* src, const int* flip_ind, int nt,
                                const double* tgt, double* r, double* dr) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npts * nt) return;
  const int p = idx / nt, t = idx - p * nt;
  const T flip = t < flip_ind[p] ? T(1) : T(-1);
  T dx, dy, dz;
  const T rr = min_image(C, (T)(tgt[3 * t] - src[3 * p]), (T)(tgt[3 * t + 1] - src[3 * p + 1]),
                         (T)(tgt[3 * t + 2] - src[3 * p + 2]), flip, dx, dy, dz);
  r[idx] = (double)rr;
  dr[3 * (size_t)idx] = (double)dx;
  dr[3 * (size_t)idx + 1] = (double)dy;
  dr[3 * (size_t)idx + 2] = (double)dz;
}

// ------------------------------------------------------------------ Jastrow from scratch
// J1 per electron over the ions and J2 per electron over its full row; the
// sums equal the lower-triangle recompute of TwoBodyJastrow.h:356-400.
template <typename T>
__global__ __launch_bounds__(1024) void k_jastrow_init(SysP<T> S, BatchP<T> B) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x, N = S.N, n = S.n, M = S.M;
  double* sR = (double*)smem;
  double* sRed = sR + 3 * N;
  Functor<T>* sFn = (Functor<T>*)(sRed + 32);
  for (int i = tid; i < 3 * N; i += nt) sR[i] = B.R[(size_t)w * 3 * N + i];
  {
    const int words = (int)(sizeof(Functor<T>) * (kMaxSpecies + 2) / 4);
    for (int i = tid; i < words; i += nt) ((int*)sFn)[i] = ((const int*)S.fn)[i];
  }
  __syncthreads();
  T* J2 = B.J2 + (size_t)w * 5 * N;
  T* J1 = B.J1 + (size_t)w * 5 * N;
  double vsum = 0, usum = 0;
  for (int i = tid; i < N; i += nt) {
    const double x = sR[i], y = sR[N + i], z = sR[2 * N + i];
    T v = 0, gx = 0, gy = 0, gz = 0, lap = 0;
    for (int ci = 0; ci < M; ++ci) {
      const Functor<T>& f = sFn[S.species[ci]];
      if (f.ncoef == 0) continue;
      T dx, dy, dz;
      const T r = min_image(S.cell, (T)(S.ions[ci] - x), (T)(S.ions[M + ci] - y), (T)(S.ions[2 * M + ci] - z), T(-1),
                            dx, dy, dz);
      if (r < f.rcut) {
        T u, du, d2u;
        functor_vgl(f, r, u, du, d2u);
        v += u;
        lap += d2u + T(2) * du;
        gx += du * dx;
        gy += du * dy;
        gz += du * dz;
      }
    }
    J1[i] = v;
    J1[N + i] = gx;
    J1[2 * N + i] = gy;
    J1[3 * N + i] = gz;
    J1[4 * N + i] = lap;
    vsum += (double)v;
    T U = 0, hx = 0, hy = 0, hz = 0, l2 = 0;
    const bool up = i < n;
    for (int j = 0; j < N; ++j) {
      if (j == i) continue;
      const Functor<T>& f = sFn[kMaxSpecies + (((j < n) == up) ? 0 : 1)];
      T dx, dy, dz;
      const T r = min_image(S.cell, (T)(sR[j] - x), (T)(sR[N + j] - y), (T)(sR[2 * N + j] - z),
                            j < i ? T(1) : T(-1), dx, dy, dz);
      if (r < f.rcut) {
        T u, du, d2u;
        functor_vgl(f, r, u, du, d2u);
        U += u;
        l2 += d2u + T(2) * du;
        hx += du * dx;
        hy += du * dy;
        hz += du * dz;
      }
    }
    J2[i] = U;
    J2[N + i] = hx;
    J2[2 * N + i] = hy;
    J2[3 * N + i] = hz;
    J2[4 * N + i] = -l2;
    usum += (double)U;
  }
  double red[2] = {vsum, usum};
  block_sum<double, 2>(red, sRed);
  if (tid == 0) {
    B.jlog[w * 2 + 0] = -red[0];
    B.jlog[w * 2 + 1] = -0.5 * red[1];
  }
}

// ------------------------------------------------------------------ evaluateGL
// determinant part, one wave per electron row (DiracDeterminant.cpp:158-187)
template <typename T>
__global__ __launch_bounds__(256) void k_det_gl(SysP<T> S, BatchP<T> B, double* G, double* Lp) {
  const int w = blockIdx.y, N = S.N, n = S.n;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int e = blockIdx.x * nw + wid;
  if (e >= N) return;
  const int det = e < n ? 0 : 1, row = e - det * n;
  const size_t moff = ((size_t)w * 2 + det) * n * n;
  const T* a = B.Ainv + moff + (size_t)row * n;
  const T* dp = B.dpsiM + ((size_t)w * 2 + det) * (size_t)n * 3 * n + (size_t)row * 3 * n;
  const T* d2 = B.d2psiM + moff + (size_t)row * n;
  T gx = 0, gy = 0, gz = 0, l = 0;
  for (int j = lane; j < n; j += 64) {
    const T v = a[j];
    gx += v * dp[j];
    gy += v * dp[n + j];
    gz += v * dp[2 * n + j];
    l += v * d2[j];
  }
  gx = wave_sum(gx);
  gy = wave_sum(gy);
  gz = wave_sum(gz);
  l = wave_sum(l);
  if (lane == 0) {
    const size_t o = (size_t)w * N + e;
    G[3 * o] = gx;
    G[3 * o + 1] = gy;
    G[3 * o + 2] = gz;
    Lp[o] = (double)l - ((double)gx * gx + (double)gy * gy + (double)gz * gz);
  }
}

// Jastrow part + LogValue (WaveFunction.cpp:261-279; J logs from the sums,
// TwoBodyJastrow.h:413-430, OneBodyJastrow.h:148-160)
template <typename T>
__global__ __launch_bounds__(256) void k_finish_gl(SysP<T> S, BatchP<T> B, double* G, double* Lp, double* logv,
                                                   double* hist, int hist_cap) {
  __shared__ double sRed[32];
  const int w = blockIdx.x, N = S.N;
  const T* J2 = B.J2 + (size_t)w * 5 * N;
  const T* J1 = B.J1 + (size_t)w * 5 * N;
  double vs = 0, us = 0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const size_t o = (size_t)w * N + i;
    G[3 * o] += (double)J1[N + i] + (double)J2[N + i];
    G[3 * o + 1] += (double)J1[2 * N + i] + (double)J2[2 * N + i];
    G[3 * o + 2] += (double)J1[3 * N + i] + (double)J2[3 * N + i];
    Lp[o] += -(double)J1[4 * N + i] + (double)J2[4 * N + i];
    vs += (double)J1[i];
    us += (double)J2[i];
  }
  double red[2] = {vs, us};
  block_sum<double, 2>(red, sRed);
  if (threadIdx.x == 0) {
    B.jlog[w * 2 + 0] = -red[0];
    B.jlog[w * 2 + 1] = -0.5 * red[1];
    const double v = B.logdet[w * 2] + B.logdet[w * 2 + 1] - red[0] - 0.5 * red[1];
    logv[w] = v;
    if (hist != nullptr) {
      const int slot = B.sweep[0] - B.sweep[1];
      if (slot >= 0 && slot < hist_cap) hist[(size_t)slot * B.W + w] = v;
    }
  }
}

__global__ void k_next_sweep(int* sweep) {
  if (threadIdx.x == 0 && blockIdx.x == 0) sweep[0] += 1;
}

}  // namespace qmcx
