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

// LDS of the set-up correction: T and T2 (E x KC each, T) and the MFMA partial tiles
// (16 x 16 floats per job, at least one job per wave)
__host__ __device__ inline int cor_jobs(int K, int E) {
  const int tiles = ((E + 15) / 16) * ((K + 15) / 16);
  return tiles > kMoveThreads / 64 ? tiles : kMoveThreads / 64;
}
template <typename T>
__host__ __device__ inline size_t cor_lds_bytes(int K, int E) {
  const int KC = (K + 15) & ~15;
  return sizeof(T) * 2 * (size_t)E * KC + sizeof(float) * 256 * (size_t)cor_jobs(K, E);
}

struct MoveLds {
  unsigned oR, oRed, oInv, oPsi, oBinv, oP, oPart, oRows, oAl, oJ1n, oFn, oCell, oProp, oJ1o, oJt, oJ2, oAc, oQ, oCor,
      oRall, total;
};
__host__ __device__ inline unsigned lds_take(unsigned& o, size_t bytes) {
  unsigned r = o;
  o = (unsigned)((o + bytes + 15) & ~size_t(15));
  return r;
}
// E = electrons per chunk; the inverse-row cache is used when K > 1 (M, the ion count, is
// not needed: the J1 terms come with the proposals).  The last two regions, for the set-up
// correction of grouped flushes, are sized only with `corr` (the kernel touches them only
// when a chunk has pending updates).
template <typename T>
__host__ __device__ inline MoveLds move_lds(int N, int M, int n, int K, int E, bool corr = true) {
  (void)M;
  MoveLds L;
  unsigned o = 0;
  const bool cache = K > 1;
  L.oR = lds_take(o, sizeof(double) * 3 * N);
  L.oRed = lds_take(o, sizeof(double) * kRed * 16);
  L.oInv = lds_take(o, sizeof(T) * n);
  L.oPsi = lds_take(o, sizeof(T) * 5 * n);
  L.oBinv = lds_take(o, cache ? sizeof(double) * K * (K + 1) : 0);   // row stride K + 1
  L.oP = lds_take(o, sizeof(double) * (K + 4));
  L.oPart = lds_take(o, cache ? 0 : sizeof(double) * n);   // Sherman-Morrison scratch
  L.oRows = lds_take(o, sizeof(int) * K);
  L.oAl = lds_take(o, sizeof(int) * E);                    // accepted moves of the chunk
  L.oJ1n = lds_take(o, sizeof(T) * 5 * E);                 // their new J1 state
  L.oFn = lds_take(o, sizeof(Functor<T>) * 2);             // J2 functors (same / opposite spin)
  L.oCell = lds_take(o, sizeof(Cell<T>));
  L.oProp = lds_take(o, sizeof(double) * kProp * E);        // the chunk's proposal records
  L.oJ1o = lds_take(o, sizeof(T) * 4 * E);                  // J1 value and gradient of its electrons
  L.oJt = lds_take(o, sizeof(T) * 5 * N);                   // J2 pair terms at the current proposal
  L.oJ2 = lds_take(o, sizeof(T) * 5 * N);
  L.oAc = lds_take(o, cache ? sizeof(T) * (size_t)E * n : 0);
  L.oQ = lds_take(o, cache ? sizeof(double) * (size_t)K * E : 0);
  L.oCor = lds_take(o, cache && corr ? cor_lds_bytes<T>(K, E) : 0);   // earlier chunks' updates at set-up
  L.oRall = lds_take(o, cache && corr ? sizeof(int) * (size_t)(K + 1) * ((n + K - 1) / K) : 0);   // row lists, counts
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


#ifdef QMCX_PHASE_TIMING
// wall_clock64 ticks per phase, taken by lane 0 of workgroup 0 into registers and added to
// this table when the kernel ends (a global read-modify-write per tick would time itself)
constexpr int kPhaseSlots = 32;
__device__ unsigned long long qmcx_phase_ticks[kPhaseSlots];
#define QMCX_TICK(slot)                                                  \
  if (w == 0 && tid == 0) {                                              \
    const unsigned long long now_ = wall_clock64();                      \
    tacc_[slot] += now_ - tick_;                                         \
    tick_ = now_;                                                        \
  }
#define QMCX_TICK_FLUSH()                                                \
  if (w == 0 && tid == 0)                                                \
    for (int i_ = 0; i_ < kPhaseSlots; ++i_) qmcx_phase_ticks[i_] += tacc_[i_];
#else
#define QMCX_TICK(slot)
#define QMCX_TICK_FLUSH()
#endif

// J2 pair terms of the items j = tid + q*nt against the electron iel at (x, y, z); items
// past N and j == iel give zeros.  Identical arithmetic to j2_pair for every valid item; the
// functor is evaluated without a branch on the cutoff.
template <typename T, int JPT>
__device__ inline void j2_items(const Cell<T>& C, const Functor<T>* fn, const double* sR, int N, int n, int iel,
                                int tid, int nt, double x, double y, double z, T (&u)[JPT], T (&du)[JPT],
                                T (&d2u)[JPT], T (&dx)[JPT], T (&dy)[JPT], T (&dz)[JPT]) {
  const bool up = iel < n;
#pragma unroll
  for (int q = 0; q < JPT; ++q) {
    u[q] = du[q] = d2u[q] = dx[q] = dy[q] = dz[q] = T(0);
    const int j = tid + q * nt;
    if (j < N) {
      const Functor<T>& f = fn[((j < n) == up) ? 0 : 1];
      const T r = min_image(C, (T)(sR[j] - x), (T)(sR[N + j] - y), (T)(sR[2 * N + j] - z), j < iel ? T(1) : T(-1),
                            dx[q], dy[q], dz[q]);
      const bool in = j != iel && r < f.rcut;
      T a, b, c2;
      functor_vgl(f, in ? r : f.rcut * T(0.5), a, b, c2);   // an interior radius keeps the lookup in range
      u[q] = in ? a : T(0);
      du[q] = in ? b : T(0);
      d2u[q] = in ? c2 : T(0);
    }
  }
}

// The chunk's stale inverse rows (sAc, E x n, in LDS) receive the delayed updates that the
// determinant's earlier chunks j0 .. j0+nj-1 of this sweep left in their state slots, in order:
// updateInvMat (DelayedUpdate.h:176-193) restricted to these rows,
//   sAc <- sAc + T2_j V_j,  T2_j = -(sAc U_j^T) B_j
// (U_j rows: the new Slater rows of chunk j's accepted moves, V_j: their stale inverse rows,
// B_j: chunk j's capacitance inverse; none of chunk j's rows is among these, so the identity
// term of the flush is absent).  sRall holds the delayed-row lists of those slots (K each),
// sNd their counts.  FP32 with n % 16 == 0 runs both products on the matrix cores with their global
// loads batched; otherwise vector code.  Ends with a barrier.
template <typename T>
__device__ inline void correct_rows(const SysP<T>& S, const BatchP<T>& B, int w, int det, int j0, int nj, int E, T* sAc,
                                    const int* sRall, const int* sNd, double* sBj, T* sT, T* sT2, float* sPart,
                                    unsigned long long* crt) {
#ifdef QMCX_PHASE_TIMING
  unsigned long long ct_ = wall_clock64();
#define CR_TICK(i)                                      \
  if (w == 0 && threadIdx.x == 0) {                     \
    const unsigned long long t_ = wall_clock64();       \
    crt[i] += t_ - ct_;                                 \
    ct_ = t_;                                           \
  }
#else
#define CR_TICK(i)
#endif
  const int n = S.n, K = B.K, KC = (K + 15) & ~15;
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  const int li = lane & 15, g = lane >> 4;
  const T* Amat = B.A + ((size_t)w * 2 + det) * n * n;
  const bool mfma = sizeof(T) == 4 && n % 16 == 0;
  for (int j = 0; j < nj; ++j) {
    const size_t sl = state_slot(B, w, det, j0 + j);
    const int cj = sNd[j];
    if (cj == 0) continue;
    const int cpj = (cj + 15) & ~15;
    const T* Vj = B.Vbuf + sl * K * n;
    const double* Bg = B.Binv + sl * K * K;
    const int* rj = sRall + j * K;
    // chunk j's capacitance inverse: loads issued now, stored to LDS after the first product
    constexpr int kBr = 4;
    double br[kBr];
    const int nb = cj * cj;
#pragma unroll
    for (int q = 0; q < kBr; ++q) {
      const int i = min(tid + q * nt, nb - 1);   // unconditional loads
      br[q] = Bg[(i / cj) * K + i % cj];
    }
    __builtin_amdgcn_sched_barrier(0);
    CR_TICK(0)
    // ---- T = sAc U_j^T (E x cj)
    if constexpr (sizeof(T) == 4) {
      if (mfma) {
        // 16x16 tiles of T, the orbital sum in slices over the waves; lane (i, g) supplies
        // A[i][g] = sAc[e][k], B[g][i] = U_j[d][k] with k = 16kt + 4g + r for MFMA r; six
        // orbital tiles' loads in flight at a time
        const int ET = (E + 15) >> 4, DT = cpj >> 4, tiles = ET * DT;
        const int ks = tiles >= nw ? 1 : nw / tiles;
        const int nkt = n >> 4, per = (nkt + ks - 1) / ks;
        for (int job = wid; job < tiles * ks; job += nw) {
          const int tile = job % tiles, sc = job / tiles;
          const int et = tile % ET, dt = tile / ET;
          const int e = et * 16 + li, d = dt * 16 + li;
          const float* ar = (const float*)sAc + (size_t)(e < E ? e : 0) * n + 4 * g;
          const float* ur = (const float*)Amat + (size_t)rj[d < cj ? d : 0] * n + 4 * g;
          f32x4_t acc = f32x4_t{0.f, 0.f, 0.f, 0.f};
          const int k1 = min(nkt, (sc + 1) * per);
          for (int kb = sc * per; kb < k1; kb += 6) {
            f32x4_t av[6], uv[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
              const int kt = min(kb + q, k1 - 1);
              uv[q] = *reinterpret_cast<const f32x4_t*>(ur + kt * 16);
              av[q] = *reinterpret_cast<const f32x4_t*>(ar + kt * 16);
            }
            __builtin_amdgcn_sched_barrier(0);   // all six loads issued before the first MFMA
#pragma unroll
            for (int q = 0; q < 6; ++q) {
              if (e >= E || kb + q >= k1) av[q] = f32x4_t{0.f, 0.f, 0.f, 0.f};
              if (d >= cj) uv[q] = f32x4_t{0.f, 0.f, 0.f, 0.f};
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][0], uv[q][0], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][1], uv[q][1], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][2], uv[q][2], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][3], uv[q][3], acc, 0, 0, 0);
            }
          }
          // register r of lane (i, g): row e = 16et + 4g + r, delayed row d = 16dt + i
#pragma unroll
          for (int r = 0; r < 4; ++r) sPart[job * 256 + (4 * g + r) * 16 + li] = acc[r];
        }
#pragma unroll
        for (int q = 0; q < kBr; ++q)
          if (tid + q * nt < nb) sBj[tid + q * nt] = br[q];
        __syncthreads();
        CR_TICK(1)
        for (int idx = tid; idx < E * cpj; idx += nt) {
          const int e = idx / cpj, d = idx - e * cpj;
          const int tile = (e >> 4) + ET * (d >> 4);
          float v = 0.f;
          for (int sc = 0; sc < ks; ++sc) v += sPart[(sc * tiles + tile) * 256 + (e & 15) * 16 + (d & 15)];
          sT[e * KC + d] = d < cj ? v : 0.f;
        }
      }
    }
    if (!mfma) {
      for (int idx = wid; idx < E * cj; idx += nw) {
        const int e = idx / cj, d = idx - e * cj;
        const T* ur = Amat + (size_t)rj[d] * n;
        double v = 0;
        for (int k = lane; k < n; k += 64) v += (double)sAc[(size_t)e * n + k] * (double)ur[k];
        v = wave_sum(v);
        if (lane == 0) sT[e * KC + d] = (T)v;
      }
#pragma unroll
      for (int q = 0; q < kBr; ++q)
        if (tid + q * nt < nb) sBj[tid + q * nt] = br[q];
    }
    for (int i = tid + kBr * nt; i < nb; i += nt) sBj[i] = Bg[(i / cj) * K + i % cj];
    __syncthreads();
    CR_TICK(2)
    // ---- T2 = -(T B_j), FP64 products
    for (int idx = tid; idx < E * cpj; idx += nt) {
      const int e = idx / cpj, d = idx - e * cpj;
      double v = 0;
      if (d < cj)
        for (int q = 0; q < cj; ++q) v += (double)sT[e * KC + q] * sBj[q * cj + d];
      sT2[e * KC + d] = (T)(-v);
    }
    __syncthreads();
    CR_TICK(3)
    // ---- sAc += T2 V_j
    if constexpr (sizeof(T) == 4) {
      if (mfma) {
        // 16x16 tiles of sAc (16 rows x 16 orbitals); a wave loads the operands of up to six of
        // its tiles (V from HBM, T2 and the tiles from LDS) for 32 delayed rows at a time before
        // their MFMAs; lane (i, g) supplies A[i][g] = T2[e][d0+g], B[g][i] = V_j[d0+g][k];
        // register r of lane (i, g) is row 16et + 4g + r, orbital 16kt + i
        const int ET = (E + 15) >> 4, nkt = n >> 4, ntile = ET * nkt;
        constexpr int kTG = 6;
        for (int t0 = wid; t0 < ntile; t0 += kTG * nw) {
          f32x4_t acc[kTG];
          for (int db = 0; db < cpj; db += 32) {
            const bool full = cpj - db >= 32;   // 8 or 4 steps of 4 delayed rows
            float bv[kTG][8], av[kTG][8];
#pragma unroll
            for (int h = 0; h < kTG; ++h) {
              const int t = min(t0 + h * nw, ntile - 1);
              const int et = t % ET, kt = t / ET;
              const int ea = min(et * 16 + li, E - 1);
#pragma unroll
              for (int s4 = 0; s4 < 8; ++s4) {
                const int dk = db + 4 * s4 + g;
                bv[h][s4] = (float)Vj[(size_t)min(dk, cj - 1) * n + kt * 16 + li];
                av[h][s4] = (float)sT2[ea * KC + min(dk, KC - 1)];
              }
              if (db == 0)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                  const int e = min(et * 16 + 4 * g + r, E - 1);
                  acc[h][r] = ((const float*)sAc)[(size_t)e * n + kt * 16 + li];
                }
            }
            __builtin_amdgcn_sched_barrier(0);   // every operand requested before the first MFMA
#pragma unroll
            for (int h = 0; h < kTG; ++h) {
              const int t = min(t0 + h * nw, ntile - 1);
              const bool arow = (t % ET) * 16 + li < E;
#pragma unroll
              for (int s4 = 0; s4 < 8; ++s4) {
                const int dk = db + 4 * s4 + g;
                const float a2 = arow && dk < cj ? av[h][s4] : 0.f;
                const float b2 = dk < cj ? bv[h][s4] : 0.f;
                if (s4 < 4 || full) acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b2, acc[h], 0, 0, 0);
              }
            }
            if (db + 32 >= cpj) {
#pragma unroll
              for (int h = 0; h < kTG; ++h) {
                const int t = t0 + h * nw;
                if (t < ntile) {
                  const int et = t % ET, kt = t / ET;
#pragma unroll
                  for (int r = 0; r < 4; ++r) {
                    const int e = et * 16 + 4 * g + r;
                    if (e < E) ((float*)sAc)[(size_t)e * n + kt * 16 + li] = acc[h][r];
                  }
                }
              }
            }
          }
        }
      }
    }
    if (!mfma) {
      for (int i = tid; i < E * n; i += nt) {
        const int e = i / n, k = i - e * n;
        double v = (double)sAc[i];
        for (int d = 0; d < cj; ++d) v += (double)sT2[e * KC + d] * (double)Vj[(size_t)d * n + k];
        sAc[i] = (T)v;
      }
    }
    __syncthreads();
    CR_TICK(4)
  }
#undef CR_TICK
}

// KD: delayed updates (K > 1: inverse-row cache, Woodbury bordering) or Sherman-Morrison (K = 1)
template <typename T, int JPT, bool KD>
__global__ __launch_bounds__(kMoveThreads) void k_moves(SysP<T> S, BatchP<T> B, MoveArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  const int N = S.N, n = S.n, K = KD ? B.K : 1;
  const int KP = K + 1;   // LDS row stride of Binv: a column walk down its rows is free of bank conflicts
  const int E = a.e_end - a.e_begin;
  const int r0 = a.e_begin - a.det * n;
  const MoveLds L = move_lds<T>(N, S.M, n, K, E);
  double* sR = (double*)(smem + L.oR);
  double* sRed = (double*)(smem + L.oRed);
  T* sInv = (T*)(smem + L.oInv);
  T* sPsi = (T*)(smem + L.oPsi);
  double* sBinv = (double*)(smem + L.oBinv);
  double* sP = (double*)(smem + L.oP);
  double* sTm = (double*)(smem + L.oPart);
  int* sRows = (int*)(smem + L.oRows);
  int* sAl = (int*)(smem + L.oAl);
  T* sJ1n = (T*)(smem + L.oJ1n);
  Functor<T>* sFn = (Functor<T>*)(smem + L.oFn);
  Cell<T>* sCell = (Cell<T>*)(smem + L.oCell);
  double* sProp = (double*)(smem + L.oProp);
  T* sJ1o = (T*)(smem + L.oJ1o);
  T* sJt = (T*)(smem + L.oJt);
  T* sJ2 = (T*)(smem + L.oJ2);
  T* sAc = (T*)(smem + L.oAc);
  double* sQ = (double*)(smem + L.oQ);
#ifdef QMCX_PHASE_TIMING
  unsigned long long tick_ = wall_clock64();
  unsigned long long tacc_[kPhaseSlots] = {};
#endif
  const size_t moff = ((size_t)w * 2 + a.det) * n * n;
  T* Ainv = B.Ainv + moff;
  T* Amat = B.A + moff;
  T* dpsiM = B.dpsiM + ((size_t)w * 2 + a.det) * (size_t)n * 3 * n;
  T* d2psiM = B.d2psiM + moff;
  T* J2 = B.J2 + (size_t)w * 5 * N;
  T* J1 = B.J1 + (size_t)w * 5 * N;
  // this chunk's delayed-update state slot
  const size_t slot = state_slot(B, w, a.det, KD ? a.chunk : 0);
  T* Vb = B.Vbuf + slot * K * n;
  // this chunk's proposals: SPO rows [e][5][npad] and records [e][kProp]
  const T* spo = (const T*)a.spo + (((size_t)w * 2 + a.det) * n + r0) * 5 * S.npad;
  const double* prop = a.prop + (((size_t)w * 2 + a.det) * n + r0) * kProp;

  // ---- chunk set-up: walker state, the chunk's proposal records and J1 state, J2 functors and
  // the cell into LDS (one round of loads)
  {
    const unsigned je = (unsigned)(sizeof(T) * E);
    const int nprev = KD ? a.chunk - a.chunk0 : 0;
    const LdsCopy cp[11] = {{sR, B.R + (size_t)w * 3 * N, (unsigned)(sizeof(double) * 3 * N)},
                            {sJ2, J2, (unsigned)(sizeof(T) * 5 * N)},
                            {sAc, Ainv + (size_t)r0 * n, KD ? (unsigned)(sizeof(T) * E * n) : 0u},
                            {sProp, prop, (unsigned)(sizeof(double) * kProp * E)},
                            {sJ1o, J1 + a.e_begin, je},
                            {sJ1o + E, J1 + N + a.e_begin, je},
                            {sJ1o + 2 * E, J1 + 2 * N + a.e_begin, je},
                            {sJ1o + 3 * E, J1 + 3 * N + a.e_begin, je},
                            {sFn, S.fn + kMaxSpecies, (unsigned)(sizeof(Functor<T>) * 2)},
                            {sCell, S.cell, (unsigned)sizeof(Cell<T>)},
                            {smem + L.oRall, nprev > 0 ? B.rows + state_slot(B, w, a.det, a.chunk0) * K : B.rows,
                             (unsigned)(sizeof(int) * nprev * K)}};
    stage_lds(cp);
  }
  QMCX_TICK(16)
  if (KD && a.chunk > a.chunk0) {
    const int* sRall = (const int*)(smem + L.oRall);
    int* sNd = (int*)(smem + L.oRall) + (size_t)K * ((n + K - 1) / K);
    for (int j = tid; j < a.chunk - a.chunk0; j += nt) sNd[j] = B.ndelay[state_slot(B, w, a.det, a.chunk0 + j)];
    __syncthreads();
    const int KC = (K + 15) & ~15;
    T* sCorT = (T*)(smem + L.oCor);
    unsigned long long* crt = nullptr;
#ifdef QMCX_PHASE_TIMING
    __shared__ unsigned long long crt_s[5];
    if (tid < 5) crt_s[tid] = 0;
    __syncthreads();
    crt = crt_s;
#endif
    correct_rows<T>(S, B, w, a.det, a.chunk0, a.chunk - a.chunk0, E, sAc, sRall, sNd, sBinv, sCorT,
                    sCorT + (size_t)E * KC,
                    (float*)(sCorT + (size_t)2 * E * KC), crt);
#ifdef QMCX_PHASE_TIMING
    if (w == 0 && tid == 0)
      for (int i = 0; i < 5; ++i) tacc_[18 + i] += crt[i];
#endif
    (void)crt;
  }
  QMCX_TICK(17)
  // a caller's accept decision (the flex stage calls) replaces the uniform: -inf accepts, +inf rejects
  const double flex_u = a.decide == nullptr ? 0.0 : (a.decide[w] != 0 ? -INFINITY : INFINITY);
  int c = 0, ca = 0;   // delayed rows; accepted moves of the chunk
  double logd = 0, pdet = 1, ph = 0, jl1 = 0, jl2 = 0;
  unsigned long long nacc = 0;
  __syncthreads();
  const Cell<T>& C = *sCell;   // read from LDS where used (a register copy costs spills)
  // Rows a move reads from HBM that no earlier move of the chunk changes: the moving electron's
  // dpsiM row and the SPO row of its proposal.  Move e+1's are loaded at the start of move e
  // (software pipelining; index clamped, so the loads are unconditional).
  const int jr = tid < n ? tid : n - 1;
  T ndp[3], np5[5];
  auto prefetch = [&](int ee) {
    const T* dq = dpsiM + (size_t)(r0 + ee) * 3 * n;
    ndp[0] = dq[jr];
    ndp[1] = dq[n + jr];
    ndp[2] = dq[2 * n + jr];
    const T* pq = spo + (size_t)ee * 5 * S.npad;
#pragma unroll
    for (int o = 0; o < 5; ++o) np5[o] = pq[(size_t)o * S.npad + jr];
  };
  prefetch(0);

  for (int e = 0; e < E; ++e) {
    const int iel = a.e_begin + e, row = r0 + e;
    QMCX_TICK(0)
    // ---- phase 1: issue the next move's global loads, then getInvRow's small algebra
    // stored state of the moving electron (J1, J2), its dpsiM row (loaded one move ahead) and
    // its proposal record: J1 at the proposal (sum u, sum (d2u + 2 du/r), gradient),
    // displacement, uniform
    const T Vat_old = sJ1o[e], g1x = sJ1o[E + e], g1y = sJ1o[2 * E + e], g1z = sJ1o[3 * E + e];
    const T Uat_old = sJ2[iel], g2x = sJ2[N + iel], g2y = sJ2[2 * N + iel], g2z = sJ2[3 * N + iel];
    const T dp0 = ndp[0], dp1 = ndp[1], dp2 = ndp[2];
    T p5[5];
#pragma unroll
    for (int o = 0; o < 5; ++o) p5[o] = np5[o];
    const double* pr = sProp + (size_t)e * kProp;
    prefetch(e + 1 < E ? e + 1 : e);
    const double x0 = sR[iel], y0 = sR[N + iel], z0 = sR[2 * N + iel];
    const double nx = x0 + pr[kPropDispl + 0];
    const double ny = y0 + pr[kPropDispl + 1];
    const double nz = z0 + pr[kPropDispl + 2];
    T inv_sm = 0;
    if constexpr (!KD) inv_sm = Ainv[(size_t)row * n + jr];   // changed by the previous accept
    const T* prow = spo + (size_t)e * 5 * S.npad;
    if (KD && c > 0) {
      // y = Binv^T p with p_d = U_d . Ainv[row] formed when row d was accepted; parked in row c of Binv
      for (int r = tid; r < c; r += nt) {
        double s = 0;
        int q = 0;
        for (; q + 4 <= c; q += 4) {
          const double b0 = sBinv[q * KP + r], b1 = sBinv[(q + 1) * KP + r], b2 = sBinv[(q + 2) * KP + r],
                       b3 = sBinv[(q + 3) * KP + r];
          const double p0 = sQ[q * E + e], p1 = sQ[(q + 1) * E + e], p2 = sQ[(q + 2) * E + e], p3 = sQ[(q + 3) * E + e];
          s += b0 * p0;
          s += b1 * p1;
          s += b2 * p2;
          s += b3 * p3;
        }
        for (; q < c; ++q) s += sBinv[q * KP + r] * sQ[q * E + e];
        sBinv[c * KP + r] = s;
      }
    }
    lds_sync();
    QMCX_TICK(1)

    // ---- phase 2: invRow, evalGrad of the determinant and ratioGrad_compute against the
    // proposal's SPO row (einspline_spo.hpp:204-223, DiracDeterminant.cpp:103-124), J2 at
    // the proposal; all partial sums reduced in one round
    double gd0 = 0, gd1 = 0, gd2 = 0, d0 = 0, d1 = 0, d2 = 0, d3 = 0;
    T vinv = 0;
    if (tid < n) {
      double v;
      if (KD) {
        v = (double)sAc[(size_t)e * n + tid];
        const double* y = sBinv + c * KP;
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
        for (int d = 0; d < c; ++d) v -= (double)sAc[(size_t)(sRows[d] - r0) * n + j] * sBinv[c * KP + d];
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
      // pair terms at the proposal: u, du dx, du dy, du dz, d2u + 2 du, parked in LDS for the
      // J2 update if the move is accepted
      T cu[JPT], cdu[JPT], cd2u[JPT], cdx[JPT], cdy[JPT], cdz[JPT];
      j2_items<T, JPT>(C, sFn, sR, N, n, iel, tid, nt, nx, ny, nz, cu, cdu, cd2u, cdx, cdy, cdz);
      T r7 = 0, r8 = 0, r9 = 0, r10 = 0, r11 = 0;
#pragma unroll
      for (int q = 0; q < JPT; ++q) {
        const T t1 = cdu[q] * cdx[q], t2 = cdu[q] * cdy[q], t3 = cdu[q] * cdz[q], t4 = cd2u[q] + T(2) * cdu[q];
        r7 += cu[q];
        r8 += t1;
        r9 += t2;
        r10 += t3;
        r11 += t4;
        const int j = tid + q * nt;
        if (j < N) {
          sJt[j] = cu[q];
          sJt[N + j] = t1;
          sJt[2 * N + j] = t2;
          sJt[3 * N + j] = t3;
          sJt[4 * N + j] = t4;
        }
      }
      QMCX_TICK(6)
      // slots: gnow 0-2, det ratio 3, its gradient 4-6, J2 7-11; after the row-wise wave sums
      // v[i] holds slot i + 3r in the lanes of row r, which lane 16r + i stores
      double v[kRed] = {gd0, gd1, gd2, d0, d1, d2, d3, r7, r8, r9, r10, r11};
      wave_sums_rows<kRed>(v);
      const int k = lane & 15;
      if (k < kRed / 4) {
        double x = v[0];
#pragma unroll
        for (int i = 1; i < kRed / 4; ++i) x = k == i ? v[i] : x;
        sRed[(k + (lane >> 4) * (kRed / 4)) * 16 + wid] = x;
      }
      QMCX_TICK(7)
    }
    lds_sync();
    QMCX_TICK(2)

    double red[kRed];
    combine_partials(sRed, nw, red);
    QMCX_TICK(3)

    // ---- phase 4: acceptance and updates
    const double rdet = red[3];
    const double ratio = rdet * exp((double)(Vat_old - (T)pr[kPropJ1]) + (double)(Uat_old - (T)red[7]));
    const double uacc = a.decide == nullptr ? pr[kPropU] : flex_u;
    const bool accept = a.mode == 0 ? (uacc < 0.5) : (uacc < ratio * ratio);
    if (B.trace != nullptr && tid == 0) {
      double* t = B.trace + ((size_t)w * N + iel) * 9;
      t[0] = ratio;
      t[1] = rdet;
      t[2] = accept ? 1.0 : 0.0;
      t[3] = red[0] + (double)g1x + (double)g2x;
      t[4] = red[1] + (double)g1y + (double)g2y;
      t[5] = red[2] + (double)g1z + (double)g2z;
      t[6] = red[4] / rdet + pr[kPropJ1 + 2] + red[8];
      t[7] = red[5] / rdet + pr[kPropJ1 + 3] + red[9];
      t[8] = red[6] / rdet + pr[kPropJ1 + 4] + red[10];
    }
    if (accept) {
      if (tid == 0) {
        // log|det| of the accepted ratios as a running product, folded into the log before it
        // can leave the double range
        pdet *= fabs(rdet);
        if (!(pdet > 0x1p-500 && pdet < 0x1p500)) {
          logd += log(pdet);
          pdet = 1.0;
        }
        ph += rdet > 0.0 ? 0.0 : M_PI;
        jl1 += (double)(Vat_old - (T)pr[kPropJ1]);
        jl2 += (double)(Uat_old - (T)red[7]);
        ++nacc;
        // the new J1 state of the electron (OneBodyJastrow.h:271-283), written at the chunk's end
        sAl[ca] = e;
        sJ1n[ca * 5 + 0] = (T)pr[kPropJ1 + 0];
        sJ1n[ca * 5 + 1] = (T)pr[kPropJ1 + 2];
        sJ1n[ca * 5 + 2] = (T)pr[kPropJ1 + 3];
        sJ1n[ca * 5 + 3] = (T)pr[kPropJ1 + 4];
        sJ1n[ca * 5 + 4] = (T)pr[kPropJ1 + 1];
      }
      ++ca;
      QMCX_TICK(12)
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
          wave_sums_rows<4>(sm);
          // row r of the wave holds the sum of row m0 + r
          if ((lane & 15) == 0 && (lane >> 4) < nr) {
            const int m = m0 + (lane >> 4);
            tm[m] = (sm[0] - (m == row ? 1.0 : 0.0)) / rdet;
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
        // chunk's later rows r', U_c . Ainv[r'] used by their getInvRow; four dot products per
        // wave share the psiV loads.  Vector t2: delayed row t2 (t2 < c), this row (t2 == c),
        // the chunk's row e + t2 - c (t2 > c).
        const int nd = c + 1 + (E - 1 - e);
        auto dvec = [&](int t2) -> const T* {
          const int ee = t2 < c ? sRows[t2] - r0 : e + (t2 - c);
          return sAc + (size_t)ee * n;
        };
        for (int t0 = wid * 4; t0 < nd; t0 += nw * 4) {
          const T* v0 = dvec(t0);
          const T* v1 = dvec(min(t0 + 1, nd - 1));
          const T* v2 = dvec(min(t0 + 2, nd - 1));
          const T* v3 = dvec(min(t0 + 3, nd - 1));
          double sv[4] = {0, 0, 0, 0};
          for (int j = lane; j < n; j += 64) {
            const double pv = (double)sPsi[j];
            sv[0] += (double)v0[j] * pv;
            sv[1] += (double)v1[j] * pv;
            sv[2] += (double)v2[j] * pv;
            sv[3] += (double)v3[j] * pv;
          }
          wave_sums_rows<4>(sv);
          // row r of the wave holds dot t0 + r
          const int t2 = t0 + (lane >> 4);
          if ((lane & 15) == 0 && t2 < nd) {
            if (t2 <= c) sP[t2] = -sv[0];
            else sQ[c * E + e + (t2 - c)] = sv[0];
          }
        }
        lds_sync();
        QMCX_TICK(13)
        if (wid == nw - 1) {
          // the capacitance inverse grows by one row and column (FP64, the last wave, which
          // has the least J2 work): z = row c (parked by phase 1), y = z.p - p_c,
          // Binv[s][c] = (Binv[s].p) / y, Binv[s][r] -= z_r Binv[s][c], row c = -z / y
          double* z = sBinv + c * KP;
          double t = 0;
          for (int i = lane; i < c; i += 64) t += z[i] * sP[i];
          const double yi = 1.0 / (wave_sum(t) - sP[c]);
          const int h = lane >> 5;   // half-waves take alternate columns r
          for (int s0 = 0; s0 < c; s0 += 32) {
            const int s2 = s0 + (lane & 31);
            double* bs = sBinv + (s2 < c ? s2 : 0) * KP;
            double acc = 0;
            if (s2 < c)
              for (int r = h; r < c; r += 2) acc += bs[r] * sP[r];
            acc += __shfl_xor(acc, 32);
            const double ys = yi * acc;
            if (s2 < c) {
              for (int r = h; r < c; r += 2) bs[r] -= z[r] * ys;
              if (h == 0) bs[c] = ys;
            }
          }
          wave_lds_fence();
          for (int i = lane; i < c; i += 64) z[i] *= -yi;
          if (lane == 0) {
            z[c] = yi;
            sRows[c] = row;
          }
        }
        ++c;
        QMCX_TICK(14)
      }
      // J2 (TwoBodyJastrow.h:307-353): the pair terms at the old position are recomputed here,
      // those at the accepted one are phase 2's
      T ou[JPT], odu[JPT], od2u[JPT], odx[JPT], ody[JPT], odz[JPT];
      j2_items<T, JPT>(C, sFn, sR, N, n, iel, tid, nt, x0, y0, z0, ou, odu, od2u, odx, ody, odz);
#pragma unroll
      for (int q = 0; q < JPT; ++q) {
        const int j = tid + q * nt;
        if (j < N && j != iel) {
          sJ2[j] += sJt[j] - ou[q];
          sJ2[4 * N + j] += od2u[q] + T(2) * odu[q] - sJt[4 * N + j];
          sJ2[N + j] -= sJt[N + j] - odu[q] * odx[q];
          sJ2[2 * N + j] -= sJt[2 * N + j] - odu[q] * ody[q];
          sJ2[3 * N + j] -= sJt[3 * N + j] - odu[q] * odz[q];
        } else if (j == iel) {
          sJ2[iel] = (T)red[7];
          sJ2[N + iel] = (T)red[8];
          sJ2[2 * N + iel] = (T)red[9];
          sJ2[3 * N + iel] = (T)red[10];
          sJ2[4 * N + iel] = (T)(-red[11]);
          sR[iel] = nx;
          sR[N + iel] = ny;
          sR[2 * N + iel] = nz;
        }
      }
      QMCX_TICK(15)
    }
    // with K = 1 the next move reads rows of Ainv that other waves just wrote
    if (!KD) __syncthreads();
    else lds_sync();
    QMCX_TICK(4)
  }

  // ---- chunk end: walker state back to HBM; the accepted rows' Slater matrix, gradient and
  // Laplacian rows (DiracDeterminant.cpp:135-141) copied from the proposals' SPO rows, the
  // stale inverse rows for the flush, the new J1 state
  for (int i = tid; i < 3 * N; i += nt) B.R[(size_t)w * 3 * N + i] = sR[i];
  for (int i = tid; i < 5 * N; i += nt) J2[i] = sJ2[i];
  for (int d = tid; d < ca; d += nt) {
    const int ie = a.e_begin + sAl[d];
    J1[ie] = sJ1n[d * 5 + 0];
    J1[N + ie] = sJ1n[d * 5 + 1];
    J1[2 * N + ie] = sJ1n[d * 5 + 2];
    J1[3 * N + ie] = sJ1n[d * 5 + 3];
    J1[4 * N + ie] = sJ1n[d * 5 + 4];
  }
  if (KD)
    for (int i = tid; i < ca * n; i += nt) {
      const int d = i / n, j = i - d * n;
      Vb[(size_t)d * n + j] = sAc[(size_t)sAl[d] * n + j];
    }
  {
    using VT = typename V16<T>::type;
    constexpr int VW = V16<T>::n;
    const bool vec = n % VW == 0 && S.npad % VW == 0;
    const int nv = vec ? n / VW : n;   // vectors (or scalars) per row
    const int tot = ca * 5 * nv;
    for (int b0 = 0; b0 < tot; b0 += 8 * nt) {
      VT v[8];
      T sv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int i = min(b0 + k * nt + tid, tot - 1);
        const int d = i / (5 * nv), rem = i - d * 5 * nv, o = rem / nv, j = rem - o * nv;
        const T* src = spo + (size_t)sAl[d] * 5 * S.npad + (size_t)o * S.npad;
        if (vec) v[k] = ((const VT*)src)[j];
        else sv[k] = src[j];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int i = b0 + k * nt + tid;
        if (i < tot) {
          const int d = i / (5 * nv), rem = i - d * 5 * nv, o = rem / nv, j = rem - o * nv;
          const int rw = r0 + sAl[d];
          T* dst = o == 0 ? Amat + (size_t)rw * n
                          : (o < 4 ? dpsiM + (size_t)rw * 3 * n + (size_t)(o - 1) * n : d2psiM + (size_t)rw * n);
          if (vec) ((VT*)dst)[j] = v[k];
          else dst[j] = sv[k];
        }
      }
    }
  }
  if (tid == 0) {
    B.logdet[w * 2 + a.det] += logd + log(pdet);
    B.phase[w * 2 + a.det] += ph;
    B.jlog[w * 2 + 0] += jl1;
    B.jlog[w * 2 + 1] += jl2;
    B.nacc[w] += nacc;
    B.ndelay[slot] = KD ? c : 0;
  }
  if (KD) {
    for (int i = tid; i < c * K; i += nt) B.Binv[slot * K * K + i] = sBinv[(i / K) * KP + i % K];
    for (int i = tid; i < c; i += nt) B.rows[slot * K + i] = sRows[i];
  }
  QMCX_TICK_FLUSH()
}

}  // namespace qmcx
