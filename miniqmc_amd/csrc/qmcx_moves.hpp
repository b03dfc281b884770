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
constexpr size_t kMoveLdsMax = 160 * 1024;   // LDS of a workgroup
// LDS row stride (elements) of the chunk's stale inverse rows: 64 bytes of padding, so four rows
// read at one column offset by the four 16-lane groups of a wave fall in different banks
template <typename T>
__host__ __device__ inline int ac_stride(int n) { return n + 64 / (int)sizeof(T); }

// LDS of the in-kernel flush between chunks (FP32): the new Slater rows U of up to kFlushOwnU
// accepted moves (row stride n + 4) and per wave a 16 x KC tile of T (in place T2)
constexpr int kFlushOwnU = 16;
__host__ __device__ inline size_t flush_own_lds(int n, int K) {
  return sizeof(float) * ((size_t)kFlushOwnU * (n + 4) + (size_t)(kMoveThreads / 64) * 16 * ((K + 15) & ~15));
}

struct MoveLds {
  unsigned oR, oRed, oInv, oPsi, oBinv, oP, oPart, oRows, oAl, oJ1n, oFn, oCell, oProp, oJ1o, oJt, oJ2, oAc, oQ, oAcc,
      oArb, oJ3, oJ3m, oJ3np, oJ3ne, oJ3k, oJ3v, oJ3r, oSt, total;
  bool binv_lds;   // streamed form: the capacitance inverse is in LDS (else in its HBM state slot)
};
// per-wave pair slots of one J3 move phase: codes [waves][kJ3Slots] and k-side terms
// [waves][5][kJ3Slots]
constexpr int kJ3Slots = 32;
template <typename T>
__host__ __device__ inline size_t j3_slots_lds(int waves) {
  return (size_t)waves * kJ3Slots * (sizeof(int) + 5 * sizeof(T));
}
// LDS of the FP64 re-evaluation of a near-tie decision (FP32 form, K > 1): three n-vectors, 2K + 32
// reduction slots and kArbGroups n-vectors of partial sums
constexpr int kArbGroups = 5;
__host__ __device__ inline size_t arb_lds(int n, int K) {
  return sizeof(double) * ((3 + kArbGroups) * (size_t)n + 2 * (size_t)K + 32);
}
__host__ __device__ inline unsigned lds_take(unsigned& o, size_t bytes) {
  unsigned r = o;
  o = (unsigned)((o + bytes + 15) & ~size_t(15));
  return r;
}
// E = electrons per chunk; the inverse-row cache is used when K > 1 (M, the ion count, is
// not needed: the J1 terms come with the proposals).  mw > 0: a three-body Jastrow with ion
// masks of mw words (its sums, masks and the chunk's near-ion records stay for the launch; its
// pair slots share the region past the J2 pair terms).  st: the streamed form of the delayed
// updates (no inverse-row cache, no products of accepted rows with later rows; the capacitance
// inverse in LDS with row stride K when binv_lds, else in HBM; p, y and reduction partials in oSt).
// arb3: the FP64 re-evaluation with a three-body Jastrow, whose pair slots stay live across the
// re-evaluation: its region then follows theirs instead of sharing it.
template <typename T>
__host__ __device__ inline MoveLds move_lds_at(int N, int n, int K, int E, int mw, bool st, bool binv_lds,
                                               bool arb3 = false) {
  MoveLds L;
  unsigned o = 0;
  const bool cache = K > 1 && !st;
  L.binv_lds = st && binv_lds;
  L.oR = lds_take(o, sizeof(double) * 3 * N);
  L.oRed = lds_take(o, sizeof(double) * kRed * 16);
  L.oInv = lds_take(o, sizeof(T) * n);
  L.oPsi = lds_take(o, sizeof(T) * 5 * n);
  L.oBinv = lds_take(o, cache ? sizeof(double) * K * (K + 1) : (L.binv_lds ? sizeof(double) * K * K : 0));
  L.oSt = lds_take(o, st ? sizeof(double) * (2 * (size_t)K + kMoveThreads) : 0);
  L.oP = lds_take(o, sizeof(double) * (K + 4));
  L.oPart = lds_take(o, K > 1 ? 0 : sizeof(double) * n);   // Sherman-Morrison scratch
  L.oRows = lds_take(o, sizeof(int) * K);
  L.oAl = lds_take(o, sizeof(int) * E);                    // accepted moves of the chunk
  L.oJ1n = lds_take(o, sizeof(T) * 5 * E);                 // their new J1 state
  L.oFn = lds_take(o, sizeof(Functor<T>) * 2);             // J2 functors (same / opposite spin)
  L.oCell = lds_take(o, sizeof(Cell<T>));
  L.oAcc = lds_take(o, sizeof(double) * 8);                 // the walker's sums over the launch
  L.oProp = lds_take(o, sizeof(double) * kProp * E);        // the chunk's proposal records
  L.oJ1o = lds_take(o, sizeof(T) * 4 * E);                  // J1 value and gradient of its electrons
  L.oJ2 = lds_take(o, sizeof(T) * 5 * N);
  L.oJ3 = lds_take(o, mw > 0 ? sizeof(T) * 5 * N : 0);
  L.oJ3m = lds_take(o, sizeof(unsigned long long) * (size_t)N * mw);
  L.oJ3np = lds_take(o, mw > 0 ? sizeof(T) * kJ3Rec * E : 0);
  L.oJ3ne = lds_take(o, mw > 0 ? sizeof(T) * kJ3Rec * E : 0);
  L.oAc = lds_take(o, cache ? sizeof(T) * (size_t)E * ac_stride<T>(n) : 0);
  L.oQ = lds_take(o, cache ? sizeof(double) * (size_t)K * E : 0);
  // J2 pair terms at the current proposal; with the region before it, the scratch of the
  // in-kernel flush between chunks
  L.oJt = lds_take(o, sizeof(T) * 5 * N);
  // the in-kernel flush (FP32) reuses the regions from oQ on, grown to its size when the
  // workgroup's LDS allows
  L.oArb = o;   // during the moves the region past oJt belongs to the FP64 re-evaluation or the J3 pairs
  {
    unsigned p = o;
    const int waves = kMoveThreads / 64;
    L.oJ3k = lds_take(p, mw > 0 ? 2 * j3_slots_lds<T>(waves) : 0);   // new and old phase slots
    L.oJ3v = L.oJ3k + (unsigned)(2 * sizeof(int) * waves * kJ3Slots);
    L.oJ3r = lds_take(p, mw > 0 ? sizeof(double) * 16 * 5 : 0);    // j-side wave partials
    if (mw > 0) o = p;
    if (mw > 0 && arb3) L.oArb = o;
  }
  if (cache && sizeof(T) == 4) {
    const size_t need = L.oQ + flush_own_lds(n, K);
    if (need > o && need <= kMoveLdsMax) o = (unsigned)need;
  }
  if (K > 1 && sizeof(T) == 4) {
    const size_t arb = L.oArb + arb_lds(n, K);
    if (arb > o) o = (unsigned)arb;
  }
  L.total = o;
  return L;
}
template <typename T>
__host__ __device__ inline MoveLds move_lds(int N, int M, int n, int K, int E, int mw = 0, bool st = false,
                                           bool arb3 = false) {
  (void)M;
  const MoveLds L = move_lds_at<T>(N, n, K, E, mw, st, st, arb3);
  return st && L.total > kMoveLdsMax ? move_lds_at<T>(N, n, K, E, mw, st, false, arb3) : L;
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
      const T ddx = (T)(sR[j] - x), ddy = (T)(sR[N + j] - y), ddz = (T)(sR[2 * N + j] - z);
      const T flip = j < iel ? T(1) : T(-1);
      const T r = C.ortho != T(0) ? min_image_ortho(C, ddx, ddy, ddz, flip, dx[q], dy[q], dz[q])
                                  : min_image(C, ddx, ddy, ddz, flip, dx[q], dy[q], dz[q]);
      const bool in = j != iel && r < f.rcut;
      T a, b, c2;
      functor_vgl(f, in ? r : f.rcut * T(0.5), a, b, c2);   // an interior radius keeps the lookup in range
      u[q] = in ? a : T(0);
      du[q] = in ? b : T(0);
      d2u[q] = in ? c2 : T(0);
    }
  }
}

// Whether the move kernel can apply each chunk's delayed updates to its walker's whole
// inverse itself (FP32, n a multiple of 16 up to 16 kFlushOwnTiles, K <= 32, the flush's LDS
// within the workgroup's).
constexpr int kFlushOwnTiles = 24;
template <typename T>
__host__ __device__ inline bool flush_own_fits(int N, int n, int K, int E) {
  if (!(sizeof(T) == 4 && K > 1 && K <= 32 && n % 16 == 0 && n <= 16 * kFlushOwnTiles)) return false;
  const MoveLds L = move_lds<T>(N, 0, n, K, E);
  return L.total - L.oQ >= flush_own_lds(n, K);
}

// T2 = -(T Binv) in FP64 for one wave's 16-row tile: T (16 x cp, row stride KC) read from sT,
// T2 written over it once every lane has read its T row
__device__ inline void flush_own_t2(int c, int cp, int KC, int KP, const double* sBinv, float* sT) {
  const int lane = threadIdx.x & 63;
  float t2v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int idx = lane + 64 * q;
    double sum = 0;
    if (idx < 16 * cp) {
      const int i = idx / cp, dd = idx - i * cp;
      if (dd < c)
        for (int e = 0; e < c; ++e) sum += (double)sT[i * KC + e] * sBinv[e * KP + dd];
    }
    t2v[q] = (float)(-sum);
  }
  wave_lds_fence();
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int idx = lane + 64 * q;
    if (idx < 16 * cp) {
      const int i = idx / cp, dd = idx - i * cp;
      sT[i * KC + dd] = t2v[q];
    }
  }
  wave_lds_fence();
}

// The same for at most 32 delayed rows on the FP64 matrix cores: T2 (16 x cp) = -T Binv in
// v_mfma_f64_16x16x4f64 steps, one 16-column block per 16 delayed rows.  Lane (j, g) = lane
// j + 16 g supplies A = T[j][4s + g] and B = Binv[4s + g][16b + j]; its result register r is
// T2[g + 4r][16b + j] (the FP64 16x16x4 layout).  Entries past c are zero (T's are stored as
// zeros, Binv's are masked here).
__device__ inline void flush_own_t2_mfma(int c, int cp, int KC, int KP, const double* sBinv, float* sT) {
  const int lane = threadIdx.x & 63, lj = lane & 15, lg = lane >> 4;
  f64x4 acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  for (int s4 = 0; s4 < cp; s4 += 4) {
    const int e = s4 + lg;
    const double a = (double)sT[lj * KC + e];
#pragma unroll
    for (int b = 0; b < 2; ++b)
      if (16 * b < cp) {
        const int dd = 16 * b + lj;
        const double bv = (e < c && dd < c) ? sBinv[e * KP + dd] : 0.0;
        acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[b], 0, 0, 0);
      }
  }
  wave_lds_fence();   // every lane has read its T entries
#pragma unroll
  for (int b = 0; b < 2; ++b)
    if (16 * b < cp)
#pragma unroll
      for (int r = 0; r < 4; ++r) sT[(lg + 4 * r) * KC + 16 * b + lj] = (float)(-acc[b][r]);
  wave_lds_fence();
}

// updateInvMat (DelayedUpdate.h:176-193) of this walker's whole inverse by its own workgroup,
// between chunks, when the new Slater rows U of the chunk's accepted moves are all in LDS:
// Ainv += T2 V with T = Ainv U^T - E, T2 = -T Binv, in 16-row tiles over the waves.  A wave loads
// its tile (all orbitals, lane (i, g): row i, orbitals 16q + 4g .. +3, the layout of
// k_flush_mfma) into registers once: it is the B operand of the first product (one pass per 16
// delayed rows) and the accumulator of the second, then stored.  U rows 0-15 come from sU (row
// stride n + 4), rows 16 .. c-1 from the slots of the chunk's rejected moves in sAc (sRej[d -
// 16], row stride sa), both staged at the chunk's end; V rows are the chunk's stale inverse rows
// (sAc rows sAl[d]), Binv in LDS (row stride KP).  c <= 16, or c <= 32 with K >= 32.
__device__ __attribute__((noinline)) void flush_own_tile(int n, int K, int KP, int c, const int* sRows, const int* sAl,
                                      const double* sBinv, const float* sAc, int sa, const float* sU,
                                      const int* sRej, float* Ainv, float* scratch) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
  const int li = lane & 15, g = lane >> 4;
  const int KC = (K + 15) & ~15, c4 = (c + 3) & ~3, cp = (c + 15) & ~15, nct = n >> 4, us = n + 4;
  float* sT = scratch + (size_t)wid * 16 * KC;
  for (int tile = wid; tile < nct; tile += nw) {
    const int r0 = tile * 16;
    float* ar = Ainv + (size_t)(r0 + li) * n + 4 * g;
    f32x4_t a[kFlushOwnTiles];
#pragma unroll
    for (int q = 0; q < kFlushOwnTiles; ++q)
      a[q] = *reinterpret_cast<const f32x4_t*>(ar + min(q, nct - 1) * 16);
    // ---- GEMM 1: T^T = U rows^T over all orbitals, 16 delayed rows per pass, two accumulators
    for (int dt = 0; dt < cp; dt += 16) {
      const int d = dt + li;
      const bool din = d < c;   // U row d of the A operand is a delayed row
      const float* ul = (dt == 0 ? sU + (size_t)(din ? li : 0) * us : sAc + (size_t)(din ? sRej[li] : 0) * sa) + 4 * g;
      f32x4_t acc0 = f32x4_t{0.f, 0.f, 0.f, 0.f}, acc1 = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q = 0; q < kFlushOwnTiles; ++q) {
        if (q < nct) {
          f32x4_t uv = *reinterpret_cast<const f32x4_t*>(ul + q * 16);
          uv = din ? uv : f32x4_t{0.f, 0.f, 0.f, 0.f};
          f32x4_t& acc = (q & 1) ? acc1 : acc0;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(uv[0], a[q][0], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(uv[1], a[q][1], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(uv[2], a[q][2], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(uv[3], a[q][3], acc, 0, 0, 0);
        }
      }
      // lane (i, g) register r: T[i][dt + 4g + r], minus E (row r0+i is delayed row dd)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int dd = dt + 4 * g + r;
        float v = acc0[r] + acc1[r];
        if (dd < c && sRows[dd] == r0 + li) v -= 1.f;
        sT[li * KC + dd] = dd < c ? v : 0.f;
      }
    }
    wave_lds_fence();
    flush_own_t2_mfma(c, cp, KC, KP, sBinv, sT);
    // ---- GEMM 2: rows^T += V^T T2^T, four delayed rows per step; V rows from LDS
    for (int d0 = 0; d0 < c4; d0 += 4) {
      const int dk = d0 + g;
      const float bv = sT[li * KC + dk];
      const float* vr = sAc + (size_t)sAl[dk < c ? dk : 0] * sa + li;
#pragma unroll
      for (int q = 0; q < kFlushOwnTiles; ++q)
        if (q < nct) {
          const float v = dk < c ? vr[q * 16] : 0.f;
          a[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(v, bv, a[q], 0, 0, 0);
        }
    }
#pragma unroll
    for (int q = 0; q < kFlushOwnTiles; ++q)
      if (q < nct) *reinterpret_cast<f32x4_t*>(ar + q * 16) = a[q];
    wave_lds_fence();   // the next tile's T overwrites this one's T2
  }
  __syncthreads();   // the next chunk reads these rows
}

// The same update when more than kFlushOwnU moves of the chunk were accepted: a tile is held
// in registers half its orbitals at a time: the first product sums both halves, the second
// updates each half and stores it.  U rows are the new Slater rows in HBM (written at the
// chunk's end, six orbital tiles' loads in flight).
constexpr int kFlushOwnHalf = kFlushOwnTiles / 2;
__device__ __attribute__((noinline)) void flush_own(int n, int K, int KP, int c, const int* sRows, const int* sAl, const double* sBinv,
                                 const float* sAc, int sa, const float* Amat, float* Ainv, float* scratch) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
  const int li = lane & 15, g = lane >> 4;
  const int KC = (K + 15) & ~15, cp = (c + 15) & ~15, nct = n >> 4;
  float* sT = scratch + (size_t)wid * 16 * KC;
  for (int tile = wid; tile < nct; tile += nw) {
    const int r0 = tile * 16;
    const float* ar = Ainv + (size_t)(r0 + li) * n + 4 * g;
    // ---- GEMM 1: T^T = U rows^T per 16 delayed rows, the orbitals in two halves
    for (int dt = 0; dt < cp / 16; ++dt) {
      const int d = dt * 16 + li;
      const float* ur = Amat + (size_t)sRows[d < c ? d : 0] * n + 4 * g;
      f32x4_t acc0 = f32x4_t{0.f, 0.f, 0.f, 0.f}, acc1 = f32x4_t{0.f, 0.f, 0.f, 0.f};
      for (int h0 = 0; h0 < nct; h0 += kFlushOwnHalf) {
        f32x4_t a[kFlushOwnHalf];
#pragma unroll
        for (int q = 0; q < kFlushOwnHalf; ++q)
          a[q] = *reinterpret_cast<const f32x4_t*>(ar + min(h0 + q, nct - 1) * 16);
#pragma unroll
        for (int kb = 0; kb < kFlushOwnHalf; kb += 6) {
          f32x4_t u[6];
#pragma unroll
          for (int q = 0; q < 6; ++q)
            u[q] = *reinterpret_cast<const f32x4_t*>(ur + min(h0 + kb + q, nct - 1) * 16);
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            const bool in = d < c && h0 + kb + q < nct;
            const f32x4_t uv = in ? u[q] : f32x4_t{0.f, 0.f, 0.f, 0.f};
            f32x4_t& acc = (q & 1) ? acc1 : acc0;
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(uv[0], a[kb + q][0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(uv[1], a[kb + q][1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(uv[2], a[kb + q][2], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(uv[3], a[kb + q][3], acc, 0, 0, 0);
          }
        }
      }
      // lane (i, g) register r: T[i][16dt + 4g + r], minus E (row r0+i is delayed row dd)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int dd = dt * 16 + 4 * g + r;
        float v = acc0[r] + acc1[r];
        if (dd < c && sRows[dd] == r0 + li) v -= 1.f;
        sT[li * KC + dd] = dd < c ? v : 0.f;
      }
    }
    wave_lds_fence();
    flush_own_t2(c, cp, KC, KP, sBinv, sT);
    // ---- GEMM 2: rows^T += V^T T2^T per half of the orbitals; V rows from LDS
    for (int h0 = 0; h0 < nct; h0 += kFlushOwnHalf) {
      f32x4_t a[kFlushOwnHalf];
#pragma unroll
      for (int q = 0; q < kFlushOwnHalf; ++q)
        a[q] = *reinterpret_cast<const f32x4_t*>(ar + min(h0 + q, nct - 1) * 16);
      for (int d0 = 0; d0 < cp; d0 += 4) {
        const int dk = d0 + g;
        const float bv = sT[li * KC + dk];
        const float* vr = sAc + (size_t)sAl[dk < c ? dk : 0] * sa + li;
#pragma unroll
        for (int q = 0; q < kFlushOwnHalf; ++q) {
          const float v = dk < c ? vr[min(h0 + q, nct - 1) * 16] : 0.f;
          a[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(v, bv, a[q], 0, 0, 0);
        }
      }
      float* aw = Ainv + (size_t)(r0 + li) * n + 4 * g;
#pragma unroll
      for (int q = 0; q < kFlushOwnHalf; ++q)
        if (h0 + q < nct) *reinterpret_cast<f32x4_t*>(aw + (h0 + q) * 16) = a[q];
    }
  }
  __syncthreads();   // the next chunk reads these rows
}

// ---- FP64 re-evaluation of a near-tie Metropolis decision (FP32 form, K > 1).
// The determinant ratio of the move, psiV64 . x with A64 x = e_row (A64: the FP64 Slater rows
// at the current positions, the chunk's accepted rows taken from their proposals' FP64 values),
// by iterative refinement from the FP32 path's inverse row x0:
//   s = e_row - A64 x,   x += X s,
// where X is the walker's current FP32 inverse: the chunk-start inverse Ainv (psiM layout,
// X s = sum_i s_i Ainv[i]) with the chunk's delayed Woodbury terms (DelayedUpdate.h:176-193):
//   X s = sigma + sum_d t_d V_d,  sigma = sum_i s_i Ainv[i],
//   t = -Binv^T q,  q_e = sigma . U_e - s[rows_e]
// (U_e the new FP32 Slater rows, V_d the stale inverse rows, Binv the capacitance inverse).  It
// stops when the ratio's last correction is below 1e-14 of it, when the corrections stop shrinking
// (the FP64 rounding floor of an ill-conditioned row), or (j2 > 0) when the error left, bounded
// from the last two corrections as a geometric tail, can no longer move (rho^2 j2) across the
// uniform u.  xv holds x0 on entry; xv, sv, tv are n doubles of LDS, red 2K + 32, part kArbGroups
// n.  Every lane returns the ratio.
constexpr int kArbMaxIter = 8;
__device__ inline double arb_det_ratio(int n, int npad, int row, int r0, int c, int KP, int sa,
                                                          const float* Ainv, const double* A64, const double* s64c,
                                                          const float* spoc, const float* sAc, const double* sBinv,
                                                          const int* sRows, const int* sAl, const double* psiv64,
                                                          double* xv, double* sv, double* tv, double* red, int K,
                                                          double* part, double u, double j2) {
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  double* q = red;           // [K]
  double* t = red + K;       // [K]
  double* bs = red + 2 * K;  // block_sum scratch
  // the chunk's accepted electrons (their FP64 Slater rows are their proposals'), a bit each
  unsigned long long m0 = 0ull, m1 = 0ull;
  for (int d = 0; d < c; ++d) {
    const int ei = sAl[d];
    if (ei < 64) m0 |= 1ull << ei;
    else m1 |= 1ull << (ei - 64);
  }
  double rho = 0.0, prev = INFINITY;
  for (int it = 0; it < kArbMaxIter; ++it) {
    // residual, four rows per wave at a time (their loads in flight together)
    for (int i0 = wid * 4; i0 < n; i0 += nw * 4) {
      const double* ar[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = min(i0 + r, n - 1), ei = i - r0;
        const bool acc = ei >= 0 && ei < 128 && (((ei < 64 ? m0 : m1) >> (ei & 63)) & 1ull) != 0ull;
        ar[r] = acc ? s64c + (size_t)ei * npad : A64 + (size_t)i * n;
      }
      double a4[4] = {0.0, 0.0, 0.0, 0.0};
      for (int j = lane; j < n; j += 64) {
        const double xj = xv[j];
#pragma unroll
        for (int r = 0; r < 4; ++r) a4[r] += ar[r][j] * xj;
      }
      wave_sums_rows<4>(a4);
      const int i = i0 + (lane >> 4);   // row r of the wave holds the sum of row i0 + r
      if ((lane & 15) == 0 && i < n) sv[i] = (i == row ? 1.0 : 0.0) - a4[0];
    }
    __syncthreads();
    // sigma = sum_i s_i Ainv[i]: a lane takes four columns (16-byte loads) and G groups of lanes
    // split the rows, their partial sums added in group order
    const int nq = n >> 2;
    if ((n & 3) == 0 && nq <= nt) {
      const int G = min(nt / nq, kArbGroups), gq = tid / nq, cq = tid - gq * nq;
      if (gq < G) {
        double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
        for (int i = gq; i < n; i += G) {
          const f32x4_t a = *reinterpret_cast<const f32x4_t*>(Ainv + (size_t)i * n + 4 * cq);
          const double si = sv[i];
#pragma unroll
          for (int r = 0; r < 4; ++r) s4[r] += si * (double)a[r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) part[(size_t)gq * n + 4 * cq + r] = s4[r];
      }
      __syncthreads();
      if (tid < n) {
        double sg = 0.0;
        for (int g = 0; g < G; ++g) sg += part[(size_t)g * n + tid];
        tv[tid] = sg;
      }
    } else if (tid < n) {
      double sg = 0.0;
#pragma unroll 16
      for (int i = 0; i < n; ++i) sg += sv[i] * (double)Ainv[(size_t)i * n + tid];
      tv[tid] = sg;
    }
    __syncthreads();
    for (int e = wid; e < c; e += nw) {
      const float* u = spoc + (size_t)(sRows[e] - r0) * 5 * npad;
      double acc = 0.0;
      for (int j = lane; j < n; j += 64) acc += tv[j] * (double)u[j];
      acc = wave_sum(acc);
      if (lane == 0) q[e] = acc - sv[sRows[e]];
    }
    __syncthreads();
    for (int d = tid; d < c; d += nt) {
      double acc = 0.0;
      for (int e = 0; e < c; ++e) acc += q[e] * sBinv[e * KP + d];
      t[d] = -acc;
    }
    __syncthreads();
    double v[2] = {0.0, 0.0};
    if (tid < n) {
      double dl = tv[tid];
      for (int d = 0; d < c; ++d) dl += t[d] * (double)sAc[(size_t)(sRows[d] - r0) * sa + tid];
      const double xn = xv[tid] + dl;
      xv[tid] = xn;
      const double pv = psiv64[tid];
      v[0] = pv * dl;
      v[1] = pv * xn;
    }
    block_sum<double, 2>(v, bs);
    rho = v[1];
    const double dr = fabs(v[0]);
    if (dr <= 1e-14 * fabs(rho) || dr >= 0.5 * prev) break;
    if (j2 > 0.0 && it > 0) {
      const double qc = dr / prev;                 // contraction of the last step
      const double rem = dr * qc / (1.0 - qc);     // geometric bound on the error left
      if (4.0 * fabs(rho) * rem * j2 < fabs(u - rho * rho * j2)) break;
    }
    prev = dr;
  }
  return rho;
}

// J1 and J2 ratio of moving electron iel from (x0, y0, z0) to (nx, ny, nz) in FP64 from scratch
// (OneBodyJastrow.h:259-268, TwoBodyJastrow.h:294-304): exp(sum_I u(old) - u(new) + sum_j ...).
// Every lane returns it; red: 64 doubles of LDS.
template <typename T>
__device__ inline double arb_jastrow_ratio(const SysP<T>& S, const double* sR, int iel, double x0,
                                                              double y0, double z0, double nx, double ny, double nz,
                                                              double* red, const unsigned long long* sMask = nullptr) {
  const int tid = threadIdx.x, nt = blockDim.x, N = S.N, n = S.n, M = S.M;
  const Cell<double>& C = *S.cell64;
  double v[4] = {0.0, 0.0, 0.0, 0.0};   // J1 old, new; J2 (+ J3) old, new
  double dx, dy, dz;
  for (int ci = tid; ci < M; ci += nt) {
    const Functor<double>& f = S.fn64[S.species[ci]];
    if (f.ncoef == 0) continue;
    const double ix = S.ions[ci], iy = S.ions[M + ci], iz = S.ions[2 * M + ci];
    const double ro = min_image(C, ix - x0, iy - y0, iz - z0, -1.0, dx, dy, dz);
    const double rn = min_image(C, ix - nx, iy - ny, iz - nz, -1.0, dx, dy, dz);
    if (ro < f.rcut) v[0] += functor_v(f, ro);
    if (rn < f.rcut) v[1] += functor_v(f, rn);
  }
  const bool up = iel < n;
  for (int j = tid; j < N; j += nt) {
    if (j == iel) continue;
    const Functor<double>& f = S.fn64[kMaxSpecies + (((j < n) == up) ? 0 : 1)];
    const double ro = min_image(C, sR[j] - x0, sR[N + j] - y0, sR[2 * N + j] - z0, -1.0, dx, dy, dz);
    const double rn = min_image(C, sR[j] - nx, sR[N + j] - ny, sR[2 * N + j] - nz, -1.0, dx, dy, dz);
    if (ro < f.rcut) v[2] += functor_v(f, ro);
    if (rn < f.rcut) v[3] += functor_v(f, rn);
  }
  if (sMask != nullptr) {
    // three-body terms (ThreeBodyJastrow.h:277-300, values only) of the electron at its old and new
    // positions in FP64: ions within L of the position, electrons k within L of the ion (the ion
    // masks of the current positions)
    const int mw = S.mw;
    for (int I = 0; I < M; ++I) {
      const int sp = S.species[I];
      const double L = S.j3_64[2 * sp].L;
      if (!(L > 0.0)) continue;
      const double ix = S.ions[I], iy = S.ions[M + I], iz = S.ions[2 * M + I];
      const double roI = min_image(C, ix - x0, iy - y0, iz - z0, -1.0, dx, dy, dz);
      const double rnI = min_image(C, ix - nx, iy - ny, iz - nz, -1.0, dx, dy, dz);
      if (!(roI < L) && !(rnI < L)) continue;
      const unsigned long long bit = 1ull << (I & 63);
      const int wd = I >> 6;
      for (int k = tid; k < N; k += nt) {
        if (k == iel || !(sMask[(size_t)k * mw + wd] & bit)) continue;
        const double kx = sR[k], ky = sR[N + k], kz = sR[2 * N + k];
        const double rkI = min_image(C, ix - kx, iy - ky, iz - kz, -1.0, dx, dy, dz);
        const Poly3<double>& f = S.j3_64[2 * sp + (((k < n) == up) ? 0 : 1)];
        if (roI < L) v[2] += poly3_v(f, min_image(C, kx - x0, ky - y0, kz - z0, -1.0, dx, dy, dz), roI, rkI);
        if (rnI < L) v[3] += poly3_v(f, min_image(C, kx - nx, ky - ny, kz - nz, -1.0, dx, dy, dz), rnI, rkI);
      }
    }
  }
  block_sum<double, 4>(v, red);
  return exp((v[0] - v[1]) + (v[2] - v[3]));
}

// ---- three-body Jastrow of one move phase (ThreeBodyJastrow::computeU3 with computeU3_engine,
// ThreeBodyJastrow.h:441-625) for the moving electron iel at position p with near-ion record rec.
// Every wave takes the pairs (k, ion q of rec) of the electrons k its own lanes hold (k = tid +
// qq nt) whose ion mask has the ion, compacts them into its kJ3Slots slots (slot = ballot rank,
// in (q, qq, lane) order), and lane r evaluates the functor for slot r: its j-side terms are
// returned in js (zero for lanes without a slot), the k-side terms go to the slot's row of sv
// [5][kJ3Slots] beside its code k | q << 16 in sk.  Returns the wave's slot count (more pairs
// than slots raise kFlagJ3Pairs).  Deterministic: the slot order depends only on the
// configuration.
template <typename T, int JPT>
__device__ inline int j3_phase(const SysP<T>& S, const Cell<T>& C, const double* sR, const unsigned long long* sMask,
                               const T* rec, int iel, double px, double py, double pz, int tid, int nt, int lane,
                               int* sk, T* sv, T js[5], int* flags) {
  const int N = S.N, n = S.n, M = S.M, mw = S.mw;
  const int cnt = (int)rec[0];
  const int nI = cnt < kJ3Near ? cnt : kJ3Near;
  int ns = 0;
  for (int q = 0; q < nI; ++q) {
    const int I = (int)rec[1 + 5 * q];
    const unsigned long long bit = 1ull << (I & 63);
    const int wd = I >> 6;
#pragma unroll
    for (int qq = 0; qq < JPT; ++qq) {
      const int k = tid + qq * nt;
      const bool in = k < N && k != iel && (sMask[(size_t)k * mw + wd] & bit) != 0ull;
      const unsigned long long bal = __ballot(in);
      const int pos = ns + __popcll(bal & ((1ull << lane) - 1ull));
      if (in && pos < kJ3Slots) sk[pos] = k | (q << 16);
      ns += __popcll(bal);
    }
  }
  if (lane == 0 && (ns > kJ3Slots || cnt > kJ3Near)) atomicOr(flags, ns > kJ3Slots ? kFlagJ3Pairs : kFlagJ3Near);
  ns = ns < kJ3Slots ? ns : kJ3Slots;
  wave_lds_fence();   // the slots are read by other lanes of the wave
#pragma unroll
  for (int c = 0; c < 5; ++c) js[c] = T(0);
  if (lane < ns) {
    const int code = sk[lane], k = code & 0xffff, q = code >> 16;
    const T* e = rec + 1 + 5 * q;
    const int I = (int)e[0];
    const T jI[3] = {e[2], e[3], e[4]};
    T kI[3], jk[3], o[9], ks[5];
    const T rkI = min_image(C, (T)(S.ions[I] - sR[k]), (T)(S.ions[M + I] - sR[N + k]),
                            (T)(S.ions[2 * M + I] - sR[2 * N + k]), T(-1), kI[0], kI[1], kI[2]);
    const T rjk = min_image(C, (T)(sR[k] - px), (T)(sR[N + k] - py), (T)(sR[2 * N + k] - pz),
                            k < iel ? T(1) : T(-1), jk[0], jk[1], jk[2]);
    poly3_vgl(S.j3[2 * S.species[I] + (((k < n) == (iel < n)) ? 0 : 1)], rjk, e[1], rkI, o);
    j3_pair_terms(o, jk, jI, kI, js, ks);
#pragma unroll
    for (int c = 0; c < 5; ++c) sv[c * kJ3Slots + lane] = ks[c];
  }
  return ns;
}

// Adds sign times the k-side terms of the wave's ns slots to the sums sJ3 [5][N], one near ion
// at a time (an electron appears at most once per ion), so the order of additions is fixed.
template <typename T>
__device__ inline void j3_apply(T* sJ3, int N, const int* sk, const T* sv, int ns, int nI, T sign, int lane) {
  for (int q = 0; q < nI; ++q) {
    if (lane < ns && (sk[lane] >> 16) == q) {
      const int k = sk[lane] & 0xffff;
#pragma unroll
      for (int c = 0; c < 5; ++c) sJ3[(size_t)c * N + k] += sign * sv[c * kJ3Slots + lane];
    }
    wave_lds_fence();
  }
}

// KD: delayed updates (K > 1: inverse-row cache, Woodbury bordering) or Sherman-Morrison (K = 1).
// ARB: the FP64 re-evaluation of near-tie Metropolis decisions is compiled in (FP32, K > 1); only
// that instantiation carries its code, whose registers would otherwise spill in every move.
// J3: the three-body Jastrow is compiled in (its ratio, gradient and update in every move).
// ST (with KD): the streamed form of the delayed updates, for delay ranks whose inverse-row cache
// does not fit the LDS and for delayed rows kept pending across launches (a.resume): the stale
// inverse rows are read from Ainv in HBM (unchanged until the flush launch), p_d = U_d . Ainv[row]
// is formed at each move (DelayedUpdate.h:108-113, U_d the new Slater row of delayed row d), and
// the capacitance inverse, the row list and the count live in the chunk's state slot (Binv in LDS
// while the chunk runs when it fits); each accepted row's stale inverse row goes to the slot's V
// rows at once.  A flush launch applies the slot.
// The workgroup moves walker w of batch (S, B) as directed by a.
template <typename T, int JPT, bool KD, bool ARB, bool J3, bool ST = false>
__device__ __forceinline__ void moves_body(const SysP<T>& S, const BatchP<T>& B, const MoveArgs& a, const int w) {
  static_assert(!ST || KD, "the streamed form is a form of the delayed updates");
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  const int N = S.N, n = S.n, K = KD ? B.K : 1;
  const int KP = K + 1;   // LDS row stride of Binv: a column walk down its rows is free of bank conflicts
  const int Emax = a.chunk_len;   // electrons per chunk (the LDS layout is sized for it)
  const MoveLds L = move_lds<T>(N, S.M, n, K, Emax, J3 ? S.mw : 0, ST, ARB && J3);
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
  const int sa = ac_stride<T>(n);   // its row stride
  double* sQ = (double*)(smem + L.oQ);
  // streamed form: p_d of the move [K], y = Binv^T p [K], reduction partials [kMoveThreads]
  double* sSp = (double*)(smem + L.oSt);
  double* sSy = sSp + K;
  double* sSg = sSy + K;
  // three-body Jastrow: sums [5][N], ion masks [N][mw], near-ion records of the chunk's
  // proposals and current positions [E][kJ3Rec], this wave's pair slots of the new and the old
  // phase, the j-side wave partials [16][5]
  constexpr int kWaves = kMoveThreads / 64;
  T* sJ3 = (T*)(smem + L.oJ3);
  unsigned long long* sMask = (unsigned long long*)(smem + L.oJ3m);
  T* sJ3np = (T*)(smem + L.oJ3np);
  T* sJ3ne = (T*)(smem + L.oJ3ne);
  int* sJk = (int*)(smem + L.oJ3k) + wid * kJ3Slots;
  int* sJkO = (int*)(smem + L.oJ3k) + (kWaves + wid) * kJ3Slots;
  T* sJv = (T*)(smem + L.oJ3v) + (size_t)wid * 5 * kJ3Slots;
  T* sJvO = (T*)(smem + L.oJ3v) + (size_t)(kWaves + wid) * 5 * kJ3Slots;
  double* sJr = (double*)(smem + L.oJ3r);
  // FP64 re-evaluation of near ties (FP32 form, Metropolis): the move's inverse row in FP64
  const bool arb_on = ARB && sizeof(T) == 4 && KD && a.band > 0.0 && a.mode == 1 && a.decide == nullptr && n <= nt;
  double* sArbX = (double*)(smem + L.oArb);
  const size_t moff = ((size_t)w * 2 + a.det) * n * n;
  T* Ainv = B.Ainv + moff;
  T* Amat = B.A + moff;
  T* dpsiM = B.dpsiM + ((size_t)w * 2 + a.det) * (size_t)n * 3 * n;
  T* d2psiM = B.d2psiM + moff;
  T* J2 = B.J2 + (size_t)w * 5 * N;
  T* J1 = B.J1 + (size_t)w * 5 * N;

  // ---- walker state that stays in LDS across the launch's chunks: positions, J2 sums, J2
  // functors, the cell
  {
    const LdsCopy cp[4] = {{sR, B.R + (size_t)w * 3 * N, (unsigned)(sizeof(double) * 3 * N)},
                           {sJ2, J2, (unsigned)(sizeof(T) * 5 * N)},
                           {sFn, S.fn + kMaxSpecies, (unsigned)(sizeof(Functor<T>) * 2)},
                           {sCell, S.cell, (unsigned)sizeof(Cell<T>)}};
    stage_lds(cp);
  }
  if constexpr (J3) {
    const T* g3 = B.J3 + (size_t)w * 5 * N;
    for (int i = tid; i < 5 * N; i += nt) sJ3[i] = g3[i];
    const unsigned long long* gm = B.j3mask + (size_t)w * N * S.mw;
    for (int i = tid; i < N * S.mw; i += nt) sMask[i] = gm[i];
  }
  // a caller's accept decision (the flex stage calls) replaces the uniform: -inf accepts, +inf rejects
  const double flex_u = a.decide == nullptr ? 0.0 : (a.decide[w] != 0 ? -INFINITY : INFINITY);
  // sums over the launch kept by thread 0 in LDS: log|det| (logd, and the running product pdet
  // folded into it before it leaves the double range), phase, J1 and J2 log changes, accepts
  double* sAcc = (double*)(smem + L.oAcc);
  enum { aLogd, aPdet, aPh, aJl1, aJl2, aNacc, aNarb, aJl3 };
  if (tid == 0) {
    sAcc[aLogd] = 0;
    sAcc[aPdet] = 1;
    sAcc[aPh] = 0;
    sAcc[aJl1] = 0;
    sAcc[aJl2] = 0;
    sAcc[aNacc] = 0;
    sAcc[aNarb] = 0;
    sAcc[aJl3] = 0;
  }
  const Cell<T>& C = *sCell;   // read from LDS where used (a register copy costs spills)
  const int jr = tid < n ? tid : n - 1;
  // FP64 recompute cadence: this walker's first electron and its accepted rows since the last
  // recompute; it stops after the move that brings them to a.cadence
  int cbn = a.e_begin, since = 0;
  bool stopped = false;
  int* cadw = nullptr;
  if (a.cadence > 0) {
    cadw = a.cad + ((size_t)w * 2 + a.det) * 2;
    cbn += cadw[0];
    since = cadw[1];
  }

  for (int ci = 0; ci < a.nchunks && cbn < a.e_end && !stopped; ++ci) {
  const int cb = cbn;                       // the chunk's first electron
  const int E = min(Emax, a.e_end - cb);
  int Eend = E;                             // moves made (fewer when the cadence stops the walker)
  const int r0 = cb - a.det * n;
  const int chunk = a.chunk + ci;           // index of the chunk in its determinant's sweep
  // this chunk's delayed-update state slot
  const size_t slot = state_slot(B, w, a.det, KD ? chunk : 0);
  T* Vb = B.Vbuf + slot * K * n;
  // this chunk's proposals: SPO rows [e][5][npad] and records [e][kProp]
  const T* spo = (const T*)a.spo + (((size_t)w * 2 + a.det) * n + r0) * 5 * S.npad;
  const double* prop = a.prop + (((size_t)w * 2 + a.det) * n + r0) * kProp;
  __syncthreads();   // the previous chunk's reads of the chunk regions are done

  // ---- chunk set-up: its stale inverse rows, proposal records and J1 state into LDS (one
  // round of loads)
  {
    const unsigned je = (unsigned)(sizeof(T) * E);
    const LdsCopy cp[6] = {{sAc, Ainv + (size_t)r0 * n, KD && !ST ? (unsigned)(sizeof(T) * E * n) : 0u,
                            (unsigned)(sizeof(T) * n), (unsigned)(sizeof(T) * sa)},
                           {sProp, prop, (unsigned)(sizeof(double) * kProp * E)},
                           {sJ1o, J1 + cb, je},
                           {sJ1o + E, J1 + N + cb, je},
                           {sJ1o + 2 * E, J1 + 2 * N + cb, je},
                           {sJ1o + 3 * E, J1 + 3 * N + cb, je}};
    stage_lds(cp);
  }
  if constexpr (J3) {
    // the near-ion records of the chunk's proposals and of its electrons' current positions
    const T* gp = B.j3prop + (((size_t)w * 2 + a.det) * n + r0) * kJ3Rec;
    const T* ge = B.j3near + ((size_t)w * N + cb) * kJ3Rec;
    for (int i = tid; i < E * kJ3Rec; i += nt) {
      sJ3np[i] = gp[i];
      sJ3ne[i] = ge[i];
    }
  }
  int c = 0, ca = 0;   // delayed rows; accepted moves of the chunk
  // streamed form: the slot's capacitance inverse (row stride K, in LDS while the chunk runs when
  // it fits) and the c0 delayed rows pending from earlier launches (a.resume)
  double* const gBinv = B.Binv + slot * K * K;
  double* const binv = ST && L.binv_lds ? sBinv : gBinv;
  int c0 = 0;
  if constexpr (ST) {
    if (a.resume) {
      c0 = B.ndelay[slot];
      for (int i = tid; i < c0; i += nt) sRows[i] = B.rows[slot * K + i];
      if (L.binv_lds)
        for (int i = tid; i < c0 * c0; i += nt) {
          const int s2 = i / c0, r = i - s2 * c0;
          sBinv[(size_t)s2 * K + r] = gBinv[(size_t)s2 * K + r];
        }
    }
    c = c0;
  }
  __syncthreads();
  // Rows a move reads from HBM that no earlier move of the chunk changes: the moving electron's
  // dpsiM row and the SPO row of its proposal.  Move e+1's are loaded at the start of move e
  // (software pipelining; index clamped, so the loads are unconditional).
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

  for (int e = 0; e < Eend; ++e) {
    const int iel = cb + e, row = r0 + e;
    // ---- phase 1: issue the next move's global loads, then getInvRow's small algebra
    // stored state of the moving electron (J1, J2), its dpsiM row (loaded one move ahead) and
    // its proposal record: J1 at the proposal (sum u, sum (d2u + 2 du/r), gradient),
    // displacement, uniform
    const T Vat_old = sJ1o[e], g1x = sJ1o[E + e], g1y = sJ1o[2 * E + e], g1z = sJ1o[3 * E + e];
    const T Uat_old = sJ2[iel], g2x = sJ2[N + iel], g2y = sJ2[2 * N + iel], g2z = sJ2[3 * N + iel];
    T u3o = 0, g3x = 0, g3y = 0, g3z = 0;
    if constexpr (J3) {
      u3o = sJ3[iel];
      g3x = sJ3[N + iel];
      g3y = sJ3[2 * N + iel];
      g3z = sJ3[3 * N + iel];
    }
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
    if constexpr (ST) {
      if (c > 0) {
        // p_d = U_d . Ainv[row] for the delayed rows, a wave per row (U_d: the chunk's proposal
        // row, or for a row pending from an earlier launch the Slater row written there)
        const T* ar = Ainv + (size_t)row * n;
        for (int d = wid; d < c; d += nw) {
          const T* u = d < c0 ? Amat + (size_t)sRows[d] * n : spo + (size_t)(sRows[d] - r0) * 5 * S.npad;
          double s = 0;
          for (int j = lane; j < n; j += 64) s += (double)u[j] * (double)ar[j];
          s = wave_sum(s);
          if (lane == 0) sSp[d] = s;
        }
        lds_sync();
        // y_r = sum_q Binv[q][r] p_q: G groups of 128 lanes take every G-th q
        const int G = nt >> 7, r = tid & 127, gq = tid >> 7;
        double s = 0;
        if (r < c)
          for (int q = gq; q < c; q += G) s += binv[(size_t)q * K + r] * sSp[q];
        sSg[gq * 128 + r] = s;
        lds_sync();
        if (tid < c) {
          double y = 0;
          for (int g = 0; g < G; ++g) y += sSg[g * 128 + tid];
          sSy[tid] = y;
        }
      }
    } else if (KD && c > 0) {
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

    // ---- phase 2: invRow, evalGrad of the determinant and ratioGrad_compute against the
    // proposal's SPO row (einspline_spo.hpp:204-223, DiracDeterminant.cpp:103-124), J2 at
    // the proposal; all partial sums reduced in one round
    double gd0 = 0, gd1 = 0, gd2 = 0, d0 = 0, d1 = 0, d2 = 0, d3 = 0;
    T vinv = 0;
    if (tid < n) {
      double v;
      if constexpr (ST) {
        // the stale row minus the delayed rows' terms, eight HBM rows in flight
        const T* ac = Ainv + tid;
        v = (double)ac[(size_t)row * n];
        int d = 0;
        for (; d + 8 <= c; d += 8) {
          T a8[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) a8[k] = ac[(size_t)sRows[d + k] * n];
#pragma unroll
          for (int k = 0; k < 8; ++k) v -= (double)a8[k] * sSy[d + k];
        }
        for (; d < c; ++d) v -= (double)ac[(size_t)sRows[d] * n] * sSy[d];
      } else if (KD) {
        v = (double)sAc[(size_t)e * sa + tid];
        const double* y = sBinv + c * KP;
        // four delayed rows' LDS reads in flight at a time; subtraction order unchanged
        int d = 0;
        for (; d + 4 <= c; d += 4) {
          const int e0 = sRows[d] - r0, e1 = sRows[d + 1] - r0, e2 = sRows[d + 2] - r0, e3 = sRows[d + 3] - r0;
          const double y0 = y[d], y1 = y[d + 1], y2 = y[d + 2], y3 = y[d + 3];
          const double a0 = (double)sAc[(size_t)e0 * sa + tid], a1 = (double)sAc[(size_t)e1 * sa + tid];
          const double a2 = (double)sAc[(size_t)e2 * sa + tid], a3 = (double)sAc[(size_t)e3 * sa + tid];
          v -= a0 * y0;
          v -= a1 * y1;
          v -= a2 * y2;
          v -= a3 * y3;
        }
        for (; d < c; ++d) v -= (double)sAc[(size_t)(sRows[d] - r0) * sa + tid] * y[d];
      } else {
        v = (double)inv_sm;
      }
      vinv = (T)v;
      sInv[tid] = vinv;
      if (arb_on) sArbX[tid] = v;
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
      double v = KD && !ST ? (double)sAc[(size_t)e * sa + j] : (double)Ainv[(size_t)row * n + j];
      if (KD)
        for (int d = 0; d < c; ++d)
          v -= ST ? (double)Ainv[(size_t)sRows[d] * n + j] * sSy[d]
                  : (double)sAc[(size_t)(sRows[d] - r0) * sa + j] * sBinv[c * KP + d];
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
    }
    int ns3 = 0;   // this wave's J3 pair slots at the proposal
    if constexpr (J3) {
      T js[5];
      ns3 = j3_phase<T, JPT>(S, C, sR, sMask, sJ3np + (size_t)e * kJ3Rec, iel, nx, ny, nz, tid, nt, lane, sJk, sJv,
                             js, B.flags);
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        const double t = wave_sum((double)js[c]);
        if (lane == 0) sJr[wid * 5 + c] = t;
      }
    }
    lds_sync();

    double red[kRed];
    combine_partials(sRed, nw, red);
    // J3 at the proposal: U, grad U, lap U of the moving electron (the wave partials in order)
    double cur3[5] = {0, 0, 0, 0, 0};
    if constexpr (J3)
      for (int q = 0; q < nw; ++q)
#pragma unroll
        for (int c = 0; c < 5; ++c) cur3[c] += sJr[q * 5 + c];

    // ---- phase 4: acceptance and updates
    double rdet = red[3];
    double ratio = rdet * exp((double)(Vat_old - (T)pr[kPropJ1]) + (double)(Uat_old - (T)red[7]) +
                              (J3 ? (double)(u3o - (T)cur3[0]) : 0.0));
    const double uacc = a.decide == nullptr ? pr[kPropU] : flex_u;
    bool accept = a.mode == 0 ? (uacc < 0.5) : (uacc < ratio * ratio);
    if constexpr (ARB && sizeof(T) == 4 && KD) {
      // a uniform within the band around ratio^2: the decision is taken on the FP64 ratio
      // a rare branch: marked cold so that the register allocator keeps its spills inside it
      if (__builtin_expect(arb_on && c0 == 0 && fabs(uacc - ratio * ratio) <= a.band * ratio * ratio, 0)) {
        double* sv = sArbX + n;
        double* tv = sv + n;
        double* ared = tv + n;
        const size_t pro = ((size_t)w * 2 + a.det) * n + r0;   // the chunk's first proposal
        // the stale rows: the LDS cache, or (streamed) Ainv's rows of the chunk in HBM
        const double jr = arb_jastrow_ratio<T>(S, sR, iel, x0, y0, z0, nx, ny, nz, ared, J3 ? sMask : nullptr);
        rdet = arb_det_ratio(n, S.npad, row, r0, c, ST ? K : KP, ST ? n : sa, (const float*)Ainv, B.A64 + moff,
                             B.spo64 + pro * S.npad, (const float*)spo,
                             ST ? (const float*)Ainv + (size_t)r0 * n : (const float*)sAc, ST ? binv : sBinv, sRows,
                             sAl, B.spo64 + (pro + e) * S.npad, sArbX, sv, tv, ared, K, ared + 2 * K + 32, uacc,
                             jr * jr);
        ratio = rdet * jr;
        accept = uacc < ratio * ratio;
        if (tid == 0) sAcc[aNarb] += 1.0;
      }
    }
    if (B.trace != nullptr && tid == 0) {
      double* t = B.trace + ((size_t)w * N + iel) * 9;
      t[0] = ratio;
      t[1] = rdet;
      t[2] = accept ? 1.0 : 0.0;
      t[3] = red[0] + (double)g1x + (double)g2x + (double)g3x;
      t[4] = red[1] + (double)g1y + (double)g2y + (double)g3y;
      t[5] = red[2] + (double)g1z + (double)g2z + (double)g3z;
      t[6] = red[4] / rdet + pr[kPropJ1 + 2] + red[8] + cur3[1];
      t[7] = red[5] / rdet + pr[kPropJ1 + 3] + red[9] + cur3[2];
      t[8] = red[6] / rdet + pr[kPropJ1 + 4] + red[10] + cur3[3];
    }
    if (accept) {
      if (tid == 0) {
        // log|det| of the accepted ratios as a running product, folded into the log before it
        // can leave the double range
        double pdet = sAcc[aPdet] * fabs(rdet);
        if (!(pdet > 0x1p-500 && pdet < 0x1p500)) {
          sAcc[aLogd] += log(pdet);
          pdet = 1.0;
        }
        sAcc[aPdet] = pdet;
        sAcc[aPh] += rdet > 0.0 ? 0.0 : M_PI;
        sAcc[aJl1] += (double)(Vat_old - (T)pr[kPropJ1]);
        sAcc[aJl2] += (double)(Uat_old - (T)red[7]);
        if constexpr (J3) sAcc[aJl3] += (double)(u3o - (T)cur3[0]);
        sAcc[aNacc] += 1.0;
        // the new J1 state of the electron (OneBodyJastrow.h:271-283), written at the chunk's end
        sAl[ca] = e;
        sJ1n[ca * 5 + 0] = (T)pr[kPropJ1 + 0];
        sJ1n[ca * 5 + 1] = (T)pr[kPropJ1 + 2];
        sJ1n[ca * 5 + 2] = (T)pr[kPropJ1 + 3];
        sJ1n[ca * 5 + 3] = (T)pr[kPropJ1 + 4];
        sJ1n[ca * 5 + 4] = (T)pr[kPropJ1 + 1];
      }
      ++ca;
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
        // Streamed: p_d only (V_d the stale rows in HBM; this move's is row `row`).
        const int nd = ST ? c + 1 : c + 1 + (E - 1 - e);
        auto dvec = [&](int t2) -> const T* {
          if constexpr (ST) return Ainv + (size_t)(t2 < c ? sRows[t2] : row) * n;
          const int ee = t2 < c ? sRows[t2] - r0 : e + (t2 - c);
          return sAc + (size_t)ee * sa;
        };
        for (int t0 = wid * 4; t0 < nd; t0 += nw * 4) {
          const T* v0 = dvec(t0);
          const T* v1 = dvec(min(t0 + 1, nd - 1));
          const T* v2 = dvec(min(t0 + 2, nd - 1));
          const T* v3 = dvec(min(t0 + 3, nd - 1));
          double sv[4] = {0, 0, 0, 0};
          using VT = typename V16<T>::type;
          constexpr int VW = V16<T>::n;
          if (n % VW == 0) {
            // 16-byte LDS reads: a lane takes VW consecutive orbitals per step
            for (int jv = lane; jv < n / VW; jv += 64) {
              const VT p4 = reinterpret_cast<const VT*>(sPsi)[jv];
              const VT a0 = reinterpret_cast<const VT*>(v0)[jv], a1 = reinterpret_cast<const VT*>(v1)[jv];
              const VT a2 = reinterpret_cast<const VT*>(v2)[jv], a3 = reinterpret_cast<const VT*>(v3)[jv];
#pragma unroll
              for (int q = 0; q < VW; ++q) {
                const double pv = (double)p4[q];
                sv[0] += (double)a0[q] * pv;
                sv[1] += (double)a1[q] * pv;
                sv[2] += (double)a2[q] * pv;
                sv[3] += (double)a3[q] * pv;
              }
            }
          } else {
            for (int j = lane; j < n; j += 64) {
              const double pv = (double)sPsi[j];
              sv[0] += (double)v0[j] * pv;
              sv[1] += (double)v1[j] * pv;
              sv[2] += (double)v2[j] * pv;
              sv[3] += (double)v3[j] * pv;
            }
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
        if constexpr (ST) {
          // the capacitance inverse grows by one row and column with every wave: acc_s = Binv[s] . p
          // (s < c) and y . p (s = c), a wave per row; then Binv[s][r] -= y_r yi acc_s,
          // Binv[s][c] = yi acc_s, Binv[c][r] = -yi y_r, Binv[c][c] = yi, yi = 1 / (y . p - p_c)
          for (int s2 = wid; s2 <= c; s2 += nw) {
            const double* br = s2 < c ? binv + (size_t)s2 * K : sSy;
            double t = 0;
            for (int r = lane; r < c; r += 64) t += br[r] * sP[r];
            t = wave_sum(t);
            if (lane == 0) sSg[s2] = t;
          }
          // the move's stale inverse row becomes the slot's V row c
          for (int j = tid; j < n; j += nt) Vb[(size_t)c * n + j] = Ainv[(size_t)row * n + j];
          __syncthreads();
          const double yi = 1.0 / (sSg[c] - sP[c]);
          for (int i = tid; i < c * c; i += nt) {
            const int s2 = i / c, r = i - s2 * c;
            binv[(size_t)s2 * K + r] -= sSy[r] * (yi * sSg[s2]);
          }
          for (int i = tid; i < c; i += nt) {
            binv[(size_t)i * K + c] = yi * sSg[i];
            binv[(size_t)c * K + i] = -sSy[i] * yi;
          }
          if (tid == 0) {
            binv[(size_t)c * K + c] = yi;
            sRows[c] = row;
          }
          __syncthreads();
        } else if (wid == nw - 1) {
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
      if constexpr (J3) {
        // J3 (ThreeBodyJastrow::acceptMove, ThreeBodyJastrow.h:302-361): the pair terms at the old
        // position, then every other electron's sums moved by new - old (each wave applies its
        // own electrons' slots), the moving electron's sums set to phase 2's and its ion mask to
        // the proposal's near ions
        T jso[5];
        const T* reco = sJ3ne + (size_t)e * kJ3Rec;
        const T* recn = sJ3np + (size_t)e * kJ3Rec;
        const int ns3o = j3_phase<T, JPT>(S, C, sR, sMask, reco, iel, x0, y0, z0, tid, nt, lane, sJkO, sJvO, jso,
                                          B.flags);
        const int nIn = min((int)recn[0], kJ3Near), nIo = min((int)reco[0], kJ3Near);
        j3_apply<T>(sJ3, N, sJk, sJv, ns3, nIn, T(1), lane);
        j3_apply<T>(sJ3, N, sJkO, sJvO, ns3o, nIo, T(-1), lane);
        if (tid == 0)
#pragma unroll
          for (int c2 = 0; c2 < 5; ++c2) sJ3[(size_t)c2 * N + iel] = (T)cur3[c2];
        if (tid < S.mw) {
          unsigned long long m = 0;
          for (int q = 0; q < nIn; ++q) {
            const int I = (int)recn[1 + 5 * q];
            if ((I >> 6) == tid) m |= 1ull << (I & 63);
          }
          sMask[(size_t)iel * S.mw + tid] = m;
        }
      }
    }
    // with K = 1 the next move reads rows of Ainv that other waves just wrote
    if (!KD) __syncthreads();
    else lds_sync();
    if (a.cadence > 0 && accept && ++since >= a.cadence) {
      stopped = true;   // the recompute follows this launch
      Eend = e + 1;
    }
  }

  // ---- chunk end: the accepted rows' Slater matrix, gradient and Laplacian rows
  // (DiracDeterminant.cpp:135-141) copied from the proposals' SPO rows, the stale inverse rows
  // for the flush, the new J1 state
  if constexpr (J3)   // the accepted electrons' near-ion records are their proposals'
    for (int i = tid; i < ca * kJ3Rec; i += nt) {
      const int d = i / kJ3Rec, q = i - d * kJ3Rec;
      B.j3near[((size_t)w * N + cb + sAl[d]) * kJ3Rec + q] = sJ3np[(size_t)sAl[d] * kJ3Rec + q];
    }
  for (int d = tid; d < ca; d += nt) {
    const int ie = cb + sAl[d];
    J1[ie] = sJ1n[d * 5 + 0];
    J1[N + ie] = sJ1n[d * 5 + 1];
    J1[2 * N + ie] = sJ1n[d * 5 + 2];
    J1[3 * N + ie] = sJ1n[d * 5 + 3];
    J1[4 * N + ie] = sJ1n[d * 5 + 4];
  }
  if (KD && !ST && !a.inflush)
    for (int i = tid; i < ca * n; i += nt) {
      const int d = i / n, j = i - d * n;
      Vb[(size_t)d * n + j] = sAc[(size_t)sAl[d] * sa + j];
    }
  {
    using VT = typename V16<T>::type;
    constexpr int VW = V16<T>::n;
    const bool vec = n % VW == 0 && S.npad % VW == 0;
    const int nv = vec ? n / VW : n;   // vectors (or scalars) per row
    const int tot = ca * 5 * nv;
    // The new Slater rows also go to LDS for the kernel's own flush (FP32): rows 0-15 to sU,
    // rows 16 on (k >= 32, while they fit) to the chunk's rejected moves' slots of sAc, whose
    // stale rows the flush does not read; their list goes to sRej (sInv's LDS, free here).
    float* sU = (float*)(smem + L.oQ);
    int* sRej = (int*)(smem + L.oInv);
    const bool stage_u = sizeof(T) == 4 && KD && a.inflush && a.flush_u > 0 &&
                         (ca <= kFlushOwnU || (K >= 32 && ca <= 32 && 2 * ca <= kFlushOwnU + E));
    if (stage_u && ca > kFlushOwnU) {
      if (wid == 0) {
        unsigned long long m = 0;
        for (int d = 0; d < ca; ++d) m |= 1ull << sAl[d];
        const bool rej = lane < E && !((m >> lane) & 1ull);
        const unsigned long long rb = __ballot(rej);
        const int rank = __popcll(rb & ((1ull << lane) - 1ull));
        if (rej) sRej[rank] = lane;
      }
      __syncthreads();
    }
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
          if constexpr (sizeof(T) == 4) {
            if (stage_u && o == 0) {
              float* ur = d < kFlushOwnU ? sU + (size_t)d * (n + 4) : (float*)sAc + (size_t)sRej[d - kFlushOwnU] * sa;
              if (vec) *reinterpret_cast<VT*>(ur + j * VW) = v[k];
              else ur[j] = sv[k];
            }
          }
        }
      }
    }
  }
  // the FP32 form's FP64 Slater rows of the accepted moves, from their proposals' FP64 values
  if (B.A64 != nullptr) {
    const double* s64 = B.spo64 + (((size_t)w * 2 + a.det) * n + r0) * S.npad;
    double* A64 = B.A64 + moff;
    for (int i = tid; i < ca * n; i += nt) {
      const int d = i / n, j = i - d * n;
      A64[(size_t)(r0 + sAl[d]) * n + j] = s64[(size_t)sAl[d] * S.npad + j];
    }
  }
  // the chunk's delayed-update state for a flush launch (none when the kernel applies it)
  if (tid == 0) B.ndelay[slot] = KD && !a.inflush ? c : 0;
  if (KD && !ST && !a.inflush)
    for (int i = tid; i < c * K; i += nt) B.Binv[slot * K * K + i] = sBinv[(i / K) * KP + i % K];
  if (ST && L.binv_lds)
    for (int i = tid; i < c * c; i += nt) {
      const int s2 = i / c, r = i - s2 * c;
      gBinv[(size_t)s2 * K + r] = sBinv[(size_t)s2 * K + r];
    }
  if (KD && !a.inflush)
    for (int i = tid; i < c; i += nt) B.rows[slot * K + i] = sRows[i];
  if constexpr (KD && sizeof(T) == 4) {
    // a stopped walker's inverse is rebuilt by the recompute instead
    if (a.inflush && c > 0 && !stopped && (cb + E < a.e_end || a.inflush == 2)) {
      // the chunk's delayed updates reach this walker's whole inverse before the next chunk
      __syncthreads();
      float* sU = (float*)(smem + L.oQ);
      const int* sRej = (const int*)(smem + L.oInv);
      float* scratch = sU + (size_t)kFlushOwnU * (n + 4);
      // the same test as the staging of the new Slater rows at the chunk's end
      if (a.flush_u > 0 && (c <= kFlushOwnU || (K >= 32 && c <= 32 && 2 * c <= kFlushOwnU + E)))
        flush_own_tile(n, K, KP, c, sRows, sAl, sBinv, (const float*)sAc, sa, sU, sRej, (float*)Ainv, scratch);
      else
        flush_own(n, K, KP, c, sRows, sAl, sBinv, (const float*)sAc, sa, (const float*)Amat, (float*)Ainv, scratch);
    }
  }
  cbn = cb + Eend;
  }   // chunks
  if (a.cadence > 0 && tid == 0) {
    cadw[0] = cbn - a.e_begin;
    cadw[1] = stopped ? 0 : since;
    a.cad_flag[w] = stopped ? 1 : 0;
  }

  // ---- walker state back to HBM
  for (int i = tid; i < 3 * N; i += nt) B.R[(size_t)w * 3 * N + i] = sR[i];
  for (int i = tid; i < 5 * N; i += nt) J2[i] = sJ2[i];
  if constexpr (J3) {
    T* g3 = B.J3 + (size_t)w * 5 * N;
    for (int i = tid; i < 5 * N; i += nt) g3[i] = sJ3[i];
    unsigned long long* gm = B.j3mask + (size_t)w * N * S.mw;
    for (int i = tid; i < N * S.mw; i += nt) gm[i] = sMask[i];
    if (tid == 0) B.jlog[w * 3 + 2] += sAcc[aJl3];
  }
  if (tid == 0) {
    B.logdet[w * 2 + a.det] += sAcc[aLogd] + log(sAcc[aPdet]);
    B.phase[w * 2 + a.det] += sAcc[aPh];
    B.jlog[w * 3 + 0] += sAcc[aJl1];
    B.jlog[w * 3 + 1] += sAcc[aJl2];
    B.nacc[w] += (unsigned long long)sAcc[aNacc];
    if (B.narb != nullptr) B.narb[w] += (unsigned long long)sAcc[aNarb];
  }
}

template <typename T, int JPT, bool KD, bool ARB, bool J3, bool ST>
__global__ __launch_bounds__(kMoveThreads) void k_moves(SysP<T> S, BatchP<T> B, MoveArgs a) {
  moves_body<T, JPT, KD, ARB, J3, ST>(S, B, a, blockIdx.x);
}

// ---- walkers of several batches (systems of different sizes) in one launch: batch g owns the
// workgroups [first[g], first[g+1]); the workgroup size and LDS are the largest any batch needs
// (a batch of a smaller system leaves the extra lanes idle).  Delayed updates with the flush
// inside the kernel (FP32), each batch at its own delay rank.
constexpr int kMaxGroups = 4;
template <typename T>
struct MoveGroup {
  int ng;
  int first[kMaxGroups + 1];
  SysP<T> S[kMaxGroups];
  BatchP<T> B[kMaxGroups];
  MoveArgs a[kMaxGroups];
};

template <typename T, int JPT>
__global__ __launch_bounds__(kMoveThreads) void k_moves_group(MoveGroup<T> G) {
  int g = 0;
  while (g + 1 < G.ng && (int)blockIdx.x >= G.first[g + 1]) ++g;
  moves_body<T, JPT, true, false, false>(G.S[g], G.B[g], G.a[g], (int)blockIdx.x - G.first[g]);
}

}  // namespace qmcx
