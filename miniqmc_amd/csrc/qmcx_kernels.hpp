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
//   logdet, phase double [W][2]; jlog double [W][3] (J1, J2, J3)
#pragma once
#include "qmcx_device.hpp"
#include "qmcx_j3.hpp"

namespace qmcx {

template <typename T>
struct SysP {
  const Cell<T>* cell;        // device copy of the minimum-image cell
  const double* spoG;         // [9] inverse of the SPO lattice
  int ng[3];
  int N, n, M, nspecies;
  int npad;
  long long xs, ys, zs;
  const T* coefs;
  const double* ions;         // [3][M]
  const int* species;         // [M]
  const Functor<T>* fn;       // [kMaxSpecies] J1 then [2] J2
  const Functor<double>* fn64;  // the same functors in FP64 (FP64 arbitration of the FP32 form)
  const Cell<double>* cell64;   // the minimum-image cell in FP64
  // three-body Jastrow: functor of (ion species, spin pair) at [2 * species + pair] (L = 0 for a
  // species without one), or null; mw = 64-bit words of an electron's ion mask, ceil(M / 64)
  const Poly3<T>* j3;
  const Poly3<double>* j3_64;   // the same functors in FP64 (FP64 re-evaluation of the FP32 form)
  int mw;
};

template <typename T>
struct BatchP {
  int W, K;
  int nch;         // delayed-update state slots per walker and determinant (one per chunk)
  double* R;
  T* A;
  T* Ainv;
  T* dpsiM;
  T* d2psiM;
  T* J2;
  T* J1;
  // delayed-update state of chunk j of determinant det of walker w, slot (w*2 + det)*nch + j:
  // V rows Vbuf[slot][K][n], Binv[slot][K][K], rows[slot][K], ndelay[slot]
  T* Vbuf;
  double* Binv;    // capacitance-matrix inverse kept in FP64 in both precisions
  int* rows;
  int* ndelay;
  double* logdet;
  double* phase;
  double* jlog;
  unsigned long long* nacc;
  double* trace;   // [W][N][9] or null
  // device sweep counters: [0] the sweep the moves are in (advanced on the batch stream after
  // evaluateGL), [1] the first sweep of the current run, [2] the sweep of the proposals being
  // evaluated (advanced on the proposals stream), [3] the history slot of the latest evaluateGL
  int* sweep;
  // FP32 form only (null in FP64): the Slater rows summed in FP64 from the same coefficients,
  // A64 [W][2][n][n] and at the sweep's proposals spo64 [W][2][n][npad]
  double* A64;
  double* spo64;
  unsigned long long* narb;   // [W] Metropolis decisions settled by the FP64 re-evaluation
  // three-body Jastrow (with S.j3): sums J3 [W][5][N] (U, grad U xyz, lap U), ion masks j3mask
  // [W][N][mw] (bit I: the electron is within L_I of ion I), near-ion records of the current
  // positions j3near [W][N][kJ3Rec] and of the sweep's proposals j3prop [W][2][n][kJ3Rec]
  T* J3;
  unsigned long long* j3mask;
  T* j3near;
  T* j3prop;
  int* flags;      // [1] error bits set by kernels (kFlag*)
};

// kernel error bits in BatchP::flags
constexpr int kFlagJ3Near = 1;    // an electron position with more than kJ3Near ions inside the J3 cutoff
constexpr int kFlagJ3Pairs = 2;   // a wave's J3 pairs of one move phase beyond its slots

// Proposal record of one electron, [W][2][n][kProp] doubles written by k_spo_proposals:
// J1 at the proposal (sum u, sum (d2u + 2 du/r), gradient), the displacement, the
// acceptance uniform.
constexpr int kProp = 10;
constexpr int kPropJ1 = 0, kPropDispl = 5, kPropU = 8;

template <typename T>
__host__ __device__ inline size_t state_slot(const BatchP<T>& B, int w, int det, int j) {
  return ((size_t)w * 2 + det) * B.nch + j;
}

struct MoveArgs {
  int e_begin, e_end, det, mode, slices;
  // The launch moves electrons [e_begin, e_end) in nchunks chunks of chunk_len; chunk is the
  // index of its first chunk in the determinant's sweep.  inflush: the kernel applies each
  // chunk's delayed updates to the walker's whole inverse before the next chunk (1), and after
  // the last chunk too (2); 0: a flush launch follows each single-chunk launch.  The kernel's
  // flush takes the new Slater rows from LDS when at most flush_u moves of the chunk were
  // accepted, else from HBM.
  int chunk, chunk_len, nchunks, inflush, flush_u;
  // streamed delayed updates: the launch continues from the delayed rows pending in its state slot
  // (flex_* calls keep them pending until the delay rank is reached or completeUpdates)
  int resume;
  // FP64 recompute cadence (cadence > 0 accepted rows): a walker's moves of the determinant start
  // at its own position cad[w][det][0] and stop after the move that brings its accepted rows since
  // the last recompute (cad[w][det][1]) to cadence; it then sets cad_flag[w] for the recompute
  int cadence;
  int* cad;
  int* cad_flag;
  const int* decide;     // [W] caller accept decisions (flex_acceptrestoreMove) or null
  // FP32 form, Metropolis: a move whose uniform u lies within band * ratio^2 of ratio^2 is
  // decided on its FP64 ratio (0: off)
  double band;
  const void* spo;       // [W][2][n][5][npad] T: SPO rows at the proposals (k_spo_proposals)
  const double* prop;    // [W][2][n][kProp] proposal records (k_spo_proposals)
};

// ------------------------------------------------------------------ spline gather
// One workgroup evaluates NO outputs (1: v; 5: v, gx, gy, gz, hxx; 10: full VGH) of
// every orbital at one point.  Lane (q, s): q = 16-byte orbital group, s =
// stencil slice owning 16/S of the (i,j) columns of the 4x4x4 stencil; each
// lane issues 4*16/S coalesced 16-byte loads along the orbital axis.  Slice
// partials go to LDS `part` [S][NO][npad] and are summed by `spline_finish`.
// weight, first and second derivative of cubic B-spline knot i (0..3) at fraction t
template <typename T>
__device__ inline void knot_weight(int i, T t, T& w, T& dw, T& d2w) {
  const T s6 = T(1) / T(6);
  // w = ((p3 t + p2) t + p1) t + p0, dw = (q2 t + q1) t + q0, d2w = r1 t + r0
  const T p3 = i == 0 ? T(-1) : i == 1 ? T(3) : i == 2 ? T(-3) : T(1);
  const T p2 = i == 0 ? T(3) : i == 1 ? T(-6) : i == 2 ? T(3) : T(0);
  const T p1 = i == 0 ? T(-3) : i == 1 ? T(0) : i == 2 ? T(3) : T(0);
  const T p0 = i == 0 ? T(1) : i == 1 ? T(4) : i == 2 ? T(1) : T(0);
  w = (((p3 * t + p2) * t + p1) * t + p0) * s6;
  dw = ((T(3) * p3 * t + T(2) * p2) * t + p1) * s6;
  d2w = (T(6) * p3 * t + T(2) * p2) * s6;
}

// V64: the value is also summed in FP64 (weights from the FP64 fraction, the same loaded
// coefficients) into part64 [S][npad]: the FP64 Slater rows of the FP32 form.
// The partial sums of lane tid (orbital quad q, stencil slice s) in registers: acc[o][e] for
// orbital q*VW + e, and with V64 the FP64 values acc64.  False for the lanes past the slices.
// UNI (one slice only): the point is the same on every lane of the wave, so its grid cell moves
// to scalar registers and every load takes a scalar address plus the lane's orbital offset.
template <typename T, int NO, int PER, bool V64, bool UNI = false>
__device__ inline bool spline_gather_regs(const SysP<T>& S, const double* spoG, double x, double y, double z, int tid,
                                          T (&acc)[NO][V16<T>::n], double (&acc64)[V16<T>::n], int& q, int& s) {
  using VT = typename V16<T>::type;
  constexpr int VW = V16<T>::n;
  constexpr int SL = 16 / PER;
  // stencil columns with loads in flight per batch: up to four in FP32 (two when the FP64 values
  // are summed as well), two in FP64
  constexpr int G = (sizeof(T) == 8 || V64) ? (PER > 2 ? 2 : PER) : (PER > 4 ? 4 : PER);
  const int nq = S.npad / VW;
  q = tid % nq;
  s = tid / nq;
  if (s >= SL) return false;
  static_assert(!UNI || SL == 1, "a wave-uniform point with one slice only");
  if constexpr (UNI) s = 0;
  int ix[3];
  double tt[3];
  spline_locate(spoG, S.ng, x, y, z, ix, tt);
  if constexpr (UNI) {
#pragma unroll
    for (int j = 0; j < 3; ++j) ix[j] = __builtin_amdgcn_readfirstlane(ix[j]);
  }
  // base0 (the cell's first coefficient) is wave-uniform with UNI; the lane's quad is added at the load
  const T* base0 = S.coefs + (size_t)ix[0] * S.xs + (size_t)ix[1] * S.ys + (size_t)ix[2] * S.zs;
  const unsigned lane_off = (unsigned)q * VW;
  const size_t zs = S.zs;
  T c[4], dc[4], d2c[4];
  bspline_weights<T>((T)tt[2], c, dc, d2c);
  const T tx = (T)tt[0], ty = (T)tt[1];
#pragma unroll
  for (int o = 0; o < NO; ++o)
#pragma unroll
    for (int e = 0; e < VW; ++e) acc[o][e] = 0;
  double c64[4];
  if constexpr (V64) {
    double dc64[4], d2c64[4];
    bspline_weights<double>(tt[2], c64, dc64, d2c64);
#pragma unroll
    for (int e = 0; e < VW; ++e) acc64[e] = 0.0;
  }
#pragma unroll
  for (int g0 = 0; g0 < PER; g0 += G) {
  VT v4[G][4];
#pragma unroll
  for (int pp = 0; pp < G; ++pp) {
    const int p = s * PER + g0 + pp;
    const T* src = base0 + (size_t)(p >> 2) * S.xs + (size_t)(p & 3) * S.ys;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      v4[pp][k] = __builtin_nontemporal_load(reinterpret_cast<const VT*>(src + k * zs + lane_off));
  }
#pragma unroll
  for (int pp = 0; pp < G; ++pp) {
    const int p = s * PER + g0 + pp;
    T ai, dai, d2ai, bj, dbj, d2bj;
    knot_weight<T>(p >> 2, tx, ai, dai, d2ai);
    knot_weight<T>(p & 3, ty, bj, dbj, d2bj);
    const T w00 = ai * bj, w10 = dai * bj, w01 = ai * dbj, w20 = d2ai * bj;
    const T w11 = dai * dbj, w02 = ai * d2bj;
    double w00d = 0.0;
    if constexpr (V64) {
      double a64, da64, d2a64, b64, db64, d2b64;
      knot_weight<double>(p >> 2, tt[0], a64, da64, d2a64);
      knot_weight<double>(p & 3, tt[1], b64, db64, d2b64);
      w00d = a64 * b64;
    }
#pragma unroll
    for (int e = 0; e < VW; ++e) {
      const T q0 = v4[pp][0][e], q1 = v4[pp][1][e], q2 = v4[pp][2][e], q3 = v4[pp][3][e];
      const T s0 = c[0] * q0 + c[1] * q1 + c[2] * q2 + c[3] * q3;
      acc[0][e] += w00 * s0;
      if constexpr (V64)
        acc64[e] += w00d * (c64[0] * (double)q0 + c64[1] * (double)q1 + c64[2] * (double)q2 + c64[3] * (double)q3);
      if constexpr (NO >= 5) {
        const T s1 = dc[0] * q0 + dc[1] * q1 + dc[2] * q2 + dc[3] * q3;
        acc[1][e] += w10 * s0;
        acc[2][e] += w01 * s0;
        acc[3][e] += w00 * s1;
        acc[4][e] += w20 * s0;
        if constexpr (NO == 10) {
          const T s2 = d2c[0] * q0 + d2c[1] * q1 + d2c[2] * q2 + d2c[3] * q3;
          acc[5][e] += w11 * s0;
          acc[6][e] += w10 * s1;
          acc[7][e] += w02 * s0;
          acc[8][e] += w01 * s1;
          acc[9][e] += w00 * s2;
        }
      }
    }
  }
  }
  return true;
}

template <typename T, int NO, int PER, bool V64 = false>
__device__ inline void spline_gather(const SysP<T>& S, const double* spoG, double x, double y, double z, T* part,
                                     int tid, double* part64 = nullptr) {
  using VT = typename V16<T>::type;
  constexpr int VW = V16<T>::n;
  T acc[NO][VW];
  double acc64[VW];
  int q, s;
  if (!spline_gather_regs<T, NO, PER, V64>(S, spoG, x, y, z, tid, acc, acc64, q, s)) return;
  // one 16-byte LDS store per output (npad is a multiple of VW)
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    VT v;
#pragma unroll
    for (int e = 0; e < VW; ++e) v[e] = acc[o][e];
    *reinterpret_cast<VT*>(part + ((size_t)s * NO + o) * S.npad + q * VW) = v;
  }
  if constexpr (V64)
#pragma unroll
    for (int e = 0; e < VW; ++e) part64[(size_t)s * S.npad + q * VW + e] = acc64[e];
}

template <typename T, int NO, bool V64 = false>
__device__ inline void spline_gather_rt(const SysP<T>& S, double x, double y, double z, int slices, T* part,
                                        double* part64 = nullptr) {
  switch (slices) {
    case 16: spline_gather<T, NO, 1, V64>(S, S.spoG, x, y, z, part, threadIdx.x, part64); break;
    case 8: spline_gather<T, NO, 2, V64>(S, S.spoG, x, y, z, part, threadIdx.x, part64); break;
    case 4: spline_gather<T, NO, 4, V64>(S, S.spoG, x, y, z, part, threadIdx.x, part64); break;
    case 2: spline_gather<T, NO, 8, V64>(S, S.spoG, x, y, z, part, threadIdx.x, part64); break;
    default: spline_gather<T, NO, 16, V64>(S, S.spoG, x, y, z, part, threadIdx.x, part64); break;
  }
}

// the FP64 value row: slice partials summed in order, out[j] for j < norb
template <typename T>
__device__ inline void spline_finish64(const SysP<T>& S, int slices, const double* part64, double* out, int norb) {
  for (int j = threadIdx.x; j < norb; j += blockDim.x) {
    double sum = 0;
    for (int s = 0; s < slices; ++s) sum += part64[(size_t)s * S.npad + j];
    out[j] = sum;
  }
}

// delta_inv powers of output o: 1, gx, gy, gz, gx^2, gx gy, gx gz, gy^2, gy gz, gz^2
template <typename T>
__device__ inline T vgh_scale(int o, T gx, T gy, T gz) {
  const T f1 = (o == 1 || o == 4 || o == 5 || o == 6) ? gx : (o == 2 || o == 7 || o == 8) ? gy : (o == 3 || o == 9) ? gz : T(1);
  const T f2 = (o == 4) ? gx : (o == 5 || o == 7) ? gy : (o == 6 || o == 8 || o == 9) ? gz : T(1);
  return f1 * f2;
}

// Sum the slice partials and scale by the grid's delta_inv powers.
// out(o, j) for j < norb is written at out[o*ostride + j].
template <typename T, int NO, typename TO>
__device__ inline void spline_finish(const SysP<T>& S, int slices, const T* part, TO* out, int ostride, int norb) {
  const T gx = (T)S.ng[0], gy = (T)S.ng[1], gz = (T)S.ng[2];
  for (int idx = threadIdx.x; idx < NO * norb; idx += blockDim.x) {
    const int o = idx / norb, j = idx - o * norb;
    T sum = 0;
    for (int s = 0; s < slices; ++s) sum += part[((size_t)s * NO + o) * S.npad + j];
    out[(size_t)o * ostride + j] = (TO)(sum * vgh_scale<T>(o, gx, gy, gz));
  }
}

// One slice (each lane sums all 64 stencil points of its orbital quad): the scaled outputs go from
// registers to out[o*npad + j] with 16-byte stores, no LDS.  The padded orbitals (zero
// coefficients) are written too.  Equal to spline_gather + spline_finish with one slice.  The
// point must be the same on every lane of the workgroup.
// With V64 the FP64 values go to out64[j] as well.
template <typename T, int NO, bool V64 = false>
__device__ inline void spline_gather_direct(const SysP<T>& S, double x, double y, double z, T* out,
                                            double* out64 = nullptr) {
  using VT = typename V16<T>::type;
  constexpr int VW = V16<T>::n;
  T acc[NO][VW];
  double acc64[VW];
  int q, s;
  if (!spline_gather_regs<T, NO, 16, V64, true>(S, S.spoG, x, y, z, threadIdx.x, acc, acc64, q, s)) return;
  if constexpr (V64) {
#pragma unroll
    for (int e = 0; e < VW; e += 2) {
      f64x2 v;
      v[0] = 0.0 + acc64[e];
      v[1] = 0.0 + acc64[e + 1];
      *reinterpret_cast<f64x2*>(out64 + q * VW + e) = v;
    }
  }
  const T gx = (T)S.ng[0], gy = (T)S.ng[1], gz = (T)S.ng[2];
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    const T f = vgh_scale<T>(o, gx, gy, gz);
    VT v;
#pragma unroll
    for (int e = 0; e < VW; ++e) v[e] = (T(0) + acc[o][e]) * f;
    *reinterpret_cast<VT*>(out + (size_t)o * S.npad + q * VW) = v;
  }
}

}  // namespace qmcx

#include "qmcx_moves.hpp"

namespace qmcx {

// ------------------------------------------------------------------ delayed flush
// updateInvMat for K > 1 (DelayedUpdate.h:176-193), rows [m0, m0+FB) of one walker:
//   T  = Ainv U^T - E          (U rows = new Slater rows, E_{m,d} = [m == rows_d])
//   T' = T Binv
//   Ainv -= T' V               (V rows = stale inverse rows)
// A launch covers rows [lo, hi) and applies the slots holding at least thr delayed rows (thr > 1:
// the walkers whose pending rows reached the delay rank, DelayedUpdate.h:151-152).
struct FlushRows {
  int lo, hi;
  int thr = 1;
  __device__ bool writes(int m) const { return m < hi; }
};
constexpr int kFlushRows = 16;
// After a flush launch over state slot 0 of determinant det with threshold thr: the applied slots
// hold no delayed rows (a thread per walker).
template <typename T>
__global__ void k_reset_delay(BatchP<T> B, int det, int thr) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= B.W) return;
  int* nd = B.ndelay + state_slot(B, w, det, 0);
  const int c = *nd;
  if (c > 0 && c >= thr) *nd = 0;
}
// FP64 recompute cadence: a determinant's sweep starts every walker at its first electron
__global__ void k_cadence_begin(int* cad, int W, int det) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < W) cad[((size_t)w * 2 + det) * 2] = 0;
}
// an FP64 inversion of determinant det of every walker: no accepted rows since (det < 0: both)
__global__ void k_cadence_reset(int* cad, int W, int det) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  for (int d = 0; d < 2; ++d)
    if (det < 0 || d == det) cad[((size_t)w * 2 + d) * 2 + 1] = 0;
}
__global__ void k_cadence_clear(int* p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}
// one state slot j per launch
template <typename T>
__global__ __launch_bounds__(256) void k_flush_generic(SysP<T> S, BatchP<T> B, int det, int j, FlushRows FR) {
  const int w = blockIdx.y, m0 = FR.lo + blockIdx.x * kFlushRows;
  const size_t sl = state_slot(B, w, det, j);
  const int c = B.ndelay[sl];
  if (c == 0 || c < FR.thr) return;
  const int n = S.n, K = B.K, tid = threadIdx.x, nt = blockDim.x;
  extern __shared__ __align__(16) unsigned char smem[];
  T* sT = (T*)smem;                     // [FB][K]
  T* sT2 = sT + kFlushRows * K;         // [FB][K]
  int* sRows = (int*)(sT2 + kFlushRows * K);
  const size_t moff = ((size_t)w * 2 + det) * n * n;
  T* Ainv = B.Ainv + moff;
  const T* Amat = B.A + moff;
  const T* Vb = B.Vbuf + sl * K * n;
  const double* Bi = B.Binv + sl * K * K;
  for (int i = tid; i < c; i += nt) sRows[i] = B.rows[sl * K + i];
  __syncthreads();
  const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  for (int idx = wid; idx < kFlushRows * c; idx += nw) {
    const int m = idx / c, d = idx - m * c;
    T s = 0;
    if (m0 + m < FR.hi) {
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
    double s = 0;
    for (int e = 0; e < c; ++e) s += (double)sT[m * K + e] * Bi[e * K + d];
    sT2[m * K + d] = (T)s;
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
      if (FR.writes(m0 + m)) Ainv[(size_t)(m0 + m) * n + j] -= acc[m];
  }
}

// FP32 flush on the matrix cores (v_mfma_f32_16x16x4f32), 16 rows per wave, kFlushWaves
// waves per workgroup (one workgroup per CU covers a walker's rows in one or two launches'
// worth of workgroups).  16x16x4 operand layout: lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator register r of lane l is D[4*(l/16)+r][l%16].
//
// A wave's 16 rows stay in registers between the two products, in the one layout both use:
// lane (i = l%16, g = l/16) holds row r0+i, columns 16ct + 4g .. 16ct + 4g + 3 of every
// 16-column tile ct (a float4 per tile; one HBM read of the rows, 64 contiguous bytes per
// row per load).  Both products are computed transposed:
//   GEMM 1:  T^T = U Ainv_rows^T     A = U (16 delayed rows x 4 orbitals), B = the rows
//            (orbital 16ct + 4g + r of MFMA r is the k index g), D[d][i] = T[i][d]
//   GEMM 2:  Ainv_rows^T += V^T T2^T  A = V^T (16 orbitals x 4 delayed rows), B = T2^T,
//            accumulators started from the rows themselves
// V (the stale inverse rows) and U (the new Slater rows) are staged in LDS per workgroup when
// the delay rank allows.
constexpr int kFlushWaves = 12;     // waves per workgroup, 16 rows each
constexpr int kFlushBM = 16 * kFlushWaves;   // rows per workgroup
constexpr int kFlushTiles = 24;     // 16-column tiles held in registers: n <= 384
constexpr int kFlushBinvLds = 64;   // Binv staged in LDS up to this delay rank
constexpr int kFlushVLds = 32;      // V and U staged in LDS up to this delay rank
__host__ __device__ inline bool flush_mfma_fits(int n) { return n % 16 == 0 && n <= 16 * kFlushTiles; }
__host__ __device__ inline int flush_v_stride(int n) { return n + 16; }   // LDS row stride of V
// LDS of k_flush_mfma: per wave T and T2 (16 x cp floats each; past kFlushBinvLds T2 is written
// over T), Binv (K x K doubles when K <= kFlushBinvLds), V and U (cp x (n + 16) floats each when
// K <= kFlushVLds), the row list
__host__ __device__ inline int flush_t_tiles(int K) { return K <= kFlushBinvLds ? 2 : 1; }
__host__ __device__ inline size_t flush_mfma_lds(int n, int K) {
  const int cp = (K + 15) & ~15;
  const size_t kb = K <= kFlushBinvLds ? (size_t)(K * K + 1) & ~(size_t)1 : 0;   // keeps V 16-byte aligned
  const size_t vb = K <= kFlushVLds ? (size_t)2 * cp * flush_v_stride(n) : 0;
  return sizeof(float) * kFlushWaves * flush_t_tiles(K) * 16 * cp + sizeof(double) * kb + sizeof(float) * vb +
         sizeof(int) * K;
}
// the largest delay rank whose T2 is held in registers while it is written over T
constexpr int kFlushT2Regs = 128;

// State slots [j0, j1) of determinant det are applied in order to the rows held in registers.
// OVER (delay ranks past kFlushBinvLds): T2 is written over T.
template <bool OVER>
__global__ __launch_bounds__(64 * kFlushWaves) void k_flush_mfma(SysP<float> S, BatchP<float> B, int det, int j0,
                                                                  int j1, FlushRows FR) {
  const int w = blockIdx.y;
  const int n = S.n, K = B.K, tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, li = lane & 15, g = lane >> 4;
  const int nct = n / 16;                  // column tiles (n % 16 == 0, nct <= kFlushTiles)
  const int r0 = FR.lo + blockIdx.x * kFlushBM + wid * 16;   // this wave's first row
  const int KC = (K + 15) & ~15;           // LDS sizing uses the delay rank, not c
  // nothing to apply for this walker: no Ainv traffic
  bool any = false;
  for (int j = j0; j < j1; ++j) {
    const int cj = B.ndelay[state_slot(B, w, det, j)];
    any = any || (cj > 0 && cj >= FR.thr);
  }
  if (!any) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tt = OVER ? 1 : 2;
  float* sT = (float*)smem + wid * tt * 16 * KC;      // [16][cp]   T = Ainv U^T - E
  float* sT2 = sT + (tt - 1) * 16 * KC;                // [16][cp]  -T Binv (over T when tt = 1)
  const bool binv_lds = K <= kFlushBinvLds;
  const bool v_lds = K <= kFlushVLds;
  double* sB = (double*)((float*)smem + kFlushWaves * tt * 16 * KC);       // [c][c]
  float* sV = (float*)(sB + (binv_lds ? (K * K + 1) & ~1 : 0));            // [cp][n + 16]
  float* sU = sV + (v_lds ? (size_t)KC * flush_v_stride(n) : 0);           // [cp][n + 16]
  int* sRows = (int*)(sU + (v_lds ? (size_t)KC * flush_v_stride(n) : 0));
  const size_t moff = ((size_t)w * 2 + det) * n * n;
  float* Ainv = B.Ainv + moff;
  const float* Amat = B.A + moff;
  const int vs = flush_v_stride(n);
  // this wave's rows into registers (row clamped into the range, zeroed after)
  const int row = r0 + li;
  const bool rok = row < FR.hi;
  const float* ar = Ainv + (size_t)(rok ? row : FR.lo) * n + 4 * g;
  f32x4_t a[kFlushTiles];
#pragma unroll
  for (int ct = 0; ct < kFlushTiles; ++ct)
    if (ct < nct) a[ct] = *reinterpret_cast<const f32x4_t*>(ar + ct * 16);
#pragma unroll
  for (int ct = 0; ct < kFlushTiles; ++ct)
    if (!rok) a[ct] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  bool changed = false;
  for (int j = j0; j < j1; ++j) {
    const size_t sl = state_slot(B, w, det, j);
    const int c = B.ndelay[sl];
    if (c == 0 || c < FR.thr) continue;
    changed = true;
    const int cp = (c + 15) & ~15;         // delayed rows padded to a 16-wide tile
    const float* Vb = B.Vbuf + sl * K * n;
    const double* Bi = B.Binv + sl * K * K;
    __syncthreads();                       // the previous slot's LDS reads are done
    for (int i = tid; i < c; i += nt) sRows[i] = B.rows[sl * K + i];
    if (binv_lds)
      for (int i = tid; i < c * c; i += nt) sB[i] = Bi[(i / c) * K + i % c];
    if (v_lds) {
      // V rows and the U rows they pair with, zero-padded to cp rows, float4 at a time; the
      // row list comes from HBM here too (it is being staged beside)
      const int q = n / 4;
      for (int i = tid; i < cp * q; i += nt) {
        const int d = i / q, jj = i - d * q;
        const int dc = d < c ? d : 0;
        f32x4_t v = *reinterpret_cast<const f32x4_t*>(Vb + (size_t)dc * n + 4 * jj);
        f32x4_t u = *reinterpret_cast<const f32x4_t*>(Amat + (size_t)B.rows[sl * K + dc] * n + 4 * jj);
        if (d >= c) {
          v = f32x4_t{0.f, 0.f, 0.f, 0.f};
          u = v;
        }
        *reinterpret_cast<f32x4_t*>(sV + (size_t)d * vs + 4 * jj) = v;
        *reinterpret_cast<f32x4_t*>(sU + (size_t)d * vs + 4 * jj) = u;
      }
    }
    __syncthreads();
    // ---- GEMM 1: T^T tile by tile of 16 delayed rows; two accumulators split the orbitals.
    // U rows from LDS (staged) or HBM, in separate loops so that each is its own kind of load.
    auto gemm1 = [&](const float* ur, int dt, int d) {
      f32x4_t acc0 = f32x4_t{0.f, 0.f, 0.f, 0.f}, acc1 = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ct = 0; ct < kFlushTiles; ++ct) {
        if (ct < nct) {
          f32x4_t u = *reinterpret_cast<const f32x4_t*>(ur + ct * 16);
          if (d >= c) u = f32x4_t{0.f, 0.f, 0.f, 0.f};
          f32x4_t& acc = (ct & 1) ? acc1 : acc0;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(u[0], a[ct][0], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(u[1], a[ct][1], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(u[2], a[ct][2], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(u[3], a[ct][3], acc, 0, 0, 0);
        }
      }
      // lane (i, g) register r: T[i][dt*16 + 4g + r]; minus E (row r0+i is delayed row dd)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int dd = dt * 16 + 4 * g + r;
        float v = acc0[r] + acc1[r];
        if (dd < c && sRows[dd] == r0 + li) v -= 1.f;
        sT[li * KC + dd] = dd < c ? v : 0.f;
      }
    };
    for (int dt = 0; dt < cp / 16; ++dt) {
      const int d = dt * 16 + li;
      if (v_lds) gemm1(sU + (size_t)d * vs + 4 * g, dt, d);
      else gemm1(Amat + (size_t)(d < c ? sRows[d] : sRows[0]) * n + 4 * g, dt, d);
    }
    wave_lds_fence();   // T and T2 are this wave's own
    // ---- T2 = -(T Binv), FP64 products of the small capacitance inverse
    if constexpr (OVER) {
      // over T: every lane's entries into registers first (cp <= kFlushT2Regs)
      float t2v[kFlushT2Regs / 4];
#pragma unroll
      for (int q = 0; q < kFlushT2Regs / 4; ++q) {
        const int idx = lane + 64 * q;
        double sum = 0;
        if (idx < 16 * cp) {
          const int i = idx / cp, dd = idx - i * cp;
          if (dd < c)
            for (int e = 0; e < c; ++e) sum += (double)sT[i * KC + e] * Bi[e * K + dd];
        }
        t2v[q] = (float)(-sum);
      }
      wave_lds_fence();
#pragma unroll
      for (int q = 0; q < kFlushT2Regs / 4; ++q) {
        const int idx = lane + 64 * q;
        if (idx < 16 * cp) {
          const int i = idx / cp, dd = idx - i * cp;
          sT[i * KC + dd] = t2v[q];
        }
      }
    } else
    for (int idx = lane; idx < 16 * cp; idx += 64) {
      const int i = idx / cp, dd = idx - i * cp;
      double sum = 0;
      if (dd < c) {
        if (binv_lds) {
          for (int e = 0; e < c; ++e) sum += (double)sT[i * KC + e] * sB[e * c + dd];
        } else {
#pragma unroll 8
          for (int e = 0; e < c; ++e) sum += (double)sT[i * KC + e] * Bi[e * K + dd];
        }
      }
      sT2[i * KC + dd] = (float)(-sum);
    }
    wave_lds_fence();
    // ---- GEMM 2: rows^T += V^T T2^T over the delayed rows, four at a time
    for (int d0 = 0; d0 < cp; d0 += 4) {
      const int dk = d0 + g;
      const float bv = sT2[li * KC + dk];
      const float* vr = v_lds ? sV + (size_t)dk * vs + li : Vb + (size_t)(dk < c ? dk : 0) * n + li;
#pragma unroll
      for (int ct = 0; ct < kFlushTiles; ++ct) {
        if (ct < nct) {
          float v = vr[ct * 16];
          if (!v_lds && dk >= c) v = 0.f;
          a[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(v, bv, a[ct], 0, 0, 0);
        }
      }
    }
  }
  if (changed && FR.writes(row)) {
    float* aw = Ainv + (size_t)row * n + 4 * g;
#pragma unroll
    for (int ct = 0; ct < kFlushTiles; ++ct)
      if (ct < nct) *reinterpret_cast<f32x4_t*>(aw + ct * 16) = a[ct];
  }
}

// ------------------------------------------------------------------ FP64 inverse (recompute)
// psiM = (A^T)^-1 = transpose of A^-1 with log|det A| and its phase, the work of
// DiracMatrix::invert_transpose (DiracMatrix.h:93-121: transpose, getrf, log det
// from the pivots, getri).  Here: X = A^T in FP64, then an in-place blocked
// Gauss-Jordan inversion with partial (row) pivoting.  Its pivots are the U
// diagonal of LU with partial pivoting, so log|det| and the sign come out the
// same way.  One workgroup per matrix; the panel of b columns (all n rows)
// lives in LDS; the other columns get the panel's rank-b update
//   X[i, c] <- (i in pivot rows ? 0 : Xsw[i, c]) + sum_t P[i, t] Xsw[k0 + t, c]
// in column chunks streamed through HBM (Xsw = rows after the panel's swaps).
// The column swaps that undo the pivoting are applied on the way out.
constexpr int kGJChunk = 64;
// Tiles per wave in a column chunk's update: three 16-row tiles by one 16-column tile (the waves
// split into four column groups).  With 512 threads the kernel then needs about 150 registers, so
// its CU keeps room for the proposals kernel's waves beside it.
template <int NT>
constexpr int gj_row_tiles() { return 3; }
template <int NT>
constexpr int gj_col_tiles() { return 1; }

struct GJLds {
  size_t oP, oXR, oXA, oF, oPr, oRed, oRi, oPiv, oSrc, oSlot, oDl, total;
};
__host__ __device__ inline GJLds gj_lds_b(int n, int b) {
  GJLds L;
  size_t o = 0;
  auto take = [&](size_t bytes) { const size_t at = o; o += (bytes + 15) & ~(size_t)15; return at; };
  L.oP = take(sizeof(double) * (size_t)n * (b + 1));      // panel, row stride b+1
  L.oXR = take(sizeof(double) * (size_t)b * kGJChunk);   // pivot rows of a column chunk
  L.oXA = take(sizeof(double) * (size_t)b * kGJChunk);   // rows displaced by the swaps
  L.oF = take(sizeof(double) * (size_t)(n > b ? n : b)); // the pivot position's old row (b entries)
  L.oPr = take(sizeof(double) * (size_t)b);              // scaled pivot row of a step
  L.oRed = take(sizeof(double) * 16);
  L.oRi = take(sizeof(int) * 16);
  L.oPiv = take(sizeof(int) * (size_t)n);
  L.oSrc = take(sizeof(int) * (size_t)n);                // row source of the panel's swaps / column source at the end
  L.oSlot = take(sizeof(int) * (size_t)n);
  L.oDl = take(sizeof(int) * (size_t)b);
  L.total = o;
  return L;
}
// panel width: a thread per row (512 threads for b = 32, 1024 below), the panel within the LDS
__host__ __device__ inline int gj_panel(int n) {
  if (n <= 512 && gj_lds_b(n, 32).total <= 160 * 1024) return 32;
  if (n <= 1024 && gj_lds_b(n, 16).total <= 160 * 1024) return 16;
  return 8;
}
__host__ __device__ inline GJLds gj_lds(int n) { return gj_lds_b(n, gj_panel(n)); }

template <typename TI, typename TO, int B, int NT>
__global__ __launch_bounds__(NT) void k_gj_invert(int n, const TI* A, TO* psiM, double* ws, double* logdet,
                                                          double* phase, int ld_stride, int mat_stride, const int* only) {
  const int mat = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  if (only != nullptr && only[mat] == 0) return;   // only the matrices whose flag is set
  const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  const int b = B, ps = b + 1;
  const GJLds L = gj_lds(n);
  extern __shared__ __align__(16) unsigned char smem[];
  double* P = (double*)(smem + L.oP);
  double* XR = (double*)(smem + L.oXR);
  double* XA = (double*)(smem + L.oXA);
  double* fcol = (double*)(smem + L.oF);
  double* prow = (double*)(smem + L.oPr);
  double* rv = (double*)(smem + L.oRed);
  int* ri = (int*)(smem + L.oRi);
  int* piv = (int*)(smem + L.oPiv);
  int* src = (int*)(smem + L.oSrc);
  int* slot = (int*)(smem + L.oSlot);
  int* dl = (int*)(smem + L.oDl);
  const size_t mo = (size_t)mat * mat_stride;   // matrix slot (every mat_stride-th matrix)
  double* X = ws + mo * 2 * n * n;
  const TI* a = A + mo * n * n;

  // X = A^T, b rows of A at a time through LDS (the panel buffer is free yet;
  // b (n + 1) <= n (b + 1) doubles)
  for (int s0 = 0; s0 < n; s0 += b) {
    const int sb = min(b, n - s0);
    double* T = P;   // [sb][n + 1]
    // eight loads in flight per thread before their LDS stores
    for (int e0 = tid; e0 < sb * n; e0 += nt * 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * nt;
        const int ec = e < sb * n ? e : 0;   // clamped: the loads stay unconditional
        v[u] = (double)a[(size_t)(s0 + ec / n) * n + ec % n];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * nt;
        if (e < sb * n) T[(e / n) * (n + 1) + e % n] = v[u];
      }
    }
    __syncthreads();
    for (int e = tid; e < sb * n; e += nt) {
      const int c = e / sb, r = e % sb;   // X[c][s0 + r] = A[s0 + r][c]
      X[(size_t)c * n + s0 + r] = T[r * (n + 1) + c];
    }
    __syncthreads();
  }

  // product of the pivots' magnitudes as mantissa and binary exponent (one log at the end)
  double pmant = 1.0;
  int pexp = 0;
  int sign = 1;
  for (int k0 = 0; k0 < n; k0 += b) {
    const int bb = min(b, n - k0);
    // ---- panel: columns [k0, k0 + bb); thread i holds row i in registers
    const bool own = tid < n;
    double pr[B];
#pragma unroll
    for (int c = 0; c < B; ++c) pr[c] = X[(size_t)(own ? tid : 0) * n + k0 + (c < bb ? c : 0)];
#pragma unroll
    for (int c = 0; c < B; ++c) pr[c] = (own && c < bb) ? pr[c] : 0.0;
    double* swp = fcol;   // old pivot-position row, for the thread that takes it
    // Steps go in groups of four: within a group step u works on pr[u] (a register fixed at
    // compile time), and after the group each thread's panel entries rotate four places left (all
    // threads alike, the pivot-row copies included), so no register is picked by a runtime index
    // and B places of rotation bring them back.
#pragma unroll 1
    for (int j0 = 0; j0 < bb; j0 += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = j0 + u;
        if (j < bb) {
          const int k = k0 + j;
          const double prj = pr[u];
          // pivot: first row of largest |P[i][j]|, i >= k (idamax semantics): the wave's largest
          // magnitude, then its first lane (rows ascend with the lanes)
          const double av = (own && tid >= k) ? fabs(prj) : -1.0;
          const double best = wave_max(av);
          const unsigned long long hit = __ballot(av == best);
          const int bi = best >= 0.0 ? wid * 64 + (__ffsll((long long)hit) - 1) : k;
          if (lane == 0) { rv[wid] = best; ri[wid] = bi; }
          __syncthreads();
          // the waves' candidates side by side in the lanes of every wave: the largest
          // magnitude, ties to the first wave (its rows are the lowest)
          const double cv = lane < nw ? rv[lane] : -2.0;
          const int cr = lane < nw ? ri[lane] : 0;
          const double cbest = wave_max(cv);
          const int qb = max(__ffsll((long long)__ballot(cv == cbest)) - 1, 0);   // 0 if a NaN hid the maximum
          const int p = __builtin_amdgcn_readlane(cr, qb);   // in [k, n)
          if (tid == p)
#pragma unroll
            for (int c = 0; c < B; ++c) prow[c] = pr[c];
          if (tid == k && p != k)
#pragma unroll
            for (int c = 0; c < B; ++c) swp[c] = pr[c];
          __syncthreads();
          const double d = prow[u];
          const double dinv = 1.0 / d;
          if (tid == 0) {
            piv[k] = p;
            if (p != k) sign = -sign;
            if (d < 0) sign = -sign;
            int e;
            pmant = frexp(pmant * fabs(d), &e);
            pexp += e;
          }
          if (own) {
            if (tid == k) {
#pragma unroll
              for (int c = 0; c < B; ++c) pr[c] = c == u ? dinv : prow[c] * dinv;
            } else {
              double f = prj;
              if (tid == p) {
#pragma unroll
                for (int c = 0; c < B; ++c) pr[c] = swp[c];
                f = swp[u];
              }
#pragma unroll
              for (int c = 0; c < B; ++c) pr[c] = c == u ? -f * dinv : pr[c] - f * (prow[c] * dinv);
            }
          }
        }
      }
      double r4[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) r4[c] = pr[c];
#pragma unroll
      for (int c = 0; c + 4 < B; ++c) pr[c] = pr[c + 4];
#pragma unroll
      for (int c = 0; c < 4; ++c) pr[B - 4 + c] = r4[c];
    }
    // a short last panel: rotate the remaining B - 4 ceil(bb / 4) places
    for (int j = (bb + 3) / 4 * 4; j < B; j += 4) {
      double r4[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) r4[c] = pr[c];
#pragma unroll
      for (int c = 0; c + 4 < B; ++c) pr[c] = pr[c + 4];
#pragma unroll
      for (int c = 0; c < 4; ++c) pr[B - 4 + c] = r4[c];
    }
    if (own)
#pragma unroll
      for (int c = 0; c < B; ++c)
        if (c < bb) P[tid * ps + c] = pr[c];
    __syncthreads();
    // ---- panel back to X; row sources of the panel's swaps; rows the swaps displaced
    for (int e = tid; e < n * bb; e += nt) {
      const int i = e / bb, j = e % bb;
      X[(size_t)i * n + k0 + j] = P[i * ps + j];
    }
    for (int i = tid; i < n; i += nt) {
      src[i] = i;
      slot[i] = -1;
    }
    __syncthreads();
    if (tid == 0) {
      for (int k = k0; k < k0 + bb; ++k) {
        const int t = src[k];
        src[k] = src[piv[k]];
        src[piv[k]] = t;
      }
      int nd = 0;
      for (int k = k0; k < k0 + bb; ++k) {
        const int p = piv[k];
        if (p >= k0 + bb && src[p] != p && slot[p] < 0) {
          slot[p] = nd;
          dl[nd++] = p;
        }
      }
      rv[0] = (double)nd;
    }
    __syncthreads();
    const int nd = (int)rv[0];
    // ---- rank-bb update of the other columns, kGJChunk at a time, on the FP64 matrix cores
    // (v_mfma_f64_16x16x4f64).  A wave takes kGJRowTiles 16-row tiles (rows 16 rt .., rt =
    // base + wr + q nrg) across kGJColTiles of the chunk's four 16-column tiles.  Operand layouts, lane
    // (j, g) = lane j + 16 g: A (16 x 4) = P[16 rt + j][t0 + g], B (4 x 16) = XR[t0 + g][16 ct + j],
    // C/D register r = row 16 rt + g + 4 r, column 16 ct + j (rows interleave over the lane
    // groups, unlike the FP32 16x16x4 layout).  The wave's X values of a chunk
    // are loaded before the chunk's pivot and displaced rows are staged, so both rounds of loads
    // are in flight together.
    constexpr int kGJRowTiles = gj_row_tiles<NT>(), kGJColTiles = gj_col_tiles<NT>();
    const int nrt = (n + 15) >> 4;
    const int lj = lane & 15, lg = lane >> 4;
    constexpr int nrg = NT / 64 / (4 / kGJColTiles);          // waves per column group
    const int wr = wid % nrg, ct0 = (wid / nrg) * kGJColTiles;   // row slot, first column tile
    for (int c0 = 0; c0 < n; c0 += kGJChunk) {
      const int cw = min(kGJChunk, n - c0);
      if (c0 >= k0 && c0 + cw <= k0 + bb) continue;   // the chunk is the panel
      f64x4 acc[kGJRowTiles][kGJColTiles];
      auto load_group = [&](int base) {
#pragma unroll
        for (int q = 0; q < kGJRowTiles; ++q)
#pragma unroll
          for (int ct = 0; ct < kGJColTiles; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int i = 16 * (base + wr + q * nrg) + lg + 4 * r, cc = 16 * (ct0 + ct) + lj;
              const bool in = i < n && cc < cw;
              acc[q][ct][r] = X[in ? (size_t)i * n + c0 + cc : 0];   // unconditional load, selected below
            }
      };
      load_group(0);
      for (int e0 = tid; e0 < (bb + nd) * cw; e0 += nt * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = e0 + u * nt;
          const int t = e / cw, col = c0 + e % cw;
          const bool ok = e < (bb + nd) * cw;
          const int srow = ok ? src[t < bb ? k0 + t : dl[t - bb]] : 0;
          v[u] = X[(size_t)srow * n + (ok ? col : 0)];   // unconditional; panel columns never read back
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = e0 + u * nt;
          const int t = e / cw, c = e % cw;
          if (e < (bb + nd) * cw) {
            if (t < bb) XR[t * kGJChunk + c] = v[u];
            else XA[(t - bb) * kGJChunk + c] = v[u];
          }
        }
      }
      __syncthreads();
      for (int base = 0; base < nrt; base += kGJRowTiles * nrg) {
        if (base > 0) load_group(base);
        // the base of each row: 0 for the pivot rows, the displaced row's old entries for rows a
        // swap moved, else the row itself
#pragma unroll
        for (int q = 0; q < kGJRowTiles; ++q)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 16 * (base + wr + q * nrg) + lg + 4 * r;
            const int sl = i < n ? slot[i] : -1;
            const bool prow_i = i >= k0 && i < k0 + bb;
#pragma unroll
            for (int ct = 0; ct < kGJColTiles; ++ct) {
              const int cc = 16 * (ct0 + ct) + lj;
              double x = acc[q][ct][r];
              if (sl >= 0) x = XA[sl * kGJChunk + (cc < cw ? cc : 0)];
              acc[q][ct][r] = prow_i ? 0.0 : x;
            }
          }
        for (int t0 = 0; t0 < bb; t0 += 4) {
          const int t = t0 + lg;
          double av[kGJRowTiles], bv[kGJColTiles];
#pragma unroll
          for (int q = 0; q < kGJRowTiles; ++q) {
            const int i = 16 * (base + wr + q * nrg) + lj;
            av[q] = (i < n && t < bb) ? P[i * ps + t] : 0.0;
          }
#pragma unroll
          for (int ct = 0; ct < kGJColTiles; ++ct) {
            const int cc = 16 * (ct0 + ct) + lj;
            bv[ct] = (t < bb && cc < cw) ? XR[t * kGJChunk + cc] : 0.0;
          }
#pragma unroll
          for (int q = 0; q < kGJRowTiles; ++q)
#pragma unroll
            for (int ct = 0; ct < kGJColTiles; ++ct)
              acc[q][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[ct], acc[q][ct], 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < kGJRowTiles; ++q)
#pragma unroll
          for (int ct = 0; ct < kGJColTiles; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int i = 16 * (base + wr + q * nrg) + lg + 4 * r, cc = 16 * (ct0 + ct) + lj, col = c0 + cc;
              if (i < n && cc < cw && !(col >= k0 && col < k0 + bb)) X[(size_t)i * n + col] = acc[q][ct][r];
            }
      }
      __syncthreads();
    }
  }
  // ---- undo the pivoting: columns swapped back in reverse order, then out in TO
  if (tid == 0) {
    for (int c = 0; c < n; ++c) src[c] = c;
    for (int k = n - 1; k >= 0; --k) {
      const int t = src[k];
      src[k] = src[piv[k]];
      src[piv[k]] = t;
    }
    logdet[(size_t)mat * ld_stride] = log(pmant) + (double)pexp * 0.69314718055994530942;
    phase[(size_t)mat * ld_stride] = sign > 0 ? 0.0 : M_PI;
  }
  __syncthreads();
  // psiM[i][j] = X[i][src[j]], b rows at a time through LDS
  TO* out = psiM + mo * n * n;
  for (int s0 = 0; s0 < n; s0 += b) {
    const int sb = min(b, n - s0);
    double* T = P;   // [sb][n + 1]
    for (int e0 = tid; e0 < sb * n; e0 += nt * 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * nt;
        const int ec = e < sb * n ? e : 0;
        v[u] = X[(size_t)(s0 + ec / n) * n + ec % n];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * nt;
        if (e < sb * n) T[(e / n) * (n + 1) + e % n] = v[u];
      }
    }
    __syncthreads();
    for (int e = tid; e < sb * n; e += nt) {
      const int r = e / n, j = e % n;
      out[(size_t)(s0 + r) * n + j] = (TO)T[r * (n + 1) + src[j]];
    }
    __syncthreads();
  }
}

// host side: the panel width is a template argument (the panel row lives in registers)
// nmat matrices at A, A + mat_stride n^2, ...; log|det| and phase of matrix m at [m * ld_stride]
template <typename TI, typename TO>
inline hipError_t launch_gj_invert(int n, int nmat, hipStream_t st, const TI* A, TO* psiM, double* ws,
                                   double* logdet, double* phase, int ld_stride, int mat_stride = 1,
                                   const int* only = nullptr) {
  // one thread per row of the panel: 512 threads up to n = 448 (32 registers pairs of panel row),
  // 1024 threads beyond
  const size_t lds = gj_lds(n).total;
  const int b = gj_panel(n);
  if (b == 32)
    hipLaunchKernelGGL((k_gj_invert<TI, TO, 32, 512>), dim3(nmat), dim3(512), lds, st, n, A, psiM, ws, logdet, phase,
                       ld_stride, mat_stride, only);
  else if (b == 16)
    hipLaunchKernelGGL((k_gj_invert<TI, TO, 16, 1024>), dim3(nmat), dim3(1024), lds, st, n, A, psiM, ws, logdet,
                       phase, ld_stride, mat_stride, only);
  else
    hipLaunchKernelGGL((k_gj_invert<TI, TO, 8, 1024>), dim3(nmat), dim3(1024), lds, st, n, A, psiM, ws, logdet, phase,
                       ld_stride, mat_stride, only);
  return hipGetLastError();
}

// ------------------------------------------------------------------ SPO rows (evaluateLog)
// SPO at every electron of every walker -> A, dpsiM, d2psiM
// (evaluate_notranspose, SPOSet.h:94; DiracDeterminant.cpp:264-280).
template <typename T>
__global__ __launch_bounds__(1024) void k_spo_rows(SysP<T> S, BatchP<T> B, int slices) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* part = (T*)smem;
  T* out = part + (size_t)slices * 5 * S.npad;  // [5][n]
  double* part64 = (double*)(smem + (((sizeof(T) * ((size_t)slices * 5 * S.npad + 5 * (size_t)S.n)) + 15) & ~(size_t)15));
  const int e = blockIdx.x, w = blockIdx.y, N = S.N, n = S.n;
  const double* R = B.R + (size_t)w * 3 * N;
  const int det = e < n ? 0 : 1, row = e - det * n;
  const size_t moff = ((size_t)w * 2 + det) * n * n;
  if (B.A64 != nullptr) {
    spline_gather_rt<T, 5, true>(S, R[e], R[N + e], R[2 * N + e], slices, part, part64);
    __syncthreads();
    spline_finish64<T>(S, slices, part64, B.A64 + moff + (size_t)row * n, n);
  } else {
    spline_gather_rt<T, 5>(S, R[e], R[N + e], R[2 * N + e], slices, part);
    __syncthreads();
  }
  spline_finish<T, 5, T>(S, slices, part, out, n, n);
  __syncthreads();
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

// SPO values (v, d/dx, d/dy, d/dz, d2/dx2; einspline_spo.hpp:204-223) at the proposals of
// electrons det*n + [e0, e0 + gridDim.x) of every walker, into spo[w][det][e][5][npad].  Within
// a sweep an electron's position changes only at its own move, so the sweep's proposals
// r_e + displacement are fixed before the first of them is accepted; evaluating them here takes
// the spline's HBM traffic out of the move kernel's accept chain.  The displacement is the
// move kernel's: sqrt(tau) times the walker's Philox normals of (sweep, electron), or the
// caller's (displ[w][3]).
// the gather of one point with the slice count fixed at compile time (SLC > 0) or chosen at run
// time (SLC == 0): a kernel that carries one variant allocates registers for that one alone
template <typename T, int NO, int SLC, bool V64 = false>
__device__ inline void spline_gather_sl(const SysP<T>& S, double x, double y, double z, int slices, T* part,
                                        double* part64 = nullptr) {
  if constexpr (SLC == 0)
    spline_gather_rt<T, NO, V64>(S, x, y, z, slices, part, part64);
  else
    spline_gather<T, NO, 16 / SLC, V64>(S, S.spoG, x, y, z, part, threadIdx.x, part64);
}

// SLC: slices fixed at compile time (0: the run-time argument); V64: the FP32 form's FP64 value
// rows are written as well (only while the batch keeps them, B.spo64 != nullptr)
template <typename T, int SLC, bool V64>
__device__ __forceinline__ void spo_proposals_body(const SysP<T>& S, const BatchP<T>& B, int det, int e0, int ne,
                                                   uint64_t seed, const int* sweep_ctr, double sqrt_tau, int walker0,
                                                   const double* displ, T* spo, double* prop, int slices, const int w) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* part = (T*)smem;
  const int N = S.N, n = S.n;
  const int sweep = displ != nullptr ? 0 : sweep_ctr[2];   // the random streams' sweep counter
  // a workgroup takes electrons e0 + blockIdx.x, + gridDim.x, ... (fewer workgroups than
  // electrons bound the gather's share of HBM beside the move kernel)
  for (int el = e0 + blockIdx.x; el < e0 + ne; el += gridDim.x) {
  const int iel = det * n + el;
  const double* R = B.R + (size_t)w * 3 * N;
  double d[3], u = 0.0;   // a caller's displacement comes with its own accept decision
  if (displ != nullptr) {
    d[0] = displ[3 * w + 0];
    d[1] = displ[3 * w + 1];
    d[2] = displ[3 * w + 2];
  } else {
    double nr[3];
    move_randoms(seed, (uint64_t)(walker0 + w), (uint64_t)sweep, (uint64_t)iel, nr, u);
    d[0] = sqrt_tau * nr[0];
    d[1] = sqrt_tau * nr[1];
    d[2] = sqrt_tau * nr[2];
  }
  const double x = R[iel] + d[0], y = R[N + iel] + d[1], z = R[2 * N + iel] + d[2];
  if constexpr (SLC == 1) {
  } else if constexpr (V64) {
    double* part64 = (double*)(smem + ((sizeof(T) * (size_t)slices * 5 * S.npad + 15) & ~(size_t)15));
    spline_gather_sl<T, 5, SLC, true>(S, x, y, z, slices, part, part64);
    __syncthreads();
    spline_finish64<T>(S, slices, part64, B.spo64 + (((size_t)w * 2 + det) * n + el) * S.npad, n);
  } else {
    spline_gather_sl<T, 5, SLC>(S, x, y, z, slices, part);
    __syncthreads();
  }
  T* out = spo + (((size_t)w * 2 + det) * n + el) * 5 * S.npad;
  if constexpr (SLC == 1) {
    spline_gather_direct<T, 5, V64>(S, x, y, z, out,
                                    V64 ? B.spo64 + (((size_t)w * 2 + det) * n + el) * S.npad : nullptr);
  } else {
    spline_finish<T, 5, T>(S, slices, part, out, S.npad, n);
    __syncthreads();
  }
  // J1 at the proposal (OneBodyJastrow::ratioGrad, OneBodyJastrow.h:259-268): it depends on the
  // proposal and the ions only, so it is evaluated here with the SPO row
  const Cell<T> C = *S.cell;
  double acc[5] = {0, 0, 0, 0, 0};
  for (int ci = threadIdx.x; ci < S.M; ci += blockDim.x) {
    const Functor<T>& f = S.fn[S.species[ci]];
    if (f.ncoef > 0) {
      T dx, dy, dz;
      const T r = min_image(C, (T)(S.ions[ci] - x), (T)(S.ions[S.M + ci] - y), (T)(S.ions[2 * S.M + ci] - z), T(-1),
                            dx, dy, dz);
      if (r < f.rcut) {
        T u, du, d2u;
        functor_vgl(f, r, u, du, d2u);
        acc[0] += (double)u;
        acc[1] += (double)(d2u + T(2) * du);
        acc[2] += (double)(du * dx);
        acc[3] += (double)(du * dy);
        acc[4] += (double)(du * dz);
      }
    }
  }
  double* red = (double*)smem;   // [waves][5]
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < 5; ++k) acc[k] = wave_sum(acc[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 5; ++k) red[wid * 5 + k] = acc[k];
  __syncthreads();
  if (threadIdx.x < 5) {
    double t = 0;
    for (int q = 0; q < nw; ++q) t += red[q * 5 + threadIdx.x];
    prop[(((size_t)w * 2 + det) * n + el) * kProp + kPropJ1 + threadIdx.x] = t;
  } else if (threadIdx.x < 9) {
    const int k = threadIdx.x - 5;
    prop[(((size_t)w * 2 + det) * n + el) * kProp + kPropDispl + k] = k < 3 ? d[k] : u;
  }
  if (S.j3 != nullptr && wid == 0) {
    // the ions within their three-body cutoff of the proposal, in ion order (ballot compaction
    // over 64 ions at a time): count, then (ion, r, ion - proposal)
    T* rec = B.j3prop + (((size_t)w * 2 + det) * n + el) * kJ3Rec;
    int cnt = 0;
    for (int c0 = 0; c0 < S.M; c0 += 64) {
      const int ci = c0 + lane;
      bool in = false;
      T r = 0, dx = 0, dy = 0, dz = 0;
      if (ci < S.M) {
        r = min_image(C, (T)(S.ions[ci] - x), (T)(S.ions[S.M + ci] - y), (T)(S.ions[2 * S.M + ci] - z), T(-1), dx,
                      dy, dz);
        in = r < S.j3[2 * S.species[ci]].L;
      }
      const unsigned long long bal = __ballot(in);
      const int pos = cnt + __popcll(bal & ((1ull << lane) - 1ull));
      if (in && pos < kJ3Near) {
        T* e = rec + 1 + 5 * pos;
        e[0] = (T)ci;
        e[1] = r;
        e[2] = dx;
        e[3] = dy;
        e[4] = dz;
      }
      cnt += __popcll(bal);
    }
    if (lane == 0) {
      rec[0] = (T)cnt;
      if (cnt > kJ3Near) atomicOr(B.flags, kFlagJ3Near);
    }
  }
  __syncthreads();   // the next electron's gather reuses the LDS
  }
}

template <typename T, int SLC, bool V64>
__global__ __launch_bounds__(1024) void k_spo_proposals(SysP<T> S, BatchP<T> B, int det, int e0, int ne, uint64_t seed,
                                                         const int* sweep_ctr, double sqrt_tau, int walker0,
                                                         const double* displ, T* spo, double* prop, int slices) {
  spo_proposals_body<T, SLC, V64>(S, B, det, e0, ne, seed, sweep_ctr, sqrt_tau, walker0, displ, spo, prop, slices,
                                  blockIdx.y);
}

// The proposals of one determinant of every walker of several batches in one launch: batch g owns
// the grid rows [first[g], first[g+1]); the workgroup size and LDS are the largest any batch needs.
constexpr int kMaxPropGroups = 4;
template <typename T>
struct PropGroup {
  int ng;
  int first[kMaxPropGroups + 1];
  SysP<T> S[kMaxPropGroups];
  BatchP<T> B[kMaxPropGroups];
  T* spo[kMaxPropGroups];
  double* prop[kMaxPropGroups];
  int walker0[kMaxPropGroups], slices[kMaxPropGroups];
};

template <typename T, bool V64>
__global__ __launch_bounds__(1024) void k_spo_proposals_group(PropGroup<T> G, int det, uint64_t seed, double sqrt_tau) {
  int g = 0;
  while (g + 1 < G.ng && (int)blockIdx.y >= G.first[g + 1]) ++g;
  spo_proposals_body<T, 0, V64>(G.S[g], G.B[g], det, 0, G.S[g].n, seed, G.B[g].sweep, sqrt_tau, G.walker0[g],
                                nullptr, G.spo[g], G.prop[g], G.slices[g], (int)blockIdx.y - G.first[g]);
}

// stand-alone VGH at arbitrary points: out[p][10][norb] (FP64)
template <typename T>
__global__ __launch_bounds__(1024) void k_spline_vgh(SysP<T> S, const double* r, double* out, int norb, int slices) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* part = (T*)smem;
  const int p = blockIdx.x;
  spline_gather_rt<T, 10>(S, r[3 * p], r[3 * p + 1], r[3 * p + 2], slices, part);
  __syncthreads();
  spline_finish<T, 10, double>(S, slices, part, out + (size_t)p * 10 * norb, norb, norb);
}

// ------------------------------------------------------------------ virtual-move ratios
// WaveFunction::evaluateRatios (WaveFunction.cpp:281-297) for electron iel of every walker at
// nvp virtual positions vpos[w][k][3], ratios[w][k]:
//   determinant: value-only SPO row at the point dotted with the electron's inverse row
//     (DiracDeterminant::evaluateRatios, DiracDeterminant.cpp:216-224; SPOSet::evaluateDetRatios,
//     SPOSet.h:72-83; MultiBspline::evaluate_v, MultiBspline.hpp:34-63);
//   J1: exp(Vat[iel] - sum over ions of u(|ion - point|))   (OneBodyJastrow.h:136-140);
//   J2: exp(Uat[iel] - sum over j != iel of u(|r_j - point|)) (TwoBodyJastrow.h:103-107);
//   J3 (with a three-body Jastrow): exp(U3[iel] - U3 at the point)   (ThreeBodyJastrow.h:262-273).
// One workgroup per (point, walker).  The inverse must carry no pending delayed rows, which
// holds between the batch's calls.
// Three-body Jastrow sum U of electron e of walker w placed at (x, y, z), the others at their
// current positions (ThreeBodyJastrow::evaluateRatios' computeU, ThreeBodyJastrow.h:262-273):
// over ions I with the point within L_I and electrons k != e within L_I of I (the batch's ion
// masks of the current positions), F(r_xk, r_xI, r_kI) with the functor of I's species and the
// (e, k) spin pair.  The lanes [tid, tid + nt, ...) take the electrons k; returns this lane's part.
template <typename T>
__device__ inline double j3_value_partial(const SysP<T>& S, const BatchP<T>& B, const Cell<T>& C, int w, int e,
                                          double x, double y, double z, int tid, int nt) {
  const int N = S.N, n = S.n, M = S.M, mw = S.mw;
  const double* R = B.R + (size_t)w * 3 * N;
  const unsigned long long* mk = B.j3mask + (size_t)w * N * mw;
  const bool up = e < n;
  double acc = 0.0;
  for (int I = 0; I < M; ++I) {
    const int sp = S.species[I];
    const T L = S.j3[2 * sp].L;
    if (!(L > T(0))) continue;   // the species has no three-body functor
    const double ix = S.ions[I], iy = S.ions[M + I], iz = S.ions[2 * M + I];
    T dx, dy, dz;
    const T rxI = min_image(C, (T)(ix - x), (T)(iy - y), (T)(iz - z), T(-1), dx, dy, dz);
    if (!(rxI < L)) continue;
    const unsigned long long bit = 1ull << (I & 63);
    const int wd = I >> 6;
    for (int k = tid; k < N; k += nt) {
      if (k == e || !(mk[(size_t)k * mw + wd] & bit)) continue;
      const double kx = R[k], ky = R[N + k], kz = R[2 * N + k];
      const T rkI = min_image(C, (T)(ix - kx), (T)(iy - ky), (T)(iz - kz), T(-1), dx, dy, dz);
      const T rxk = min_image(C, (T)(kx - x), (T)(ky - y), (T)(kz - z), T(-1), dx, dy, dz);
      acc += (double)poly3_v(S.j3[2 * sp + (((k < n) == up) ? 0 : 1)], rxk, rxI, rkI);
    }
  }
  return acc;
}

// Ratio at one point for walker w, electron iel; the workgroup's lanes share the work and
// every lane returns the ratio.  part: LDS [slices][npad] T; red: LDS 48 doubles.
template <typename T>
__device__ inline double virtual_ratio(const SysP<T>& S, const BatchP<T>& B, const Cell<T>& C, int w, int iel,
                                       double x, double y, double z, int slices, T* part, double* red) {
  const int N = S.N, n = S.n, M = S.M;
  spline_gather_rt<T, 1>(S, x, y, z, slices, part);
  __syncthreads();
  const int det = iel < n ? 0 : 1, row = iel - det * n;
  const T* arow = B.Ainv + ((size_t)w * 2 + det) * n * n + (size_t)row * n;
  double v[3] = {0.0, 0.0, 0.0};   // det ratio, J1 sum, J2 and J3 sum (both enter the exponent alike)
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    T s = 0;
    for (int sl = 0; sl < slices; ++sl) s += part[(size_t)sl * S.npad + j];
    v[0] += (double)s * (double)arow[j];
  }
  T dx, dy, dz;   // values only
  for (int ci = threadIdx.x; ci < M; ci += blockDim.x) {
    const Functor<T>& f = S.fn[S.species[ci]];
    if (f.ncoef == 0) continue;
    const T r = min_image(C, (T)(S.ions[ci] - x), (T)(S.ions[M + ci] - y), (T)(S.ions[2 * M + ci] - z), T(-1), dx,
                          dy, dz);
    if (r < f.rcut) v[1] += (double)functor_v(f, r);
  }
  const double* R = B.R + (size_t)w * 3 * N;
  const bool up = iel < n;
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    if (j == iel) continue;
    const Functor<T>& f = S.fn[kMaxSpecies + (((j < n) == up) ? 0 : 1)];
    const T r = min_image(C, (T)(R[j] - x), (T)(R[N + j] - y), (T)(R[2 * N + j] - z), T(-1), dx, dy, dz);
    if (r < f.rcut) v[2] += (double)functor_v(f, r);
  }
  if (S.j3 != nullptr) v[2] += j3_value_partial<T>(S, B, C, w, iel, x, y, z, threadIdx.x, blockDim.x);
  block_sum<double, 3>(v, red);   // ends with a barrier: part and red are free again
  const size_t o = (size_t)w * 5 * N + iel;
  const double u23 = (double)B.J2[o] + (S.j3 != nullptr ? (double)B.J3[o] : 0.0);
  return v[0] * exp(((double)B.J1[o] - v[1]) + (u23 - v[2]));
}

template <typename T>
__global__ __launch_bounds__(1024) void k_virtual_ratios(SysP<T> S, BatchP<T> B, int iel, int nvp, const double* vpos,
                                                         double* ratios, int slices) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* part = (T*)smem;                                  // [slices][npad]
  double* red = (double*)(part + (size_t)slices * S.npad);
  const int k = blockIdx.x, w = blockIdx.y;
  const double* p = vpos + ((size_t)w * nvp + k) * 3;
  const Cell<T> C = *S.cell;
  const double r = virtual_ratio<T>(S, B, C, w, iel, p[0], p[1], p[2], slices, part, red);
  if (threadIdx.x == 0) ratios[(size_t)w * nvp + k] = r;
}

// NonLocalPP::evaluate (NonLocalPP.hpp:115-140) for every electron of every walker.
// k_nlpp_pairs: one lane per (walker, electron) lists the ions within rmax in ion order
// (minimum image, ion - electron as the BA table's displacement) into
// near[(w*N + e)*maxp + p][4] = dist, displacement and pkey[...] = ion * kNlppRadialBins + the
// distance's bin in [0, rmax), counts them in npairs[w][e] (the full count, also beyond maxp) and
// per key in icnt[key].  k_nlpp_offsets and k_nlpp_scatter order the listed pairs by key into the
// work list items[] (count[0] entries).  k_nlpp_ratios walks knot-major over that list, each
// XCD over a contiguous run of ion keys.  Every walker has its own randomly rotated knots
// (spheres[w][12][3], one rotation per walker and evaluation as the reference's per-mover
// NonLocalPP::randomize, NonLocalPP.hpp:100-113, miniqmc.cpp:470); knot k of a pair sits at
// dist * spheres[w][k] + displ + r_e (NonLocalPP.hpp:133); ratios[w][e][maxp][kNlppKnots].
constexpr int kNlppKnots = 12;
constexpr int kNlppRadialBins = 16;
struct NlppKnots {
  double p[kNlppKnots][3];
};

__global__ void k_zero_ints(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_nlpp_pairs(SysP<T> S, BatchP<T> B, double rmax, int maxp, double* near,
                                                    int* pkey, int* npairs, int* icnt) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x, N = S.N, M = S.M;
  if (idx >= B.W * N) return;
  const int w = idx / N, e = idx - w * N;
  const Cell<T> C = *S.cell;
  const double* R = B.R + (size_t)w * 3 * N;
  const double x = R[e], y = R[N + e], z = R[2 * N + e];
  int c = 0;
  for (int ci = 0; ci < M; ++ci) {
    T dx, dy, dz;
    const T r = min_image(C, (T)(S.ions[ci] - x), (T)(S.ions[M + ci] - y), (T)(S.ions[2 * M + ci] - z), T(-1), dx, dy,
                          dz);
    if ((double)r < rmax) {
      if (c < maxp) {
        double* q = near + ((size_t)idx * maxp + c) * 4;
        q[0] = (double)r;
        q[1] = (double)dx;
        q[2] = (double)dy;
        q[3] = (double)dz;
        const int bin = (int)((double)r / rmax * kNlppRadialBins);
        const int key = ci * kNlppRadialBins + (bin < kNlppRadialBins - 1 ? bin : kNlppRadialBins - 1);
        pkey[(size_t)idx * maxp + c] = key;
        atomicAdd(icnt + key, 1);
      }
      ++c;
    }
  }
  npairs[idx] = c;
}

// exclusive prefix of the per-key counts into the per-key cursors; count[0] = listed pairs.
// One workgroup of 256 lanes: each lane sums a contiguous run of keys, lane 0 scans the 256 run
// sums in LDS, and each lane writes its run's cursors.
__global__ __launch_bounds__(256) void k_nlpp_offsets(const int* icnt, int* icur, int nkeys, int* count) {
  __shared__ int run[256];
  const int t = threadIdx.x, per = (nkeys + 255) / 256, a = t * per;
  const int b = a + per < nkeys ? a + per : nkeys;
  int s = 0;
  for (int c = a; c < b; ++c) s += icnt[c];
  run[t] = s;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int i = 0; i < 256; ++i) {
      const int v = run[i];
      run[i] = acc;
      acc += v;
    }
    count[0] = acc;
  }
  __syncthreads();
  int o = run[t];
  for (int c = a; c < b; ++c) {
    icur[c] = o;
    o += icnt[c];
  }
}

__global__ __launch_bounds__(256) void k_nlpp_scatter(int WN, int maxp, const int* npairs, const int* pkey, int* icur,
                                                      int* items) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= WN) return;
  const int c = npairs[idx] < maxp ? npairs[idx] : maxp;
  for (int p = 0; p < c; ++p) items[atomicAdd(icur + pkey[(size_t)idx * maxp + p], 1)] = idx * maxp + p;
}

// A work unit is P of a pair's twelve knots.  Lanes [g*L, (g+1)*L), L = (16/PER)*nq, gather
// knot g's value row (a lane: one 16-byte orbital group, PER of the 16 stencil columns), then
// every lane takes a share of the P dot products and of the P Jastrow sums, reading each
// electron's position once for all P knots.  J3: the system has a three-body Jastrow (its sums
// at the knots join the J2 sums, as both enter the exponent alike).
constexpr int kNlppThreads = 768;
template <typename T, int P, int PER, bool J3>
__global__ __launch_bounds__(kNlppThreads) void k_nlpp_ratios(SysP<T> S, BatchP<T> B, const double* spheres, int maxp,
                                                              const double* near, const int* items, const int* count,
                                                              double* ratios) {
  constexpr int SL = 16 / PER, KG = kNlppKnots / P;
  extern __shared__ __align__(16) unsigned char smem[];
  T* part = (T*)smem;                                        // [P][SL][npad]
  double* red = (double*)(part + (size_t)P * SL * S.npad);   // [3P][16]
  const int N = S.N, n = S.n, M = S.M;
  const int tid = threadIdx.x, nt = blockDim.x;
  const Cell<T> C = *S.cell;
  const int L = SL * (S.npad / V16<T>::n);
  const int g = tid / L;
  // XCD-aware walk: workgroup b runs on XCD b % 8, and XCD x takes the x-th eighth of the
  // knot-major units, so each XCD's L2 sees a contiguous run of keys (gridDim.x % 8 == 0)
  const int nitems = count[0];
  const int nunits = nitems * KG;
  const int xcd = blockIdx.x & 7, per = gridDim.x >> 3;
  const int lo = (int)((long long)nunits * xcd / 8), hi = (int)((long long)nunits * (xcd + 1) / 8);
  for (int un = lo + (int)(blockIdx.x >> 3); un < hi; un += per) {
    const int kg = un / nitems;
    const int item = items[un - kg * nitems];
    const int k0 = kg * P;
    const int idx = item / maxp;
    const int w = idx / N, e = idx - w * N;
    const double* R = B.R + (size_t)w * 3 * N;
    const double* sp = spheres + (size_t)w * 3 * kNlppKnots;   // this walker's knots
    const double* q = near + (size_t)item * 4;
    const double d = q[0];
    const double ox = R[e] + q[1], oy = R[N + e] + q[2], oz = R[2 * N + e] + q[3];
    const int det = e < n ? 0 : 1, row = e - det * n;
    const T* arow = B.Ainv + ((size_t)w * 2 + det) * n * n + (size_t)row * n;
    const size_t o = (size_t)w * 5 * N + e;
    const double j1o = (double)B.J1[o], j2o = (double)B.J2[o] + (J3 ? (double)B.J3[o] : 0.0);
    const bool up = e < n;
    double* out = ratios + (size_t)item * kNlppKnots;
    {
      if (g < P) {
        const int k = k0 + g;
        spline_gather<T, 1, PER>(S, S.spoG, d * sp[3 * k] + ox, d * sp[3 * k + 1] + oy, d * sp[3 * k + 2] + oz,
                                 part + (size_t)g * SL * S.npad, tid - g * L);
      }
      double v[3 * P];   // det ratios, J1 sums, J2 (+ J3) sums of the P knots
#pragma unroll
      for (int i = 0; i < 3 * P; ++i) v[i] = 0.0;
      T dx, dy, dz;   // the values only: each knot needs u, not its derivatives
      for (int ci = tid; ci < M; ci += nt) {
        const Functor<T>& f = S.fn[S.species[ci]];
        if (f.ncoef == 0) continue;
        const double ix = S.ions[ci], iy = S.ions[M + ci], iz = S.ions[2 * M + ci];
#pragma unroll
        for (int gg = 0; gg < P; ++gg) {
          const double px = d * sp[3 * (k0 + gg)] + ox, py = d * sp[3 * (k0 + gg) + 1] + oy, pz = d * sp[3 * (k0 + gg) + 2] + oz;
          const T r = min_image(C, (T)(ix - px), (T)(iy - py), (T)(iz - pz), T(-1), dx, dy, dz);
          if (r < f.rcut) v[P + gg] += (double)functor_v(f, r);
        }
      }
      for (int j = tid; j < N; j += nt) {
        if (j == e) continue;
        const Functor<T>& f = S.fn[kMaxSpecies + (((j < n) == up) ? 0 : 1)];
        const double rx = R[j], ry = R[N + j], rz = R[2 * N + j];
#pragma unroll
        for (int gg = 0; gg < P; ++gg) {
          const double px = d * sp[3 * (k0 + gg)] + ox, py = d * sp[3 * (k0 + gg) + 1] + oy, pz = d * sp[3 * (k0 + gg) + 2] + oz;
          const T r = min_image(C, (T)(rx - px), (T)(ry - py), (T)(rz - pz), T(-1), dx, dy, dz);
          if (r < f.rcut) v[2 * P + gg] += (double)functor_v(f, r);
        }
      }
      if constexpr (J3)
#pragma unroll
        for (int gg = 0; gg < P; ++gg) {
          const double px = d * sp[3 * (k0 + gg)] + ox, py = d * sp[3 * (k0 + gg) + 1] + oy, pz = d * sp[3 * (k0 + gg) + 2] + oz;
          v[2 * P + gg] += j3_value_partial<T>(S, B, C, w, e, px, py, pz, tid, nt);
        }
      __syncthreads();   // the P value rows are in LDS
      for (int ii = tid; ii < P * n; ii += nt) {
        const int gi = ii / n, j = ii - gi * n;
        T sv = 0;
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) sv += part[((size_t)gi * SL + sl) * S.npad + j];
        const double t = (double)sv * (double)arow[j];
#pragma unroll
        for (int gg = 0; gg < P; ++gg)
          if (gg == gi) v[gg] += t;
      }
      block_sum<double, 3 * P>(v, red);   // ends with a barrier: part and red are free again
      if (tid < P) {
#pragma unroll
        for (int gg = 0; gg < P; ++gg)
          if (gg == tid) out[k0 + gg] = v[gg] * exp((j1o - v[P + gg]) + (j2o - v[2 * P + gg]));
      }
    }
  }
}

// stand-alone distance rows: lane per (source, target)
template <typename T>
__global__ void k_distance_rows(Cell<T> C, int npts, const double* src, const int* flip_ind, int nt,
                                const double* tgt, double* r, double* dr) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)npts * nt) return;
  const int p = (int)(idx / nt), t = (int)(idx - (size_t)p * nt);
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
// sums equal the lower-triangle recompute of TwoBodyJastrow.h:356-400.  With a
// three-body Jastrow (ThreeBodyJastrow::recompute, ThreeBodyJastrow.h:363-401): every
// electron's near-ion record and ion mask, then every electron's sums over its pairs
// (each pair is evaluated by both of its electrons, so each writes only its own sums).
// LDS: positions [3N], reduction [48], functors, cell, then the ion masks [N][mw].
template <typename T>
__host__ __device__ inline size_t jastrow_init_lds(int N, int mw) {
  return sizeof(double) * (3 * (size_t)N + 48) + sizeof(Functor<T>) * (kMaxSpecies + 2) + sizeof(Cell<T>) + 16 +
         sizeof(unsigned long long) * (size_t)N * mw;
}
template <typename T>
__global__ __launch_bounds__(1024) void k_jastrow_init(SysP<T> S, BatchP<T> B) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x, N = S.N, n = S.n, M = S.M;
  double* sR = (double*)smem;
  double* sRed = sR + 3 * N;
  Functor<T>* sFn = (Functor<T>*)(sRed + 48);
  Cell<T>* sCell = (Cell<T>*)(sFn + kMaxSpecies + 2);
  unsigned long long* sMask =
      (unsigned long long*)(smem + ((sizeof(double) * (3 * (size_t)N + 48) + sizeof(Functor<T>) * (kMaxSpecies + 2) +
                                      sizeof(Cell<T>) + 15) & ~(size_t)15));
  const bool j3 = S.j3 != nullptr;
  const int mw = S.mw;
  if (tid == 0) *sCell = *S.cell;
  for (int i = tid; i < 3 * N; i += nt) sR[i] = B.R[(size_t)w * 3 * N + i];
  {
    const int words = (int)(sizeof(Functor<T>) * (kMaxSpecies + 2) / 4);
    for (int i = tid; i < words; i += nt) ((int*)sFn)[i] = ((const int*)S.fn)[i];
  }
  __syncthreads();
  T* J2 = B.J2 + (size_t)w * 5 * N;
  T* J1 = B.J1 + (size_t)w * 5 * N;
  double vsum = 0, usum = 0, tsum = 0;
  for (int i = tid; i < N; i += nt) {
    const double x = sR[i], y = sR[N + i], z = sR[2 * N + i];
    T v = 0, gx = 0, gy = 0, gz = 0, lap = 0;
    T* rec = j3 ? B.j3near + ((size_t)w * N + i) * kJ3Rec : nullptr;
    unsigned long long mk[kJ3MaxIons / 64] = {0, 0, 0, 0};
    int cnt = 0;
    for (int ci = 0; ci < M; ++ci) {
      const Functor<T>& f = sFn[S.species[ci]];
      const T L3 = j3 ? S.j3[2 * S.species[ci]].L : T(0);
      if (f.ncoef == 0 && !(L3 > T(0))) continue;
      T dx, dy, dz;
      const T r = min_image(*sCell, (T)(S.ions[ci] - x), (T)(S.ions[M + ci] - y), (T)(S.ions[2 * M + ci] - z), T(-1),
                            dx, dy, dz);
      if (f.ncoef > 0 && r < f.rcut) {
        T u, du, d2u;
        functor_vgl(f, r, u, du, d2u);
        v += u;
        lap += d2u + T(2) * du;
        gx += du * dx;
        gy += du * dy;
        gz += du * dz;
      }
      if (r < L3) {
        if (cnt < kJ3Near) {
          T* e = rec + 1 + 5 * cnt;
          e[0] = (T)ci;
          e[1] = r;
          e[2] = dx;
          e[3] = dy;
          e[4] = dz;
        }
        ++cnt;
        mk[ci >> 6] |= 1ull << (ci & 63);
      }
    }
    if (j3) {
      rec[0] = (T)cnt;
      if (cnt > kJ3Near) atomicOr(B.flags, kFlagJ3Near);
      for (int q = 0; q < mw; ++q) {
        sMask[(size_t)i * mw + q] = mk[q];
        B.j3mask[((size_t)w * N + i) * mw + q] = mk[q];
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
      const T r = min_image(*sCell, (T)(sR[j] - x), (T)(sR[N + j] - y), (T)(sR[2 * N + j] - z),
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
  if (j3) {
    __syncthreads();   // every electron's mask and near-ion record are in place
    T* J3 = B.J3 + (size_t)w * 5 * N;
    for (int j = tid; j < N; j += nt) {
      const double x = sR[j], y = sR[N + j], z = sR[2 * N + j];
      const T* rec = B.j3near + ((size_t)w * N + j) * kJ3Rec;
      const int cnt = min((int)rec[0], kJ3Near);
      const bool jup = j < n;
      T acc[5] = {0, 0, 0, 0, 0};
      for (int q = 0; q < cnt; ++q) {
        const T* e = rec + 1 + 5 * q;
        const int I = (int)e[0];
        const T rjI = e[1], jI[3] = {e[2], e[3], e[4]};
        const unsigned long long bit = 1ull << (I & 63);
        const int wd = I >> 6, sp = S.species[I];
        for (int k = 0; k < N; ++k) {
          if (k == j || !(sMask[(size_t)k * mw + wd] & bit)) continue;
          T kI[3], jk[3];
          const T rkI = min_image(*sCell, (T)(S.ions[I] - sR[k]), (T)(S.ions[M + I] - sR[N + k]),
                                  (T)(S.ions[2 * M + I] - sR[2 * N + k]), T(-1), kI[0], kI[1], kI[2]);
          const T rjk = min_image(*sCell, (T)(sR[k] - x), (T)(sR[N + k] - y), (T)(sR[2 * N + k] - z),
                                  k < j ? T(1) : T(-1), jk[0], jk[1], jk[2]);
          T o[9], js[5], ks[5];
          poly3_vgl(S.j3[2 * sp + (((k < n) == jup) ? 0 : 1)], rjk, rjI, rkI, o);
          j3_pair_terms(o, jk, jI, kI, js, ks);
#pragma unroll
          for (int c = 0; c < 5; ++c) acc[c] += js[c];
        }
      }
#pragma unroll
      for (int c = 0; c < 5; ++c) J3[(size_t)c * N + j] = acc[c];
      tsum += (double)acc[0];
    }
  }
  double red[3] = {vsum, usum, tsum};
  block_sum<double, 3>(red, sRed);
  if (tid == 0) {
    B.jlog[w * 3 + 0] = -red[0];
    B.jlog[w * 3 + 1] = -0.5 * red[1];
    B.jlog[w * 3 + 2] = -0.5 * red[2];
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
  __shared__ double sRed[48];
  const int w = blockIdx.x, N = S.N;
  const T* J2 = B.J2 + (size_t)w * 5 * N;
  const T* J1 = B.J1 + (size_t)w * 5 * N;
  const T* J3 = S.j3 != nullptr ? B.J3 + (size_t)w * 5 * N : nullptr;
  double vs = 0, us = 0, ts = 0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const size_t o = (size_t)w * N + i;
    G[3 * o] += (double)J1[N + i] + (double)J2[N + i];
    G[3 * o + 1] += (double)J1[2 * N + i] + (double)J2[2 * N + i];
    G[3 * o + 2] += (double)J1[3 * N + i] + (double)J2[3 * N + i];
    Lp[o] += -(double)J1[4 * N + i] + (double)J2[4 * N + i];
    vs += (double)J1[i];
    us += (double)J2[i];
    if (J3 != nullptr) {   // ThreeBodyJastrow::evaluateGL (ThreeBodyJastrow.h:627-645)
      G[3 * o] += (double)J3[N + i];
      G[3 * o + 1] += (double)J3[2 * N + i];
      G[3 * o + 2] += (double)J3[3 * N + i];
      Lp[o] += (double)J3[4 * N + i];
      ts += (double)J3[i];
    }
  }
  double red[3] = {vs, us, ts};
  block_sum<double, 3>(red, sRed);
  if (threadIdx.x == 0) {
    B.jlog[w * 3 + 0] = -red[0];
    B.jlog[w * 3 + 1] = -0.5 * red[1];
    B.jlog[w * 3 + 2] = -0.5 * red[2];
    const double v = B.logdet[w * 2] + B.logdet[w * 2 + 1] - red[0] - 0.5 * red[1] - 0.5 * red[2];
    logv[w] = v;
    if (hist != nullptr) {
      const int slot = B.sweep[0] - B.sweep[1];
      if (slot >= 0 && slot < hist_cap) hist[(size_t)slot * B.W + w] = v;
      if (w == 0) B.sweep[3] = slot;
    }
  }
}

// Holds its stream for `ticks` of the constant 100 MHz wall clock (one lane, a bounded wait that
// always ends).  Side-stream work released by the same event as a move launch waits behind it, so
// the move kernel's workgroups, which each need a whole CU, are placed before the side kernels'
// workgroups take a share of every CU.
__global__ void k_hold(unsigned long long ticks) {
  if (threadIdx.x == 0) {
    const unsigned long long t0 = wall_clock64();
    while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(8);
  }
}

__global__ void k_count(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += 1;
}

// The second determinant's FP64 refresh lands after that sweep's evaluateGL (it runs beside
// the next sweep's first-determinant moves): swap in the fresh log|det| and phase and move the
// sweep's recorded log|Psi| by the same difference.
// The sweep's history slot is the one its evaluateGL wrote (sweep[3]): this kernel runs after
// that evaluateGL and before the next one.
__global__ void k_refresh_det(int W, int det, const double* ld_fresh, const double* ph_fresh, double* logdet,
                              double* phase, double* logv, double* hist, int hist_cap, const int* sweep) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const double delta = ld_fresh[w] - logdet[w * 2 + det];
  logdet[w * 2 + det] = ld_fresh[w];
  phase[w * 2 + det] = ph_fresh[w];
  logv[w] += delta;
  const int slot = sweep[3];
  if (hist != nullptr && slot >= 0 && slot < hist_cap) hist[(size_t)slot * W + w] += delta;
}

}  // namespace qmcx

namespace qmcx {

// ------------------------------------------------------------------ stand-alone delayed-update engine
// One determinant, device-resident inverse (the DU_TYPE plug point of
// DiracDeterminant<DU_TYPE>, DiracDeterminant.cpp:339-341).  The flush reuses the
// batch flush kernels with one walker.
//
// getInvRow (DelayedUpdate.h:96-114): invRow = Ainv[row] - V^T Binv^T U Ainv[row];
// Binv^T U Ainv[row] is parked in row c of Binv, where acceptRow reads it.
template <typename T>
__global__ __launch_bounds__(256) void k_du_get_inv_row(BatchP<T> B, int n, int row, double* out) {
  __shared__ double sp[128];
  const int c = B.ndelay[0], K = B.K, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const T* ar = B.Ainv + (size_t)row * n;
  for (int d = wid; d < c; d += 4) {
    const T* u = B.A + (size_t)B.rows[d] * n;
    double s = 0;
    for (int j = lane; j < n; j += 64) s += (double)u[j] * (double)ar[j];
    s = wave_sum(s);
    if (lane == 0) sp[d] = s;
  }
  __syncthreads();
  for (int r = tid; r < c; r += 256) {
    double s = 0;
    for (int q = 0; q < c; ++q) s += B.Binv[q * K + r] * sp[q];
    B.Binv[c * K + r] = s;
  }
  __syncthreads();
  for (int j = tid; j < n; j += 256) {
    double v = (double)ar[j];
    for (int d = 0; d < c; ++d) v -= (double)B.Vbuf[(size_t)d * n + j] * B.Binv[c * K + d];
    out[j] = v;
  }
}

// acceptRow (DelayedUpdate.h:124-153): V_c = Ainv[row], U_c = psiV (kept as row `row` of A),
// the capacitance inverse grows by one row and column (FP64).
template <typename T>
__global__ __launch_bounds__(256) void k_du_accept_row(BatchP<T> B, int n, int row, const double* psiv) {
  __shared__ double sp[130];
  const int c = B.ndelay[0], K = B.K, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const T* ar = B.Ainv + (size_t)row * n;
  for (int d = wid; d <= c; d += 4) {
    const T* v = d == c ? ar : B.Vbuf + (size_t)d * n;
    double s = 0;
    for (int j = lane; j < n; j += 64) s += (double)v[j] * psiv[j];
    s = wave_sum(s);
    if (lane == 0) sp[d] = -s;
  }
  __syncthreads();
  for (int j = tid; j < n; j += 256) {
    B.Vbuf[(size_t)c * n + j] = ar[j];
    B.A[(size_t)row * n + j] = (T)psiv[j];
  }
  double y = -sp[c];
  for (int i = 0; i < c; ++i) y += B.Binv[c * K + i] * sp[i];
  const double yi = 1.0 / y;
  for (int s2 = tid; s2 < c; s2 += 256) {
    double t = 0;
    for (int r = 0; r < c; ++r) t += B.Binv[s2 * K + r] * sp[r];
    B.Binv[s2 * K + c] = yi * t;
  }
  __syncthreads();
  for (int idx = tid; idx < c * c; idx += 256) {
    const int s2 = idx / c, r = idx - s2 * c;
    B.Binv[s2 * K + r] -= B.Binv[c * K + r] * B.Binv[s2 * K + c];
  }
  __syncthreads();
  for (int i = tid; i < c; i += 256) B.Binv[c * K + i] *= -yi;
  if (tid == 0) {
    B.Binv[c * K + c] = yi;
    B.rows[c] = row;
    B.ndelay[0] = c + 1;
  }
}

__global__ void k_du_reset(int* ndelay) {
  if (threadIdx.x == 0) ndelay[0] = 0;
}

}  // namespace qmcx
