// Device-side building blocks of the VMC particle-move engine (gfx950).
//
// Everything here runs inside one workgroup that owns one walker, or is a
// per-lane helper.  Wavefronts are 64 lanes; block reductions combine the
// per-wave partials through LDS in a fixed order so every lane of the
// workgroup sees bit-identical sums (the accept/reject decision is evaluated
// redundantly by all lanes).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qmcx {

constexpr int kMaxFunctorCoefs = 24;
constexpr int kMaxSpecies = 4;
constexpr int kWave = 64;

// ------------------------------------------------------------------ RNG
// Philox4x64-10 (Salmon et al., SC'11); the walker move stream draws one block
// for the three Gaussian components and one for the acceptance uniform.
__device__ inline void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += W0; k1 += W1; }
    const uint64_t hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
    const uint64_t hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
    const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

__device__ inline double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

// Gaussian displacement components and acceptance uniform of move
// (walker, sweep, electron): counter {iel, sweep, walker, 0|1}, key {seed, 0x51A7}.
__device__ inline void move_randoms(uint64_t seed, uint64_t walker, uint64_t sweep, uint64_t iel, double nrm[3],
                                    double& uacc) {
  const double two_pi = 6.283185307179586476925286766559;
  uint64_t c[4] = {iel, sweep, walker, 0};
  philox4x64_10(c, seed, 0x51A7ULL);
  const double r0 = sqrt(-2.0 * log(1.0 - u01(c[0])));
  const double r1 = sqrt(-2.0 * log(1.0 - u01(c[2])));
  double s1, c1, s3, c3;
  sincos(two_pi * u01(c[1]), &s1, &c1);
  sincos(two_pi * u01(c[3]), &s3, &c3);
  nrm[0] = r0 * c1;
  nrm[1] = r0 * s1;
  nrm[2] = r1 * c3;
  uint64_t d[4] = {iel, sweep, walker, 1};
  philox4x64_10(d, seed, 0x51A7ULL);
  uacc = u01(d[0]);
}

// ------------------------------------------------------------------ reductions
// Wave-wide sum through DPP lane moves (quad_perm, row mirrors, row broadcasts);
// the total is read from lane 63 into a scalar register, so every lane holds
// the same bits.
template <int CTRL, int ROWMASK>
__device__ inline float dpp_add(float x) {
  const int y = __builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWMASK, 0xf, false);
  return x + __int_as_float(y);
}
template <int CTRL, int ROWMASK>
__device__ inline double dpp_add(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
  return x + __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline float lane63(float x) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63)); }
__device__ inline double lane63(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T>
__device__ inline T wave_sum(T v) {
  v = dpp_add<0xB1, 0xf>(v);    // quad_perm [1,0,3,2]
  v = dpp_add<0x4E, 0xf>(v);    // quad_perm [2,3,0,1]
  v = dpp_add<0x141, 0xf>(v);   // row_half_mirror
  v = dpp_add<0x140, 0xf>(v);   // row_mirror: every lane holds its row (16 lanes) sum
  v = dpp_add<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
  v = dpp_add<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3: lane 63 holds the total
  return lane63(v);
}

// Wave maximum through DPP lane moves (the same pattern as wave_sum; rows outside a move's
// row mask keep their own value); every lane gets lane 63's result.
template <int CTRL, int ROWMASK>
__device__ inline double dpp_max(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp((int)(b & 0xffffffff), (int)(b & 0xffffffff), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
  return fmax(x, __longlong_as_double(((long long)hi << 32) | (unsigned int)lo));
}
__device__ inline double wave_max(double v) {
  v = dpp_max<0xB1, 0xf>(v);
  v = dpp_max<0x4E, 0xf>(v);
  v = dpp_max<0x141, 0xf>(v);
  v = dpp_max<0x140, 0xf>(v);
  v = dpp_max<0x142, 0xa>(v);
  v = dpp_max<0x143, 0xc>(v);
  return lane63(v);
}

// Exchange of wave halves between two values (v_permlane32_swap: lanes 32-63 of x
// trade places with lanes 0-31 of y): x <- [x lanes 0-31 | y lanes 0-31],
// y <- [x lanes 32-63 | y lanes 32-63].
__device__ inline void swap_halves(double& x, double& y) {
  const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
  const auto lo = __builtin_amdgcn_permlane32_swap((int)(bx & 0xffffffff), (int)(by & 0xffffffff), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((int)(bx >> 32), (int)(by >> 32), false, false);
  x = __longlong_as_double(((long long)hi[0] << 32) | (unsigned int)lo[0]);
  y = __longlong_as_double(((long long)hi[1] << 32) | (unsigned int)lo[1]);
}
// Exchange of 16-lane rows (v_permlane16_swap: odd rows of x trade places with even
// rows of y): x <- [x0 y0 x2 y2], y <- [x1 y1 x3 y3] by rows.
__device__ inline void swap_rows(double& x, double& y) {
  const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
  const auto lo = __builtin_amdgcn_permlane16_swap((int)(bx & 0xffffffff), (int)(by & 0xffffffff), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((int)(bx >> 32), (int)(by >> 32), false, false);
  x = __longlong_as_double(((long long)hi[0] << 32) | (unsigned int)lo[0]);
  y = __longlong_as_double(((long long)hi[1] << 32) | (unsigned int)lo[1]);
}

// Wave sums of NV values (NV a multiple of 4) with NV/2 + NV/4 lane exchanges and NV/4 row
// reductions instead of NV full reductions: halves are traded between value pairs (i, i+NV/2),
// then rows between (i, i+NV/4).  On return v[i] (i < NV/4) holds, in every lane of row r
// (lanes 16r .. 16r+15), the wave total of value i + r*NV/4.
template <int NV>
__device__ inline void wave_sums_rows(double (&v)[NV]) {
  static_assert(NV % 4 == 0, "four values per row group");
#pragma unroll
  for (int i = 0; i < NV / 2; ++i) {
    swap_halves(v[i], v[i + NV / 2]);
    v[i] += v[i + NV / 2];
  }
#pragma unroll
  for (int i = 0; i < NV / 4; ++i) {
    swap_rows(v[i], v[i + NV / 4]);
    v[i] += v[i + NV / 4];
  }
#pragma unroll
  for (int i = 0; i < NV / 4; ++i) {
    v[i] = dpp_add<0xB1, 0xf>(v[i]);    // quad_perm [1,0,3,2]
    v[i] = dpp_add<0x4E, 0xf>(v[i]);    // quad_perm [2,3,0,1]
    v[i] = dpp_add<0x141, 0xf>(v[i]);   // row_half_mirror
    v[i] = dpp_add<0x140, 0xf>(v[i]);   // row_mirror
  }
}

// Global -> LDS copy of NR regions by the whole workgroup: every thread issues up to eight
// loads before its first LDS store, so the regions arrive in one or two memory round trips.
// 16-byte vectors when every address and size allows, 4-byte words otherwise.
// A copy of `bytes` contiguous bytes into LDS; with row > 0 the source is rows of `row` bytes,
// each landing `pitch` bytes after the previous one in LDS.
struct LdsCopy {
  void* dst;
  const void* src;
  unsigned bytes;
  unsigned row = 0, pitch = 0;
};
template <typename V, int NR>
__device__ inline void stage_lds_v(const LdsCopy (&c)[NR]) {
  const int tid = threadIdx.x, nt = blockDim.x;
  unsigned pre[NR + 1];
  pre[0] = 0;
#pragma unroll
  for (int r = 0; r < NR; ++r) pre[r + 1] = pre[r] + c[r].bytes / (unsigned)sizeof(V);
  const unsigned tot = pre[NR];
  for (unsigned b = 0; b < tot; b += 8u * nt) {
    V v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const unsigned i = min(b + (unsigned)(k * nt + tid), tot - 1);
      const V* s = (const V*)c[0].src + i;
#pragma unroll
      for (int r = 1; r < NR; ++r)
        if (i >= pre[r]) s = (const V*)c[r].src + (i - pre[r]);
      v[k] = *s;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const unsigned i = b + (unsigned)(k * nt + tid);
      auto at = [&](const LdsCopy& cr, unsigned l) -> V* {
        if (cr.row == 0) return (V*)cr.dst + l;
        const unsigned rv = cr.row / (unsigned)sizeof(V);
        return (V*)cr.dst + (l / rv) * (cr.pitch / (unsigned)sizeof(V)) + l % rv;
      };
      V* d = at(c[0], i);
#pragma unroll
      for (int r = 1; r < NR; ++r)
        if (i >= pre[r]) d = at(c[r], i - pre[r]);
      if (i < tot) *d = v[k];
    }
  }
}
template <int NR>
__device__ inline void stage_lds(const LdsCopy (&c)[NR]) {
  bool vec = true;
#pragma unroll
  for (int r = 0; r < NR; ++r)
    vec = vec && (((uintptr_t)c[r].dst | (uintptr_t)c[r].src | (uintptr_t)c[r].bytes | c[r].row | c[r].pitch) & 15) == 0;
  if (vec) stage_lds_v<int4, NR>(c);
  else stage_lds_v<int, NR>(c);
}

// LDS data exchanged between the lanes of one wave: its LDS operations have completed.
__device__ inline void wave_lds_fence() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Barrier for data exchanged through LDS only: waits for this wave's LDS
// operations, not for its outstanding global loads.
__device__ inline void lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Sum NV values over the workgroup; result is broadcast to every lane.
// scratch: LDS of at least NV*16 elements.  Ends with a barrier so scratch can
// be reused immediately.
template <typename T, int NV>
__device__ inline void block_sum(T (&v)[NV], T* scratch) {
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) scratch[i * 16 + wid] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    T s = 0;
    for (int q = 0; q < nw; ++q) s += scratch[i * 16 + q];
    v[i] = s;
  }
  __syncthreads();
}

// ------------------------------------------------------------------ radial functor
// Uniform cubic B-spline u(r) with cutoff (BsplineFunctor.h:93-339): spline coefficients
// c[0] = p1 - 2 dr cusp, c[1..] = p0, p1, ..., three trailing zeros.  Kept as a cubic per knot
// interval: u = ((a3 t + a2) t + a1) t + a0 with t the fraction in interval i and (a0 .. a3) =
// poly[i], the power form of the basis over c_i .. c_i+3 computed on the host in FP64; one 16-byte
// (FP32) read per evaluation.
template <typename T>
struct Functor {
  T rcut, drinv;
  int ncoef;
  int pad;
  alignas(16) T poly[kMaxFunctorCoefs - 3][4];
};

template <typename T>
__device__ inline void functor_poly(const Functor<T>& f, T r, T& t, T& a0, T& a1, T& a2, T& a3) {
  const T x = r * f.drinv;
  const int i = (int)x;
  t = x - (T)i;
  if constexpr (sizeof(T) == 4) {
    const float4 p = *reinterpret_cast<const float4*>(f.poly[i]);
    a0 = p.x, a1 = p.y, a2 = p.z, a3 = p.w;
  } else {
    a0 = f.poly[i][0], a1 = f.poly[i][1], a2 = f.poly[i][2], a3 = f.poly[i][3];
  }
}

// u, (du/dr)/r and d2u/dr2; caller guarantees r < rcut.
template <typename T>
__device__ inline void functor_vgl(const Functor<T>& f, T r, T& u, T& du, T& d2u) {
  T t, a0, a1, a2, a3;
  functor_poly(f, r, t, a0, a1, a2, a3);
  u = ((a3 * t + a2) * t + a1) * t + a0;
  const T d = (T(3) * a3 * t + T(2) * a2) * t + a1;   // du/dt
  const T dd = T(6) * a3 * t + T(2) * a2;              // d2u/dt2
  if constexpr (sizeof(T) == 4)
    du = f.drinv * d * __builtin_amdgcn_rcpf(r);   // v_rcp_f32 (1 ulp) in the FP32 form
  else
    du = f.drinv * d / r;
  d2u = f.drinv * f.drinv * dd;
}

template <typename T>
__device__ inline T functor_v(const Functor<T>& f, T r) {
  T t, a0, a1, a2, a3;
  functor_poly(f, r, t, a0, a1, a2, a3);
  return ((a3 * t + a2) * t + a1) * t + a0;
}

// ------------------------------------------------------------------ minimum image
// Periodic general cell: reduced basis rb (rows), its inverse g, eight corner
// shifts (ParticleBConds.h:99-217).  d is (target - source) before the flip.
template <typename T>
struct alignas(16) Cell {   // a whole number of 16-byte vectors (copied to LDS as such)
  T g[9];        // g[3*i+j] = g(i,j)
  T rb[9];       // rb[3*i+j] = rb[i][j]
  T corner[24];
  T ortho;       // 1 when the reduced basis vectors are orthogonal (min_image_ortho applies)
};

template <typename T>
__device__ inline T min_image(const Cell<T>& C, T d0, T d1, T d2, T flip, T& ox, T& oy, T& oz) {
  d0 *= flip;
  d1 *= flip;
  d2 *= flip;
  const T s0 = -floor(d0 * C.g[0] + d1 * C.g[3] + d2 * C.g[6]);
  const T s1 = -floor(d0 * C.g[1] + d1 * C.g[4] + d2 * C.g[7]);
  const T s2 = -floor(d0 * C.g[2] + d1 * C.g[5] + d2 * C.g[8]);
  const T ex = d0 + s0 * C.rb[0] + s1 * C.rb[3] + s2 * C.rb[6];
  const T ey = d1 + s0 * C.rb[1] + s1 * C.rb[4] + s2 * C.rb[7];
  const T ez = d2 + s0 * C.rb[2] + s1 * C.rb[5] + s2 * C.rb[8];
  // corner 0 is the origin; the winner's coordinates are carried by selects (no indexing by
  // the winning corner, so a cell held in registers stays there)
  T best = ex * ex + ey * ey + ez * ez, bx = ex + C.corner[0], by = ey + C.corner[1], bz = ez + C.corner[2];
#pragma unroll
  for (int c = 1; c < 8; ++c) {
    const T x = ex + C.corner[3 * c], y = ey + C.corner[3 * c + 1], z = ez + C.corner[3 * c + 2];
    const T q = x * x + y * y + z * z;
    const bool lt = q < best;
    bx = lt ? x : bx;
    by = lt ? y : by;
    bz = lt ? z : bz;
    best = lt ? q : best;
  }
  ox = flip * bx;
  oy = flip * by;
  oz = flip * bz;
  return sqrt(best);
}

// min_image for an orthogonal reduced basis: over the eight corners the squared distance splits
// into one term per axis, so the winner subtracts axis a exactly when the fractional coordinate
// after the floor exceeds one half; its corner shift is the same table entry min_image adds.
// The image is min_image's wherever the nearest image is unique beyond rounding, in particular
// for every distance under half the shortest axis (the radial functors' range).
template <typename T>
__device__ inline T min_image_ortho(const Cell<T>& C, T d0, T d1, T d2, T flip, T& ox, T& oy, T& oz) {
  d0 *= flip;
  d1 *= flip;
  d2 *= flip;
  const T f0 = d0 * C.g[0] + d1 * C.g[3] + d2 * C.g[6];
  const T f1 = d0 * C.g[1] + d1 * C.g[4] + d2 * C.g[7];
  const T f2 = d0 * C.g[2] + d1 * C.g[5] + d2 * C.g[8];
  const T s0 = -floor(f0), s1 = -floor(f1), s2 = -floor(f2);
  const T ex = d0 + s0 * C.rb[0] + s1 * C.rb[3] + s2 * C.rb[6];
  const T ey = d1 + s0 * C.rb[1] + s1 * C.rb[4] + s2 * C.rb[7];
  const T ez = d2 + s0 * C.rb[2] + s1 * C.rb[5] + s2 * C.rb[8];
  // corner order: 0, -b0, -b1, -b2, -(b0+b1), -(b0+b2), -(b1+b2), -(b0+b1+b2)
  const int bits = (f0 + s0 > T(0.5) ? 1 : 0) | (f1 + s1 > T(0.5) ? 2 : 0) | (f2 + s2 > T(0.5) ? 4 : 0);
  const int c = bits == 3 ? 4 : (bits == 4 ? 3 : bits);
  const T x = ex + C.corner[3 * c], y = ey + C.corner[3 * c + 1], z = ez + C.corner[3 * c + 2];
  ox = flip * x;
  oy = flip * y;
  oz = flip * z;
  if constexpr (sizeof(T) == 4)
    return __builtin_amdgcn_sqrtf(x * x + y * y + z * z);   // v_sqrt_f32 (1 ulp) in the FP32 form
  else
    return sqrt(x * x + y * y + z * z);
}

// ------------------------------------------------------------------ spline location
// toUnit_floor with the SPO lattice, then the grid cell and fraction per axis
// (CrystalLattice.h:185-196, MultiBsplineEvalHelper.hpp:43-64); kept in FP64.
__device__ inline void spline_locate(const double* Ginv, const int* ng, double x, double y, double z, int ix[3],
                                     double t[3]) {
  const double r[3] = {x, y, z};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double v = r[0] * Ginv[j] + r[1] * Ginv[3 + j] + r[2] * Ginv[6 + j];
    if (-2.220446049250313e-16 < v && v < 0.0) v = 0.0;
    else v -= floor(v);
    const double xs = v * (double)ng[j];
    if (xs < 0) {
      ix[j] = 0;
      t[j] = 0.0;
    } else {
      const double ip = floor(xs);
      t[j] = xs - ip;
      ix[j] = (int)ip;
      if (ix[j] > ng[j] - 1) {
        ix[j] = ng[j] - 1;
        t[j] = 1.0 - 2.220446049250313e-16;
      }
    }
  }
}

// cubic B-spline weights a, da, d2a at fraction t (MultiBsplineData.hpp:52-80)
template <typename T>
__device__ inline void bspline_weights(T t, T a[4], T da[4], T d2a[4]) {
  const T s6 = T(1) / T(6);
  a[0] = ((T(-1) * t + T(3)) * t - T(3)) * t * s6 + s6;
  a[1] = ((T(3) * t - T(6)) * t) * t * s6 + T(4) * s6;
  a[2] = ((T(-3) * t + T(3)) * t + T(3)) * t * s6 + s6;
  a[3] = t * t * t * s6;
  da[0] = (T(-0.5) * t + T(1)) * t - T(0.5);
  da[1] = (T(1.5) * t - T(2)) * t;
  da[2] = (T(-1.5) * t + T(1)) * t + T(0.5);
  da[3] = T(0.5) * t * t;
  d2a[0] = T(1) - t;
  d2a[1] = T(3) * t - T(2);
  d2a[2] = T(1) - T(3) * t;
  d2a[3] = t;
}

// 16-byte vector of T (clang ext_vector_type so the load can be non-temporal)
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef f32x4 f32x4_t;
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));   // FP64 MFMA accumulator (16x16x4)
template <typename T>
struct V16;
template <>
struct V16<float> {
  using type = f32x4;
  static constexpr int n = 4;
};
template <>
struct V16<double> {
  using type = f64x2;
  static constexpr int n = 2;
};

}  // namespace qmcx
