// C ABI of the VMC particle-move engine: host-side orchestration of the
// kernels in qmcx_kernels.hpp.  See include/qmcx.h for the contract.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qmcx.h"
#include "qmcx_kernels.hpp"

using namespace qmcx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return fail(QMCX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device buffer freed when it goes out of scope, so an early return on a HIP error leaks nothing
struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes > 0 ? bytes : 1); }
  template <typename P>
  P* as() const { return (P*)p; }
};

// The library's three HIP streams on a device: work of a batch (moves, evaluateGL, stand-alone
// kernels), the FP64 refresh, and the proposals' SPO rows.  They are created once per device,
// on the first system created there, and kept until the process ends; every system, batch and
// DU engine on the device shares them.  The runtime gives each stream a hardware queue while it
// has queues free and then shares the least-used one (HIP's default limit is four per process,
// one of them the null stream's): three streams made once get three queues, whereas streams made
// per batch could land two of a batch's side streams on one queue, where they run one after the
// other instead of beside the moves.
constexpr int kStreamDevices = 64;
std::mutex g_stream_mu;
hipStream_t g_streams[kStreamDevices][3] = {};

int device_streams(hipStream_t out[3]) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kStreamDevices) return fail(QMCX_ERR_ARG, "device index beyond the stream table");
  std::lock_guard<std::mutex> lock(g_stream_mu);
  for (int i = 0; i < 3; ++i)
    if (g_streams[dev][i] == nullptr) HIPCHK(hipStreamCreate(&g_streams[dev][i]));
  for (int i = 0; i < 3; ++i) out[i] = g_streams[dev][i];
  return 0;
}

// ------------------------------------------------------------------ lattice (host)
void inv3(const double a[9], double o[9]) {
  const double c00 = a[4] * a[8] - a[5] * a[7];
  const double c01 = a[5] * a[6] - a[3] * a[8];
  const double c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  const double id = 1.0 / det;
  o[0] = c00 * id;
  o[3] = c01 * id;
  o[6] = c02 * id;
  o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}
double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Replace the longest basis vector by a shorter +-combination when one exists
// (LatticeAnalyzer.h:134-158); repeated until no change (:160-184).
bool shorten(double rb[9]) {
  const double eps = 10.0 * std::numeric_limits<float>::epsilon();
  int imax = 0;
  double r2max = dot3(rb, rb);
  for (int i = 1; i < 3; ++i) {
    const double r2 = dot3(rb + 3 * i, rb + 3 * i);
    if (r2 - r2max > eps) {
      r2max = r2;
      imax = i;
    }
  }
  double cand[4][3];
  for (int d = 0; d < 3; ++d) {
    cand[0][d] = rb[d] + rb[3 + d] - rb[6 + d];
    cand[1][d] = rb[d] + rb[6 + d] - rb[3 + d];
    cand[2][d] = rb[3 + d] + rb[6 + d] - rb[d];
    cand[3][d] = rb[d] + rb[3 + d] + rb[6 + d];
  }
  for (int i = 0; i < 4; ++i)
    if (dot3(cand[i], cand[i]) - r2max < -eps) {
      for (int d = 0; d < 3; ++d) rb[3 * imax + d] = cand[i][d];
      return true;
    }
  return false;
}
void reduce_basis(double rb[9]) {
  for (;;) {
    double saved[9];
    std::memcpy(saved, rb, sizeof(saved));
    bool changed = false;
    for (int i = 0; i < 3; ++i) {
      rb[3 * i] = rb[3 * i + 1] = rb[3 * i + 2] = 0.0;
      changed = shorten(rb);
      for (int d = 0; d < 3; ++d) rb[3 * i + d] = saved[3 * i + d];
      if (changed) break;
    }
    if (!changed && !shorten(rb)) return;
  }
}

template <typename T>
Cell<T> make_cell(const double lat[9]) {
  double rb[9], g[9];
  std::memcpy(rb, lat, sizeof(rb));
  reduce_basis(rb);
  inv3(rb, g);
  Cell<T> C;
  for (int i = 0; i < 9; ++i) {
    C.g[i] = (T)g[i];
    C.rb[i] = (T)rb[i];
  }
  const double* b0 = rb;
  const double* b1 = rb + 3;
  const double* b2 = rb + 6;
  for (int d = 0; d < 3; ++d) {
    C.corner[0 + d] = T(0);
    C.corner[3 + d] = (T)(-b0[d]);
    C.corner[6 + d] = (T)(-b1[d]);
    C.corner[9 + d] = (T)(-b2[d]);
    C.corner[12 + d] = (T)(-(b0[d] + b1[d]));
    C.corner[15 + d] = (T)(-(b0[d] + b2[d]));
    C.corner[18 + d] = (T)(-(b1[d] + b2[d]));
    C.corner[21 + d] = (T)(-(b0[d] + b1[d] + b2[d]));
  }
  // orthogonal reduced basis: every pair of vectors at right angles to 1e-12 relative
  const double d01 = dot3(b0, b1), d02 = dot3(b0, b2), d12 = dot3(b1, b2);
  const double n0 = std::sqrt(dot3(b0, b0)), n1 = std::sqrt(dot3(b1, b1)), n2 = std::sqrt(dot3(b2, b2));
  const bool ortho = std::fabs(d01) <= 1e-12 * n0 * n1 && std::fabs(d02) <= 1e-12 * n0 * n2 &&
                     std::fabs(d12) <= 1e-12 * n1 * n2;
  C.ortho = ortho ? T(1) : T(0);
  return C;
}

// radial functor coefficients (BsplineFunctor.h:93-130)
template <typename T>
Functor<T> make_functor(int np, double rcut, double cusp, const double* p) {
  Functor<T> f;
  std::memset(&f, 0, sizeof(f));
  const double dr = rcut / (double)(np + 1);
  std::vector<double> c(np + 4, 0.0);
  c[1] = p[0];
  c[2] = p[1];
  c[0] = p[1] - 2.0 * dr * cusp;
  for (int i = 2; i < np; ++i) c[i + 1] = p[i];
  f.rcut = (T)rcut;
  f.drinv = (T)(1.0 / dr);
  f.ncoef = np + 4;
  // interval i (r / dr in [i, i + 1), i <= np): the power form of the cubic B-spline basis over
  // c_i .. c_i+3 (BsplineFunctor.h:240-300 evaluates the same spline)
  for (int i = 0; i <= np; ++i) {
    const double c0 = c[i], c1 = c[i + 1], c2 = c[i + 2], c3 = c[i + 3];
    f.poly[i][0] = (T)((c0 + 4.0 * c1 + c2) / 6.0);
    f.poly[i][1] = (T)(0.5 * (c2 - c0));
    f.poly[i][2] = (T)(0.5 * (c0 - 2.0 * c1 + c2));
    f.poly[i][3] = (T)((-c0 + 3.0 * c1 - 3.0 * c2 + c3) / 6.0);
  }
  return f;
}

}  // namespace

// ====================================================================== handles
struct qmcx_system_s {
  int dev = 0;
  double lattice[9];
  int nions = 0, nspecies = 0, nup = 0, ndn = 0, N = 0, n = 0;
  double* d_ions = nullptr;
  int* d_species = nullptr;
  Functor<double> fnd[kMaxSpecies + 2];
  Functor<float> fnf[kMaxSpecies + 2];
  Functor<double>* d_fnd = nullptr;
  Functor<float>* d_fnf = nullptr;
  bool fn_dirty = true;
  // three-body Jastrow: functor of (ion species, spin pair) at [2 * species + pair] in both
  // precisions, the cutoff radius each species was given (0: none yet), ion-mask words
  Poly3<double> j3d[2 * kMaxSpecies];
  Poly3<float> j3f[2 * kMaxSpecies];
  double j3rcut[kMaxSpecies] = {0, 0, 0, 0};
  Poly3<double>* d_j3d = nullptr;
  Poly3<float>* d_j3f = nullptr;
  bool has_j3 = false, j3_dirty = false;
  int mw = 0;
  int nbatches = 0;   // live batches of this system
  int ng[3] = {0, 0, 0};
  int norb = 0, npad = 0;
  double spoG[9];
  float* d_coef_f = nullptr;
  double* d_coef_d = nullptr;
  Cell<float>* d_cell_f = nullptr;
  Cell<double>* d_cell_d = nullptr;
  double* d_spoG = nullptr;
  hipStream_t stream = nullptr;   // the stand-alone kernels (qmcx_spline_vgh, qmcx_distance_rows)

  template <typename T>
  SysP<T> params() const {
    SysP<T> S;
    S.spoG = d_spoG;
    S.ng[0] = ng[0];
    S.ng[1] = ng[1];
    S.ng[2] = ng[2];
    S.N = N;
    S.n = n;
    S.M = nions;
    S.nspecies = nspecies;
    S.npad = npad;
    S.zs = npad;
    S.ys = (long long)(ng[2] + 3) * npad;
    S.xs = (long long)(ng[1] + 3) * S.ys;
    S.ions = d_ions;
    S.species = d_species;
    S.fn64 = d_fnd;
    S.cell64 = d_cell_d;
    S.mw = has_j3 ? mw : 0;
    S.j3_64 = has_j3 ? d_j3d : nullptr;
    if constexpr (sizeof(T) == 4) {
      S.coefs = d_coef_f;
      S.fn = d_fnf;
      S.cell = d_cell_f;
      S.j3 = has_j3 ? d_j3f : nullptr;
    } else {
      S.coefs = d_coef_d;
      S.fn = d_fnd;
      S.cell = d_cell_d;
      S.j3 = has_j3 ? d_j3d : nullptr;
    }
    return S;
  }
  int sync_functors() {
    if (j3_dirty) {
      HIPCHK(hipMemcpy(d_j3d, j3d, sizeof(j3d), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(d_j3f, j3f, sizeof(j3f), hipMemcpyHostToDevice));
      j3_dirty = false;
    }
    if (!fn_dirty) return 0;
    HIPCHK(hipMemcpy(d_fnd, fnd, sizeof(fnd), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_fnf, fnf, sizeof(fnf), hipMemcpyHostToDevice));
    fn_dirty = false;
    return 0;
  }
};


// the parameters a captured run depends on
struct GraphKey {
  int nsweeps;
  uint64_t seed;
  double tau;
  int mode, recompute_every;
  bool operator==(const GraphKey& o) const {
    return nsweeps == o.nsweeps && seed == o.seed && tau == o.tau && mode == o.mode &&
           recompute_every == o.recompute_every;
  }
};

struct qmcx_batch_s {
  qmcx_system_s* sys = nullptr;
  int W = 0, K = 1, prec = QMCX_FP64, walker0 = 0, slices = 0, nt = 0;
  int nch = 1;   // delayed-update state slots per walker and determinant: one per chunk
  int flush_kind = 0;   // 0: matrix cores when possible, 1: vector FMA, 2: matrix cores, U rows from HBM
  // streamed form of the delayed updates (stale rows and capacitance inverse in HBM, a flush launch
  // per chunk): needed when the LDS-resident form does not fit the delay rank (st_need), or asked
  // for (st_ask); flex_pending: flex_* moves left delayed rows pending in state slot 0
  bool st_need = false, st_ask = false, flex_pending = false;
  bool streamed() const { return K > 1 && (st_need || st_ask); }
  // FP64 recompute every `cadence` accepted rows of a walker's determinant (0: off): per walker and
  // determinant its position in the sweep and accepted rows since the last recompute, cad [W][2][2];
  // the walkers a round stopped for their recompute, cad_flag [W]
  int cadence = 0;
  int* cad = nullptr;
  int* cad_flag = nullptr;
  // FP32 Metropolis: near-tie band of the FP64 re-evaluation (0: off).  While it is on the batch
  // keeps FP64 Slater rows (A64, spo64), which the re-evaluation and the FP64 refresh read;
  // rows64_stale until an evaluateLog has filled them
  double band = 0.0;
  bool rows64 = false, rows64_stale = false;
  bool fp64_rows = false;   // FP64 Slater rows asked for on their own (rows64 also follows the band)
  double* flex_displ = nullptr;   // [W][3] proposals of the pending flex move
  int* flex_decide = nullptr;     // [W]
  int flex_iel = -1;
  double* vr_pos = nullptr;       // [W][vr_cap][3] virtual positions of qmcx_flex_evaluate_ratios
  double* vr_out = nullptr;       // [W][vr_cap] their ratios
  int vr_cap = 0;
  double* nl_out = nullptr;       // [W][N][nl_cap][12] NonLocalPP ratios
  int* nl_cnt = nullptr;          // [W][N] electron-ion pairs inside rmax
  double* nl_near = nullptr;      // [W][N][nl_cap][4] listed pairs: distance, displacement
  int* nl_items = nullptr;        // [W*N*nl_cap] work list of listed pairs, then its length
  int nl_cap = 0;
  double* nl_sph = nullptr;       // [W][12][3] each walker's rotated quadrature knots
  int nl_unit = 0;                // knots per NonLocalPP work unit (0: by size and precision)
  hipStream_t stream = nullptr;
  // FP64 recompute stream: a determinant's refresh runs there, on the CUs the move
  // kernel leaves idle, while the other determinant's electrons move
  hipStream_t aux = nullptr;
  hipEvent_t ev_done[2] = {nullptr, nullptr}, ev_fresh[2] = {nullptr, nullptr}, ev_gl = nullptr;
  // proposals stream: the SPO rows of a determinant's next proposals are evaluated there
  // as soon as its electrons have moved, beside the other determinant's moves
  hipStream_t pre = nullptr;
  hipEvent_t ev_pre[2] = {nullptr, nullptr}, ev_start = nullptr;
  bool pending_dn = false;
  double* ld_fresh = nullptr;   // [W] log|det| and phase of the refreshed second determinant
  double* ph_fresh = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.f;
  std::vector<void*> allocs;
  double* R = nullptr;
  void* A = nullptr;
  void* Ainv = nullptr;
  void* dpsiM = nullptr;
  void* d2psiM = nullptr;
  void* J2 = nullptr;
  void* J1 = nullptr;
  void* Vbuf = nullptr;
  void* Binv = nullptr;
  void* spo = nullptr;   // [W][2][n][5][npad] SPO rows at the sweep's proposals
  // three-body Jastrow state (systems with one): sums [W][5][N], ion masks [W][N][mw], near-ion
  // records of the current positions [W][N][kJ3Rec] and of the proposals [W][2][n][kJ3Rec]
  void* J3 = nullptr;
  unsigned long long* j3mask = nullptr;
  void* j3near = nullptr;
  void* j3prop = nullptr;
  bool j3 = false;
  bool counted = false;   // counted in its system's live batches
  // graph replay of whole runs (qmcx_batch_set_graph): the instantiated graph of the last run
  // and the run parameters it was captured with
  bool graph_on = false;
  hipGraphExec_t gexec = nullptr;
  GraphKey gkey{};
  unsigned epoch = 0;   // counts setting changes (a group's graph compares it)
  void drop_graph() {
    if (gexec) hipGraphExecDestroy(gexec);
    gexec = nullptr;
    ++epoch;
  }
  int* flags = nullptr;   // [1] kernel error bits
  double* A64 = nullptr;   // FP32 form: [W][2][n][n] FP64 Slater rows (the FP64 refresh inverts these)
  double* spo64 = nullptr; // FP32 form: [W][2][n][npad] FP64 values at the sweep's proposals
  double* prop = nullptr; // [W][2][n][kProp] proposal records of the sweep
  int* rows = nullptr;
  int* ndelay = nullptr;
  double* logdet = nullptr;
  double* phase = nullptr;
  double* jlog = nullptr;
  unsigned long long* nacc = nullptr;
  unsigned long long* narb = nullptr;
  double* trace = nullptr;
  int* sweep = nullptr;
  double* G = nullptr;
  double* L = nullptr;
  double* logv = nullptr;
  double* hist = nullptr;
  int hist_cap = 0;
  double* lu_ws = nullptr;
  bool trace_on = false;
  // optional per-kernel timing: HIP events bracket every move / flush launch
  bool prof_on = false;
  std::vector<hipEvent_t> prof_ev;
  std::vector<int> prof_kind;   // 0 = move kernel, 1 = flush kernel
  size_t prof_used = 0;
  double prof_ms[5] = {0, 0, 0, 0, 0};   // move, flush, FP64 recompute, evaluateGL, proposals
  long long prof_n[5] = {0, 0, 0, 0, 0};
  int prof_record(int kind, hipStream_t st = nullptr) {
    if (!prof_on) return 0;
    if (prof_used + 2 > prof_ev.size()) {
      for (int i = 0; i < 64; ++i) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        prof_ev.push_back(e);
        prof_kind.push_back(0);
      }
    }
    prof_kind[prof_used] = kind;
    HIPCHK(hipEventRecord(prof_ev[prof_used++], st ? st : stream));
    return 0;
  }
  int prof_close(hipStream_t st = nullptr) {
    if (!prof_on) return 0;
    HIPCHK(hipEventRecord(prof_ev[prof_used++], st ? st : stream));
    return 0;
  }
  int prof_collect() {
    if (!prof_on) return 0;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipStreamSynchronize(aux));
    HIPCHK(hipStreamSynchronize(pre));
    for (size_t i = 0; i + 1 < prof_used; i += 2) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, prof_ev[i], prof_ev[i + 1]));
      prof_ms[prof_kind[i]] += ms;
      prof_n[prof_kind[i]] += 1;
    }
    prof_used = 0;
    return 0;
  }

  template <typename T>
  BatchP<T> params() const {
    BatchP<T> B;
    B.W = W;
    B.K = K;
    B.nch = nch;
    B.R = R;
    B.A = (T*)A;
    B.Ainv = (T*)Ainv;
    B.dpsiM = (T*)dpsiM;
    B.d2psiM = (T*)d2psiM;
    B.J2 = (T*)J2;
    B.J1 = (T*)J1;
    B.Vbuf = (T*)Vbuf;
    B.Binv = (double*)Binv;
    B.rows = rows;
    B.ndelay = ndelay;
    B.logdet = logdet;
    B.phase = phase;
    B.jlog = jlog;
    B.nacc = nacc;
    B.trace = trace_on ? trace : nullptr;
    B.sweep = sweep;
    B.A64 = sizeof(T) == 4 && rows64 ? A64 : nullptr;
    B.spo64 = sizeof(T) == 4 && rows64 ? spo64 : nullptr;
    B.narb = narb;
    B.J3 = (T*)J3;
    B.j3mask = j3mask;
    B.j3near = (T*)j3near;
    B.j3prop = (T*)j3prop;
    B.flags = flags;
    return B;
  }
  template <typename P>
  int alloc(P*& p, size_t count) {
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(P)));
    HIPCHK(hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(P)));
    allocs.push_back(q);
    p = (P*)q;
    return 0;
  }
};

namespace {

// workgroup size of the move kernel: a lane per orbital, at most two electrons per lane for
// the J2 rows, at most kMoveThreads lanes; a whole number of waves per SIMD (multiples of 256
// lanes), so the four SIMDs of the CU share the J2 rows and the in-kernel flush evenly
int choose_shape(const qmcx_system_s* s, int& nt) {
  const int need = std::max(s->n, (s->N + 1) / 2);
  nt = std::max(256, (need + 255) / 256 * 256);
  if (nt > kMoveThreads) return fail(QMCX_ERR_ARG, "workgroup too large for this system");
  return 0;
}

// spline slices (16/slices stencil columns per lane) of the proposals kernel.  One by default:
// each lane sums all 64 stencil points of its orbital quad and stores its outputs from
// registers, so the workgroups need no LDS for partials and fit beside the move kernel's and
// the FP64 refresh's workgroups on their CUs, using bandwidth those leave idle.
template <typename T>
int proposal_slices(const qmcx_system_s* s, int want) {
  (void)s;
  return want > 0 ? want : 1;
}

template <typename T, int JPT, bool KD, bool ARB, bool J3, bool ST>
int launch_moves_kd(qmcx_batch_t b, int nt, size_t lds, const MoveArgs& a) {
  const qmcx_system_s* s = b->sys;
  hipLaunchKernelGGL((k_moves<T, JPT, KD, ARB, J3, ST>), dim3(b->W), dim3(nt), lds, b->stream, s->params<T>(),
                     b->params<T>(), a);
  HIPCHK(hipGetLastError());
  return 0;
}

// the near-tie FP64 re-evaluation exists for FP32 delayed updates under Metropolis acceptance
// (the kernel with ARB); ST: the streamed form of the delayed updates
template <typename T>
bool arb_launch(const qmcx_batch_t b, int mode, double band, const void* decide) {
  return sizeof(T) == 4 && b->K > 1 && mode == QMCX_ACCEPT_METROPOLIS && band > 0.0 && decide == nullptr;
}
template <typename T, int JPT, bool ST>
int launch_moves_jpt(qmcx_batch_t b, int nt, size_t lds, const MoveArgs& a) {
  if (b->K == 1)
    return b->j3 ? launch_moves_kd<T, JPT, false, false, true, false>(b, nt, lds, a)
                 : launch_moves_kd<T, JPT, false, false, false, false>(b, nt, lds, a);
  if constexpr (sizeof(T) == 4)
    if (arb_launch<T>(b, a.mode, a.band, a.decide))
      return b->j3 ? launch_moves_kd<T, JPT, true, true, true, ST>(b, nt, lds, a)
                   : launch_moves_kd<T, JPT, true, true, false, ST>(b, nt, lds, a);
  return b->j3 ? launch_moves_kd<T, JPT, true, false, true, ST>(b, nt, lds, a)
               : launch_moves_kd<T, JPT, true, false, false, ST>(b, nt, lds, a);
}

template <typename T>
int launch_moves(qmcx_batch_t b, int nt, size_t lds, const MoveArgs& a, bool st) {
  if (st)
    return (b->sys->N <= nt) ? launch_moves_jpt<T, 1, true>(b, nt, lds, a) : launch_moves_jpt<T, 2, true>(b, nt, lds, a);
  return (b->sys->N <= nt) ? launch_moves_jpt<T, 1, false>(b, nt, lds, a) : launch_moves_jpt<T, 2, false>(b, nt, lds, a);
}

// SPO rows at the proposals of electrons det*n + [e0, e0 + ne) of every walker
constexpr int kPropWorkgroups = 1024;
// virtual positions per walker in one qmcx_flex_evaluate_ratios call (grid x dimension)
constexpr int kMaxVirtualPoints = 1 << 16;
// electron-ion pairs per electron held by qmcx_nlpp_evaluate
constexpr int kMaxNlppPairs = 64;
// workgroups of the NonLocalPP ratio launch (they walk the pair list)
constexpr int kNlppWorkgroups = 8192;

template <typename T>
int launch_proposals(qmcx_batch_t b, hipStream_t st, int det, int e0, int ne, uint64_t seed, double tau,
                     const double* displ) {
  const qmcx_system_s* s = b->sys;
  const int nq = s->npad / (16 / (int)sizeof(T));
  const int slices = proposal_slices<T>(s, b->slices);
  const int nt = std::max(64, (nq * slices + 63) / 64 * 64);
  if (nt > 1024) return fail(QMCX_ERR_ARG, "the orbital table is too wide for one proposals workgroup");
  const bool v64 = sizeof(T) == 4 && b->rows64;   // BatchP::spo64 is set exactly then
  // LDS: the J1 wave sums, the slice partials and the FP64 values of the FP32 form's FP64 rows
  // (neither with one slice: its lanes store from registers)
  size_t lds = sizeof(double) * 16 * 5;
  if (slices != 1) lds = std::max(lds, sizeof(T) * (size_t)slices * 5 * s->npad);
  if (v64 && slices != 1)
    lds = std::max(lds, ((sizeof(T) * (size_t)slices * 5 * s->npad + 15) & ~(size_t)15) +
                            sizeof(double) * (size_t)slices * s->npad);
  // about 1024 workgroups (four per CU), each looping over its electrons: the gather's burst of
  // HBM traffic stays below what slows the latency-bound move kernel beside it
  const int gx = std::max(1, std::min(ne, kPropWorkgroups / b->W));
  auto kern = [&](auto kslc, auto kv64) {
    hipLaunchKernelGGL((k_spo_proposals<T, decltype(kslc)::value, decltype(kv64)::value>), dim3(gx, b->W), dim3(nt),
                       lds, st, s->params<T>(), b->params<T>(), det, e0, ne, seed, b->sweep, std::sqrt(tau),
                       b->walker0, displ, (T*)b->spo, b->prop, slices);
  };
  auto by_slices = [&](auto kv64) {
    switch (slices) {
      case 1: kern(std::integral_constant<int, 1>{}, kv64); break;
      case 2: kern(std::integral_constant<int, 2>{}, kv64); break;
      case 4: kern(std::integral_constant<int, 4>{}, kv64); break;
      case 8: kern(std::integral_constant<int, 8>{}, kv64); break;
      default: kern(std::integral_constant<int, 16>{}, kv64); break;
    }
  };
  if (v64)
    by_slices(std::true_type{});
  else
    by_slices(std::false_type{});
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T>
int spline_slices_for(const qmcx_system_s* s, int nt) {
  const int vw = 16 / (int)sizeof(T);
  const int nq = s->npad / vw;
  int S = 1;
  while (S < 16 && nq * S * 2 <= nt) S *= 2;
  return S;
}

template <typename T>
int launch_spo_rows(qmcx_batch_t b) {
  const qmcx_system_s* s = b->sys;
  const int vw = 16 / (int)sizeof(T);
  const int nq = s->npad / vw;
  const int slices = spline_slices_for<T>(s, 1024);
  const int nt = std::max(64, (nq * slices + 63) / 64 * 64);
  size_t lds = sizeof(T) * ((size_t)slices * 5 * s->npad + 5 * (size_t)s->n);
  if (sizeof(T) == 4) lds = ((lds + 15) & ~(size_t)15) + sizeof(double) * (size_t)slices * s->npad;   // FP64 values
  hipLaunchKernelGGL(k_spo_rows<T>, dim3(s->N, b->W), dim3(nt), lds, b->stream, s->params<T>(), b->params<T>(),
                     slices);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T>
int launch_lu_batch(qmcx_batch_t b) {
  const qmcx_system_s* s = b->sys;
  const int n = s->n;
  if (sizeof(T) == 4 && b->rows64)   // the FP32 form with FP64 Slater rows inverts those
    HIPCHK((launch_gj_invert<double, T>(n, b->W * 2, b->stream, b->A64, (T*)b->Ainv, b->lu_ws, b->logdet, b->phase, 1)));
  else
    HIPCHK((launch_gj_invert<T, T>(n, b->W * 2, b->stream, (const T*)b->A, (T*)b->Ainv, b->lu_ws, b->logdet, b->phase, 1)));
  return 0;
}

template <typename T>
int launch_gl(qmcx_batch_t b, bool record) {
  const qmcx_system_s* s = b->sys;
  hipLaunchKernelGGL(k_det_gl<T>, dim3((s->N + 3) / 4, b->W), dim3(256), 0, b->stream, s->params<T>(),
                     b->params<T>(), b->G, b->L);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_finish_gl<T>, dim3(b->W), dim3(256), 0, b->stream, s->params<T>(), b->params<T>(), b->G, b->L,
                     b->logv, record ? b->hist : nullptr, b->hist_cap);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T>
int evaluate_log_impl(qmcx_batch_t b) {
  const qmcx_system_s* s = b->sys;
  if (launch_spo_rows<T>(b)) return QMCX_ERR_HIP;
  if (launch_lu_batch<T>(b)) return QMCX_ERR_HIP;
  // positions, reduction scratch, functors, the minimum-image cell and the J3 ion masks
  const size_t lds = jastrow_init_lds<T>(s->N, s->has_j3 ? s->mw : 0);
  hipLaunchKernelGGL(k_jastrow_init<T>, dim3(b->W), dim3(1024), lds, b->stream, s->params<T>(), b->params<T>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemsetAsync(b->ndelay, 0, sizeof(int) * b->W * 2 * b->nch, b->stream));
  return launch_gl<T>(b, false);
}

// the matrix-core flush needs FP32 and n a multiple of 16 (row blocks of 32 are
// clipped); it can be switched off for comparisons
template <typename T>
bool use_mfma_flush(const qmcx_batch_t b) {
  return sizeof(T) == 4 && flush_mfma_fits(b->sys->n) && b->flush_kind != 1;
}

// Electrons per chunk of a move launch: the delay rank (one flush per chunk), or up to 64 with
// Sherman-Morrison updates (K = 1; 16 with a three-body Jastrow), which bounds the chunk's LDS
// records.
static int move_chunk(int n, int K, bool j3) { return K == 1 ? std::min(n, j3 ? 16 : 64) : K; }

// updateInvMat of the delayed rows of state slots [j0, j1) of determinant det, in order; thr > 1:
// only the slots holding at least thr rows
template <typename T>
int launch_flush(qmcx_batch_t b, hipStream_t st, int det, int j0, int j1, int thr = 1) {
  const qmcx_system_s* s = b->sys;
  const FlushRows FR{0, s->n, thr};
  if (use_mfma_flush<T>(b)) {
    const size_t lds = flush_mfma_lds(s->n, b->K);
    if (flush_t_tiles(b->K) == 1)
      hipLaunchKernelGGL(k_flush_mfma<true>, dim3((s->n + kFlushBM - 1) / kFlushBM, b->W), dim3(64 * kFlushWaves), lds,
                         st, s->params<float>(), b->params<float>(), det, j0, j1, FR);
    else
      hipLaunchKernelGGL(k_flush_mfma<false>, dim3((s->n + kFlushBM - 1) / kFlushBM, b->W), dim3(64 * kFlushWaves),
                         lds, st, s->params<float>(), b->params<float>(), det, j0, j1, FR);
    HIPCHK(hipGetLastError());
  } else {
    const size_t lds = sizeof(T) * 2 * kFlushRows * b->K + sizeof(int) * b->K;
    for (int j = j0; j < j1; ++j) {
      hipLaunchKernelGGL(k_flush_generic<T>, dim3((s->n + kFlushRows - 1) / kFlushRows, b->W), dim3(256), lds, st,
                         s->params<T>(), b->params<T>(), det, j, FR);
      HIPCHK(hipGetLastError());
    }
  }
  return 0;
}

template <typename T>
int invert_det(qmcx_batch_t b, int det, bool into_fresh, hipStream_t st, const int* only);

// The moves of one determinant's electrons in one sweep.  With K > 1 each chunk's delayed
// updates (recorded in its state slot) reach the whole inverse before the next chunk: inside
// the move kernel when its LDS holds the flush (FP32, n a multiple of 16 up to 384), else by a
// flush launch after the chunk's move launch.  The last chunk is flushed only when `flush` is
// set; otherwise the inverse is left for the FP64 refresh to rebuild from the Slater matrix.
template <typename T>
int moves_det(qmcx_batch_t b, int det, int mode, bool flush) {
  const qmcx_system_s* s = b->sys;
  int nt = 0;
  if (choose_shape(s, nt)) return QMCX_ERR_ARG;
  const int n = s->n;
  const int chunk = move_chunk(n, b->K, b->j3);
  const bool st = b->streamed();
  // the launch takes the FP64 re-evaluation's kernel (launch_moves_jpt) exactly when arb3 holds for
  // a three-body Jastrow: its LDS layout then keeps the re-evaluation apart from the J3 pair slots
  const bool arb3 = b->j3 && arb_launch<T>(b, mode, b->rows64 ? b->band : 0.0, nullptr);
  const MoveLds L = move_lds<T>(s->N, s->nions, n, b->K, chunk, b->j3 ? s->mw : 0, st, arb3);
  MoveArgs a;
  std::memset(&a, 0, sizeof(a));
  a.mode = mode;
  a.slices = 0;
  a.spo = b->spo;
  a.prop = b->prop;
  a.decide = nullptr;
  a.band = b->rows64 ? b->band : 0.0;
  const int nchunks = (n + chunk - 1) / chunk;
  a.det = det;
  a.chunk_len = chunk;
  a.inflush = 0;
  a.flush_u = 0;
  // one launch for all the chunks when no flush launch is needed between them: Sherman-Morrison
  // moves, or each chunk's updates applied by the move kernel itself
  const bool inflush = b->K > 1 && !st && use_mfma_flush<T>(b) && flush_own_fits<T>(s->N, n, b->K, chunk);
  if (b->cadence > 0) {
    // FP64 recompute every b->cadence accepted rows of a walker: rounds of one move launch (each
    // walker from its own position until its cadence or the determinant's end) and one FP64
    // inversion of the walkers that stopped; a walker stops at most n / cadence times
    if (!(b->K == 1 || inflush))
      return fail(QMCX_ERR_STATE, "the recompute cadence needs Sherman-Morrison moves or delayed updates applied "
                                  "inside the move kernel");
    hipLaunchKernelGGL(k_cadence_begin, dim3((b->W + 255) / 256), dim3(256), 0, b->stream, b->cad, b->W, det);
    HIPCHK(hipGetLastError());
    a.e_begin = det * n;
    a.e_end = det * n + n;
    a.chunk = 0;
    a.nchunks = n;   // bounded by the determinant's end
    a.inflush = inflush ? (flush ? 2 : 1) : 0;
    a.flush_u = b->flush_kind == 2 ? 0 : kFlushOwnU;
    a.cadence = b->cadence;
    a.cad = b->cad;
    a.cad_flag = b->cad_flag;
    for (int round = 0; round <= n / b->cadence; ++round) {
      if (b->prof_record(0)) return QMCX_ERR_HIP;
      if (launch_moves<T>(b, nt, L.total, a, false)) return QMCX_ERR_HIP;
      if (b->prof_close()) return QMCX_ERR_HIP;
      if (b->prof_record(2)) return QMCX_ERR_HIP;
      if (int rc = invert_det<T>(b, det, false, b->stream, b->cad_flag)) return rc;
      if (b->prof_close()) return QMCX_ERR_HIP;
    }
    return 0;
  }
  if (b->K == 1 || inflush) {
    a.e_begin = det * n;
    a.e_end = det * n + n;
    a.chunk = 0;
    a.nchunks = nchunks;
    a.inflush = inflush ? (flush ? 2 : 1) : 0;
    a.flush_u = b->flush_kind == 2 ? 0 : kFlushOwnU;
    if (b->prof_record(0)) return QMCX_ERR_HIP;
    if (launch_moves<T>(b, nt, L.total, a, false)) return QMCX_ERR_HIP;
    if (b->prof_close()) return QMCX_ERR_HIP;
    return 0;
  }
  a.nchunks = 1;
  for (int ci = 0; ci < nchunks; ++ci) {
    const int e = ci * chunk;
    a.e_begin = det * n + e;
    a.e_end = det * n + std::min(n, e + chunk);
    a.chunk = ci;
    if (b->prof_record(0)) return QMCX_ERR_HIP;
    if (launch_moves<T>(b, nt, L.total, a, st)) return QMCX_ERR_HIP;
    if (b->prof_close()) return QMCX_ERR_HIP;
    if (b->K > 1 && (ci + 1 < nchunks || flush)) {
      if (b->prof_record(1)) return QMCX_ERR_HIP;
      if (int rc = launch_flush<T>(b, b->stream, det, ci, ci + 1)) return rc;
      if (b->prof_close()) return QMCX_ERR_HIP;
    }
  }
  return 0;
}

// The delayed rows flex_* moves left pending (state slot 0 of each determinant) reach the inverse
// (DiracDeterminant::completeUpdates, DiracDeterminant.cpp:294-321), or, with thr > 1, only those of
// the walkers whose pending rows reached thr; the applied slots are emptied.
template <typename T>
int flush_pending_t(qmcx_batch_t b, int thr) {
  for (int det = 0; det < 2; ++det) {
    if (int rc = launch_flush<T>(b, b->stream, det, 0, 1, thr)) return rc;
    hipLaunchKernelGGL(k_reset_delay<T>, dim3((b->W + 255) / 256), dim3(256), 0, b->stream, b->params<T>(), det, thr);
    HIPCHK(hipGetLastError());
  }
  return 0;
}
int flush_pending(qmcx_batch_t b, int thr = 1) {
  if (!b->flex_pending) return 0;
  const int rc = b->prec == QMCX_FP32 ? flush_pending_t<float>(b, thr) : flush_pending_t<double>(b, thr);
  if (rc) return rc;
  if (thr <= 1) b->flex_pending = false;
  return 0;
}

// FP64 inverse of one determinant of every walker (only: of the walkers whose flag is set) from
// its Slater rows on stream st: the inverse in place; log|det| and phase into the batch arrays
// (into_fresh false) or into ld_fresh / ph_fresh
template <typename T>
int invert_det(qmcx_batch_t b, int det, bool into_fresh, hipStream_t st, const int* only) {
  const int n = b->sys->n;
  const size_t nn = (size_t)n * n;
  double* ld = into_fresh ? b->ld_fresh : b->logdet + det;
  double* ph = into_fresh ? b->ph_fresh : b->phase + det;
  const int lds = into_fresh ? 1 : 2;
  if (sizeof(T) == 4 && b->rows64)   // the FP32 form with FP64 Slater rows inverts those
    HIPCHK((launch_gj_invert<double, T>(n, b->W, st, b->A64 + det * nn, (T*)b->Ainv + det * nn,
                                        b->lu_ws + det * 2 * nn, ld, ph, lds, 2, only)));
  else
    HIPCHK((launch_gj_invert<T, T>(n, b->W, st, (const T*)b->A + det * nn, (T*)b->Ainv + det * nn,
                                   b->lu_ws + det * 2 * nn, ld, ph, lds, 2, only)));
  return 0;
}

// FP64 refresh of one determinant of every walker on the recompute stream (it also restarts the
// count of the recompute cadence)
template <typename T>
int refresh_det(qmcx_batch_t b, int det, bool into_fresh) {
  if (b->prof_record(2, b->aux)) return QMCX_ERR_HIP;
  if (int rc = invert_det<T>(b, det, into_fresh, b->aux, nullptr)) return rc;
  if (b->prof_close(b->aux)) return QMCX_ERR_HIP;
  if (b->cad) {
    hipLaunchKernelGGL(k_cadence_reset, dim3((b->W + 255) / 256), dim3(256), 0, b->aux, b->cad, b->W, det);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

// Side-stream holds (microseconds) after the event that releases the side work together with a
// move launch: the FP64 refresh, and the proposals of each determinant (the second's are released
// before evaluateGL, so they wait longer).
#ifndef QMCX_HOLD_REFRESH_US
#define QMCX_HOLD_REFRESH_US 30
#endif
#ifndef QMCX_HOLD_PROP0_US
#define QMCX_HOLD_PROP0_US 100
#endif
#ifndef QMCX_HOLD_PROP1_US
#define QMCX_HOLD_PROP1_US 500
#endif
int hold_stream(hipStream_t st, int us) {
  if (us <= 0) return 0;
  hipLaunchKernelGGL(k_hold, dim3(1), dim3(64), 0, st, (unsigned long long)us * 100ull);
  HIPCHK(hipGetLastError());
  return 0;
}

// SPO rows of determinant det's proposals on the proposals stream, after the event `after`
// (positions final), announced by ev_pre[det]; `hold` when a move launch is released by the same
// event.  Their sweep number (the random streams' counter) is the device counter sweep[2], which
// only this stream reads and advances, after the second determinant's proposals.
template <typename T>
int issue_proposals(qmcx_batch_t b, int det, hipEvent_t after, uint64_t seed, double tau, bool hold = false) {
  if (after) HIPCHK(hipStreamWaitEvent(b->pre, after, 0));
  if (hold && hold_stream(b->pre, det == 0 ? QMCX_HOLD_PROP0_US : QMCX_HOLD_PROP1_US)) return QMCX_ERR_HIP;
  if (b->prof_record(4, b->pre)) return QMCX_ERR_HIP;
  if (int rc = launch_proposals<T>(b, b->pre, det, 0, b->sys->n, seed, tau, nullptr)) return rc;
  if (b->prof_close(b->pre)) return QMCX_ERR_HIP;
  if (det == 1) {
    hipLaunchKernelGGL(k_count, dim3(1), dim3(64), 0, b->pre, b->sweep + 2);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(b->ev_pre[det], b->pre));
  return 0;
}

// One sweep.  A determinant's moves wait
// for its proposals' SPO rows; as soon as they are done, the next sweep's rows for that
// determinant are evaluated on the proposals stream (unless this is the run's last sweep).
// With `recompute`, each determinant's inverse and log|det| are refreshed in FP64 on the
// recompute stream as soon as its electrons have moved: the first determinant's refresh runs
// beside the second determinant's moves and is in place before evaluateGL; the second
// determinant's runs beside the next sweep's first-determinant moves and, when it lands,
// corrects this sweep's recorded log|Psi| (k_refresh_det).  The next sweep's
// second-determinant moves wait for it.
template <typename T>
int sweep_impl(qmcx_batch_t b, uint64_t seed, double tau, int mode, bool recompute, bool last) {
  HIPCHK(hipStreamWaitEvent(b->stream, b->ev_pre[0], 0));
  if (int rc = moves_det<T>(b, 0, mode, !recompute)) return rc;
  HIPCHK(hipEventRecord(b->ev_done[0], b->stream));
  if (!last)
    if (int rc = issue_proposals<T>(b, 0, b->ev_done[0], seed, tau, true)) return rc;
  if (recompute) {
    HIPCHK(hipStreamWaitEvent(b->aux, b->ev_done[0], 0));
    if (hold_stream(b->aux, QMCX_HOLD_REFRESH_US)) return QMCX_ERR_HIP;
    if (int rc = refresh_det<T>(b, 0, false)) return rc;
    HIPCHK(hipEventRecord(b->ev_fresh[0], b->aux));
  }
  if (b->pending_dn) {
    HIPCHK(hipStreamWaitEvent(b->stream, b->ev_fresh[1], 0));
    b->pending_dn = false;
  }
  HIPCHK(hipStreamWaitEvent(b->stream, b->ev_pre[1], 0));
  if (int rc = moves_det<T>(b, 1, mode, true)) return rc;
  HIPCHK(hipEventRecord(b->ev_done[1], b->stream));
  if (!last)
    if (int rc = issue_proposals<T>(b, 1, b->ev_done[1], seed, tau, true)) return rc;
  if (recompute) HIPCHK(hipStreamWaitEvent(b->stream, b->ev_fresh[0], 0));
  if (b->prof_record(3)) return QMCX_ERR_HIP;
  if (int rc = launch_gl<T>(b, true)) return rc;
  if (b->prof_close()) return QMCX_ERR_HIP;
  if (recompute) {
    HIPCHK(hipEventRecord(b->ev_gl, b->stream));
    HIPCHK(hipStreamWaitEvent(b->aux, b->ev_gl, 0));
    if (hold_stream(b->aux, QMCX_HOLD_REFRESH_US)) return QMCX_ERR_HIP;
    if (int rc = refresh_det<T>(b, 1, true)) return rc;
    hipLaunchKernelGGL(k_refresh_det, dim3((b->W + 63) / 64), dim3(64), 0, b->aux, b->W, 1, b->ld_fresh, b->ph_fresh,
                       b->logdet, b->phase, b->logv, b->hist, b->hist_cap, b->sweep);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(b->ev_fresh[1], b->aux));
    b->pending_dn = true;
  }
  return 0;
}

// wait for a pending refresh of the second determinant (end of a run, before any other use)
int settle_refresh(qmcx_batch_t b) {
  if (b->pending_dn) {
    HIPCHK(hipStreamWaitEvent(b->stream, b->ev_fresh[1], 0));
    b->pending_dn = false;
  }
  return 0;
}

// ---- allocation-free FP64 inversion (qmcx_invert_transpose_noalloc): its Gauss-Jordan
// workspace is this library-owned device array, 32 MiB, which every GPU the library runs on
// holds from the code object's load until the process ends.  Host code never names the array
// in a launch or copy (on the host the name is only a stand-in): it uses the device address
// that hipGetSymbolAddress gives for the current device, looked up once per device.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
// largest order it takes: one matrix needs 2 n^2 doubles of it
constexpr int kMaxNoallocOrder = 1024;
static_assert(sizeof(double) * 2 * kMaxNoallocOrder * kMaxNoallocOrder <= kScratchBytes, "scratch too small");
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};

int scratch_of_current_device(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) return fail(QMCX_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

// The launches of one run of nsweeps sweeps, on the batch's three streams: the first sweep's
// proposals, the sweeps, and the wait for the last refresh.  Nothing here synchronizes with the
// host or clears memory, so the same sequence can be captured into a graph.
static int enqueue_run(qmcx_batch_t b, int nsweeps, uint64_t seed, double tau, int accept_mode, int recompute_every) {
  const bool f32 = b->prec == QMCX_FP32;
  // the first sweep's proposals, after whatever the batch stream did before this run
  HIPCHK(hipEventRecord(b->ev_start, b->stream));
  for (int det = 0; det < 2; ++det) {
    int rc = f32 ? issue_proposals<float>(b, det, b->ev_start, seed, tau)
                 : issue_proposals<double>(b, det, b->ev_start, seed, tau);
    if (rc) return rc;
  }
  for (int sw = 0; sw < nsweeps; ++sw) {
    const bool rec = recompute_every > 0 && (sw + 1) % recompute_every == 0;
    const bool last = sw + 1 == nsweeps;
    int rc = f32 ? sweep_impl<float>(b, seed, tau, accept_mode, rec, last)
                 : sweep_impl<double>(b, seed, tau, accept_mode, rec, last);
    if (rc) return rc;
    if (b->prof_on && b->prof_used > 4096 && b->prof_collect()) return QMCX_ERR_HIP;
    hipLaunchKernelGGL(k_count, dim3(1), dim3(64), 0, b->stream, b->sweep);
    HIPCHK(hipGetLastError());
  }
  return settle_refresh(b);
}

// ====================================================================== mixed-size groups
// Walkers of several batches (systems of different sizes, FP32, delayed updates) advanced together:
// one grouped launch per move stage and per proposals stage covers every batch's walkers; the FP64
// refresh runs per batch (one launch per matrix size); evaluateGL and the sweep counters per batch.
struct qmcx_group_s {
  std::vector<qmcx_batch_t> b;
  int Wtot = 0, jpt = 1, nt = 256, pnt = 64, pgx = 1;
  size_t lds = 0, plds = 0;
  hipStream_t stream = nullptr, aux = nullptr, pre = nullptr;
  hipEvent_t ev_pre[2] = {nullptr, nullptr}, ev_done[2] = {nullptr, nullptr}, ev_fresh[2] = {nullptr, nullptr};
  hipEvent_t ev_gl = nullptr, ev_start = nullptr, ev0 = nullptr, ev1 = nullptr;
  bool pending_dn = false;
  float last_ms = 0.f;
  bool graph_on = false;
  hipGraphExec_t gexec = nullptr;
  GraphKey gkey{};
  std::vector<unsigned> epochs;   // the batches' setting epochs at capture
  void drop_graph() {
    if (gexec) hipGraphExecDestroy(gexec);
    gexec = nullptr;
  }
  bool graph_current(const GraphKey& key) const {
    if (!gexec || !(gkey == key) || epochs.size() != b.size()) return false;
    for (size_t i = 0; i < b.size(); ++i)
      if (epochs[i] != b[i]->epoch) return false;
    return true;
  }
};

namespace {

// the moves of determinant det of every batch in one launch; each batch's chunks are flushed by
// the kernel, the last one only when `flush`
int group_moves(qmcx_group_s* g, int det, int mode, bool flush) {
  MoveGroup<float> G;
  std::memset(&G, 0, sizeof(G));
  G.ng = (int)g->b.size();
  int first = 0;
  for (int i = 0; i < G.ng; ++i) {
    qmcx_batch_t b = g->b[i];
    const qmcx_system_s* s = b->sys;
    G.first[i] = first;
    first += b->W;
    G.S[i] = s->params<float>();
    G.B[i] = b->params<float>();
    MoveArgs& a = G.a[i];
    a.mode = mode;
    a.spo = b->spo;
    a.prop = b->prop;
    a.det = det;
    a.chunk_len = b->K;
    a.e_begin = det * s->n;
    a.e_end = det * s->n + s->n;
    a.chunk = 0;
    a.nchunks = (s->n + b->K - 1) / b->K;
    a.inflush = flush ? 2 : 1;
    a.flush_u = b->flush_kind == 2 ? 0 : kFlushOwnU;
  }
  G.first[G.ng] = first;
  if (g->jpt == 1)
    hipLaunchKernelGGL((k_moves_group<float, 1>), dim3(g->Wtot), dim3(g->nt), g->lds, g->stream, G);
  else
    hipLaunchKernelGGL((k_moves_group<float, 2>), dim3(g->Wtot), dim3(g->nt), g->lds, g->stream, G);
  HIPCHK(hipGetLastError());
  return 0;
}

// the proposals of determinant det of every batch, on the proposals stream after `after`
int group_proposals(qmcx_group_s* g, int det, hipEvent_t after, uint64_t seed, double tau, bool hold) {
  if (after) HIPCHK(hipStreamWaitEvent(g->pre, after, 0));
  if (hold && hold_stream(g->pre, det == 0 ? QMCX_HOLD_PROP0_US : QMCX_HOLD_PROP1_US)) return QMCX_ERR_HIP;
  PropGroup<float> P;
  std::memset(&P, 0, sizeof(P));
  P.ng = (int)g->b.size();
  int first = 0;
  for (int i = 0; i < P.ng; ++i) {
    qmcx_batch_t b = g->b[i];
    P.first[i] = first;
    first += b->W;
    P.S[i] = b->sys->params<float>();
    P.B[i] = b->params<float>();
    P.spo[i] = (float*)b->spo;
    P.prop[i] = b->prop;
    P.walker0[i] = b->walker0;
    P.slices[i] = proposal_slices<float>(b->sys, b->slices);
  }
  P.first[P.ng] = first;
  if (g->b[0]->rows64)   // every batch of the group agrees (checked at each run)
    hipLaunchKernelGGL((k_spo_proposals_group<float, true>), dim3(g->pgx, g->Wtot), dim3(g->pnt), g->plds, g->pre, P,
                       det, seed, std::sqrt(tau));
  else
    hipLaunchKernelGGL((k_spo_proposals_group<float, false>), dim3(g->pgx, g->Wtot), dim3(g->pnt), g->plds, g->pre, P,
                       det, seed, std::sqrt(tau));
  HIPCHK(hipGetLastError());
  if (det == 1)
    for (qmcx_batch_t b : g->b) {
      hipLaunchKernelGGL(k_count, dim3(1), dim3(64), 0, g->pre, b->sweep + 2);
      HIPCHK(hipGetLastError());
    }
  HIPCHK(hipEventRecord(g->ev_pre[det], g->pre));
  return 0;
}

// one sweep of the group, the DAG of sweep_impl with grouped move and proposals launches
int group_sweep(qmcx_group_s* g, uint64_t seed, double tau, int mode, bool recompute, bool last) {
  HIPCHK(hipStreamWaitEvent(g->stream, g->ev_pre[0], 0));
  if (int rc = group_moves(g, 0, mode, !recompute)) return rc;
  HIPCHK(hipEventRecord(g->ev_done[0], g->stream));
  if (!last)
    if (int rc = group_proposals(g, 0, g->ev_done[0], seed, tau, true)) return rc;
  if (recompute) {
    HIPCHK(hipStreamWaitEvent(g->aux, g->ev_done[0], 0));
    if (hold_stream(g->aux, QMCX_HOLD_REFRESH_US)) return QMCX_ERR_HIP;
    for (qmcx_batch_t b : g->b)
      if (int rc = refresh_det<float>(b, 0, false)) return rc;
    HIPCHK(hipEventRecord(g->ev_fresh[0], g->aux));
  }
  if (g->pending_dn) {
    HIPCHK(hipStreamWaitEvent(g->stream, g->ev_fresh[1], 0));
    g->pending_dn = false;
  }
  HIPCHK(hipStreamWaitEvent(g->stream, g->ev_pre[1], 0));
  if (int rc = group_moves(g, 1, mode, true)) return rc;
  HIPCHK(hipEventRecord(g->ev_done[1], g->stream));
  if (!last)
    if (int rc = group_proposals(g, 1, g->ev_done[1], seed, tau, true)) return rc;
  if (recompute) HIPCHK(hipStreamWaitEvent(g->stream, g->ev_fresh[0], 0));
  for (qmcx_batch_t b : g->b)
    if (int rc = launch_gl<float>(b, true)) return rc;
  if (recompute) {
    HIPCHK(hipEventRecord(g->ev_gl, g->stream));
    HIPCHK(hipStreamWaitEvent(g->aux, g->ev_gl, 0));
    if (hold_stream(g->aux, QMCX_HOLD_REFRESH_US)) return QMCX_ERR_HIP;
    for (qmcx_batch_t b : g->b) {
      if (int rc = refresh_det<float>(b, 1, true)) return rc;
      hipLaunchKernelGGL(k_refresh_det, dim3((b->W + 63) / 64), dim3(64), 0, g->aux, b->W, 1, b->ld_fresh, b->ph_fresh,
                         b->logdet, b->phase, b->logv, b->hist, b->hist_cap, b->sweep);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(g->ev_fresh[1], g->aux));
    g->pending_dn = true;
  }
  for (qmcx_batch_t b : g->b) {
    hipLaunchKernelGGL(k_count, dim3(1), dim3(64), 0, g->stream, b->sweep);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

int enqueue_group_run(qmcx_group_s* g, int nsweeps, uint64_t seed, double tau, int mode, int recompute_every) {
  HIPCHK(hipEventRecord(g->ev_start, g->stream));
  for (int det = 0; det < 2; ++det)
    if (int rc = group_proposals(g, det, g->ev_start, seed, tau, false)) return rc;
  for (int sw = 0; sw < nsweeps; ++sw) {
    const bool rec = recompute_every > 0 && (sw + 1) % recompute_every == 0;
    if (int rc = group_sweep(g, seed, tau, mode, rec, sw + 1 == nsweeps)) return rc;
  }
  if (g->pending_dn) {
    HIPCHK(hipStreamWaitEvent(g->stream, g->ev_fresh[1], 0));
    g->pending_dn = false;
  }
  return 0;
}

}  // namespace

// ====================================================================== C ABI
// A batch keeps three-body Jastrow state only when its system had one at its creation.
static int check_j3(const qmcx_batch_t b) {
  if (b->sys->has_j3 != b->j3)
    return fail(QMCX_ERR_STATE, "the three-body Jastrow was set after the batch was created: create it again");
  return 0;
}

// The kernels' error bits since the last check (the call that reads them has synchronized).
static int check_flags(qmcx_batch_t b) {
  int f = 0;
  HIPCHK(hipMemcpy(&f, b->flags, sizeof(int), hipMemcpyDeviceToHost));
  if (f == 0) return 0;
  HIPCHK(hipMemset(b->flags, 0, sizeof(int)));
  if (f & kFlagJ3Near)
    return fail(QMCX_ERR_STATE, "an electron position has more than " + std::to_string(kJ3Near) +
                                    " ions within their three-body Jastrow cutoff");
  return fail(QMCX_ERR_STATE, "more than " + std::to_string(kJ3Slots) + " three-body Jastrow pairs in one wave of a move");
}

template <typename T>
static int launch_virtual_ratios(qmcx_batch_t b, int iel, int nvp) {
  const qmcx_system_s* s = b->sys;
  const int nq = s->npad / (16 / (int)sizeof(T));
  const int slices = proposal_slices<T>(s, b->slices);
  const int nt = std::max(64, (nq * slices + 63) / 64 * 64);
  if (nt > 1024) return fail(QMCX_ERR_ARG, "spline row does not fit a workgroup");
  const size_t lds = sizeof(T) * (size_t)slices * s->npad + sizeof(double) * 16 * 3;
  hipLaunchKernelGGL(k_virtual_ratios<T>, dim3(nvp, b->W), dim3(nt), lds, b->stream, s->params<T>(), b->params<T>(),
                     iel, nvp, b->vr_pos, b->vr_out, slices);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T, int P, int PER>
static int launch_nlpp_p(qmcx_batch_t b, double rmax, int maxp) {
  const qmcx_system_s* s = b->sys;
  const int WN = b->W * s->N;
  // nl_items: work list [WN*maxp], its length, per-key counts and cursors [M*kNlppRadialBins each],
  // pair keys [WN*maxp]
  const int nkeys = s->nions * kNlppRadialBins;
  int* count = b->nl_items + (size_t)WN * maxp;
  int* icnt = count + 1;
  int* icur = icnt + nkeys;
  int* pkey = icur + nkeys;
  hipLaunchKernelGGL(k_zero_ints, dim3(4), dim3(256), 0, b->stream, icnt, nkeys);
  hipLaunchKernelGGL(k_nlpp_pairs<T>, dim3((WN + 255) / 256), dim3(256), 0, b->stream, s->params<T>(), b->params<T>(),
                     rmax, maxp, b->nl_near, pkey, b->nl_cnt, icnt);
  hipLaunchKernelGGL(k_nlpp_offsets, dim3(1), dim3(256), 0, b->stream, icnt, icur, nkeys, count);
  hipLaunchKernelGGL(k_nlpp_scatter, dim3((WN + 255) / 256), dim3(256), 0, b->stream, WN, maxp, b->nl_cnt, pkey, icur,
                     b->nl_items);
  constexpr int SL = 16 / PER;
  const int nq = s->npad / (16 / (int)sizeof(T));
  const int nt = std::max(64, (P * SL * nq + 63) / 64 * 64);
  const size_t lds = sizeof(T) * (size_t)P * SL * s->npad + sizeof(double) * 3 * P * 16;
  // workgroups walk the work list (P knots of a pair per unit); its length is only known on the device
  const int grid = (std::max(1, std::min(WN * maxp * (kNlppKnots / P), kNlppWorkgroups)) + 7) / 8 * 8;
  if (b->j3)
    hipLaunchKernelGGL((k_nlpp_ratios<T, P, PER, true>), dim3(grid), dim3(nt), lds, b->stream, s->params<T>(),
                       b->params<T>(), b->nl_sph, maxp, b->nl_near, b->nl_items, count, b->nl_out);
  else
    hipLaunchKernelGGL((k_nlpp_ratios<T, P, PER, false>), dim3(grid), dim3(nt), lds, b->stream, s->params<T>(),
                       b->params<T>(), b->nl_sph, maxp, b->nl_near, b->nl_items, count, b->nl_out);
  HIPCHK(hipGetLastError());
  return 0;
}

// Work-unit shape: P knots of a pair per unit, PER of the 16 stencil columns per lane.  With the
// shared sphere of an earlier form, NiO-64 (384 orbitals, 128 walkers, 33k electron-ion pairs) ran
// fastest in FP32 at one knot per unit with two stencil slices, and in FP64 with the most knots
// whose value rows fit (four).  `unit` (qmcx_batch_set_nlpp_unit) forces P.
template <typename T>
static int launch_nlpp(qmcx_batch_t b, double rmax, int maxp) {
  const int nq = b->sys->npad / (16 / (int)sizeof(T));
  const int unit = b->nl_unit;
  auto fits = [&](int p, int per) { return p * (16 / per) * nq <= kNlppThreads; };
  if (unit == 0) {
    if (sizeof(T) == 4 && fits(1, 8)) return launch_nlpp_p<T, 1, 8>(b, rmax, maxp);
    if (fits(6, 16)) return launch_nlpp_p<T, 6, 16>(b, rmax, maxp);
    if (fits(4, 16)) return launch_nlpp_p<T, 4, 16>(b, rmax, maxp);
    if (fits(2, 16)) return launch_nlpp_p<T, 2, 16>(b, rmax, maxp);
    if (fits(1, 16)) return launch_nlpp_p<T, 1, 16>(b, rmax, maxp);
    return fail(QMCX_ERR_ARG, "spline row does not fit a workgroup");
  }
  switch (unit) {
    case 1:
      if (fits(1, 8)) return launch_nlpp_p<T, 1, 8>(b, rmax, maxp);
      if (fits(1, 16)) return launch_nlpp_p<T, 1, 16>(b, rmax, maxp);
      break;
    case 2:
      if (fits(2, 16)) return launch_nlpp_p<T, 2, 16>(b, rmax, maxp);
      break;
    case 4:
      if (fits(4, 16)) return launch_nlpp_p<T, 4, 16>(b, rmax, maxp);
      break;
    case 6:
      if (fits(6, 16)) return launch_nlpp_p<T, 6, 16>(b, rmax, maxp);
      break;
    default:
      break;
  }
  return fail(QMCX_ERR_ARG, "NonLocalPP work unit of " + std::to_string(unit) + " knots does not fit a workgroup");
}

extern "C" {

const char* qmcx_last_error(void) { return g_err.c_str(); }

int qmcx_device_count(int* count) {
  if (!count) return fail(QMCX_ERR_ARG, "count is null");
  HIPCHK(hipGetDeviceCount(count));
  return 0;
}

int qmcx_system_create(qmcx_system_t* out, const double* lattice9, int nions, const double* ion_pos,
                       const int* ion_species, int nspecies, int nup, int ndn) {
  if (!out || !lattice9 || nions <= 0 || !ion_pos || !ion_species) return fail(QMCX_ERR_ARG, "null argument");
  if (nspecies <= 0 || nspecies > kMaxSpecies) return fail(QMCX_ERR_ARG, "nspecies out of range");
  if (nup != ndn || nup <= 0) return fail(QMCX_ERR_ARG, "nup must equal ndn");
  for (int i = 0; i < nions; ++i)
    if (ion_species[i] < 0 || ion_species[i] >= nspecies) return fail(QMCX_ERR_ARG, "ion species out of range");
  auto* s = new qmcx_system_s;
  HIPCHK(hipGetDevice(&s->dev));
  std::memcpy(s->lattice, lattice9, sizeof(s->lattice));
  s->nions = nions;
  s->nspecies = nspecies;
  s->nup = nup;
  s->ndn = ndn;
  s->N = nup + ndn;
  s->n = nup;
  std::vector<double> soa(3 * (size_t)nions);
  for (int i = 0; i < nions; ++i)
    for (int d = 0; d < 3; ++d) soa[(size_t)d * nions + i] = ion_pos[3 * i + d];
  std::memset(s->fnd, 0, sizeof(s->fnd));
  std::memset(s->fnf, 0, sizeof(s->fnf));
  HIPCHK(hipMalloc(&s->d_ions, sizeof(double) * soa.size()));
  HIPCHK(hipMemcpy(s->d_ions, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&s->d_species, sizeof(int) * nions));
  HIPCHK(hipMemcpy(s->d_species, ion_species, sizeof(int) * nions, hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&s->d_fnd, sizeof(s->fnd)));
  HIPCHK(hipMalloc(&s->d_j3d, sizeof(s->j3d)));
  HIPCHK(hipMalloc(&s->d_j3f, sizeof(s->j3f)));
  std::memset(s->j3d, 0, sizeof(s->j3d));
  std::memset(s->j3f, 0, sizeof(s->j3f));
  HIPCHK(hipMalloc(&s->d_fnf, sizeof(s->fnf)));
  {
    const Cell<float> cf = make_cell<float>(lattice9);
    const Cell<double> cd = make_cell<double>(lattice9);
    HIPCHK(hipMalloc(&s->d_cell_f, sizeof(cf)));
    HIPCHK(hipMalloc(&s->d_cell_d, sizeof(cd)));
    HIPCHK(hipMemcpy(s->d_cell_f, &cf, sizeof(cf), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->d_cell_d, &cd, sizeof(cd), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&s->d_spoG, sizeof(double) * 9));
  }
  {
    hipStream_t st[3];
    if (int rc = device_streams(st)) return rc;
    s->stream = st[0];
  }
  *out = s;
  return 0;
}

int qmcx_system_destroy(qmcx_system_t s) {
  if (!s) return 0;
  hipFree(s->d_ions);
  hipFree(s->d_species);
  hipFree(s->d_fnd);
  hipFree(s->d_j3d);
  hipFree(s->d_j3f);
  hipFree(s->d_fnf);
  hipFree(s->d_coef_f);
  hipFree(s->d_coef_d);
  hipFree(s->d_cell_f);
  hipFree(s->d_cell_d);
  hipFree(s->d_spoG);
  delete s;   // its stream is the device's shared one
  return 0;
}

static int set_functor(qmcx_system_t s, int slot, int nparams, double rcut, double cusp, const double* params) {
  if (!s || !params) return fail(QMCX_ERR_ARG, "null argument");
  if (nparams < 2 || nparams + 4 > kMaxFunctorCoefs) return fail(QMCX_ERR_ARG, "nparams out of range");
  if (!(rcut > 0)) return fail(QMCX_ERR_ARG, "rcut must be positive");
  s->fnd[slot] = make_functor<double>(nparams, rcut, cusp, params);
  s->fnf[slot] = make_functor<float>(nparams, rcut, cusp, params);
  s->fn_dirty = true;
  return 0;
}

// PolynomialFunctor3D::resize + reset_gamma (PolynomialFunctor3D.h:65-260): the unknowns gamma(l,m,n),
// m <= l (gamma is symmetric in l, m), numbered m-major then l then n; the constraints
//   e-e no cusp, k = 0..2 neI:    sum_{l+m=k} gamma(l,m,1) = 0          (each unordered l, m once)
//   e-I no cusp, k = 0..neI+nee:  sum_{m+n=k} 3 gamma(0,m,n) - L gamma(1,m,n) = 0
// reduced to row echelon form scanning the unknowns in order (largest pivot of each column,
// columns below 1e-6 skipped); the unknowns without a pivot take the parameters in order and the
// pivot unknowns follow from their rows.  g[(l*4 + m)*4 + n], zero past the shape.
static int poly3_gamma(int neI, int nee, double rcut, const double* params, int nparams, double g[kJ3G]) {
  const double L = 0.5 * rcut;
  const int D = kJ3Deg + 1;
  std::vector<int> uid(D * D * D, -1);
  auto U = [&](int l, int m, int n) -> int& { return uid[(l * D + m) * D + n]; };
  int nu = 0;
  for (int m = 0; m <= neI; ++m)
    for (int l = m; l <= neI; ++l)
      for (int n = 0; n <= nee; ++n) U(l, m, n) = U(m, l, n) = nu++;
  const int nee_rows = 2 * neI + 1, nc = nee_rows + neI + nee + 1;
  if (nparams != nu - nc) return fail(QMCX_ERR_ARG, "a (" + std::to_string(neI) + ", " + std::to_string(nee) +
                                                         ") three-body functor takes " + std::to_string(nu - nc) +
                                                         " parameters");
  std::vector<double> A((size_t)nc * nu, 0.0);
  for (int k = 0; k <= 2 * neI; ++k)
    for (int l = k; l >= 0; --l) {
      const int m = k - l;
      if (l <= neI && m <= l) A[(size_t)k * nu + U(l, m, 1)] = l > m ? 2.0 : 1.0;
    }
  for (int k = 0; k <= neI + nee; ++k)
    for (int n = 0; n <= std::min(k, nee); ++n) {
      const int m = k - n;
      if (m > neI) continue;
      A[(size_t)(nee_rows + k) * nu + U(0, m, n)] = 3.0;
      A[(size_t)(nee_rows + k) * nu + U(1, m, n)] = -L;
    }
  std::vector<int> pc(nc);
  std::vector<char> is_pivot(nu, 0);
  int col = 0;
  for (int r = 0; r < nc; ++r) {
    int best = -1;
    for (; col < nu; ++col) {
      double bmax = 0.0;
      for (int q = r; q < nc; ++q) {
        const double v = std::fabs(A[(size_t)q * nu + col]);
        if (v > bmax) { bmax = v; best = q; }
      }
      if (bmax >= 1e-6) break;
    }
    if (col == nu) return fail(QMCX_ERR_ARG, "three-body functor constraints are rank deficient");
    for (int c = 0; c < nu; ++c) std::swap(A[(size_t)r * nu + c], A[(size_t)best * nu + c]);
    const double inv = 1.0 / A[(size_t)r * nu + col];
    for (int c = 0; c < nu; ++c) A[(size_t)r * nu + c] *= inv;
    for (int q = 0; q < nc; ++q) {
      const double f = A[(size_t)q * nu + col];
      if (q == r || f == 0.0) continue;
      for (int c = 0; c < nu; ++c) A[(size_t)q * nu + c] -= f * A[(size_t)r * nu + c];
    }
    pc[r] = col;
    is_pivot[col] = 1;
    ++col;
  }
  std::vector<double> x(nu, 0.0);
  for (int c = 0, p = 0; c < nu; ++c)
    if (!is_pivot[c]) x[c] = params[p++];
  for (int r = 0; r < nc; ++r) {
    double v = 0.0;
    for (int c = 0; c < nu; ++c)
      if (c != pc[r]) v -= A[(size_t)r * nu + c] * x[c];
    x[pc[r]] = v;
  }
  for (int l = 0; l < D; ++l)
    for (int m = 0; m < D; ++m)
      for (int n = 0; n < D; ++n)
        g[(l * D + m) * D + n] = (l <= neI && m <= neI && n <= nee) ? x[U(l, m, n)] : 0.0;
  return 0;
}

int qmcx_system_set_j3(qmcx_system_t s, int species, int pair, int neI, int nee, double rcut, const double* params,
                       int nparams) {
  if (!s || !params) return fail(QMCX_ERR_ARG, "null argument");
  if (species < 0 || species >= s->nspecies) return fail(QMCX_ERR_ARG, "species out of range");
  if (pair < 0 || pair > 1) return fail(QMCX_ERR_ARG, "pair must be 0 (same spin) or 1 (opposite spin)");
  if (neI < 1 || nee < 1 || neI > kJ3Deg || nee > kJ3Deg)
    return fail(QMCX_ERR_ARG, "three-body functor degrees must be in [1, " + std::to_string(kJ3Deg) + "]");
  if (!(rcut > 0)) return fail(QMCX_ERR_ARG, "rcut must be positive");
  if (s->j3rcut[species] != 0.0 && s->j3rcut[species] != rcut)
    return fail(QMCX_ERR_ARG, "the three-body functors of one ion species need one cutoff (ThreeBodyJastrow.h:196-218)");
  if (s->nions > kJ3MaxIons) return fail(QMCX_ERR_ARG, "a three-body Jastrow takes at most 256 ions");
  if (s->nbatches > 0) return fail(QMCX_ERR_STATE, "set the three-body Jastrow before creating batches of the system");
  double g[kJ3G];
  if (int rc = poly3_gamma(neI, nee, rcut, params, nparams, g)) return rc;
  Poly3<double>& fd = s->j3d[2 * species + pair];
  Poly3<float>& ff = s->j3f[2 * species + pair];
  fd.L = 0.5 * rcut;
  ff.L = (float)fd.L;
  for (int i = 0; i < kJ3G; ++i) {
    fd.g[i] = g[i];
    ff.g[i] = (float)g[i];
  }
  s->j3rcut[species] = rcut;
  s->has_j3 = true;
  s->j3_dirty = true;
  s->mw = (s->nions + 63) / 64;
  return 0;
}

int qmcx_system_set_j1(qmcx_system_t s, int species, int nparams, double rcut, double cusp, const double* params) {
  if (!s || species < 0 || species >= s->nspecies) return fail(QMCX_ERR_ARG, "species out of range");
  return set_functor(s, species, nparams, rcut, cusp, params);
}

int qmcx_system_set_j2(qmcx_system_t s, int pair, int nparams, double rcut, double cusp, const double* params) {
  if (!s || pair < 0 || pair > 1) return fail(QMCX_ERR_ARG, "pair must be 0 (same spin) or 1 (opposite spin)");
  return set_functor(s, kMaxSpecies + pair, nparams, rcut, cusp, params);
}

int qmcx_system_set_spline(qmcx_system_t s, const int* ng, int norb, int npad, const float* coefs,
                           const double* spo_lattice9, int precision_mask) {
  if (!s || !ng || !coefs || !spo_lattice9) return fail(QMCX_ERR_ARG, "null argument");
  if (norb != s->n) return fail(QMCX_ERR_ARG, "norb must equal the number of electrons per spin");
  if (npad < norb || npad % 4 != 0) return fail(QMCX_ERR_ARG, "npad must be >= norb and a multiple of 4");
  if (ng[0] < 1 || ng[1] < 1 || ng[2] < 1) return fail(QMCX_ERR_ARG, "grid must be positive");
  if ((precision_mask & 3) == 0) return fail(QMCX_ERR_ARG, "precision_mask selects no table");
  s->ng[0] = ng[0];
  s->ng[1] = ng[1];
  s->ng[2] = ng[2];
  s->norb = norb;
  s->npad = npad;
  inv3(spo_lattice9, s->spoG);
  HIPCHK(hipMemcpy(s->d_spoG, s->spoG, sizeof(double) * 9, hipMemcpyHostToDevice));
  const size_t cnt = (size_t)(ng[0] + 3) * (ng[1] + 3) * (ng[2] + 3) * npad;
  hipFree(s->d_coef_f);
  hipFree(s->d_coef_d);
  s->d_coef_f = nullptr;
  s->d_coef_d = nullptr;
  HIPCHK(hipMalloc(&s->d_coef_f, sizeof(float) * cnt));
  HIPCHK(hipMemcpy(s->d_coef_f, coefs, sizeof(float) * cnt, hipMemcpyHostToDevice));
  if (precision_mask & 2) {
    HIPCHK(hipMalloc(&s->d_coef_d, sizeof(double) * cnt));
    // widen on the device: FP32 values are exact in FP64
    const size_t chunk = 1 << 24;
    std::vector<double> tmp(std::min(cnt, chunk));
    for (size_t o = 0; o < cnt; o += chunk) {
      const size_t m = std::min(chunk, cnt - o);
      for (size_t i = 0; i < m; ++i) tmp[i] = (double)coefs[o + i];
      HIPCHK(hipMemcpy(s->d_coef_d + o, tmp.data(), sizeof(double) * m, hipMemcpyHostToDevice));
    }
  }
  if (!(precision_mask & 1)) {
    hipFree(s->d_coef_f);
    s->d_coef_f = nullptr;
  }
  return 0;
}

int qmcx_spline_vgh(qmcx_system_t s, int precision, int npts, const double* r, double* out) {
  if (!s || !r || !out || npts <= 0) return fail(QMCX_ERR_ARG, "bad argument");
  const bool f32 = precision == QMCX_FP32;
  if ((f32 && !s->d_coef_f) || (!f32 && !s->d_coef_d)) return fail(QMCX_ERR_STATE, "no spline table of that precision");
  const size_t nout = 10 * (size_t)npts * s->norb;
  DevBuf dr, dout;
  HIPCHK(dr.alloc(sizeof(double) * 3 * npts));
  HIPCHK(dout.alloc(sizeof(double) * nout));
  HIPCHK(hipMemcpyAsync(dr.p, r, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, s->stream));
  if (f32) {
    const int slices = spline_slices_for<float>(s, 1024);
    const int nt = std::max(64, (s->npad / 4 * slices + 63) / 64 * 64);
    hipLaunchKernelGGL(k_spline_vgh<float>, dim3(npts), dim3(nt), sizeof(float) * slices * 10 * s->npad, s->stream,
                       s->params<float>(), dr.as<double>(), dout.as<double>(), s->norb, slices);
  } else {
    const int slices = spline_slices_for<double>(s, 1024);
    const int nt = std::max(64, (s->npad / 2 * slices + 63) / 64 * 64);
    hipLaunchKernelGGL(k_spline_vgh<double>, dim3(npts), dim3(nt), sizeof(double) * slices * 10 * s->npad, s->stream,
                       s->params<double>(), dr.as<double>(), dout.as<double>(), s->norb, slices);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout.p, sizeof(double) * nout, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int qmcx_distance_rows(qmcx_system_t s, int precision, int npts, const double* src, const int* flip_ind, int nt,
                       const double* targets, double* r, double* dr) {
  if (!s || !src || !flip_ind || !targets || !r || !dr || npts <= 0 || nt <= 0) return fail(QMCX_ERR_ARG, "bad argument");
  const size_t np = (size_t)npts * nt;
  if ((np + 255) / 256 > (size_t)INT_MAX) return fail(QMCX_ERR_ARG, "npts x nt too large");
  DevBuf d_src, d_tgt, d_r, d_dr, d_flip;
  HIPCHK(d_src.alloc(sizeof(double) * 3 * npts));
  HIPCHK(d_tgt.alloc(sizeof(double) * 3 * nt));
  HIPCHK(d_r.alloc(sizeof(double) * np));
  HIPCHK(d_dr.alloc(sizeof(double) * 3 * np));
  HIPCHK(d_flip.alloc(sizeof(int) * npts));
  HIPCHK(hipMemcpyAsync(d_src.p, src, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(d_tgt.p, targets, sizeof(double) * 3 * nt, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(d_flip.p, flip_ind, sizeof(int) * npts, hipMemcpyHostToDevice, s->stream));
  const int blocks = (int)((np + 255) / 256);
  if (precision == QMCX_FP32)
    hipLaunchKernelGGL(k_distance_rows<float>, dim3(blocks), dim3(256), 0, s->stream, make_cell<float>(s->lattice),
                       npts, d_src.as<double>(), d_flip.as<int>(), nt, d_tgt.as<double>(), d_r.as<double>(),
                       d_dr.as<double>());
  else
    hipLaunchKernelGGL(k_distance_rows<double>, dim3(blocks), dim3(256), 0, s->stream, make_cell<double>(s->lattice),
                       npts, d_src.as<double>(), d_flip.as<int>(), nt, d_tgt.as<double>(), d_r.as<double>(),
                       d_dr.as<double>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(r, d_r.p, sizeof(double) * np, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(dr, d_dr.p, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int qmcx_invert_transpose(int n, int nmat, const double* a, double* psiM, double* logdet, double* phase) {
  if (n <= 0 || n > 1024 || nmat <= 0 || !a || !psiM || !logdet || !phase) return fail(QMCX_ERR_ARG, "bad argument");
  const size_t nn = (size_t)n * n * nmat;
  DevBuf d_a, d_p, d_ws, d_ld, d_ph;
  HIPCHK(d_a.alloc(sizeof(double) * nn));
  HIPCHK(d_p.alloc(sizeof(double) * nn));
  HIPCHK(d_ws.alloc(sizeof(double) * 2 * nn));
  HIPCHK(d_ld.alloc(sizeof(double) * nmat));
  HIPCHK(d_ph.alloc(sizeof(double) * nmat));
  hipStream_t st = hipStreamPerThread;
  HIPCHK(hipMemcpyAsync(d_a.p, a, sizeof(double) * nn, hipMemcpyHostToDevice, st));
  HIPCHK((launch_gj_invert<double, double>(n, nmat, st, d_a.as<double>(), d_p.as<double>(), d_ws.as<double>(),
                                          d_ld.as<double>(), d_ph.as<double>(), 1)));
  HIPCHK(hipMemcpyAsync(psiM, d_p.p, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(logdet, d_ld.p, sizeof(double) * nmat, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(phase, d_ph.p, sizeof(double) * nmat, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int qmcx_invert_transpose_noalloc(int n, int nmat, const double* d_a, double* d_psiM, double* d_logdet,
                                  double* d_phase, void* stream) {
  if (n <= 0 || nmat <= 0 || !d_a || !d_psiM || !d_logdet || !d_phase) return fail(QMCX_ERR_ARG, "bad argument");
  if (n > kMaxNoallocOrder)
    return fail(QMCX_ERR_ARG, "qmcx_invert_transpose_noalloc: n = " + std::to_string(n) + " exceeds " +
                                  std::to_string(kMaxNoallocOrder) + " (the scratch holds 16 n^2 bytes per matrix)");
  // matrices per launch: each takes 2 n^2 doubles of the scratch; launches on one stream run in
  // order, so each reuses the whole array
  const size_t per = kScratchBytes / (sizeof(double) * 2 * (size_t)n * n);
  double* ws = nullptr;
  if (int rc = scratch_of_current_device(&ws)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const size_t nn = (size_t)n * n;
  for (int m0 = 0; m0 < nmat; m0 += (int)per) {
    const int m = (int)std::min<size_t>(per, (size_t)(nmat - m0));
    HIPCHK((launch_gj_invert<double, double>(n, m, st, d_a + m0 * nn, d_psiM + m0 * nn, ws, d_logdet + m0,
                                            d_phase + m0, 1)));
  }
  return 0;
}

int qmcx_batch_create(qmcx_batch_t* out, qmcx_system_t s, int nwalkers, int delay_rank, int precision) {
  if (!out || !s || nwalkers <= 0) return fail(QMCX_ERR_ARG, "bad argument");
  if (precision != QMCX_FP32 && precision != QMCX_FP64) return fail(QMCX_ERR_ARG, "unknown precision");
  if (delay_rank < 1 || delay_rank > 128 || delay_rank > s->n) return fail(QMCX_ERR_ARG, "delay rank out of range");
  if (s->npad == 0) return fail(QMCX_ERR_STATE, "spline table not set");
  if (precision == QMCX_FP32 && !s->d_coef_f) return fail(QMCX_ERR_STATE, "no FP32 spline table");
  if (precision == QMCX_FP64 && !s->d_coef_d) return fail(QMCX_ERR_STATE, "no FP64 spline table");
  if (s->N > 2048) return fail(QMCX_ERR_ARG, "at most 2048 electrons per walker");
  if (int rc = s->sync_functors()) return rc;
  auto* b = new qmcx_batch_s;
  b->sys = s;
  b->W = nwalkers;
  b->K = delay_rank;
  b->nch = delay_rank > 1 ? (s->n + delay_rank - 1) / delay_rank : 1;
  b->prec = precision;
  const size_t es = precision == QMCX_FP32 ? 4 : 8;
  const size_t W = nwalkers, n = s->n, N = s->N, K = delay_rank;
  // LDS check for the move kernel
  {
    int nt;
    int rc = choose_shape(s, nt);
    if (rc) {
      delete b;
      return rc;
    }
    const int chunk = move_chunk(s->n, delay_rank, s->has_j3);
    const int mw = s->has_j3 ? s->mw : 0;
    auto lds_of = [&](bool st) {
      return precision == QMCX_FP32 ? move_lds<float>(s->N, s->nions, s->n, delay_rank, chunk, mw, st).total
                                    : move_lds<double>(s->N, s->nions, s->n, delay_rank, chunk, mw, st).total;
    };
    size_t lds = lds_of(false);
    if (lds > kMoveLdsMax && delay_rank > 1) {
      // the delay rank's inverse-row cache does not fit: the streamed form of the delayed updates
      b->st_need = true;
      lds = lds_of(true);
    }
    if (lds > kMoveLdsMax) {
      delete b;
      return fail(QMCX_ERR_ARG, "walker state (" + std::to_string(lds / 1024) +
                                    " KB) does not fit in the 160 KB of LDS; use a smaller delay rank");
    }
  }
  int rc = 0;
  unsigned char* p8 = nullptr;
  rc |= b->alloc(b->R, 3 * W * N);
  rc |= b->alloc(p8, es * W * 2 * n * n);
  b->A = p8;
  rc |= b->alloc(p8, es * W * 2 * n * n);
  b->Ainv = p8;
  rc |= b->alloc(p8, es * W * 2 * n * 3 * n);
  b->dpsiM = p8;
  rc |= b->alloc(p8, es * W * 2 * n * n);
  b->d2psiM = p8;
  rc |= b->alloc(p8, es * W * 5 * N);
  b->J2 = p8;
  rc |= b->alloc(p8, es * W * 5 * N);
  b->J1 = p8;
  const size_t S = W * 2 * (size_t)b->nch;   // delayed-update state slots
  rc |= b->alloc(p8, es * S * K * n);
  b->Vbuf = p8;
  rc |= b->alloc(p8, sizeof(double) * S * K * K);
  b->Binv = p8;
  rc |= b->alloc(b->rows, S * K);
  rc |= b->alloc(b->ndelay, S);
  rc |= b->alloc(b->logdet, W * 2);
  rc |= b->alloc(b->phase, W * 2);
  rc |= b->alloc(b->jlog, W * 3);
  rc |= b->alloc(b->flags, 1);
  if (s->has_j3) {
    b->j3 = true;
    rc |= b->alloc(p8, es * W * 5 * N);
    b->J3 = p8;
    rc |= b->alloc(b->j3mask, W * N * (size_t)s->mw);
    rc |= b->alloc(p8, es * W * N * kJ3Rec);
    b->j3near = p8;
    rc |= b->alloc(p8, es * W * 2 * n * kJ3Rec);
    b->j3prop = p8;
  }
  rc |= b->alloc(b->nacc, W);
  rc |= b->alloc(b->narb, W);
  rc |= b->alloc(b->sweep, 4);
  rc |= b->alloc(b->G, W * N * 3);
  rc |= b->alloc(b->L, W * N);
  rc |= b->alloc(b->logv, W);
  rc |= b->alloc(b->lu_ws, W * 2 * 2 * n * n);
  rc |= b->alloc(p8, es * W * 2 * n * 5 * s->npad);
  b->spo = p8;
  if (precision == QMCX_FP32) {
    rc |= b->alloc(b->A64, W * 2 * n * n);
    rc |= b->alloc(b->spo64, W * 2 * n * (size_t)s->npad);
  }
  rc |= b->alloc(b->prop, W * 2 * n * kProp);
  rc |= b->alloc(b->ld_fresh, W);
  rc |= b->alloc(b->ph_fresh, W);
  if (rc) {
    qmcx_batch_destroy(b);
    return QMCX_ERR_NOMEM;
  }
  {
    hipStream_t st[3];
    if (int rc = device_streams(st)) {
      qmcx_batch_destroy(b);
      return rc;
    }
    b->stream = st[0];
    b->aux = st[1];
    b->pre = st[2];
  }
  for (int d = 0; d < 2; ++d) HIPCHK(hipEventCreateWithFlags(&b->ev_pre[d], hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&b->ev_start, hipEventDisableTiming));
  for (int d = 0; d < 2; ++d) {
    HIPCHK(hipEventCreateWithFlags(&b->ev_done[d], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->ev_fresh[d], hipEventDisableTiming));
  }
  HIPCHK(hipEventCreateWithFlags(&b->ev_gl, hipEventDisableTiming));
  HIPCHK(hipEventCreate(&b->ev0));
  HIPCHK(hipEventCreate(&b->ev1));
  s->nbatches += 1;
  b->counted = true;
  *out = b;
  return 0;
}

int qmcx_batch_destroy(qmcx_batch_t b) {
  if (!b) return 0;
  if (b->counted) b->sys->nbatches -= 1;
  b->drop_graph();
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->aux) hipStreamSynchronize(b->aux);
  if (b->pre) hipStreamSynchronize(b->pre);
  for (void* p : b->allocs) hipFree(p);
  if (b->trace) hipFree(b->trace);
  if (b->hist) hipFree(b->hist);
  if (b->flex_displ) hipFree(b->flex_displ);
  if (b->flex_decide) hipFree(b->flex_decide);
  if (b->cad) hipFree(b->cad);
  if (b->cad_flag) hipFree(b->cad_flag);
  if (b->vr_pos) hipFree(b->vr_pos);
  if (b->vr_out) hipFree(b->vr_out);
  if (b->nl_out) hipFree(b->nl_out);
  if (b->nl_cnt) hipFree(b->nl_cnt);
  if (b->nl_near) hipFree(b->nl_near);
  if (b->nl_items) hipFree(b->nl_items);
  if (b->nl_sph) hipFree(b->nl_sph);
  // the streams are the device's shared ones and stay
  for (int d = 0; d < 2; ++d)
    if (b->ev_pre[d]) hipEventDestroy(b->ev_pre[d]);
  if (b->ev_start) hipEventDestroy(b->ev_start);
  for (int d = 0; d < 2; ++d) {
    if (b->ev_done[d]) hipEventDestroy(b->ev_done[d]);
    if (b->ev_fresh[d]) hipEventDestroy(b->ev_fresh[d]);
  }
  if (b->ev_gl) hipEventDestroy(b->ev_gl);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  for (hipEvent_t e : b->prof_ev) hipEventDestroy(e);
  delete b;
  return 0;
}

int qmcx_batch_set_positions(qmcx_batch_t b, const double* R) {
  if (!b || !R) return fail(QMCX_ERR_ARG, "null argument");
  const int N = b->sys->N;
  std::vector<double> soa(3 * (size_t)N * b->W);
  for (int w = 0; w < b->W; ++w)
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < 3; ++d) soa[((size_t)w * 3 + d) * N + i] = R[((size_t)w * N + i) * 3 + d];
  HIPCHK(hipMemcpy(b->R, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  b->flex_pending = false;   // the state is rebuilt by the next evaluateLog; pending rows are void
  return 0;
}

int qmcx_batch_get_positions(qmcx_batch_t b, double* R) {
  if (!b || !R) return fail(QMCX_ERR_ARG, "null argument");
  const int N = b->sys->N;
  std::vector<double> soa(3 * (size_t)N * b->W);
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(soa.data(), b->R, sizeof(double) * soa.size(), hipMemcpyDeviceToHost));
  for (int w = 0; w < b->W; ++w)
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < 3; ++d) R[((size_t)w * N + i) * 3 + d] = soa[((size_t)w * 3 + d) * N + i];
  return 0;
}

int qmcx_batch_set_walker_offset(qmcx_batch_t b, int walker0) {
  if (!b || walker0 < 0) return fail(QMCX_ERR_ARG, "bad argument");
  b->drop_graph();   // a captured run holds the old setting
  b->walker0 = walker0;
  return 0;
}

int qmcx_batch_set_recompute_rows(qmcx_batch_t b, int rows) {
  if (!b || rows < 0) return fail(QMCX_ERR_ARG, "bad argument");
  if (rows > 0 && !b->cad) {
    HIPCHK(hipMalloc(&b->cad, sizeof(int) * 4 * (size_t)b->W));
    HIPCHK(hipMalloc(&b->cad_flag, sizeof(int) * (size_t)b->W));
    hipLaunchKernelGGL(k_cadence_clear, dim3((4 * b->W + 255) / 256), dim3(256), 0, b->stream, b->cad, 4 * b->W);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_cadence_clear, dim3((b->W + 255) / 256), dim3(256), 0, b->stream, b->cad_flag, b->W);
    HIPCHK(hipGetLastError());
  }
  b->drop_graph();   // a captured run holds the old setting
  b->cadence = rows;
  return 0;
}

int qmcx_batch_set_delay_form(qmcx_batch_t b, int form) {
  if (!b || form < 0 || form > 1)
    return fail(QMCX_ERR_ARG, "delay form must be 0 (automatic) or 1 (streamed)");
  if (int rc = flush_pending(b)) return rc;
  b->drop_graph();   // a captured run holds the old setting
  b->st_ask = form == 1;
  return 0;
}

int qmcx_batch_get_delay_form(qmcx_batch_t b, int* form) {
  if (!b || !form) return fail(QMCX_ERR_ARG, "null argument");
  *form = b->streamed() ? 1 : 0;
  return 0;
}

int qmcx_batch_set_flush(qmcx_batch_t b, int kind) {
  if (!b || kind < 0 || kind > 2)
    return fail(QMCX_ERR_ARG, "flush kind must be 0 (matrix cores), 1 (vector) or 2 (matrix cores, U rows from HBM)");
  b->drop_graph();   // a captured run holds the old setting
  b->flush_kind = kind;
  return 0;
}

int qmcx_batch_set_tile(qmcx_batch_t b, int spline_slices) {
  if (!b || spline_slices < 0 || spline_slices > 16 || (spline_slices && 16 % spline_slices)) return fail(QMCX_ERR_ARG, "bad tile");
  b->drop_graph();   // a captured run holds the old setting
  const int nq = b->sys->npad / (b->prec == QMCX_FP32 ? 4 : 2);
  if (nq * spline_slices > 1024) return fail(QMCX_ERR_ARG, "spline tile does not fit a workgroup");
  b->slices = spline_slices;
  return 0;
}

int qmcx_evaluate_log(qmcx_batch_t b, double* log_out) {
  if (!b) return fail(QMCX_ERR_ARG, "null batch");
  if (int rc = check_j3(b)) return rc;
  if (int rc = b->sys->sync_functors()) return rc;
  int rc = b->prec == QMCX_FP32 ? evaluate_log_impl<float>(b) : evaluate_log_impl<double>(b);
  if (rc) return rc;
  b->flex_pending = false;   // the inverse is rebuilt from scratch: pending delayed rows are void
  if (b->cad) {
    hipLaunchKernelGGL(k_cadence_reset, dim3((b->W + 255) / 256), dim3(256), 0, b->stream, b->cad, b->W, -1);
    HIPCHK(hipGetLastError());
  }
  b->rows64_stale = false;   // the FP64 Slater rows are filled with the SPO rows when kept
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->j3)
    if (int rc2 = check_flags(b)) return rc2;
  if (log_out) HIPCHK(hipMemcpy(log_out, b->logv, sizeof(double) * b->W, hipMemcpyDeviceToHost));
  return 0;
}

int qmcx_set_trace(qmcx_batch_t b, int enable) {
  if (!b) return fail(QMCX_ERR_ARG, "null batch");
  b->drop_graph();   // a captured run holds the old setting
  if (enable && !b->trace) HIPCHK(hipMalloc(&b->trace, sizeof(double) * 9 * (size_t)b->W * b->sys->N));
  b->trace_on = enable != 0;
  return 0;
}

int qmcx_get_trace(qmcx_batch_t b, double* trace) {
  if (!b || !trace || !b->trace) return fail(QMCX_ERR_STATE, "trace not enabled");
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(trace, b->trace, sizeof(double) * 9 * (size_t)b->W * b->sys->N, hipMemcpyDeviceToHost));
  return 0;
}

int qmcx_run_sweeps(qmcx_batch_t b, int nsweeps, uint64_t seed, double tau, int accept_mode, int recompute_every,
                    double* log_out) {
  if (!b || nsweeps < 0 || !(tau > 0)) return fail(QMCX_ERR_ARG, "bad argument");
  if (accept_mode != QMCX_ACCEPT_REFERENCE && accept_mode != QMCX_ACCEPT_METROPOLIS)
    return fail(QMCX_ERR_ARG, "unknown accept mode");
  if (b->rows64 && b->rows64_stale)
    return fail(QMCX_ERR_STATE, "the accept band was switched on after qmcx_evaluate_log: call it again");
  if (int rc = check_j3(b)) return rc;
  if (b->graph_on && b->prof_on) return fail(QMCX_ERR_STATE, "kernel timing and graph replay exclude each other");
  if (int rc = flush_pending(b)) return rc;
  if (nsweeps > b->hist_cap) {
    if (b->hist) hipFree(b->hist);
    b->hist = nullptr;
    HIPCHK(hipMalloc(&b->hist, sizeof(double) * (size_t)nsweeps * b->W));
    b->hist_cap = nsweeps;
    b->drop_graph();   // the graph's kernels hold the old history buffer
  }
  // this run's counters: evaluateGL writes its history from slot 0 (sweep[1] = the first sweep),
  // the first proposals belong to that sweep (sweep[2])
  int cur[4];
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(cur, b->sweep, sizeof(cur), hipMemcpyDeviceToHost));
  cur[1] = cur[0];
  cur[2] = cur[0];
  HIPCHK(hipMemcpy(b->sweep, cur, sizeof(cur), hipMemcpyHostToDevice));
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  if (nsweeps > 0) {
    const GraphKey key{nsweeps, seed, tau, accept_mode, recompute_every};
    if (b->graph_on) {
      // the whole run as one graph, captured on its first use and replayed while the run's
      // parameters stay the same (the sweep numbers come from the device counters)
      if (!b->gexec || !(b->gkey == key)) {
        b->drop_graph();
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_run(b, nsweeps, seed, tau, accept_mode, recompute_every);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
          if (g) hipGraphDestroy(g);
          return rc;
        }
        HIPCHK(e);
        const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        HIPCHK(ei);
        b->gkey = key;
      }
      HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    } else if (int rc = enqueue_run(b, nsweeps, seed, tau, accept_mode, recompute_every)) {
      return rc;
    }
  }
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  HIPCHK(hipEventSynchronize(b->ev1));
  HIPCHK(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
  if (b->prof_collect()) return QMCX_ERR_HIP;
  if (b->j3)
    if (int rc = check_flags(b)) return rc;
  if (log_out && nsweeps > 0)
    HIPCHK(hipMemcpy(log_out, b->hist, sizeof(double) * (size_t)nsweeps * b->W, hipMemcpyDeviceToHost));
  return 0;
}

int qmcx_batch_set_graph(qmcx_batch_t b, int enable) {
  if (!b) return fail(QMCX_ERR_ARG, "null batch");
  b->graph_on = enable != 0;
  if (!b->graph_on) b->drop_graph();
  return 0;
}

int qmcx_get_gl(qmcx_batch_t b, double* G, double* L, double* log_out) {
  if (!b) return fail(QMCX_ERR_ARG, "null batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  const size_t WN = (size_t)b->W * b->sys->N;
  if (G) HIPCHK(hipMemcpy(G, b->G, sizeof(double) * 3 * WN, hipMemcpyDeviceToHost));
  if (L) HIPCHK(hipMemcpy(L, b->L, sizeof(double) * WN, hipMemcpyDeviceToHost));
  if (log_out) HIPCHK(hipMemcpy(log_out, b->logv, sizeof(double) * b->W, hipMemcpyDeviceToHost));
  return 0;
}

int qmcx_get_accepts(qmcx_batch_t b, long long* naccept) {
  if (!b || !naccept) return fail(QMCX_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<unsigned long long> tmp(b->W);
  HIPCHK(hipMemcpy(tmp.data(), b->nacc, sizeof(unsigned long long) * b->W, hipMemcpyDeviceToHost));
  for (int i = 0; i < b->W; ++i) naccept[i] = (long long)tmp[i];
  return 0;
}

int qmcx_batch_set_accept_band(qmcx_batch_t b, double band) {
  if (!b || !(band >= 0.0) || band > 1.0) return fail(QMCX_ERR_ARG, "band must be in [0, 1]");
  b->drop_graph();   // a captured run holds the old setting
  if (band > 0.0 && b->j3 && b->prec == QMCX_FP32 && b->K > 1) {
    // with a three-body Jastrow the re-evaluation takes LDS of its own beside the J3 pair slots
    const qmcx_system_s* s = b->sys;
    const size_t lds = move_lds<float>(s->N, s->nions, s->n, b->K, move_chunk(s->n, b->K, true), s->mw,
                                       b->streamed(), true).total;
    if (lds > kMoveLdsMax)
      return fail(QMCX_ERR_ARG, "the near-tie band with a three-body Jastrow needs " + std::to_string(lds / 1024) +
                                    " KB of LDS at this delay rank; use a smaller one");
  }
  b->band = band;
  const bool on = b->prec == QMCX_FP32 && (band > 0.0 || b->fp64_rows);
  if (on && !b->rows64) b->rows64_stale = true;
  b->rows64 = on;
  return 0;
}

int qmcx_batch_set_fp64_rows(qmcx_batch_t b, int enable) {
  if (!b) return fail(QMCX_ERR_ARG, "null batch");
  b->drop_graph();   // a captured run holds the old setting
  b->fp64_rows = enable != 0;
  const bool on = b->prec == QMCX_FP32 && (b->band > 0.0 || b->fp64_rows);
  if (on && !b->rows64) b->rows64_stale = true;
  b->rows64 = on;
  return 0;
}

int qmcx_get_arbitrations(qmcx_batch_t b, long long* narb) {
  if (!b || !narb) return fail(QMCX_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<unsigned long long> tmp(b->W);
  HIPCHK(hipMemcpy(tmp.data(), b->narb, sizeof(unsigned long long) * b->W, hipMemcpyDeviceToHost));
  for (int i = 0; i < b->W; ++i) narb[i] = (long long)tmp[i];
  return 0;
}

int qmcx_get_inverse(qmcx_batch_t b, int walker, int det, double* psiM) {
  if (!b || !psiM || walker < 0 || walker >= b->W || det < 0 || det > 1) return fail(QMCX_ERR_ARG, "bad argument");
  if (int rc = flush_pending(b)) return rc;
  HIPCHK(hipStreamSynchronize(b->stream));
  const size_t n = b->sys->n, nn = n * n, off = ((size_t)walker * 2 + det) * nn;
  if (b->prec == QMCX_FP32) {
    std::vector<float> tmp(nn);
    HIPCHK(hipMemcpy(tmp.data(), (float*)b->Ainv + off, sizeof(float) * nn, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nn; ++i) psiM[i] = tmp[i];
  } else {
    HIPCHK(hipMemcpy(psiM, (double*)b->Ainv + off, sizeof(double) * nn, hipMemcpyDeviceToHost));
  }
  return 0;
}

// ---- flex_* stage calls: one electron of every walker per call, through the move kernel
static int flex_launch(qmcx_batch_t b, int iel, bool apply) {
  const qmcx_system_s* s = b->sys;
  const bool f32 = b->prec == QMCX_FP32;
  int nt = 0;
  int rc = choose_shape(s, nt);
  if (rc) return rc;
  const int det = iel < s->n ? 0 : 1;
  rc = f32 ? launch_proposals<float>(b, b->stream, det, iel - det * s->n, 1, 0, 1.0, b->flex_displ)
           : launch_proposals<double>(b, b->stream, det, iel - det * s->n, 1, 0, 1.0, b->flex_displ);
  if (rc) return rc;
  MoveArgs a;
  std::memset(&a, 0, sizeof(a));
  a.spo = b->spo;
  a.prop = b->prop;
  a.det = iel < s->n ? 0 : 1;
  a.e_begin = iel;
  a.e_end = iel + 1;
  a.chunk_len = 1;
  a.nchunks = 1;
  a.mode = QMCX_ACCEPT_METROPOLIS;
  a.slices = 0;
  a.decide = b->flex_decide;
  // delay rank > 1: the streamed form keeps each walker's accepted rows pending in state slot 0
  // of the determinant (DelayedUpdate::acceptRow, DelayedUpdate.h:124-153) until the walker's
  // count reaches the delay rank (flushed after the accept) or completeUpdates
  const bool st = b->K > 1;
  a.resume = b->flex_pending ? 1 : 0;
  const bool trace_was = b->trace_on;
  b->trace_on = true;
  if (f32) {
    const MoveLds L = move_lds<float>(s->N, s->nions, s->n, b->K, 1, b->j3 ? s->mw : 0, st);
    rc = launch_moves<float>(b, nt, L.total, a, st);
  } else {
    const MoveLds L = move_lds<double>(s->N, s->nions, s->n, b->K, 1, b->j3 ? s->mw : 0, st);
    rc = launch_moves<double>(b, nt, L.total, a, st);
  }
  b->trace_on = trace_was;
  if (rc) return rc;
  if (st) {
    b->flex_pending = true;
    if (apply)
      if (int rc2 = flush_pending(b, b->K)) return rc2;
  }
  return 0;
}

int qmcx_flex_ratio_grad(qmcx_batch_t b, int iel, const double* displ, double* ratio, double* grad_now,
                         double* grad_new) {
  if (!b || !displ || iel < 0 || iel >= b->sys->N) return fail(QMCX_ERR_ARG, "bad argument");
  if (int rc = check_j3(b)) return rc;
  if (!b->flex_displ) {
    HIPCHK(hipMalloc(&b->flex_displ, sizeof(double) * 3 * b->W));
    HIPCHK(hipMalloc(&b->flex_decide, sizeof(int) * b->W));
  }
  if (!b->trace) HIPCHK(hipMalloc(&b->trace, sizeof(double) * 9 * (size_t)b->W * b->sys->N));
  std::vector<int> none(b->W, 0);
  HIPCHK(hipMemcpyAsync(b->flex_displ, displ, sizeof(double) * 3 * b->W, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(b->flex_decide, none.data(), sizeof(int) * b->W, hipMemcpyHostToDevice, b->stream));
  if (int rc = flex_launch(b, iel, false)) return rc;
  // the electron's [W] slice of the trace ([W][N][9]) in one strided copy
  std::vector<double> rec(9 * (size_t)b->W);
  HIPCHK(hipMemcpy2DAsync(rec.data(), sizeof(double) * 9, b->trace + (size_t)iel * 9, sizeof(double) * 9 * b->sys->N,
                          sizeof(double) * 9, b->W, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (int w = 0; w < b->W; ++w) {
    const double* q = rec.data() + 9 * (size_t)w;
    if (ratio) ratio[w] = q[0];
    for (int d = 0; d < 3; ++d) {
      if (grad_now) grad_now[3 * w + d] = q[3 + d];
      if (grad_new) grad_new[3 * w + d] = q[6 + d];
    }
  }
  b->flex_iel = iel;
  return 0;
}

int qmcx_flex_accept_restore(qmcx_batch_t b, int iel, const int* accepted) {
  if (!b || !accepted) return fail(QMCX_ERR_ARG, "null argument");
  if (iel != b->flex_iel) return fail(QMCX_ERR_STATE, "accept_restore must follow ratio_grad of the same electron");
  HIPCHK(hipMemcpyAsync(b->flex_decide, accepted, sizeof(int) * b->W, hipMemcpyHostToDevice, b->stream));
  if (int rc = flex_launch(b, iel, true)) return rc;
  HIPCHK(hipStreamSynchronize(b->stream));
  b->flex_iel = -1;
  return 0;
}

int qmcx_flex_complete_updates(qmcx_batch_t b) {
  if (!b) return fail(QMCX_ERR_ARG, "null batch");
  if (int rc = flush_pending(b)) return rc;
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int qmcx_flex_evaluate_gl(qmcx_batch_t b, double* log_out) {
  if (!b) return fail(QMCX_ERR_ARG, "null batch");
  if (int rc = flush_pending(b)) return rc;   // evaluateGL reads the up-to-date inverse
  int rc = b->prec == QMCX_FP32 ? launch_gl<float>(b, false) : launch_gl<double>(b, false);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(b->stream));
  if (log_out) HIPCHK(hipMemcpy(log_out, b->logv, sizeof(double) * b->W, hipMemcpyDeviceToHost));
  return 0;
}

int qmcx_flex_evaluate_ratios(qmcx_batch_t b, int iel, int nvp, const double* vpos, double* ratios) {
  if (!b || !vpos || !ratios || iel < 0 || iel >= b->sys->N || nvp < 1 || nvp > kMaxVirtualPoints)
    return fail(QMCX_ERR_ARG, "bad argument");
  if (int rc = check_j3(b)) return rc;   // the kernel reads the batch's J3 state when the system has one
  if (int rc = flush_pending(b)) return rc;   // the ratios read the up-to-date inverse rows
  if (nvp > b->vr_cap) {
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->vr_pos) hipFree(b->vr_pos);
    if (b->vr_out) hipFree(b->vr_out);
    b->vr_pos = b->vr_out = nullptr;
    b->vr_cap = 0;
    HIPCHK(hipMalloc(&b->vr_pos, sizeof(double) * 3 * (size_t)nvp * b->W));
    HIPCHK(hipMalloc(&b->vr_out, sizeof(double) * (size_t)nvp * b->W));
    b->vr_cap = nvp;
  }
  const size_t np = (size_t)nvp * b->W;
  HIPCHK(hipMemcpyAsync(b->vr_pos, vpos, sizeof(double) * 3 * np, hipMemcpyHostToDevice, b->stream));
  const int rc = b->prec == QMCX_FP32 ? launch_virtual_ratios<float>(b, iel, nvp)
                                      : launch_virtual_ratios<double>(b, iel, nvp);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(ratios, b->vr_out, sizeof(double) * np, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int qmcx_batch_set_nlpp_unit(qmcx_batch_t b, int knots_per_unit) {
  if (!b || !(knots_per_unit == 0 || knots_per_unit == 1 || knots_per_unit == 2 || knots_per_unit == 4 ||
              knots_per_unit == 6))
    return fail(QMCX_ERR_ARG, "knots per NonLocalPP work unit must be 0 (automatic), 1, 2, 4 or 6");
  b->nl_unit = knots_per_unit;
  return 0;
}

int qmcx_nlpp_evaluate(qmcx_batch_t b, double rmax, const double* spheres, int max_pairs, double* ratios,
                       int* npairs) {
  if (!b || !spheres || !(rmax > 0) || max_pairs < 1 || max_pairs > kMaxNlppPairs) return fail(QMCX_ERR_ARG, "bad argument");
  if (int rc = check_j3(b)) return rc;   // the kernel reads the batch's J3 state when the system has one
  if (int rc = flush_pending(b)) return rc;   // the ratios read the up-to-date inverse rows
  const size_t WN = (size_t)b->W * b->sys->N;
  // work-list entries and knot units are int indices on the device
  if (WN * max_pairs * kNlppKnots >= (size_t)INT_MAX) return fail(QMCX_ERR_ARG, "walkers x electrons x max_pairs too large");
  if (max_pairs > b->nl_cap) {
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->nl_out) hipFree(b->nl_out);
    if (b->nl_cnt) hipFree(b->nl_cnt);
    b->nl_out = nullptr;
    b->nl_cnt = nullptr;
    b->nl_cap = 0;
    if (b->nl_near) hipFree(b->nl_near);
    if (b->nl_items) hipFree(b->nl_items);
    b->nl_near = nullptr;
    b->nl_items = nullptr;
    HIPCHK(hipMalloc(&b->nl_out, sizeof(double) * WN * max_pairs * kNlppKnots));
    HIPCHK(hipMalloc(&b->nl_cnt, sizeof(int) * WN));
    HIPCHK(hipMalloc(&b->nl_near, sizeof(double) * 4 * WN * max_pairs));
    HIPCHK(hipMalloc(&b->nl_items,
                     sizeof(int) * (2 * WN * max_pairs + 1 + 2 * (size_t)b->sys->nions * kNlppRadialBins)));
    b->nl_cap = max_pairs;
  }
  if (!b->nl_sph) HIPCHK(hipMalloc(&b->nl_sph, sizeof(double) * 3 * kNlppKnots * (size_t)b->W));
  HIPCHK(hipMemcpyAsync(b->nl_sph, spheres, sizeof(double) * 3 * kNlppKnots * (size_t)b->W, hipMemcpyHostToDevice,
                        b->stream));
  const int rc = b->prec == QMCX_FP32 ? launch_nlpp<float>(b, rmax, max_pairs) : launch_nlpp<double>(b, rmax, max_pairs);
  if (rc) return rc;
  std::vector<int> cnt(WN);
  HIPCHK(hipMemcpyAsync(cnt.data(), b->nl_cnt, sizeof(int) * WN, hipMemcpyDeviceToHost, b->stream));
  if (ratios)
    HIPCHK(hipMemcpyAsync(ratios, b->nl_out, sizeof(double) * WN * max_pairs * kNlppKnots, hipMemcpyDeviceToHost,
                          b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  if (npairs) std::copy(cnt.begin(), cnt.end(), npairs);
  for (size_t i = 0; i < WN; ++i)
    if (cnt[i] > max_pairs) return fail(QMCX_ERR_ARG, "more ions within rmax of an electron than max_pairs");
  return 0;
}

int qmcx_set_kernel_timing(qmcx_batch_t b, int enable) {
  if (!b) return fail(QMCX_ERR_ARG, "null batch");
  b->prof_on = enable != 0;
  for (int i = 0; i < 5; ++i) {
    b->prof_ms[i] = 0;
    b->prof_n[i] = 0;
  }
  b->prof_used = 0;
  return 0;
}

int qmcx_get_kernel_timing(qmcx_batch_t b, double* ms5, long long* launches5) {
  if (!b || !ms5 || !launches5) return fail(QMCX_ERR_ARG, "null argument");
  for (int i = 0; i < 5; ++i) {
    ms5[i] = b->prof_ms[i];
    launches5[i] = b->prof_n[i];
  }
  return 0;
}

int qmcx_last_sweep_ms(qmcx_batch_t b, double* ms) {
  if (!b || !ms) return fail(QMCX_ERR_ARG, "null argument");
  *ms = b->last_ms;
  return 0;
}

}  // extern "C"


// ====================================================================== delayed-update engine (DU_TYPE)
struct qmcx_du_s {
  int n = 0, K = 1, prec = QMCX_FP64, count = 0;
  hipStream_t stream = nullptr;
  void* Ainv = nullptr;
  void* A = nullptr;
  void* Vbuf = nullptr;
  double* Binv = nullptr;
  int* rows = nullptr;
  int* ndelay = nullptr;
  double* vec = nullptr;    // n doubles: invRow out / psiV in
  double* mat = nullptr;    // n x n doubles: LU input
  double* lu_ws = nullptr;
  double* ldph = nullptr;   // log|det|, phase
  std::vector<void*> allocs;

  template <typename T>
  BatchP<T> params() const {
    BatchP<T> B;
    std::memset(&B, 0, sizeof(B));
    B.W = 1;
    B.K = K;
    B.nch = 1;
    B.A = (T*)A;
    B.Ainv = (T*)Ainv;
    B.Vbuf = (T*)Vbuf;
    B.Binv = Binv;
    B.rows = rows;
    B.ndelay = ndelay;
    return B;
  }
  template <typename T>
  SysP<T> sys() const {
    SysP<T> S;
    std::memset(&S, 0, sizeof(S));
    S.n = n;
    return S;
  }
};
typedef qmcx_du_s* qmcx_du_t;

namespace {
template <typename T>
int du_flush(qmcx_du_t h) {
  if (h->count == 0) return 0;
  if (sizeof(T) == 4 && flush_mfma_fits(h->n)) {
    const size_t lds = flush_mfma_lds(h->n, h->K);
    if (flush_t_tiles(h->K) == 1)
      hipLaunchKernelGGL(k_flush_mfma<true>, dim3((h->n + kFlushBM - 1) / kFlushBM, 1), dim3(64 * kFlushWaves), lds,
                         h->stream, h->sys<float>(), h->params<float>(), 0, 0, 1, FlushRows{0, h->n});
    else
      hipLaunchKernelGGL(k_flush_mfma<false>, dim3((h->n + kFlushBM - 1) / kFlushBM, 1), dim3(64 * kFlushWaves), lds,
                         h->stream, h->sys<float>(), h->params<float>(), 0, 0, 1, FlushRows{0, h->n});
  } else {
    const size_t lds = sizeof(T) * 2 * kFlushRows * h->K + sizeof(int) * h->K;
    hipLaunchKernelGGL(k_flush_generic<T>, dim3((h->n + kFlushRows - 1) / kFlushRows, 1), dim3(256), lds, h->stream,
                       h->sys<T>(), h->params<T>(), 0, 0, FlushRows{0, h->n});
  }
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_du_reset, dim3(1), dim3(64), 0, h->stream, h->ndelay);
  HIPCHK(hipGetLastError());
  h->count = 0;
  return 0;
}
}  // namespace

extern "C" {

int qmcx_du_create(qmcx_du_t* out, int norb, int delay, int precision) {
  if (!out || norb <= 0 || norb > 1024 || delay < 1 || delay > 128 || delay > norb)
    return fail(QMCX_ERR_ARG, "norb in [1,1024] and delay in [1, min(128, norb)] required");
  if (precision != QMCX_FP32 && precision != QMCX_FP64) return fail(QMCX_ERR_ARG, "unknown precision");
  auto* h = new qmcx_du_s;
  h->n = norb;
  h->K = delay;
  h->prec = precision;
  const size_t es = precision == QMCX_FP32 ? 4 : 8, n = norb, K = delay;
  auto grab = [&](void** p, size_t bytes) -> int {
    HIPCHK(hipMalloc(p, std::max<size_t>(bytes, 16)));
    HIPCHK(hipMemset(*p, 0, std::max<size_t>(bytes, 16)));
    h->allocs.push_back(*p);
    return 0;
  };
  int rc = 0;
  rc |= grab(&h->Ainv, es * n * n);
  rc |= grab(&h->A, es * n * n);
  rc |= grab(&h->Vbuf, es * K * n);
  rc |= grab((void**)&h->Binv, sizeof(double) * K * K);
  rc |= grab((void**)&h->rows, sizeof(int) * K);
  rc |= grab((void**)&h->ndelay, sizeof(int));
  rc |= grab((void**)&h->vec, sizeof(double) * n);
  rc |= grab((void**)&h->mat, sizeof(double) * n * n);
  rc |= grab((void**)&h->lu_ws, sizeof(double) * 2 * n * n);
  rc |= grab((void**)&h->ldph, sizeof(double) * 2);
  if (rc) {
    for (void* p : h->allocs) hipFree(p);
    delete h;
    return QMCX_ERR_NOMEM;
  }
  {
    hipStream_t st[3];
    if (int rc2 = device_streams(st)) {
      for (void* p : h->allocs) hipFree(p);
      delete h;
      return rc2;
    }
    h->stream = st[0];
  }
  *out = h;
  return 0;
}

int qmcx_du_destroy(qmcx_du_t h) {
  if (!h) return 0;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (void* p : h->allocs) hipFree(p);
  delete h;   // its stream is the device's shared one
  return 0;
}

int qmcx_du_invert_transpose(qmcx_du_t h, const double* logdetT, double* Ainv, double* log_value, double* phase) {
  if (!h || !logdetT) return fail(QMCX_ERR_ARG, "null argument");
  const size_t nn = (size_t)h->n * h->n;
  HIPCHK(hipMemcpyAsync(h->mat, logdetT, sizeof(double) * nn, hipMemcpyHostToDevice, h->stream));
  if (h->prec == QMCX_FP32)
    HIPCHK((launch_gj_invert<double, float>(h->n, 1, h->stream, h->mat, (float*)h->Ainv, h->lu_ws, h->ldph,
                                           h->ldph + 1, 1)));
  else
    HIPCHK((launch_gj_invert<double, double>(h->n, 1, h->stream, h->mat, (double*)h->Ainv, h->lu_ws, h->ldph,
                                            h->ldph + 1, 1)));
  hipLaunchKernelGGL(k_du_reset, dim3(1), dim3(64), 0, h->stream, h->ndelay);
  HIPCHK(hipGetLastError());
  h->count = 0;
  double lp[2];
  HIPCHK(hipMemcpyAsync(lp, h->ldph, sizeof(lp), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (log_value) *log_value = lp[0];
  if (phase) *phase = lp[1];
  if (Ainv) return qmcx_du_get_inverse(h, Ainv);
  return 0;
}

int qmcx_du_set_inverse(qmcx_du_t h, const double* Ainv) {
  if (!h || !Ainv) return fail(QMCX_ERR_ARG, "null argument");
  const size_t nn = (size_t)h->n * h->n;
  if (h->prec == QMCX_FP32) {
    std::vector<float> tmp(Ainv, Ainv + nn);
    HIPCHK(hipMemcpy(h->Ainv, tmp.data(), sizeof(float) * nn, hipMemcpyHostToDevice));
  } else {
    HIPCHK(hipMemcpy(h->Ainv, Ainv, sizeof(double) * nn, hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(k_du_reset, dim3(1), dim3(64), 0, h->stream, h->ndelay);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  h->count = 0;
  return 0;
}

int qmcx_du_get_inverse(qmcx_du_t h, double* Ainv) {
  if (!h || !Ainv) return fail(QMCX_ERR_ARG, "null argument");
  const size_t nn = (size_t)h->n * h->n;
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->prec == QMCX_FP32) {
    std::vector<float> tmp(nn);
    HIPCHK(hipMemcpy(tmp.data(), h->Ainv, sizeof(float) * nn, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nn; ++i) Ainv[i] = tmp[i];
  } else {
    HIPCHK(hipMemcpy(Ainv, h->Ainv, sizeof(double) * nn, hipMemcpyDeviceToHost));
  }
  return 0;
}

int qmcx_du_get_inv_row(qmcx_du_t h, int row, double* inv_row) {
  if (!h || !inv_row || row < 0 || row >= h->n) return fail(QMCX_ERR_ARG, "bad argument");
  if (h->prec == QMCX_FP32)
    hipLaunchKernelGGL(k_du_get_inv_row<float>, dim3(1), dim3(256), 0, h->stream, h->params<float>(), h->n, row, h->vec);
  else
    hipLaunchKernelGGL(k_du_get_inv_row<double>, dim3(1), dim3(256), 0, h->stream, h->params<double>(), h->n, row,
                       h->vec);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(inv_row, h->vec, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int qmcx_du_accept_row(qmcx_du_t h, int row, const double* psiV) {
  if (!h || !psiV || row < 0 || row >= h->n) return fail(QMCX_ERR_ARG, "bad argument");
  HIPCHK(hipMemcpyAsync(h->vec, psiV, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
  if (h->prec == QMCX_FP32)
    hipLaunchKernelGGL(k_du_accept_row<float>, dim3(1), dim3(256), 0, h->stream, h->params<float>(), h->n, row, h->vec);
  else
    hipLaunchKernelGGL(k_du_accept_row<double>, dim3(1), dim3(256), 0, h->stream, h->params<double>(), h->n, row,
                       h->vec);
  HIPCHK(hipGetLastError());
  h->count += 1;
  if (h->count == h->K) {
    int rc = h->prec == QMCX_FP32 ? du_flush<float>(h) : du_flush<double>(h);
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int qmcx_du_update_inv_mat(qmcx_du_t h) {
  if (!h) return fail(QMCX_ERR_ARG, "null argument");
  int rc = h->prec == QMCX_FP32 ? du_flush<float>(h) : du_flush<double>(h);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int qmcx_du_delay_count(qmcx_du_t h, int* count) {
  if (!h || !count) return fail(QMCX_ERR_ARG, "null argument");
  *count = h->count;
  return 0;
}

}  // extern "C"

// ====================================================================== mixed-size groups (C ABI)
extern "C" {

int qmcx_group_create(qmcx_group_t* out, const qmcx_batch_t* batches, int nbatches) {
  if (!out || !batches || nbatches < 1 || nbatches > kMaxGroups)
    return fail(QMCX_ERR_ARG, "a group takes 1 to " + std::to_string(kMaxGroups) + " batches");
  auto* g = new qmcx_group_s;
  int nt = 256, need_jpt2 = 0;
  size_t lds = 0, plds = 0;
  int pnt = 64, nmax = 1;
  for (int i = 0; i < nbatches; ++i) {
    qmcx_batch_t b = batches[i];
    const qmcx_system_s* s = b ? b->sys : nullptr;
    std::string why;
    if (!b) why = "null batch";
    else if (b->prec != QMCX_FP32) why = "FP32 batches only";
    else if (b->K < 2) why = "delay rank 2 or more";
    else if (b->j3) why = "no three-body Jastrow";
    else if (b->band > 0.0) why = "no near-tie band";
    else if (b->rows64 != batches[0]->rows64) why = "all batches with FP64 Slater rows or none";
    else if (b->cadence > 0) why = "no recompute cadence";
    else if (b->streamed() || !use_mfma_flush<float>(b) || !flush_own_fits<float>(s->N, s->n, b->K, b->K))
      why = "each batch's delayed updates must be applied inside the move kernel (n a multiple of 16, delay <= 32)";
    else if (b->sys->dev != batches[0]->sys->dev) why = "batches of one device";
    for (int j = 0; j < i && why.empty(); ++j)
      if (batches[j] == b) why = "a batch appears twice";
    if (!why.empty()) {
      delete g;
      return fail(QMCX_ERR_ARG, "group batch " + std::to_string(i) + ": " + why);
    }
    int bnt = 0;
    if (int rc = choose_shape(s, bnt)) {
      delete g;
      return rc;
    }
    nt = std::max(nt, bnt);
    g->b.push_back(b);
    g->Wtot += b->W;
    nmax = std::max(nmax, s->n);
  }
  for (qmcx_batch_t b : g->b) {
    const qmcx_system_s* s = b->sys;
    if (s->N > nt) need_jpt2 = 1;
    lds = std::max(lds, (size_t)move_lds<float>(s->N, s->nions, s->n, b->K, b->K).total);
    const int nq = s->npad / 4, slices = proposal_slices<float>(s, b->slices);
    pnt = std::max(pnt, (nq * slices + 63) / 64 * 64);
    // slice partials, then the FP64 values when the batches keep FP64 Slater rows
    plds = std::max(plds, std::max(((sizeof(float) * (size_t)slices * 5 * s->npad + 15) & ~(size_t)15) +
                                       sizeof(double) * (size_t)slices * s->npad,
                                   sizeof(double) * 16 * 5));
  }
  g->nt = nt;
  g->jpt = need_jpt2 ? 2 : 1;
  g->lds = lds;
  g->pnt = pnt;
  g->plds = plds;
  g->pgx = std::max(1, std::min(nmax, kPropWorkgroups / g->Wtot));
  if (lds > kMoveLdsMax || pnt > 1024) {
    delete g;
    return fail(QMCX_ERR_ARG, "group workgroup does not fit");
  }
  g->stream = g->b[0]->stream;
  g->aux = g->b[0]->aux;
  g->pre = g->b[0]->pre;
  for (int d = 0; d < 2; ++d) {
    HIPCHK(hipEventCreateWithFlags(&g->ev_pre[d], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&g->ev_done[d], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&g->ev_fresh[d], hipEventDisableTiming));
  }
  HIPCHK(hipEventCreateWithFlags(&g->ev_gl, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&g->ev_start, hipEventDisableTiming));
  HIPCHK(hipEventCreate(&g->ev0));
  HIPCHK(hipEventCreate(&g->ev1));
  *out = g;
  return 0;
}

int qmcx_group_destroy(qmcx_group_t g) {
  if (!g) return 0;
  if (g->stream) hipStreamSynchronize(g->stream);
  g->drop_graph();
  for (hipEvent_t e : {g->ev_pre[0], g->ev_pre[1], g->ev_done[0], g->ev_done[1], g->ev_fresh[0], g->ev_fresh[1],
                       g->ev_gl, g->ev_start, g->ev0, g->ev1})
    if (e) hipEventDestroy(e);
  delete g;
  return 0;
}

int qmcx_group_set_graph(qmcx_group_t g, int enable) {
  if (!g) return fail(QMCX_ERR_ARG, "null group");
  g->graph_on = enable != 0;
  if (!g->graph_on) g->drop_graph();
  return 0;
}

int qmcx_group_run_sweeps(qmcx_group_t g, int nsweeps, uint64_t seed, double tau, int accept_mode, int recompute_every,
                          double* log_out) {
  if (!g || nsweeps < 0 || !(tau > 0)) return fail(QMCX_ERR_ARG, "bad argument");
  if (accept_mode != QMCX_ACCEPT_REFERENCE && accept_mode != QMCX_ACCEPT_METROPOLIS)
    return fail(QMCX_ERR_ARG, "unknown accept mode");
  HIPCHK(hipStreamSynchronize(g->stream));
  for (qmcx_batch_t b : g->b) {
    if (b->band > 0.0 || b->prof_on)
      return fail(QMCX_ERR_STATE, "a group batch took the near-tie band or kernel timing");
    if (b->rows64 != g->b[0]->rows64)
      return fail(QMCX_ERR_STATE, "the group's batches must all keep FP64 Slater rows or none");
    if (b->rows64 && b->rows64_stale)
      return fail(QMCX_ERR_STATE, "FP64 Slater rows were switched on after qmcx_evaluate_log: call it again");
    if (int rc = check_j3(b)) return rc;
    if (b->cadence > 0 || b->streamed())
      return fail(QMCX_ERR_STATE, "a group batch took the recompute cadence or the streamed delay form");
    if (int rc = flush_pending(b)) return rc;
    if (nsweeps > b->hist_cap) {
      if (b->hist) hipFree(b->hist);
      b->hist = nullptr;
      HIPCHK(hipMalloc(&b->hist, sizeof(double) * (size_t)nsweeps * b->W));
      b->hist_cap = nsweeps;
      b->drop_graph();
      g->drop_graph();   // the group's graph holds the old history buffer
    }
    int cur[4];
    HIPCHK(hipMemcpy(cur, b->sweep, sizeof(cur), hipMemcpyDeviceToHost));
    cur[1] = cur[0];
    cur[2] = cur[0];
    HIPCHK(hipMemcpy(b->sweep, cur, sizeof(cur), hipMemcpyHostToDevice));
  }
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  if (nsweeps > 0) {
    const GraphKey key{nsweeps, seed, tau, accept_mode, recompute_every};
    if (g->graph_on) {
      if (!g->graph_current(key)) {
        g->drop_graph();
        HIPCHK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_group_run(g, nsweeps, seed, tau, accept_mode, recompute_every);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(g->stream, &gr);
        if (rc) {
          if (gr) hipGraphDestroy(gr);
          return rc;
        }
        HIPCHK(e);
        const hipError_t ei = hipGraphInstantiate(&g->gexec, gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        HIPCHK(ei);
        g->gkey = key;
        g->epochs.clear();
        for (qmcx_batch_t b : g->b) g->epochs.push_back(b->epoch);
      }
      HIPCHK(hipGraphLaunch(g->gexec, g->stream));
    } else if (int rc = enqueue_group_run(g, nsweeps, seed, tau, accept_mode, recompute_every)) {
      return rc;
    }
  }
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  HIPCHK(hipEventSynchronize(g->ev1));
  HIPCHK(hipEventElapsedTime(&g->last_ms, g->ev0, g->ev1));
  if (log_out && nsweeps > 0) {
    // [nsweeps][all walkers], the batches' walkers in group order
    std::vector<double> h;
    size_t off = 0;
    for (qmcx_batch_t b : g->b) {
      h.resize((size_t)nsweeps * b->W);
      HIPCHK(hipMemcpy(h.data(), b->hist, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
      for (int sw = 0; sw < nsweeps; ++sw)
        std::memcpy(log_out + (size_t)sw * g->Wtot + off, h.data() + (size_t)sw * b->W, sizeof(double) * b->W);
      off += b->W;
    }
  }
  return 0;
}

int qmcx_group_last_ms(qmcx_group_t g, double* ms) {
  if (!g || !ms) return fail(QMCX_ERR_ARG, "null argument");
  *ms = g->last_ms;
  return 0;
}

}  // extern "C"
