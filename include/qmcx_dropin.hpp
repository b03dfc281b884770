// qmcx_dropin.hpp — reference-shaped C++ classes over the C ABI of libqmcx.so.
//
// DelayedUpdateHIP<T, T_FP> has the interface of the reference's delayed-update
// engine DelayedUpdate<T, T_FP> (src/QMCWaveFunctions/DelayedUpdate.h:29-231) and
// plugs into DiracDeterminant<DU_TYPE> the way the reference's GPU engine is
// meant to (src/QMCWaveFunctions/DiracDeterminant.cpp:339-341):
//
//     template class DiracDeterminant<qmcx::DelayedUpdateHIP<ValueType, mValueType>>;
//
// The inverse lives on the GPU.  The caller's Ainv matrix is refreshed from the
// device after invert_transpose and after every flush (acceptRow reaching the
// delay, updateInvMat); between flushes it is stale, exactly as the reference's
// Ainv is until updateInvMat.  MAT is any row-major matrix with rows(), cols(),
// data(); VVT any vector with size(), data() (the reference's Matrix<T>/Vector<T>).
//
// SPOSetHIP (SPOSet / einspline_spo), DistanceTableHIP (DistanceTableData, the AA
// and BA kinds) and WaveFunctionHIP / WaveFunctionBatchHIP (WaveFunction, single
// walker and flex_* over a list of walkers) are the object API of the particle-move
// path; each cites the reference class it stands in for.  PS is any particle set
// with getTotalNum(), R[i][d] (positions) and activeR(iat) (the proposed position
// of the active particle, else R[iat]), as the reference's ParticleSet has.
//
// Errors: the reference aborts (APP_ABORT, src/Utilities/Configuration.h:35-39);
// here a failing call throws std::runtime_error carrying qmcx_last_error().
#ifndef QMCX_DROPIN_HPP
#define QMCX_DROPIN_HPP

#include <array>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "qmcx.h"

namespace qmcx {

inline void check(int rc, const char* what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + qmcx_last_error());
}

template <typename T, typename T_FP>
class DelayedUpdateHIP {
 public:
  using real_type = T;

  DelayedUpdateHIP() = default;
  DelayedUpdateHIP(const DelayedUpdateHIP&) = delete;
  DelayedUpdateHIP& operator=(const DelayedUpdateHIP&) = delete;
  ~DelayedUpdateHIP() {
    if (h_) qmcx_du_destroy(h_);
  }

  // DelayedUpdate::resize (DelayedUpdate.h:65-74)
  void resize(int norb, int delay) {
    if (h_) qmcx_du_destroy(h_);
    h_ = nullptr;
    check(qmcx_du_create(&h_, norb, delay, std::is_same<T, float>::value ? QMCX_FP32 : QMCX_FP64), "qmcx_du_create");
    n_ = norb;
    row_.resize(n_);
    full_.resize((size_t)n_ * n_);
  }

  // DelayedUpdate::invert_transpose -> DiracMatrix::invert_transpose (DelayedUpdate.h:75-82);
  // the inversion runs in FP64 (T_FP) on the device
  template <typename MAT>
  void invert_transpose(const MAT& logdetT, MAT& Ainv, real_type& LogValue, real_type& PhaseValue) {
    for (int i = 0; i < n_; ++i)
      for (int j = 0; j < n_; ++j) full_[(size_t)i * n_ + j] = (double)logdetT.data()[(size_t)i * logdetT.cols() + j];
    double lv = 0, ph = 0;
    check(qmcx_du_invert_transpose(h_, full_.data(), full_.data(), &lv, &ph), "qmcx_du_invert_transpose");
    store(Ainv);
    LogValue = (real_type)lv;
    PhaseValue = (real_type)ph;
  }

  // DelayedUpdate::initializeInv (DelayedUpdate.h:87-91): the device copy is refreshed too
  template <typename MAT>
  void initializeInv(const MAT& Ainv) {
    for (int i = 0; i < n_; ++i)
      for (int j = 0; j < n_; ++j) full_[(size_t)i * n_ + j] = (double)Ainv.data()[(size_t)i * Ainv.cols() + j];
    check(qmcx_du_set_inverse(h_, full_.data()), "qmcx_du_set_inverse");
  }

  // DelayedUpdate::getInvRow (DelayedUpdate.h:96-114); Ainv is the device copy
  template <typename MAT, typename VVT>
  void getInvRow(const MAT& /*Ainv*/, int rowchanged, VVT& invRow) {
    check(qmcx_du_get_inv_row(h_, rowchanged, row_.data()), "qmcx_du_get_inv_row");
    for (int j = 0; j < n_; ++j) invRow.data()[j] = (typename std::remove_reference<decltype(invRow.data()[0])>::type)row_[j];
  }

  // DelayedUpdate::acceptRow (DelayedUpdate.h:124-153); flushes at the delay rank
  template <typename MAT, typename VVT>
  void acceptRow(MAT& Ainv, int rowchanged, const VVT& psiV) {
    for (int j = 0; j < n_; ++j) row_[j] = (double)psiV.data()[j];
    int before = 0, after = 0;
    check(qmcx_du_delay_count(h_, &before), "qmcx_du_delay_count");
    check(qmcx_du_accept_row(h_, rowchanged, row_.data()), "qmcx_du_accept_row");
    check(qmcx_du_delay_count(h_, &after), "qmcx_du_delay_count");
    (void)before;
    if (after == 0) refresh(Ainv);   // the delay was reached and Ainv was updated
  }

  // DelayedUpdate::updateInvMat (DelayedUpdate.h:158-231)
  template <typename MAT>
  void updateInvMat(MAT& Ainv) {
    int c = 0;
    check(qmcx_du_delay_count(h_, &c), "qmcx_du_delay_count");
    if (c == 0) return;
    check(qmcx_du_update_inv_mat(h_), "qmcx_du_update_inv_mat");
    refresh(Ainv);
  }

 private:
  template <typename MAT>
  void refresh(MAT& Ainv) {
    check(qmcx_du_get_inverse(h_, full_.data()), "qmcx_du_get_inverse");
    store(Ainv);
  }
  template <typename MAT>
  void store(MAT& Ainv) {
    using E = typename std::remove_reference<decltype(Ainv.data()[0])>::type;
    for (int i = 0; i < n_; ++i)
      for (int j = 0; j < n_; ++j) Ainv.data()[(size_t)i * Ainv.cols() + j] = (E)full_[(size_t)i * n_ + j];
  }

  qmcx_du_t h_ = nullptr;
  int n_ = 0;
  std::vector<double> row_, full_;
};

// ---- SPOSet (src/QMCWaveFunctions/SPOSet.h:29-127; einspline_spo.hpp:150-223) ----------------
// The orbitals of a system's spline table, evaluated on the device in precision T (float: the FP32
// table, double: the FP64 copy; the system must keep that table).  Outputs follow einspline_spo:
// gradients of the table in the SPO lattice's coordinates and d2psi = h_xx (einspline_spo.hpp:215-220).
template <typename T>
class SPOSetHIP {
 public:
  SPOSetHIP(qmcx_system_t sys, int norb) : sys_(sys), norb_(norb) {}
  int size() const { return norb_; }

  // einspline_spo::evaluate_v through SPOSet::evaluate(P, iat, psi_v) (SPOSet.h:55)
  template <typename PS, typename VV>
  void evaluate(const PS& P, int iat, VV& psi_v) {
    const auto& r = P.activeR(iat);
    run(1, &r);
    for (int j = 0; j < norb_; ++j) psi_v[j] = (T)out_[j];
  }
  // SPOSet::evaluate(P, iat, psi_v, dpsi_v, d2psi_v) (SPOSet.h:64; einspline_spo.hpp:204-223)
  template <typename PS, typename VV, typename GV>
  void evaluate(const PS& P, int iat, VV& psi_v, GV& dpsi_v, VV& d2psi_v) {
    const auto& r = P.activeR(iat);
    run(1, &r);
    unpack(0, psi_v, dpsi_v, d2psi_v);
  }
  // einspline_spo::evaluate_vgh (einspline_spo.hpp:194-202): psi[j], grad[j][3], hess[j][6]
  // (xx, xy, xz, yy, yz, zz) at a Cartesian point
  void evaluate_vgh(const double r[3], T* psi, T* grad, T* hess) {
    const double pt[3] = {r[0], r[1], r[2]};
    run_points(1, pt);
    for (int j = 0; j < norb_; ++j) {
      psi[j] = (T)out_[j];
      for (int d = 0; d < 3; ++d) grad[3 * j + d] = (T)out_[(size_t)(1 + d) * norb_ + j];
      for (int d = 0; d < 6; ++d) hess[6 * j + d] = (T)out_[(size_t)(4 + d) * norb_ + j];
    }
  }
  // SPOSet::multi_evaluate (SPOSet.h:111-127): every walker's active particle in one launch
  template <typename PS, typename VV, typename GV>
  static void multi_evaluate(const std::vector<SPOSetHIP*>& spo_list, const std::vector<PS*>& P_list, int iat,
                             std::vector<VV*>& psi_v_list, std::vector<GV*>& dpsi_v_list,
                             std::vector<VV*>& d2psi_v_list) {
    if (spo_list.empty()) return;
    SPOSetHIP& s0 = *spo_list[0];
    std::vector<double> pts(3 * P_list.size());
    for (size_t w = 0; w < P_list.size(); ++w) {
      if (spo_list[w]->sys_ != s0.sys_) throw std::runtime_error("multi_evaluate: SPO sets of different tables");
      const auto& r = P_list[w]->activeR(iat);
      for (int d = 0; d < 3; ++d) pts[3 * w + d] = (double)r[d];
    }
    s0.run_points((int)P_list.size(), pts.data());
    for (size_t w = 0; w < P_list.size(); ++w) s0.unpack(w, *psi_v_list[w], *dpsi_v_list[w], *d2psi_v_list[w]);
  }

 private:
  template <typename POS>
  void run(int npts, const POS* r) {
    double pt[3] = {(double)(*r)[0], (double)(*r)[1], (double)(*r)[2]};
    run_points(npts, pt);
  }
  void run_points(int npts, const double* pts) {
    out_.resize((size_t)npts * 10 * norb_);
    check(qmcx_spline_vgh(sys_, std::is_same<T, float>::value ? QMCX_FP32 : QMCX_FP64, npts, pts, out_.data()),
          "qmcx_spline_vgh");
  }
  template <typename VV, typename GV>
  void unpack(size_t p, VV& psi_v, GV& dpsi_v, VV& d2psi_v) const {
    const double* o = out_.data() + p * 10 * norb_;
    for (int j = 0; j < norb_; ++j) {
      psi_v[j] = (T)o[j];
      for (int d = 0; d < 3; ++d) dpsi_v[j][d] = (T)o[(size_t)(1 + d) * norb_ + j];
      d2psi_v[j] = (T)o[(size_t)4 * norb_ + j];
    }
  }
  qmcx_system_t sys_;
  int norb_;
  std::vector<double> out_;
};

// ---- DistanceTableData (src/Particle/DistanceTableData.h:59-155) ---------------------------------
// AA: targets are the particle set's own particles (DistanceTableAA.h:28-113; row i holds the
// minimum-image distances from particle i to every particle, the flip index i as in
// DTD_BConds::computeDistances, ParticleBConds.h:154-217, and the particle's own entry the largest
// value).  BA: targets are fixed sources, e.g. the ions (DistanceTableBA.h:28-105, flip index 0).  Distances[i][j], Displacements[i] as [3][Ntargets] (structure of
// arrays, as the reference's RowContainer), Temp_r / Temp_dr for the active particle's proposal.
template <typename T>
class DistanceTableHIP {
 public:
  // AA table of a particle set of n particles
  DistanceTableHIP(qmcx_system_t sys, int n) : sys_(sys), aa_(true), nt_(n) { alloc(n); }
  // BA table: nsources fixed sources (positions [nsources][3]) for a set of n particles
  DistanceTableHIP(qmcx_system_t sys, int n, int nsources, const double* sources)
      : sys_(sys), aa_(false), nt_(nsources), src_(sources, sources + 3 * (size_t)nsources) {
    alloc(n);
  }
  int targets() const { return nt_; }
  std::vector<T> Distances;       // [N][Ntargets]
  std::vector<T> Displacements;   // [N][3][Ntargets]
  std::vector<T> Temp_r;          // [Ntargets]
  std::vector<T> Temp_dr;         // [3][Ntargets]

  // DistanceTableData::evaluate(P): every row
  template <typename PS>
  void evaluate(PS& P) {
    const int n = P.getTotalNum();
    std::vector<double> pos(3 * (size_t)n);
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) pos[3 * i + d] = (double)P.R[i][d];
    std::vector<int> flip(n);
    for (int i = 0; i < n; ++i) flip[i] = aa_ ? i : 0;
    rows(n, pos.data(), flip.data(), aa_ ? pos.data() : src_.data(), Distances.data(), Displacements.data());
    if (aa_)   // DistanceTableAA::evaluate gives a particle's own entry the largest distance
      for (int i = 0; i < n; ++i) Distances[(size_t)i * nt_ + i] = std::numeric_limits<T>::max();
  }
  // DistanceTableData::evaluate(P, jat): row jat
  template <typename PS>
  void evaluate(PS& P, int jat) {
    const double p[3] = {(double)P.R[jat][0], (double)P.R[jat][1], (double)P.R[jat][2]};
    const int flip = aa_ ? jat : 0;
    std::vector<double> tg;
    rows(1, p, &flip, targets_of(P, tg), Distances.data() + (size_t)jat * nt_,
         Displacements.data() + (size_t)jat * 3 * nt_);
  }
  // DistanceTableData::move(P, rnew): the proposal's row into Temp_r, Temp_dr (the active
  // particle's own entry is its flip index)
  template <typename PS, typename POS>
  void move(const PS& P, const POS& rnew, int iat) {
    const double p[3] = {(double)rnew[0], (double)rnew[1], (double)rnew[2]};
    const int flip = aa_ ? iat : 0;
    std::vector<double> tg;
    rows(1, p, &flip, targets_of(P, tg), Temp_r.data(), Temp_dr.data());
  }
  // DistanceTableData::update(jat): the accepted proposal's row replaces row jat
  void update(int jat) {
    std::copy(Temp_r.begin(), Temp_r.end(), Distances.begin() + (size_t)jat * nt_);
    std::copy(Temp_dr.begin(), Temp_dr.end(), Displacements.begin() + (size_t)jat * 3 * nt_);
  }

 private:
  void alloc(int n) {
    Distances.assign((size_t)n * nt_, T(0));
    Displacements.assign((size_t)n * 3 * nt_, T(0));
    Temp_r.assign(nt_, T(0));
    Temp_dr.assign(3 * (size_t)nt_, T(0));
  }
  template <typename PS>
  const double* targets_of(const PS& P, std::vector<double>& tg) const {
    if (!aa_) return src_.data();
    tg.resize(3 * (size_t)nt_);
    for (int i = 0; i < nt_; ++i)
      for (int d = 0; d < 3; ++d) tg[3 * i + d] = (double)P.R[i][d];
    return tg.data();
  }
  // npts rows from the device: r[npts][nt], dr[npts][nt][3] -> Distances / SoA displacements
  void rows(int npts, const double* pts, const int* flip, const double* tg, T* r_out, T* dr_out) {
    r_.resize((size_t)npts * nt_);
    dr_.resize((size_t)npts * nt_ * 3);
    check(qmcx_distance_rows(sys_, std::is_same<T, float>::value ? QMCX_FP32 : QMCX_FP64, npts, pts, flip, nt_, tg,
                             r_.data(), dr_.data()),
          "qmcx_distance_rows");
    for (int p = 0; p < npts; ++p)
      for (int j = 0; j < nt_; ++j) {
        r_out[(size_t)p * nt_ + j] = (T)r_[(size_t)p * nt_ + j];
        for (int d = 0; d < 3; ++d) dr_out[((size_t)p * 3 + d) * nt_ + j] = (T)dr_[((size_t)p * nt_ + j) * 3 + d];
      }
  }
  qmcx_system_t sys_;
  bool aa_;
  int nt_;
  std::vector<double> src_, r_, dr_;
};

// ---- WaveFunction (src/QMCWaveFunctions/WaveFunction.h:38-127, WaveFunction.cpp:178-499) ---------
// The product of the up and down DiracDeterminant with the system's J1, J2 (and three-body) Jastrows
// for the walkers of one device batch.  WaveFunctionBatchHIP holds W walkers and takes the flex_*
// calls over them (walker w = position w of the lists); WaveFunctionHIP is one walker with the
// single-walker calls.  Ratios and gradients follow the reference: ratio Psi(R')/Psi(R), gradients of
// log Psi (the determinant's part in the SPO lattice's coordinates, as einspline_spo produces them).
// precision: QMCX_FP64 (the plain drop-in) or QMCX_FP32 (FP32 splines, Jastrows and delayed updates,
// FP64 log|det| and ratio product).
class WaveFunctionBatchHIP {
 public:
  WaveFunctionBatchHIP(qmcx_system_t sys, int nwalkers, int nel, int delay_rank, int precision)
      : W_(nwalkers), N_(nel) {
    check(qmcx_batch_create(&b_, sys, nwalkers, delay_rank, precision), "qmcx_batch_create");
    R_.resize((size_t)W_ * N_ * 3);
    displ_.resize(3 * (size_t)W_);
    ratio_.resize(W_);
    gnow_.resize(3 * (size_t)W_);
    gnew_.resize(3 * (size_t)W_);
    log_.resize(W_);
  }
  WaveFunctionBatchHIP(const WaveFunctionBatchHIP&) = delete;
  WaveFunctionBatchHIP& operator=(const WaveFunctionBatchHIP&) = delete;
  ~WaveFunctionBatchHIP() {
    if (b_) qmcx_batch_destroy(b_);
  }
  int size() const { return W_; }
  qmcx_batch_t handle() const { return b_; }
  double getLogValue(int w) const { return log_[w]; }

  // WaveFunction::flex_evaluateLog (WaveFunction.cpp:306-318): positions from P_list
  template <typename PS>
  void flex_evaluateLog(const std::vector<PS*>& P_list) {
    take_positions(P_list);
    check(qmcx_batch_set_positions(b_, R_.data()), "qmcx_batch_set_positions");
    check(qmcx_evaluate_log(b_, log_.data()), "qmcx_evaluate_log");
  }
  // WaveFunction::flex_evalGrad (WaveFunction.cpp:319-342): grad_now[w] = grad_iat log Psi at R
  template <typename PS, typename POS>
  void flex_evalGrad(const std::vector<PS*>& P_list, int iat, std::vector<POS>& grad_now) {
    std::fill(displ_.begin(), displ_.end(), 0.0);
    check(qmcx_flex_ratio_grad(b_, iat, displ_.data(), ratio_.data(), gnow_.data(), gnew_.data()),
          "qmcx_flex_ratio_grad");
    for (size_t w = 0; w < P_list.size(); ++w)
      for (int d = 0; d < 3; ++d) grad_now[w][d] = gnow_[3 * w + d];
  }
  // WaveFunction::flex_ratioGrad (WaveFunction.cpp:343-417): the proposals are each P's activeR(iat)
  template <typename PS, typename VT, typename POS>
  void flex_ratioGrad(const std::vector<PS*>& P_list, int iat, std::vector<VT>& ratio_list,
                      std::vector<POS>& grad_new) {
    for (size_t w = 0; w < P_list.size(); ++w) {
      const auto& rn = P_list[w]->activeR(iat);
      for (int d = 0; d < 3; ++d) displ_[3 * w + d] = (double)rn[d] - R_[((size_t)w * N_ + iat) * 3 + d];
    }
    check(qmcx_flex_ratio_grad(b_, iat, displ_.data(), ratio_.data(), gnow_.data(), gnew_.data()),
          "qmcx_flex_ratio_grad");
    for (size_t w = 0; w < P_list.size(); ++w) {
      ratio_list[w] = (VT)ratio_[w];
      for (int d = 0; d < 3; ++d) grad_new[w][d] = gnew_[3 * w + d];
    }
  }
  // WaveFunction::flex_acceptrestoreMove (WaveFunction.cpp:418-485) with the ParticleSet's
  // acceptMove/rejectMove (miniqmc_sync_move.cpp:496-505): the accepted proposals become R
  template <typename PS>
  void flex_acceptrestoreMove(const std::vector<PS*>& P_list, const std::vector<bool>& isAccepted, int iat) {
    std::vector<int> acc(W_, 0);
    for (size_t w = 0; w < P_list.size(); ++w) acc[w] = isAccepted[w] ? 1 : 0;
    accept_restore(iat, acc.data());
  }
  // accepted[w] != 0: walker w takes its last proposal of iat
  void accept_restore(int iat, const int* accepted) {
    for (int w = 0; w < W_; ++w)
      if (accepted[w])
        for (int d = 0; d < 3; ++d) R_[((size_t)w * N_ + iat) * 3 + d] += displ_[3 * w + d];
    check(qmcx_flex_accept_restore(b_, iat, accepted), "qmcx_flex_accept_restore");
  }
  // WaveFunction::flex_completeUpdates (WaveFunction.cpp:486-499)
  void flex_completeUpdates() { check(qmcx_flex_complete_updates(b_), "qmcx_flex_complete_updates"); }
  // WaveFunction::flex_evaluateGL (WaveFunction.cpp:459-485): log Psi, and P.G, P.L when the
  // particle sets carry them (G[i][d], L[i])
  template <typename PS>
  void flex_evaluateGL(const std::vector<PS*>& P_list) {
    check(qmcx_flex_evaluate_gl(b_, log_.data()), "qmcx_flex_evaluate_gl");
    G_.resize((size_t)W_ * N_ * 3);
    L_.resize((size_t)W_ * N_);
    check(qmcx_get_gl(b_, G_.data(), L_.data(), nullptr), "qmcx_get_gl");
    for (size_t w = 0; w < P_list.size(); ++w)
      for (int i = 0; i < N_; ++i) {
        for (int d = 0; d < 3; ++d) P_list[w]->G[i][d] = G_[((size_t)w * N_ + i) * 3 + d];
        P_list[w]->L[i] = L_[(size_t)w * N_ + i];
      }
  }

 private:
  template <typename PS>
  void take_positions(const std::vector<PS*>& P_list) {
    if ((int)P_list.size() != W_) throw std::runtime_error("flex call: the list must hold every walker of the batch");
    for (int w = 0; w < W_; ++w)
      for (int i = 0; i < N_; ++i)
        for (int d = 0; d < 3; ++d) R_[((size_t)w * N_ + i) * 3 + d] = (double)P_list[w]->R[i][d];
  }
  qmcx_batch_t b_ = nullptr;
  int W_, N_;
  std::vector<double> R_, displ_, ratio_, gnow_, gnew_, log_, G_, L_;
};

// One walker: WaveFunction::evaluateLog, evalGrad, ratioGrad, acceptMove, restore, completeUpdates,
// evaluateGL (WaveFunction.h:63-70) over a batch of one.
class WaveFunctionHIP {
 public:
  WaveFunctionHIP(qmcx_system_t sys, int nel, int delay_rank, int precision)
      : wb_(sys, 1, nel, delay_rank, precision) {}
  double getLogValue() const { return wb_.getLogValue(0); }
  template <typename PS>
  void evaluateLog(PS& P) {
    std::vector<PS*> l{&P};
    wb_.flex_evaluateLog(l);
  }
  template <typename PS>
  std::array<double, 3> evalGrad(PS& P, int iat) {
    std::vector<PS*> l{&P};
    std::vector<std::array<double, 3>> g(1);
    wb_.flex_evalGrad(l, iat, g);
    return g[0];
  }
  template <typename PS, typename POS>
  double ratioGrad(PS& P, int iat, POS& grad) {
    std::vector<PS*> l{&P};
    std::vector<double> r(1);
    std::vector<std::array<double, 3>> g(1);
    wb_.flex_ratioGrad(l, iat, r, g);
    for (int d = 0; d < 3; ++d) grad[d] = g[0][d];
    return r[0];
  }
  template <typename PS>
  void acceptMove(PS& /*P*/, int iat) {
    const int one = 1;
    wb_.accept_restore(iat, &one);
  }
  // WaveFunction::restore: the proposal of iat is dropped
  void restore(int iat) {
    const int zero = 0;
    wb_.accept_restore(iat, &zero);
  }
  void completeUpdates() { wb_.flex_completeUpdates(); }
  template <typename PS>
  void evaluateGL(PS& P) {
    std::vector<PS*> l{&P};
    wb_.flex_evaluateGL(l);
  }

 private:
  WaveFunctionBatchHIP wb_;
};

}  // namespace qmcx

#endif
