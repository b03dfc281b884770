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
// Errors: the reference aborts (APP_ABORT, src/Utilities/Configuration.h:35-39);
// here a failing call throws std::runtime_error carrying qmcx_last_error().
#ifndef QMCX_DROPIN_HPP
#define QMCX_DROPIN_HPP

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

}  // namespace qmcx

#endif
