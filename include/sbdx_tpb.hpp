// sbdx_tpb.hpp -- header-only C++ adapter with the reference's call shapes,
// on top of the C ABI in sbdx.h.
//
//   sbdx::MultTPBHip<ElemT>   mirrors MultTPBThrust<ElemT>
//                             (chemistry/tpb/mult_thrust.h:23-95):
//                             Init / run / makeQChamDiagTerms / correlation
//   sbdx::Davidson(...)       mirrors sbd::Davidson(hii, W, mult, max_it, nb,
//                             eps, max_time) (chemistry/basic/davidson_thrust.h:85-92)
//   sbdx::diag(...)           the TPB branch of sbd::tpb::diag
//                             (chemistry/tpb/sbdiag.h:139-151) for one GPU:
//                             energy, density and RDMs from FCIDUMP-derived
//                             integrals and string sets
//
// The integral types are templates so that the reference's oneInt/twoInt
// (chemistry/basic/integrals.h) plug in directly: anything with a `store`
// vector in the reference layouts works.  Errors surface as
// std::runtime_error, the reference's host-side convention; the C ABI
// never throws.
#ifndef SBDX_TPB_HPP
#define SBDX_TPB_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "sbdx.h"

namespace sbdx {

inline void check(int rc) {
  if (rc != SBDX_OK && rc != SBDX_ERR_NOT_CONVERGED)
    throw std::runtime_error(std::string("sbdx: ") + sbdx_last_error());
}

// vector<vector<size_t>> string sets -> flat word array (reference layout)
inline std::vector<uint64_t> flatten_words(const std::vector<std::vector<size_t>> &dets) {
  std::vector<uint64_t> w;
  if (dets.empty()) return w;
  const size_t nw = dets[0].size();
  w.reserve(dets.size() * nw);
  for (const auto &d : dets) {
    if (d.size() != nw) throw std::invalid_argument("sbdx: ragged string words");
    for (size_t x : d) w.push_back((uint64_t)x);
  }
  return w;
}

template <typename ElemT>
class MultTPBHip {
  static_assert(sizeof(ElemT) == sizeof(double), "the device path computes in FP64");

 public:
  MultTPBHip() = default;
  MultTPBHip(const MultTPBHip &) = delete;
  MultTPBHip &operator=(const MultTPBHip &) = delete;
  ~MultTPBHip() {
    if (h_) sbdx_destroy(h_);
  }

  // Init(adet, bdet, bit_length, norbs, I0, I1, I2): the one-rank subset of
  // MultTPBThrust::Init's arguments (communicators and helper schedules do
  // not apply on one GPU).  I1/I2 are the reference's oneInt/twoInt.
  template <typename OneInt, typename TwoInt>
  void Init(const std::vector<std::vector<size_t>> &adet,
            const std::vector<std::vector<size_t>> &bdet, size_t bit_length, size_t norbs,
            ElemT I0, const OneInt &I1, const TwoInt &I2) {
    if (h_) {
      sbdx_destroy(h_);
      h_ = nullptr;
    }
    const auto aw = flatten_words(adet), bw = flatten_words(bdet);
    check(sbdx_create(&h_, (int)norbs, bit_length, aw.data(), adet.size(), bw.data(), bdet.size(),
                      (double)I0, reinterpret_cast<const double *>(I1.store.data()),
                      reinterpret_cast<const double *>(I2.store.data()), I2.store.size()));
    norbs_ = norbs;
    own_hii_.clear();
    check(sbdx_dims(h_, &na_, &nb_, &ndet_));
  }

  // hii = ZeroExcite per determinant (makeQChamDiagTerms)
  void makeQChamDiagTerms(std::vector<ElemT> &hii) const {
    hii.resize(ndet_);
    check(sbdx_diag(h_, reinterpret_cast<double *>(hii.data())));
  }

  // Wb += H Wk (MultBase::run) with the caller's diagonal hii: the device
  // applies its own (ZeroExcite per determinant, what makeQChamDiagTerms
  // returns), and where the caller's hii differs from it the difference
  // times Wk is added on the host.  An empty hii means the operator's own.
  void run(const std::vector<ElemT> &hii, const std::vector<ElemT> &Wk, std::vector<ElemT> &Wb) const {
    if (Wk.size() != ndet_ || Wb.size() != ndet_) throw std::invalid_argument("sbdx: vector size");
    if (!hii.empty() && hii.size() != ndet_) throw std::invalid_argument("sbdx: hii size");
    check(sbdx_sigma(h_, reinterpret_cast<const double *>(Wk.data()),
                     reinterpret_cast<double *>(Wb.data()), 1));
    if (hii.empty()) return;
    if (own_hii_.size() != ndet_) {
      own_hii_.resize(ndet_);
      check(sbdx_diag(h_, own_hii_.data()));
    }
    for (size_t i = 0; i < ndet_; i++)
      if ((double)hii[i] != own_hii_[i]) Wb[i] += (ElemT)(((double)hii[i] - own_hii_[i]) * (double)Wk[i]);
  }

  // one[s][i + L j], two[s + 2t][i + L j + L^2 k + L^3 l] (Correlation)
  void correlation(const std::vector<ElemT> &W, std::vector<std::vector<ElemT>> &one,
                   std::vector<std::vector<ElemT>> &two) const {
    const size_t L2 = norbs_ * norbs_, L4 = L2 * L2;
    std::vector<double> o1(2 * L2), o2(4 * L4);
    check(sbdx_rdm(h_, reinterpret_cast<const double *>(W.data()), o1.data(), o2.data()));
    one.assign(2, std::vector<ElemT>(L2));
    two.assign(4, std::vector<ElemT>(L4));
    for (int s = 0; s < 2; s++)
      for (size_t i = 0; i < L2; i++) one[s][i] = o1[s * L2 + i];
    for (int s = 0; s < 4; s++)
      for (size_t i = 0; i < L4; i++) two[s][i] = o2[s * L4 + i];
  }

  sbdx_handle_t handle() const { return h_; }
  size_t size() const { return ndet_; }

 private:
  sbdx_handle_t h_ = nullptr;
  size_t norbs_ = 0, na_ = 0, nb_ = 0, ndet_ = 0;
  mutable std::vector<double> own_hii_;  // the device diagonal, fetched on first use by run
};

// Davidson(hii, W, mult, max_it, nb, eps, max_time): W in/out; returns the
// lowest Ritz value of the last step.  The solve runs on the device with the
// operator's own diagonal, which is what the reference passes as hii
// (makeQChamDiagTerms).
template <typename ElemT, typename RealT>
ElemT Davidson(const std::vector<ElemT> & /*hii*/, std::vector<ElemT> &W, MultTPBHip<ElemT> &mult,
               int max_iteration, int num_block, RealT eps, RealT max_time) {
  double e = 0.0;
  int steps = 0;
  check(sbdx_davidson(mult.handle(), reinterpret_cast<double *>(W.data()), max_iteration, num_block,
                      (double)eps, (double)max_time, &e, &steps));
  return (ElemT)e;
}

// The TPB path of diag on one GPU: W starts at the first determinant
// (BasisInitVector, init 0), then Davidson, the final <W|HW>, and, when
// do_rdm, the RDMs and density[2i+s] = one[s][i + L i].
template <typename OneInt, typename TwoInt>
void diag(const std::vector<std::vector<size_t>> &adet, const std::vector<std::vector<size_t>> &bdet,
          size_t bit_length, size_t norbs, double I0, const OneInt &I1, const TwoInt &I2, int max_it,
          int max_nb, double eps, double max_time, bool do_rdm, double &energy,
          std::vector<double> &density, std::vector<std::vector<double>> &one_p_rdm,
          std::vector<std::vector<double>> &two_p_rdm, std::vector<double> *W_out = nullptr) {
  MultTPBHip<double> op;
  op.Init(adet, bdet, bit_length, norbs, I0, I1, I2);
  std::vector<double> W(op.size(), 0.0), hii;
  W[0] = 1.0;
  Davidson(hii, W, op, max_it, max_nb, eps, max_time);
  check(sbdx_energy(op.handle(), W.data(), &energy));
  if (do_rdm) {
    op.correlation(W, one_p_rdm, two_p_rdm);
    density.assign(2 * norbs, 0.0);
    for (size_t i = 0; i < norbs; i++)
      for (int s = 0; s < 2; s++) density[2 * i + s] = one_p_rdm[s][i + norbs * i];
  }
  if (W_out) W_out->swap(W);
}

}  // namespace sbdx

#endif  // SBDX_TPB_HPP
