// qmcb_functions.hpp -- C++ host interface with the shape of the reference's
// QMCFunctions plugin (QMcBeaver/include/QMCFunctions.h:26-105), implemented
// over the C ABI of libqmcb_amd.so (qmcbeaver_amd.h).  Header-only; link with
// -lqmcb_amd.
//
// Reference:                                        Here:
//   QMCFunctionsFactory::functionsFactory(...)      QMCFunctions(sys, opts)
//   evaluate(Array2D<double>& X, QMCWalkerData&)    evaluate(X, data)
//   evaluate(Array1D<QMCWalkerData*>&,              evaluate(walkerData, xData, num)
//            Array1D<Array2D<double>*>&, int num)
//   (rejected 1-e move: re-evaluate at old R)       accept(walkerData, accepted)
//   getNumTimers/aggregateTimers                    kernelTimes()
//
// QMCWalkerData's evaluate outputs (QMCSCFJastrow.cpp:156-169) map to
// WalkerData below; the Slater/Jastrow caches the reference keeps in
// QMCWalkerData live on the device, keyed by WalkerData::slot.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "qmcbeaver_amd.h"

namespace qmcb {

struct WalkerData {
  int slot = -1;            // resident walker slot on the device
  int whichE = -1;          // -1: all-electron evaluate, k: single-electron move of k
  // outputs (QMCWalkerData fields)
  double localEnergy = 0, kineticEnergy = 0, potentialEnergy = 0;
  double neEnergy = 0, eeEnergy = 0;
  double logPsi = 0;        // ln|psi| (the reference keeps psi as a QMCDouble)
  int signPsi = 1;
  double lapPsiRatio = 0;
  double detRatio = 1;
  double U = 0;
  double modificationRatio = 1;  // |modifiedGradPsiRatio| / |gradPsiRatio|
  double x2 = 0, y2 = 0, z2 = 0;  // sums over the electrons of the squared coordinates
  bool singular = false;
  std::vector<double> gradPsiRatio;          // [N][3]
  std::vector<double> modifiedGradPsiRatio;  // [N][3]
};

class QMCFunctions {
 public:
  QMCFunctions(const qmcb_system &sys, const qmcb_options &opt)
      : N_(sys.n_up + sys.n_dn) {
    qmcb_status st = qmcb_create(&sys, &opt, &ctx_);
    if (st != QMCB_OK) throw std::runtime_error(std::string("qmcb_create: ") + qmcb_last_error(nullptr));
  }
  QMCFunctions(const QMCFunctions &) = delete;
  QMCFunctions &operator=(const QMCFunctions &) = delete;
  virtual ~QMCFunctions() { qmcb_destroy(ctx_); }

  int numElectrons() const { return N_; }

  // One walker (QMCSCFJastrow::evaluate(X, data), QMCSCFJastrow.cpp:183-192).
  // X: [N][3] row-major.
  void evaluate(const std::vector<double> &X, WalkerData &data) {
    std::vector<WalkerData *> w(1, &data);
    std::vector<const double *> x(1, X.data());
    evaluate(w, x, 1);
  }

  // A pass of walkers (QMCSCFJastrow::evaluate(walkerData, xData, num),
  // QMCSCFJastrow.cpp:91-181).  whichE is shared by the pass, as in the
  // reference (walkerData(0)->whichE).
  void evaluate(std::vector<WalkerData *> &walkerData, std::vector<const double *> &xData,
                int num) {
    if (num <= 0) return;
    const int whichE = walkerData[0]->whichE;
    ids_.resize(num);
    R_.resize((size_t)num * N_ * 3);
    for (int i = 0; i < num; i++) {
      ids_[i] = walkerData[i]->slot;
      std::copy(xData[i], xData[i] + 3 * N_, R_.begin() + (size_t)i * N_ * 3);
    }
    alloc_(num);
    qmcb_results r = results_();
    check_(qmcb_evaluate(ctx_, num, ids_.data(), R_.data(), whichE, &r), "qmcb_evaluate");
    for (int i = 0; i < num; i++) {
      WalkerData &d = *walkerData[i];
      d.localEnergy = el_[i];
      d.kineticEnergy = ke_[i];
      d.potentialEnergy = pe_[i];
      d.neEnergy = vne_[i];
      d.eeEnergy = vee_[i];
      d.logPsi = lp_[i];
      d.signPsi = sg_[i];
      d.lapPsiRatio = lap_[i];
      d.detRatio = dr_[i];
      d.U = u_[i];
      d.modificationRatio = mr_[i];
      d.x2 = r2_[3 * i]; d.y2 = r2_[3 * i + 1]; d.z2 = r2_[3 * i + 2];
      d.singular = sing_[i] != 0;
      d.gradPsiRatio.assign(grad_.begin() + (size_t)i * N_ * 3, grad_.begin() + (size_t)(i + 1) * N_ * 3);
      d.modifiedGradPsiRatio.assign(mgrad_.begin() + (size_t)i * N_ * 3,
                                    mgrad_.begin() + (size_t)(i + 1) * N_ * 3);
    }
    if (whichE >= 0) staged_.assign(ids_.begin(), ids_.end());
  }

  // Commit or discard the single-electron proposals of the last pass; the
  // reference instead re-evaluates a rejected walker at its old position
  // (QMCWalker.cpp:254-266).
  void accept(const std::vector<WalkerData *> &walkerData, const std::vector<bool> &accepted) {
    std::vector<int32_t> ids(walkerData.size());
    std::vector<uint8_t> mask(walkerData.size());
    for (size_t i = 0; i < walkerData.size(); i++) {
      ids[i] = walkerData[i]->slot;
      mask[i] = accepted[i] ? 1 : 0;
    }
    check_(qmcb_accept(ctx_, (int32_t)ids.size(), ids.data(), mask.data()), "qmcb_accept");
  }

  // Per-kernel device time of the last device-VMC call (the reference's
  // getNumTimers/aggregateTimers stopwatches, QMCFunctions.h:81-87).
  std::map<std::string, double> kernelTimes() {
    const char *names[32];
    double ms[32];
    int64_t n[32];
    int k = qmcb_kernel_times(ctx_, 32, names, ms, n);
    std::map<std::string, double> out;
    for (int i = 0; i < k && i < 32; i++) out[names[i]] = ms[i];
    return out;
  }

  qmcb_ctx *handle() { return ctx_; }

 private:
  void check_(qmcb_status st, const char *what) {
    if (st != QMCB_OK) throw std::runtime_error(std::string(what) + ": " + qmcb_last_error(ctx_));
  }
  void alloc_(int num) {
    el_.resize(num); ke_.resize(num); pe_.resize(num); vne_.resize(num); vee_.resize(num);
    lp_.resize(num); lap_.resize(num); dr_.resize(num); u_.resize(num); mr_.resize(num);
    r2_.resize((size_t)num * 3);
    sg_.resize(num); sing_.resize(num);
    grad_.resize((size_t)num * N_ * 3);
    mgrad_.resize((size_t)num * N_ * 3);
  }
  qmcb_results results_() {
    qmcb_results r{};
    r.log_psi = lp_.data(); r.sign_psi = sg_.data(); r.grad_lnpsi = grad_.data();
    r.mod_drift = mgrad_.data(); r.lap_ratio = lap_.data(); r.e_local = el_.data();
    r.kinetic = ke_.data(); r.potential = pe_.data(); r.v_ne = vne_.data(); r.v_ee = vee_.data();
    r.det_ratio = dr_.data(); r.jastrow_u = u_.data(); r.singular = sing_.data();
    r.mod_ratio = mr_.data(); r.r2 = r2_.data();
    return r;
  }

  qmcb_ctx *ctx_ = nullptr;
  int N_;
  std::vector<int32_t> ids_, staged_;
  std::vector<double> R_, el_, ke_, pe_, vne_, vee_, lp_, lap_, dr_, u_, mr_, r2_, grad_, mgrad_;
  std::vector<int8_t> sg_, sing_;
};

}  // namespace qmcb
