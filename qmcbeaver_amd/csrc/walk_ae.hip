// walk_ae.hip -- all-electron importance-sampled moves of the device walk
// (the reference's default, one_e_per_iter 0): every electron of a walker
// moves at once, the whole evaluate runs at the proposal, and the move is
// accepted or the walker's record restored.
//
// Reference: QMCWalker::initializePropagation / moveElectronsImportanceSampling
// (QMCWalker.cpp:179-198, 343-398: R' = R + tau F(R) + sqrt(tau) xi for all
// electrons, forward G = (2 pi tau)^(-3N/2) exp(-sum |sqrt(tau) xi|^2 /
// (2 tau))), calculateReverseGreensFunctionImportanceSampling (:756-780:
// exp(-sum |R - R' - tau F(R')|^2 / (2 tau)) with the same prefactor),
// calculateMoveAcceptanceProbability / acceptOrRejectMove (:1578-1685),
// processPropagation (:200-266: on rejection the saved record is restored,
// :259-262; reweighting at every all-electron step, :226-228) and
// reweight_walker (:1060-1185).
//
// The walker's current record lives in `drift` (the modified drift F at R,
// all electrons), `logpsi`, `sgn`, `mr` (|modified drift| / |drift|, the
// umrigar93 energy modification's ratio) and cur_el.  A move writes R' over
// R (saving R in Rsave), the ordinary all-electron evaluate runs at R' and
// its batch-row outputs `prop` decide the move: accepted, `prop` becomes the
// record; rejected, R is restored and the record kept.  The Slater and
// Jastrow state of a rejected walker then describes R', which no
// all-electron move reads; the host re-evaluates every walker at R after its
// last step, so the resident state describes R between calls.
#include "qmcb_kernels.hpp"

namespace qmcb {

// One wave (64 lanes) per walker; lanes stride over the electrons.
__global__ void __launch_bounds__(64) ae_propose_kernel(AeArgs A, const int *walkers, int num) {
  const int b = blockIdx.x;
  if (b >= num) return;
  const int w = walkers ? walkers[b] : b;
  const int t = threadIdx.x, N = A.N;
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  const uint32_t s0 = (uint32_t)A.step, s1 = (uint32_t)((uint64_t)A.step >> 32);
  const double st = sqrt(A.tau), two_pi = 6.283185307179586;
  double sq = 0.0;
  for (int e = t; e < N; e += 64) {
    // counter (step, walker, tag): tags 0 and 1 are the single-electron
    // proposal's, kAeTag + e this electron's, kAeTag - 1 the accept draw
    const u4 r = philox4x32(u4{s0, s1, (uint32_t)w, kAeTag + (uint32_t)e}, k0, k1);
    const double u0 = u01(r.x, r.y), u1 = u01(r.y, r.z), u2 = u01(r.z, r.w), u3 = u01(r.w, r.x);
    const double r1 = sqrt(-2.0 * log(u0)), r2 = sqrt(-2.0 * log(u2));
    const double g[3] = {r1 * cos(two_pi * u1), r1 * sin(two_pi * u1), r2 * cos(two_pi * u3)};
    double *x = A.R + ((long)w * N + e) * 3;
    double *xs = A.Rsave + ((long)w * N + e) * 3;
    const double *f = A.drift + ((long)w * N + e) * 3;
    for (int k = 0; k < 3; k++) {
      const double chi = st * g[k];
      xs[k] = x[k];
      x[k] = x[k] + A.tau * f[k] + chi;
      sq += chi * chi;
    }
  }
  sq = wave_sum(sq);
  if (t == 0) {
    A.fwd[w] = sq;
    const u4 r = philox4x32(u4{s0, s1, (uint32_t)w, kAeTag - 1u}, k0, k1);
    A.unif[b] = u01(r.x, r.y);
  }
}

__global__ void __launch_bounds__(64) ae_accept_kernel(AeArgs A, const int *walkers, int num) {
  const int b = blockIdx.x;
  if (b >= num) return;
  const int w = walkers ? walkers[b] : b;
  const int t = threadIdx.x, N = A.N;
  const qmcb_results &P = A.prop;
  const double *dn = P.mod_drift + (long)b * N * 3;  // drift at the proposal
  const double *gn = P.grad_lnpsi + (long)b * N * 3;
  // |modified drift|^2, |drift|^2 at the proposal; reverse Green's function
  double msq = 0.0, gsq = 0.0, rev = 0.0;
  for (int e = t; e < N; e += 64) {
    const double *x = A.R + ((long)w * N + e) * 3;
    const double *xs = A.Rsave + ((long)w * N + e) * 3;
    for (int k = 0; k < 3; k++) {
      const double d = dn[3 * e + k], g = gn[3 * e + k];
      msq += d * d;
      gsq += g * g;
      if (A.mode != 0) {
        const double tt = xs[k] - x[k] - A.tau * d;
        rev += tt * tt;
      }
    }
  }
  msq = wave_sum(msq);
  gsq = wave_sum(gsq);
  rev = wave_sum(rev);
  const double mr_t = gsq > 0.0 ? sqrt(msq) / sqrt(gsq) : 1.0;
  const double EL = P.e_local[b], lp_new = P.log_psi[b];
  const double sg_new = P.sign_psi[b] < 0 ? -1.0 : 1.0;
  if (A.mode == 0) {
    // the record of the current positions (the resident state's evaluate)
    double *f = A.drift + (long)w * N * 3;
    for (int i = t; i < 3 * N; i += 64) f[i] = dn[i];
    if (t == 0) {
      A.logpsi[w] = lp_new;
      A.sgn[w] = sg_new;
      A.mr[w] = mr_t;
      A.cur_el[w] = EL;
    }
    return;
  }
  const bool sing = P.singular[b] != 0 || !isfinite(lp_new) || !isfinite(EL);
  const double lp_old = A.logpsi[w], el_old = A.cur_el[w], mr_o = A.mr[w];
  const int age_old = A.walk.age ? A.walk.age[w] : 0;
  double p = sing ? 0.0 : exp(2.0 * (lp_new - lp_old) - (rev - A.fwd[w]) / (2.0 * A.tau));
  if (!(p == p)) p = 0.0;
  else if (!sing) p = old_walker_boost(p, age_old, A.walk.old_walker);
  if (A.mode == 2 && sg_new * A.sgn[w] < 0) p = 0.0;  // fixed node (QMCWalker.cpp:1641-1646)
  if (p > 1.0) p = 1.0;
  const bool accept = p > A.unif[b];
  const int age_new = accept ? 0 : age_old + 1;
  if (accept) {
    double *f = A.drift + (long)w * N * 3;
    for (int i = t; i < 3 * N; i += 64) f[i] = dn[i];
  } else {
    double *x = A.R + (long)w * N * 3;
    const double *xs = A.Rsave + (long)w * N * 3;
    for (int i = t; i < 3 * N; i += 64) x[i] = xs[i];
  }
  if (t != 0) return;
  if (accept) {
    A.logpsi[w] = lp_new;
    A.sgn[w] = sg_new;
    A.mr[w] = mr_t;
    A.cur_el[w] = EL;
  }
  const double el = accept ? EL : el_old;
  if (A.walk.age) A.walk.age[w] = age_new;
  double wt = 1.0;
  if (A.mode == 2) {
    // umrigar93_probability_weighted reweighting at every all-electron step
    // (UNR93 eq. 25): dW = exp(tau_eff (p (S' + S) / 2 + (1 - p) S))
    wt = A.weight[w];
    if (A.reweight) {
      double so_, st_;
      if (A.energy_mod) {
        so_ = (A.e_trial - A.e_est) + (A.e_est - el_old) * mr_o;
        st_ = (A.e_trial - A.e_est) + (A.e_est - EL) * mr_t;
      } else {
        so_ = A.e_trial - el_old;
        st_ = A.e_trial - EL;
      }
      if (!(p > 0.0)) st_ = 0.0;
      const double dW = exp(A.tau_eff * (p * 0.5 * (so_ + st_) + (1.0 - p) * so_));
      wt = reweight_walker(wt, dW, age_new, el, A.walk);
      A.weight[w] = wt;
    }
    A.acc_w[w] += wt;
  }
  A.acc_el[w] += wt * el;
  A.acc_el2[w] += wt * el * el;
  A.acc_n[w] += accept ? 1.0 : 0.0;
  A.acc_prop_el[w] += sing ? 0.0 : EL;
  if (A.obs) {
    // estimator p E_L(R') + (1 - p) E_L(R) (QMCWalker.cpp:2096-2104)
    A.obs[3L * b] = p > 0.0 ? p * EL + (1.0 - p) * el_old : el_old;
    A.obs[3L * b + 1] = wt;
    A.obs[3L * b + 2] = p;
  }
}

hipError_t launch_ae_propose(const AeArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  hipLaunchKernelGGL(ae_propose_kernel, dim3(num), dim3(64), 0, st, a, walkers, num);
  return hipGetLastError();
}

hipError_t launch_ae_accept(const AeArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  hipLaunchKernelGGL(ae_accept_kernel, dim3(num), dim3(64), 0, st, a, walkers, num);
  return hipGetLastError();
}

}  // namespace qmcb
