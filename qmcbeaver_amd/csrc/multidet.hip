// multidet.hip -- several determinants: Psi = sum_d c_d det_up(d) det_dn(d) e^U.
//
// Reference: QMCSCFJastrow::update_SCF (QMCSCFJastrow.cpp:281-327) combines
// the per-determinant ratios into D, D_x, D_xx with the weights
// termPR(d) = c_d det_up(d) det_dn(d) / D; QMCSlater::calculate_DerivativeRatios
// (QMCSlater.cpp:648-764) keeps one inverse and one set of gradient /
// Laplacian ratios per determinant, updated by the row Sherman-Morrison
// formula in single-electron mode.  The determinants of a spin share the
// orbital rows of the active orbitals (one GEMM); determinant d takes its
// columns through the occupation map (QMCWavefunction.cpp:873-905).
//
// Kernels:
//   mdet_move_kernel      one workgroup per (walker, determinant): ratio,
//                         traces and the updated inverse of the proposal
//                         (row replacement), written to staging;
//   mdet_assemble_kernel  one workgroup per walker: the determinant sum with a
//                         shared exponent (ln|D| = M + ln|sum s_d e^{L_d - M}|),
//                         grad ln Psi of all electrons, Laplacian, E_L, drift;
//                         in VMC mode also the current configuration, the
//                         Green's function ratio and the Metropolis decision;
//   mdet_commit_kernel    one workgroup per (walker, determinant): accepted
//                         proposals become the resident state.
#include "qmcb_kernels.hpp"

namespace qmcb {

namespace {

constexpr int kMaxDet = 64;

// one spin's determinant record for the combination: staged (the moved spin of
// a proposal, indexed b * D + d) or resident (indexed w * D + d)
struct DetView {
  const double *logdet, *sgn, *lap, *grad;  // grad: [.][n][3]
  const int *sing;
  long base;                                // record index of determinant 0
};

__device__ inline DetView resident_view(const SlaterArgs &S, long w, int D) {
  return DetView{S.logdet, S.sgn, S.lapPR, S.gradPR, S.singular, w * D};
}

// ln|c_d det_up(d) det_dn(d)| and its sign; singular flags OR-ed into sing
__device__ inline double det_term(const MdetArgs &A, const DetView &up, const DetView &dn, int nup,
                                  int ndn, int d, double &sg, bool &sing) {
  const double c = A.ci[d];
  double l = c != 0.0 ? log(fabs(c)) : -1e300;
  sg = c < 0 ? -1.0 : 1.0;
  if (nup > 0) {
    sing = sing || up.sing[up.base + d];
    l += up.logdet[up.base + d];
    sg *= up.sgn[up.base + d];
  }
  if (ndn > 0) {
    sing = sing || dn.sing[dn.base + d];
    l += dn.logdet[dn.base + d];
    sg *= dn.sgn[dn.base + d];
  }
  return l;
}

// weights t_d = s_d e^{L_d - M} / sum_d s_d e^{L_d - M} (termPR), ln|D| and
// its sign, with M the largest L_d; false if singular.  t is in LDS.
__device__ inline bool combine(const MdetArgs &A, const DetView &up, const DetView &dn, int nup,
                               int ndn, double *t, double &lnD, double &sgnD) {
  const int D = A.D;
  double M = -1e300, sg;
  bool sing = false;
  for (int d = 0; d < D; d++) {
    const double l = det_term(A, up, dn, nup, ndn, d, sg, sing);
    if (A.ci[d] != 0.0 && l > M) M = l;
  }
  double S = 0.0;
  for (int d = 0; d < D; d++) {
    const double l = det_term(A, up, dn, nup, ndn, d, sg, sing);
    t[d] = A.ci[d] != 0.0 ? sg * exp(l - M) : 0.0;
    S += t[d];
  }
  if (sing || !(S != 0.0) || !isfinite(S)) {
    for (int d = 0; d < D; d++) t[d] = 0.0;
    lnD = 0.0;
    sgnD = 0.0;
    return false;
  }
  for (int d = 0; d < D; d++) t[d] /= S;
  lnD = M + log(fabs(S));
  sgnD = S < 0 ? -1.0 : 1.0;
  return true;
}

template <int NT>
__device__ inline void block_sum5(double v[5], double (*scratch)[NT / 64], double *out) {
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const double s = wave_sum(v[k]);
    if ((t & 63) == 0) scratch[k][t >> 6] = s;
  }
  __syncthreads();
  if (t == 0)
    for (int k = 0; k < 5; k++) {
      double s = 0.0;
      for (int q = 0; q < NT / 64; q++) s += scratch[k][q];
      out[k] = s;
    }
  __syncthreads();
}

}  // namespace

// ---------------------------------------------------------------------------
template <typename TS, typename TO>
__global__ void __launch_bounds__(128)
mdet_move_kernel(MdetArgs A, const int *walkers, int num) {
  const int sp = A.moved < A.s[0].n ? 0 : 1;
  const SlaterArgs &S = A.s[sp];
  const int D = A.D, d = blockIdx.x % D, b = blockIdx.x / D;
  if (b >= num) return;
  const long w0 = walkers ? walkers[b] : b, w = w0 * D + d, bd = (long)b * D + d;
  const int n = S.n, nact = act_count(S), r = A.moved - S.spin_start;
  const long cs = (long)nact * n, nn2 = (long)n * n;
  const int t = threadIdx.x;
  __shared__ double v[kComp][128];
  __shared__ double colr[128];
  __shared__ double wv[128];
  __shared__ double s_new[4];
  __shared__ double red[2];
  const TO *vr = (const TO *)A.vrow + b * A.v_s_walker;
  for (int i = t; i < kComp * n; i += 128) {
    const int k = i / n, o = i % n;
    const double x = (double)vr[(long)k * A.v_s_comp + occ_row(S, d, o)];
    v[k][o] = k == 0 ? x : as_stored<TS>(x);
  }
  const double *inv = S.inv + w * nn2;
  for (int o = t; o < n; o += 128) colr[o] = inv[(long)o * n + r];
  __syncthreads();
  const TS *dm = (const TS *)S.dmat + w0 * kDer * cs;
  const int c = t;
  double wc = 0.0, T1[4] = {0, 0, 0, 0}, T2[4] = {0, 0, 0, 0};
  if (c < n) {
    for (int o = 0; o < n; o++) {
      const double a = inv[(long)o * n + c];
      const long ro = (long)occ_row(S, d, o) * n + c;
      wc += v[0][o] * a;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const double x = (double)dm[dm_at<TS>(k, ro, cs)];
        T1[k] += x * a;
        T2[k] += x * colr[o];
      }
    }
    wv[c] = wc;
  }
  if (t == 0) {
    // derivative row of the moved electron times the new column r (times ratio)
    for (int k = 0; k < 4; k++) {
      double s = 0.0;
      for (int o = 0; o < n; o++) s += v[1 + k][o] * colr[o];
      s_new[k] = s;
    }
  }
  __syncthreads();
  const double ratio = wv[r];
  const bool ok = ratio != 0.0 && isfinite(ratio);
  const double qi = 1.0 / ratio;
  double lapc = 0.0;
  if (c < n) {
    double g[4];
#pragma unroll
    for (int k = 0; k < 4; k++) g[k] = (c == r) ? s_new[k] * qi : T1[k] - wv[c] * qi * T2[k];
    double *sg = A.st_grad + (bd * n + c) * 3;
    sg[0] = g[0]; sg[1] = g[1]; sg[2] = g[2];
    lapc = g[3];
    // staged inverse A'^-1 = A^-1 - A^-1(:,r) z^T, z(c) = (w(c) - delta_cr) / ratio
    if (ok) {
      const double z = (wv[c] - (c == r ? 1.0 : 0.0)) * qi;
      double *st = A.st_inv + bd * nn2 + c;
      for (int o = 0; o < n; o++) st[(long)o * n] = inv[(long)o * n + c] - colr[o] * z;
    }
  }
  const double s = wave_sum(lapc);
  if ((t & 63) == 0) red[t >> 6] = s;
  __syncthreads();
  if (t == 0) {
    A.st_lap[bd] = red[0] + red[1];
    A.st_logdet[bd] = S.logdet[w] + (ok ? log(fabs(ratio)) : 0.0);
    A.st_sgn[bd] = S.sgn[w] * (ratio < 0 ? -1.0 : 1.0);
    A.st_singular[bd] = (!ok || S.singular[w]) ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
mdet_assemble_kernel(MdetArgs A, const int *walkers, int num) {
  constexpr int NT = 256;
  const int b = blockIdx.x;
  if (b >= num) return;
  const long w0 = walkers ? walkers[b] : b;
  const int D = A.D, N = A.N, t = threadIdx.x;
  const SlaterArgs &Su = A.s[0], &Sd = A.s[1];
  const int nu = Su.n, nd = Sd.n;
  const bool staged = A.moved >= 0;
  const int ms = staged ? (A.moved < nu ? 0 : 1) : -1;
  __shared__ double t_new[kMaxDet], t_cur[kMaxDet];
  __shared__ double s_sc[8];
  __shared__ double scratch[5][NT / 64];
  __shared__ double s_red[5];
  const DetView ru = resident_view(Su, w0, D), rd = resident_view(Sd, w0, D);
  const DetView st{A.st_logdet, A.st_sgn, A.st_lap, A.st_grad, A.st_singular, (long)b * D};
  const DetView &nu_v = (ms == 0) ? st : ru;
  const DetView &nd_v = (ms == 1) ? st : rd;
  if (t == 0) {
    double lnD, sgD;
    const bool okn = combine(A, nu_v, nd_v, nu, nd, t_new, lnD, sgD);
    s_sc[0] = lnD; s_sc[1] = sgD; s_sc[2] = okn ? 0.0 : 1.0;
    if (staged) {
      const bool okc = combine(A, ru, rd, nu, nd, t_cur, lnD, sgD);
      s_sc[3] = lnD; s_sc[4] = sgD; s_sc[5] = okc ? 0.0 : 1.0;
    }
  }
  __syncthreads();
  const double *Uxn = staged ? A.pUx + (long)b * N * 3 : A.Ux + w0 * N * 3;
  const double *Uxc = A.Ux + w0 * N * 3;
  const double *Rw = A.R + w0 * N * 3;
  const double tau = A.tau;
  const bool qf = staged && A.mode >= 1 && A.drift_mod == QMCB_DRIFT_UMRIGAR93_UNEQUAL;
  double lsum = 0.0, gsq = 0.0, fwd = 0.0, rev = 0.0, dummy = 0.0;
  double vbt = 0.0, vbo = 0.0, vo = 0.0;  // |modified drift|^2 (trial, current), |v|^2 current
  for (int e = t; e < N; e += NT) {
    const int sp = e < nu ? 0 : 1;
    const int le = e - (sp == 0 ? 0 : nu);
    const int ne = sp == 0 ? nu : nd;
    const DetView &vn = sp == 0 ? nu_v : nd_v;
    const DetView &vc = sp == 0 ? ru : rd;
    double gd[3] = {0, 0, 0}, gc[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) {
      const double *g = vn.grad + ((vn.base + d) * ne + le) * 3;
      gd[0] += t_new[d] * g[0]; gd[1] += t_new[d] * g[1]; gd[2] += t_new[d] * g[2];
      if (staged && A.mode >= 1) {
        const double *h = vc.grad + ((vc.base + d) * ne + le) * 3;
        gc[0] += t_cur[d] * h[0]; gc[1] += t_cur[d] * h[1]; gc[2] += t_cur[d] * h[2];
      }
    }
    double *cg = staged ? A.st_cgrad + ((long)b * N + e) * 3 : A.cgrad + (w0 * N + e) * 3;
    cg[0] = gd[0]; cg[1] = gd[1]; cg[2] = gd[2];
    double g[3];
    for (int k = 0; k < 3; k++) {
      const double ux = Uxn[3 * e + k];
      g[k] = gd[k] + ux;
      lsum += 2.0 * ux * gd[k] + ux * ux;
      gsq += g[k] * g[k];
    }
    if (A.out.grad_lnpsi) {
      double *o = A.out.grad_lnpsi + ((long)b * N + e) * 3;
      o[0] = g[0]; o[1] = g[1]; o[2] = g[2];
    }
    if (staged && A.mode >= 1) {
      // Green's function terms (QMCWalker.cpp:343-398, 756-780) with the
      // drift gradPsiRatio or its umrigar93_unequalelectrons modification
      // (QMCSCFJastrow.cpp:368-443) at the trial (ft) and current (fo) states
      double c[3];
      for (int k = 0; k < 3; k++) c[k] = gc[k] + Uxc[3 * e + k];
      double ft = 1.0, fo = 1.0;
      if (qf) {
        const double xo[3] = {Rw[3 * e], Rw[3 * e + 1], Rw[3 * e + 2]};
        const int co = nearest_nucleus(xo, A.n_nuc, A.nuc_xyz);
        fo = drift_factor_near(c, xo, co, A.nuc_xyz, A.nuc_z, tau);
        if (e != A.moved) {
          ft = drift_factor_near(g, xo, co, A.nuc_xyz, A.nuc_z, tau);
        } else {
          const double xt[3] = {A.propR[3L * b], A.propR[3L * b + 1], A.propR[3L * b + 2]};
          ft = drift_factor_unequal(g, xt, A.n_nuc, A.nuc_xyz, A.nuc_z, tau);
        }
        double g2 = 0.0, c2 = 0.0;
        for (int k = 0; k < 3; k++) { g2 += g[k] * g[k]; c2 += c[k] * c[k]; }
        vbt += ft * ft * g2;
        vbo += fo * fo * c2;
        vo += c2;
      }
      if (e != A.moved) {
        for (int k = 0; k < 3; k++) {
          fwd += tau * fo * c[k] * tau * fo * c[k];
          rev += tau * ft * g[k] * tau * ft * g[k];
        }
      } else {
        for (int k = 0; k < 3; k++) {
          const double tt = Rw[3 * e + k] - A.propR[3L * b + k] - tau * ft * g[k];
          rev += tt * tt;
          fwd += A.gauss[3L * b + k] * A.gauss[3L * b + k];
        }
      }
    }
  }
  double vals[5] = {lsum, gsq, fwd, rev, dummy};
  block_sum5<NT>(vals, scratch, s_red);
  // |modified drift| / |drift| of the trial and current states (umrigar93
  // energy modification of the DMC weights)
  double mr_t = 1.0, mr_o = 1.0;
  if (qf) {
    __shared__ double s_red2[5];
    double v2[5] = {vbt, vbo, vo, 0.0, 0.0};
    block_sum5<NT>(v2, scratch, s_red2);
    if (s_red[1] > 0.0) mr_t = sqrt(s_red2[0]) / sqrt(s_red[1]);
    if (s_red2[2] > 0.0) mr_o = sqrt(s_red2[1]) / sqrt(s_red2[2]);
  }
  bool sing = s_sc[2] != 0.0;
  const double gsum = s_red[1];
  if (!(gsum > 0.0) || !isfinite(gsum)) sing = true;
  // drift modification (QMCSCFJastrow.cpp:329-472)
  if (A.out.mod_drift) {
    double feq = 1.0;
    if (A.drift_mod == QMCB_DRIFT_UMRIGAR93_EQUAL)
      feq = (-1.0 + sqrt(1.0 + 2 * A.umrigar_a * gsum * A.dt)) / (A.umrigar_a * gsum * A.dt);
    for (int e = t; e < N; e += NT) {
      const double *src = A.out.grad_lnpsi + ((long)b * N + e) * 3;
      const double g[3] = {src[0], src[1], src[2]};
      double f = feq;
      if (A.drift_mod == QMCB_DRIFT_UMRIGAR93_UNEQUAL) {
        double x[3];
        for (int k = 0; k < 3; k++) x[k] = (e == A.moved) ? A.propR[3L * b + k] : Rw[3 * e + k];
        f = drift_factor_unequal(g, x, A.n_nuc, A.nuc_xyz, A.nuc_z, A.dt);
      }
      double *o = A.out.mod_drift + ((long)b * N + e) * 3;
      o[0] = g[0] * f; o[1] = g[1] * f; o[2] = g[2] * f;
    }
  }
  if (t != 0) return;
  // Laplacian ratio: sum_d t_d (lap_up(d) + lap_dn(d)) + Jastrow terms
  double dlap = 0.0;
  for (int d = 0; d < D; d++) {
    double l = 0.0;
    if (nu > 0) l += nu_v.lap[nu_v.base + d];
    if (nd > 0) l += nd_v.lap[nd_v.base + d];
    dlap += t_new[d] * l;
  }
  const double Uxx = staged ? A.pUxx[b] : A.Uxx[w0];
  const double U = staged ? A.pU[b] : A.U[w0];
  const double Vee = staged ? A.pVee[b] : A.Vee[w0];
  const double Ven = staged ? A.pVen[b] : A.Ven[w0];
  const double lapR = dlap + Uxx + s_red[0];
  const double V = A.v_nn + Vee + Ven;
  const double EL = -0.5 * lapR + V;
  const double logpsi = s_sc[0] + U;
  const qmcb_results &o = A.out;
  if (o.log_psi) o.log_psi[b] = logpsi;
  if (o.sign_psi) o.sign_psi[b] = (int8_t)(s_sc[1] < 0 ? -1 : 1);
  if (o.lap_ratio) o.lap_ratio[b] = lapR;
  if (o.e_local) o.e_local[b] = EL;
  if (o.kinetic) o.kinetic[b] = -0.5 * lapR;
  if (o.potential) o.potential[b] = V;
  if (o.v_ne) o.v_ne[b] = Ven;
  if (o.v_ee) o.v_ee[b] = Vee;
  if (o.det_ratio)
    o.det_ratio[b] = staged ? s_sc[1] * s_sc[4] * exp(s_sc[0] - s_sc[3]) : 1.0;
  if (o.jastrow_u) o.jastrow_u[b] = U;
  if (o.singular) o.singular[b] = sing ? 1 : 0;
  if (staged) A.st_row_singular[b] = sing ? 1 : 0;
  if (!staged) {
    if (A.cur_el) A.cur_el[w0] = EL;
    if (A.walk.age) A.walk.age[w0] = 0;
    return;
  }
  if (A.mode < 1) return;
  // ---- Metropolis with the Green's function ratio (QMCWalker.cpp:1578-1685);
  // DMC also rejects a change of the sign of the determinant sum (fixed
  // node, :1641-1646) and reweights once per sweep as the one-determinant
  // path does (QMCWalker.cpp:1060-1185)
  const double logpsi_old = s_sc[3] + A.U[w0];
  double p = (sing || s_sc[5] != 0.0)
                 ? 0.0
                 : exp(2.0 * (logpsi - logpsi_old) - (s_red[3] - s_red[2]) / (2.0 * tau));
  const bool bad = sing || s_sc[5] != 0.0;
  const int age_old = A.walk.age ? A.walk.age[w0] : 0;
  if (!(p == p)) p = 0.0;
  else if (!bad) p = old_walker_boost(p, age_old, A.walk.old_walker);
  if (A.mode == 2 && s_sc[1] * s_sc[4] < 0.0) p = 0.0;
  if (p > 1.0) p = 1.0;
  const bool accept = p > A.unif[b];
  const int age_new = accept ? 0 : age_old + 1;  // QMCWalker.cpp:1656-1672
  if (A.walk.age) A.walk.age[w0] = age_new;
  const double el_old = A.cur_el[w0];
  A.mask[b] = accept ? 1 : 0;
  if (accept) A.cur_el[w0] = EL;
  const double el = accept ? EL : el_old;
  double wt = 1.0;
  if (A.mode == 2) {
    wt = A.weight[w0];
    if (A.reweight) {
      double so_, st_;
      if (A.energy_mod) {  // umrigar93 (QMCWalker.cpp:1121-1132)
        so_ = (A.e_trial - A.e_est) + (A.e_est - el_old) * mr_o;
        st_ = (A.e_trial - A.e_est) + (A.e_est - EL) * mr_t;
      } else {
        so_ = A.e_trial - el_old;
        st_ = A.e_trial - EL;
      }
      if (!(p > 0.0)) st_ = 0.0;
      const double dW = exp(A.tau_eff * (p * 0.5 * (so_ + st_) + (1.0 - p) * so_));
      wt = reweight_walker(wt, dW, age_new, el, A.walk);
      A.weight[w0] = wt;
    }
    A.acc_w[w0] += wt;
  }
  A.acc_el[w0] += wt * el;
  A.acc_el2[w0] += wt * el * el;
  A.acc_n[w0] += accept ? 1.0 : 0.0;
  A.acc_prop_el[w0] += sing ? 0.0 : EL;
  if (A.obs) {
    A.obs[3L * b] = p > 0.0 ? p * EL + (1.0 - p) * el_old : el_old;
    A.obs[3L * b + 1] = wt;
    A.obs[3L * b + 2] = p;
  }
}

// ---------------------------------------------------------------------------
template <typename TS, typename TO>
__global__ void __launch_bounds__(256)
mdet_commit_kernel(MdetArgs A, const int *walkers, const uint8_t *mask, int num) {
  const int sp = A.moved < A.s[0].n ? 0 : 1;
  const SlaterArgs &S = A.s[sp];
  const int D = A.D, d = blockIdx.x % D, b = blockIdx.x / D;
  if (b >= num || !mask[b] || A.st_row_singular[b]) return;
  const long w0 = walkers ? walkers[b] : b, w = w0 * D + d, bd = (long)b * D + d;
  const int n = S.n, nact = act_count(S), N = A.N, r = A.moved - S.spin_start;
  const long nn2 = (long)n * n, cs = (long)nact * n;
  const int t = threadIdx.x;
  double *inv = S.inv + w * nn2;
  const double *st = A.st_inv + bd * nn2;
  for (long i = t; i < nn2; i += blockDim.x) inv[i] = st[i];
  double *g = S.gradPR + w * n * 3;
  const double *sg = A.st_grad + bd * n * 3;
  for (int i = t; i < 3 * n; i += blockDim.x) g[i] = sg[i];
  if (t == 0) {
    S.lapPR[w] = A.st_lap[bd];
    S.logdet[w] = A.st_logdet[bd];
    S.sgn[w] = A.st_sgn[bd];
    S.singular[w] = 0;
    S.dcount[w] = 0;
  }
  if (d != 0) return;
  // walker-level state, once: the new column r of the shared orbital rows
  const TO *vr = (const TO *)A.vrow + b * A.v_s_walker;
  TS *dmw = (TS *)S.dmat + w0 * kDer * cs;
  double *dvw = S.dval + w0 * cs;
  for (int i = t; i < kComp * nact; i += blockDim.x) {
    const int k = i / nact, o = i % nact;
    const double x = (double)vr[(long)k * A.v_s_comp + o];
    if (k == 0) dvw[(long)o * n + r] = x;
    else dmw[dm_at<TS>(k - 1, (long)o * n + r, cs)] = (TS)x;
  }
  double *Uxw = A.Ux + w0 * N * 3;
  const double *pUx = A.pUx + (long)b * N * 3;
  for (int i = t; i < 3 * N; i += blockDim.x) Uxw[i] = pUx[i];
  double *cg = A.cgrad + w0 * N * 3;
  const double *scg = A.st_cgrad + (long)b * N * 3;
  for (int i = t; i < 3 * N; i += blockDim.x) cg[i] = scg[i];
  if (t == 0) {
    double *Rm = A.R + (w0 * N + A.moved) * 3;
    Rm[0] = A.propR[3L * b]; Rm[1] = A.propR[3L * b + 1]; Rm[2] = A.propR[3L * b + 2];
    A.U[w0] = A.pU[b];
    A.Uxx[w0] = A.pUxx[b];
    A.Vee[w0] = A.pVee[b];
    A.Ven[w0] = A.pVen[b];
  }
}

// ---------------------------------------------------------------------------
template <typename TS, typename TO>
hipError_t launch_mdet_move(const MdetArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  if (a.D > kMaxDet || a.s[0].n > 128 || a.s[1].n > 128) return hipErrorInvalidValue;
  hipLaunchKernelGGL((mdet_move_kernel<TS, TO>), dim3((unsigned)num * a.D), dim3(128), 0, st, a,
                     walkers, num);
  return hipGetLastError();
}

hipError_t launch_mdet_assemble(const MdetArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  if (a.D > kMaxDet) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mdet_assemble_kernel, dim3(num), dim3(256), 0, st, a, walkers, num);
  return hipGetLastError();
}

template <typename TS, typename TO>
hipError_t launch_mdet_commit(const MdetArgs &a, const int *walkers, const uint8_t *mask, int num,
                              hipStream_t st) {
  if (num <= 0) return hipSuccess;
  hipLaunchKernelGGL((mdet_commit_kernel<TS, TO>), dim3((unsigned)num * a.D), dim3(256), 0, st,
                     a, walkers, mask, num);
  return hipGetLastError();
}

#define INST(TS, TO)                                                                           \
  template hipError_t launch_mdet_move<TS, TO>(const MdetArgs &, const int *, int, hipStream_t); \
  template hipError_t launch_mdet_commit<TS, TO>(const MdetArgs &, const int *, const uint8_t *, \
                                                 int, hipStream_t);
INST(double, double)
INST(float, float)
INST(float, double)
INST(bf16s, float)
#undef INST

}  // namespace qmcb
