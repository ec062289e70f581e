// jastrow.hip -- kernel (d): two-body Pade Jastrow (e-e, e-n), Coulomb
// potential and local-energy assembly, FP64.
//
// Reference: QMCJastrow::evaluate (QMCJastrow.cpp:134-187),
// QMCJastrowElectronElectron::updateAll/updateOne/collectForPair
// (QMCJastrowElectronElectron.cpp:56-298), QMCJastrowElectronNuclear::
// evaluate (QMCJastrowElectronNuclear.cpp:54-187), QMCPotential_Energy::
// calc_P_en/calc_P_ee (QMCPotential_Energy.cpp:242-333) and the assembly in
// QMCSCFJastrow::calculate_Psi_quantities / calculate_E_Local
// (QMCSCFJastrow.cpp:196-282, QMCSCFJastrow.h:311-314).
//
// Per pair: u, u' and u'' of the Pade function; U gets u, the gradient of
// electron i gets u' (x_i - x_j)/r, the Laplacian gets 2(2u'/r + u'') per
// e-e pair (both electrons) and 2u'/r + u'' per e-n pair.  The reference
// caches every pair term in N x N arrays; here the single-electron update
// recomputes the moved electron's old pair terms from its old position
// instead (same values, no O(N^2) per-walker cache in HBM).
#include "qmcb_kernels.hpp"

namespace qmcb {


template <int NT>
__device__ __forceinline__ void block_sum(double *vals, int nv, double *scratch /*[nv][NT/64]*/) {
  constexpr int NW = NT / 64;
  const int t = threadIdx.x;
  for (int k = 0; k < nv; k++) {
    const double s = wave_sum(vals[k]);
    if ((t & 63) == 0) scratch[k * NW + (t >> 6)] = s;
  }
  __syncthreads();
  for (int k = 0; k < nv; k++) {
    double s = 0.0;
    for (int q = 0; q < NW; q++) s += scratch[k * NW + q];
    vals[k] = s;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// all-electron Jastrow + potential, one workgroup per walker
constexpr int kJT = 256;
constexpr int kMaxN = 512;

__global__ void __launch_bounds__(kJT) jastrow_all_kernel(JastrowArgs A, const int *walkers, int num) {
  const int b = blockIdx.x;
  if (b >= num) return;
  const int w = walkers ? walkers[b] : b;
  const DevSystem &S = A.sys;
  const int N = S.N, nn = S.nn, nu = S.n[0];
  __shared__ double x[kMaxN * 3];
  __shared__ double scratch[4 * (kJT / 64)];
  const double *Rw = A.R + (long)w * N * 3;
  for (int i = threadIdx.x; i < 3 * N; i += kJT) x[i] = Rw[i];
  __syncthreads();
  double U = 0.0, Uxx = 0.0, Vee = 0.0, Ven = 0.0;
  for (int i = threadIdx.x; i < N; i += kJT) {
    double g[3] = {0.0, 0.0, 0.0};
    const int sp = i < nu ? 0 : 1;
    for (int A_ = 0; A_ < nn; A_++) {
      double d[3], r2 = 0.0;
      for (int k = 0; k < 3; k++) { d[k] = x[3 * i + k] - S.nuc_xyz[3 * A_ + k]; r2 += d[k] * d[k]; }
      const double r = sqrt(r2), ir = 1.0 / r;
      double u, du, d2u;
      pade_eval(S.en[sp][A_], r, u, du, d2u);
      U += u;
      Uxx += 2.0 * ir * du + d2u;
      for (int k = 0; k < 3; k++) g[k] += du * ir * d[k];
      Ven -= S.nuc_z[A_] * ir;
    }
    for (int j = 0; j < N; j++) {
      if (j == i) continue;
      double d[3], r2 = 0.0;
      for (int k = 0; k < 3; k++) { d[k] = x[3 * i + k] - x[3 * j + k]; r2 += d[k] * d[k]; }
      const double r = sqrt(r2), ir = 1.0 / r;
      double u, du, d2u;
      pade_eval(S.ee[ee_type(i, j, nu)], r, u, du, d2u);
      for (int k = 0; k < 3; k++) g[k] += du * ir * d[k];
      if (j < i) {
        U += u;
        Uxx += 2.0 * (2.0 * ir * du + d2u);
        Vee += ir;
      }
    }
    double *ux = A.Ux + ((long)w * N + i) * 3;
    ux[0] = g[0]; ux[1] = g[1]; ux[2] = g[2];
  }
  double v[4] = {U, Uxx, Vee, Ven};
  block_sum<kJT>(v, 4, scratch);
  if (threadIdx.x == 0) {
    A.U[w] = v[0];
    A.Uxx[w] = v[1];
    A.Vee[w] = v[2];
    A.Ven[w] = v[3];
  }
}

hipError_t launch_jastrow_all(const JastrowArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  if (a.sys.N > kMaxN) return hipErrorInvalidValue;
  hipLaunchKernelGGL(jastrow_all_kernel, dim3(num), dim3(kJT), 0, st, a, walkers, num);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// single-electron Jastrow + potential at the proposal (updateOne,
// QMCJastrowElectronElectron.cpp:68-152; e-n of the moved electron only,
// QMCJastrowElectronNuclear.cpp:117-120; Coulomb difference of its pairs).
// Writes U', grad U' for all N electrons, Lap U', V_ee', V_en' to batch row b.
constexpr int kMT = 128;

// At most 76 VGPRs (6 waves per SIMD, no spills): 10 % faster alone than at
// its free register count (122, 4 waves), and its waves fit beside sm_move's
// (3 or 2 waves per SIMD) when walker groups overlap the two kernels.
#ifndef QMCB_JMOVE_WAVES
#define QMCB_JMOVE_WAVES 6
#endif
__global__ void __launch_bounds__(kMT) __attribute__((amdgpu_waves_per_eu(QMCB_JMOVE_WAVES, 8)))
jastrow_move_kernel(JastrowArgs A, const double *cU, const double *cUx, const double *cUxx,
                    const double *cVee, const double *cVen, const int *walkers, int num) {
  const int b = blockIdx.x;
  if (b >= num) return;
  const int w = walkers ? walkers[b] : b;
  const DevSystem &S = A.sys;
  const int N = S.N, nn = S.nn, nu = S.n[0], m = A.moved;
  __shared__ double scratch[7 * (kMT / 64)];
  const double *Rw = A.R + (long)w * N * 3;
  const double xo[3] = {Rw[3 * m], Rw[3 * m + 1], Rw[3 * m + 2]};
  const double xn[3] = {A.propR[3L * b], A.propR[3L * b + 1], A.propR[3L * b + 2]};
  const double *ux = cUx + (long)w * N * 3;
  double *pux = A.Ux + (long)b * N * 3;
  double dU = 0.0, dUxx = 0.0, dVee = 0.0, dVen = 0.0, gm[3] = {0.0, 0.0, 0.0};
  for (int j = threadIdx.x; j < N; j += kMT) {
    if (j == m) continue;
    const double xj[3] = {Rw[3 * j], Rw[3 * j + 1], Rw[3 * j + 2]};
    double dn[3], dold[3], rn2 = 0.0, ro2 = 0.0;
    for (int k = 0; k < 3; k++) {
      dn[k] = xn[k] - xj[k]; rn2 += dn[k] * dn[k];
      dold[k] = xo[k] - xj[k]; ro2 += dold[k] * dold[k];
    }
    const double rn = sqrt(rn2), ro = sqrt(ro2), irn = 1.0 / rn, iro = 1.0 / ro;
    const qmcb_pade &f = S.ee[ee_type(m, j, nu)];
    double un, dun, d2un, uo, duo, d2uo;
    pade_eval(f, rn, un, dun, d2un);
    pade_eval(f, ro, uo, duo, d2uo);
    dU += un - uo;
    dUxx += 2.0 * (2.0 * irn * dun + d2un) - 2.0 * (2.0 * iro * duo + d2uo);
    dVee += irn - iro;
    const double gn = dun * irn, go = duo * iro;
    for (int k = 0; k < 3; k++) {
      // gradient on j: u'(r) (x_j - x_m)/r
      pux[3 * j + k] = ux[3 * j + k] - gn * dn[k] + go * dold[k];
      gm[k] += gn * dn[k];
    }
  }
  const int sp = m < nu ? 0 : 1;
  for (int a_ = threadIdx.x; a_ < nn; a_ += kMT) {
    double dn[3], dold[3], rn2 = 0.0, ro2 = 0.0;
    for (int k = 0; k < 3; k++) {
      dn[k] = xn[k] - S.nuc_xyz[3 * a_ + k]; rn2 += dn[k] * dn[k];
      dold[k] = xo[k] - S.nuc_xyz[3 * a_ + k]; ro2 += dold[k] * dold[k];
    }
    const double rn = sqrt(rn2), ro = sqrt(ro2), irn = 1.0 / rn, iro = 1.0 / ro;
    double un, dun, d2un, uo, duo, d2uo;
    pade_eval(S.en[sp][a_], rn, un, dun, d2un);
    pade_eval(S.en[sp][a_], ro, uo, duo, d2uo);
    dU += un - uo;
    dUxx += (2.0 * irn * dun + d2un) - (2.0 * iro * duo + d2uo);
    dVen += S.nuc_z[a_] * (iro - irn);
    const double gn = dun * irn;
    for (int k = 0; k < 3; k++) gm[k] += gn * dn[k];
  }
  double v[7] = {dU, dUxx, dVee, dVen, gm[0], gm[1], gm[2]};
  block_sum<kMT>(v, 7, scratch);
  if (threadIdx.x == 0) {
    pux[3 * m] = v[4];
    pux[3 * m + 1] = v[5];
    pux[3 * m + 2] = v[6];
  }
  if (threadIdx.x == 1) {
    // outputs of the proposal, batch row b (A.U etc. point at the proposal arrays)
    A.U[b] = cU[w] + v[0];
    A.Uxx[b] = cUxx[w] + v[1];
    A.Vee[b] = cVee[w] + v[2];
    A.Ven[b] = cVen[w] + v[3];
  }
}

hipError_t launch_jastrow_move(const JastrowArgs &a, const double *U, const double *Ux,
                               const double *Uxx, const double *Vee, const double *Ven,
                               const int *walkers, int num, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  hipLaunchKernelGGL(jastrow_move_kernel, dim3(num), dim3(kMT), 0, st, a, U, Ux, Uxx, Vee, Ven,
                     walkers, num);
  return hipGetLastError();
}


// all-electron assembly, one workgroup per walker
__global__ void __launch_bounds__(kJT) assemble_kernel(AssembleArgs A, const int *walkers, int num) {
  const int b = blockIdx.x;
  if (b >= num) return;
  const int w = walkers ? walkers[b] : b;
  const DevSystem &S = A.sys;
  const int N = S.N, nu = S.n[0];
  __shared__ double scratch[2 * (kJT / 64)];
  const double *ux = A.Ux + (long)w * N * 3;
  const double *Rw = A.R + (long)w * N * 3;
  double lsum = 0.0, gsq = 0.0;
  for (int e = threadIdx.x; e < N; e += kJT) {
    const int sp = e < nu ? 0 : 1;
    const int le = e - S.start[sp];
    const double *gp = A.s[sp].gradPR + ((long)w * S.n[sp] + le) * 3;
    double g[3];
    for (int k = 0; k < 3; k++) {
      g[k] = gp[k] + ux[3 * e + k];
      lsum += ux[3 * e + k] * gp[k] * 2.0;
      lsum += ux[3 * e + k] * ux[3 * e + k];
      gsq += g[k] * g[k];
    }
    if (A.out.grad_lnpsi) {
      double *o = A.out.grad_lnpsi + ((long)b * N + e) * 3;
      o[0] = g[0]; o[1] = g[1]; o[2] = g[2];
    }
  }
  double v[2] = {lsum, gsq};
  block_sum<kJT>(v, 2, scratch);
  bool sing = false;
  double logdet = 0.0, sgn = 1.0, lapR = A.Uxx[w] + v[0];
  for (int sp = 0; sp < 2; sp++) {
    if (S.n[sp] == 0) continue;
    sing = sing || A.s[sp].singular[w];
    logdet += A.s[sp].logdet[w];
    sgn *= A.s[sp].sgn[w];
    lapR += A.s[sp].lapPR[w];
  }
  if (!(v[1] > 0.0) || !isfinite(v[1])) sing = true;
  if (A.out.mod_drift) {
    double feq = 1.0;
    if (A.drift_mod == QMCB_DRIFT_UMRIGAR93_EQUAL)
      feq = (-1.0 + sqrt(1.0 + 2 * A.umrigar_a * v[1] * A.dt)) / (A.umrigar_a * v[1] * A.dt);
    for (int e = threadIdx.x; e < N; e += kJT) {
      const int sp = e < nu ? 0 : 1;
      const int le = e - S.start[sp];
      const double *gp = A.s[sp].gradPR + ((long)w * S.n[sp] + le) * 3;
      double g[3];
      for (int k = 0; k < 3; k++) g[k] = gp[k] + ux[3 * e + k];
      double f = feq;
      if (A.drift_mod == QMCB_DRIFT_UMRIGAR93_UNEQUAL) {
        const double x[3] = {Rw[3 * e], Rw[3 * e + 1], Rw[3 * e + 2]};
        f = drift_factor_unequal(g, x, S.nn, S.nuc_xyz, S.nuc_z, A.dt);
      }
      double *o = A.out.mod_drift + ((long)b * N + e) * 3;
      o[0] = g[0] * f; o[1] = g[1] * f; o[2] = g[2] * f;
    }
  }
  if (threadIdx.x == 0) {
    const double V = S.v_nn + A.Vee[w] + A.Ven[w];
    const double EL = -0.5 * lapR + V;
    const qmcb_results &o = A.out;
    if (o.log_psi) o.log_psi[b] = logdet + A.U[w];
    if (o.sign_psi) o.sign_psi[b] = (int8_t)(sgn < 0 ? -1 : 1);
    if (o.lap_ratio) o.lap_ratio[b] = lapR;
    if (o.e_local) o.e_local[b] = EL;
    if (o.kinetic) o.kinetic[b] = -0.5 * lapR;
    if (o.potential) o.potential[b] = V;
    if (o.v_ne) o.v_ne[b] = A.Ven[w];
    if (o.v_ee) o.v_ee[b] = A.Vee[w];
    if (o.det_ratio) o.det_ratio[b] = 1.0;
    if (o.jastrow_u) o.jastrow_u[b] = A.U[w];
    if (o.singular) o.singular[b] = sing ? 1 : 0;
    if (A.cur_el) A.cur_el[w] = EL;
    if (A.age) A.age[w] = 0;
  }
}

hipError_t launch_assemble(const AssembleArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  hipLaunchKernelGGL(assemble_kernel, dim3(num), dim3(kJT), 0, st, a, walkers, num);
  return hipGetLastError();
}

// modificationRatio and x2/y2/z2 of an evaluate (QMCSCFJastrow.cpp:164-169,
// 450-458), one wave per row, from the drift outputs already written and the
// evaluated configuration
__global__ void __launch_bounds__(64)
results_extra_kernel(qmcb_results out, const double *R, const int *walkers, const double *propR,
                     int whichE, int N, int num) {
  const int b = blockIdx.x;
  if (b >= num) return;
  const long w = walkers ? walkers[b] : b;
  const double *Rw = R + w * N * 3;
  double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // |F|^2, |F_mod|^2, x2, y2, z2
  for (int e = threadIdx.x; e < N; e += 64) {
    if (out.mod_ratio) {
      const double *g = out.grad_lnpsi + ((long)b * N + e) * 3;
      const double *f = out.mod_drift + ((long)b * N + e) * 3;
      for (int k = 0; k < 3; k++) { s[0] += g[k] * g[k]; s[1] += f[k] * f[k]; }
    }
    const double *x = e == whichE ? propR + 3L * b : Rw + 3 * e;
    for (int k = 0; k < 3; k++) s[2 + k] += x[k] * x[k];
  }
  for (int k = 0; k < 5; k++) s[k] = wave_sum(s[k]);
  if (threadIdx.x == 0) {
    if (out.mod_ratio) out.mod_ratio[b] = sqrt(s[1]) / sqrt(s[0]);
    if (out.r2)
      for (int k = 0; k < 3; k++) out.r2[3L * b + k] = s[2 + k];
  }
}

hipError_t launch_results_extra(const qmcb_results &out, const double *R, const int *walkers,
                                const double *propR, int whichE, int N, int num, hipStream_t st) {
  if (num <= 0 || (!out.mod_ratio && !out.r2)) return hipSuccess;
  if (out.mod_ratio && (!out.grad_lnpsi || !out.mod_drift)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(results_extra_kernel, dim3(num), dim3(64), 0, st, out, R, walkers, propR,
                     whichE, N, num);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Importance-sampled proposal of electron `moved` (QMCWalker.cpp:343-398):
//   x' = x + tau F + sqrt(tau) g,  g ~ N(0,1)^3,  F = grad ln Psi (drift)
// Random numbers: Philox4x32-10 keyed by (seed), counter (step, walker),
// so every walker's stream is independent of the batch it runs in.
__global__ void __launch_bounds__(256) propose_kernel(ProposeArgs A, const int *walkers, int num) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= num) return;
  const int w = walkers ? walkers[b] : b;
  const int m = A.moved;
  const int sp = m < A.n_up ? 0 : 1;
  const int le = sp == 0 ? m : m - A.n_up;
  const int nsp = sp == 0 ? A.n_up : A.N - A.n_up;
  const double *gp = A.cgrad ? A.cgrad + ((long)w * A.N + m) * 3
                             : A.gradPR[sp] + ((long)w * nsp + le) * 3;
  const double *ux = A.Ux + ((long)w * A.N + m) * 3;
  const double *xp = A.R + ((long)w * A.N + m) * 3;
  const double v[3] = {gp[0] + ux[0], gp[1] + ux[1], gp[2] + ux[2]};
  const double x[3] = {xp[0], xp[1], xp[2]};
  propose_row(A, A.row_from_walker ? (long)w - A.row_base : (long)b, w, v, x);
}

hipError_t launch_propose(const ProposeArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  hipLaunchKernelGGL(propose_kernel, dim3((num + 255) / 256), dim3(256), 0, st, a, walkers, num);
  return hipGetLastError();
}

}  // namespace qmcb
