// qmcb_common.hpp -- device-side data structures shared by the kernels.
//
// Layout of the resident walker state in HBM (one context, capacity Wc):
//   R        [Wc][N][3]           FP64 positions, alpha electrons first
//   inv[s]   [Wc][n][n]           FP64 inverse Slater matrix, element (o,e)
//   dval[s]  [Wc][n][n]           FP64 value matrix, dval(o, e) = D(e, o)
//   dmat[s]  [Wc][n][n][4]        Tstore, the four components k in {d/dx,
//                                 d/dy, d/dz, Laplacian} of an element
//                                 together: dmat(o, e, k) = d_k D(e, o)
//   logdet[s], sgn[s] [Wc]        FP64 ln|det|, sign (+1/-1)
//   gradPR[s] [Wc][n][3]          FP64 grad det / det per electron
//   lapPR[s]  [Wc]                FP64 Laplacian det / det
//   U, Uxx, Vee, Ven [Wc]; Ux [Wc][N][3]   FP64 Jastrow / potential sums
// The transposed derivative storage lets one thread own one electron column
// of A^-1 and stream the matching column of the derivative matrices with
// coalesced loads (64 lanes = 64 consecutive electrons), all four components
// of an element in one load.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/qmcbeaver_amd.h"

namespace qmcb {

constexpr int kComp = 5;  // value, x, y, z, Laplacian

struct DevSystem {
  int N, n[2], start[2], nn, B, nshell, nprim;
  const double *nuc_xyz, *nuc_z;
  const int *sh_atom, *sh_prim_off, *sh_nprim, *sh_bf0, *sh_ncart;
  const int *bf_pow;  // [B][3]
  const int *chunk_beg;  // [nchunk+1] shell ranges of the basis kernel's grid.y
  int nchunk;
  const double *prim_alpha, *prim_coef;
  qmcb_pade ee[3];                    // up-dn, up-up, dn-dn
  const qmcb_pade *en[2];             // [nn] per spin
  double v_nn;
};

// ---- Pade u(r) = P(r)/Q(r), P(0)=0, Q(0)=1 (PadeCorrelationFunction.cpp:64-83)
// The quotient-rule derivatives follow from u Q = P differentiated once and
// twice: u' = (P' - u Q') / Q, u'' = (P'' - 2 u' Q' - u Q'') / Q, so one
// reciprocal of Q serves all three.
__device__ __forceinline__ void pade_eval(const qmcb_pade &f, double r, double &u,
                                          double &du, double &d2u) {
  // Horner for value, first and second derivative (Polynomial.cpp:127-149)
  double a = 0.0, ap = 0.0, app = 0.0;
  if (f.n_num > 0) {
    a = f.num[f.n_num - 1];
    for (int i = f.n_num - 2; i >= -1; i--) {
      double c = i >= 0 ? f.num[i] : 0.0;
      app = app * r + ap;
      ap = ap * r + a;
      a = a * r + c;
    }
    app *= 2.0;
  }
  double b = 1.0, bp = 0.0, bpp = 0.0;
  if (f.n_den > 0) {
    b = f.den[f.n_den - 1];
    for (int i = f.n_den - 2; i >= -1; i--) {
      double c = i >= 0 ? f.den[i] : 1.0;
      bpp = bpp * r + bp;
      bp = bp * r + b;
      b = b * r + c;
    }
    bpp *= 2.0;
  }
  const double ib = 1.0 / b;
  u = a * ib;
  du = (ap - u * bp) * ib;
  d2u = (app - 2.0 * du * bp - u * bpp) * ib;
}

// e-e Pade type of the pair (i, j): 0 up-dn, 1 up-up, 2 dn-dn (DevSystem::ee)
__device__ __forceinline__ int ee_type(int i, int j, int n_up) {
  const bool ui = i < n_up, uj = j < n_up;
  return (ui != uj) ? 0 : (ui ? 1 : 2);
}

// ---- walker bookkeeping of the reference's QMCWalker ------------------------
// Acceptance of an old walker (QMCWalker.cpp:1578-1600): after more than
// old_walker consecutive rejected moves the probability is raised by
// 1.1^(age - old_walker), and set to 1 beyond 500 more.
__device__ __forceinline__ double old_walker_boost(double p, int age, int old_walker) {
  if (old_walker <= 0) return p;
  const int a = age - old_walker;
  if (a > 500) return 1.0;
  if (a > 0) return p * pow(1.1, (double)a);
  return p;
}

// DMC weight after one sweep's reweighting factor dW (QMCWalker.cpp:1180-1395):
// an old walker decays by 0.9 instead of dW; then the reference's deletion
// rules (weight 0, removed at the next branching).  Equilibration
// (iteration < -100): heavy (> 2 branch_threshold), light (< 0.1), dW outside
// [0.5, 1.5], a bad energy, or age >= 30.  Production: dW^(age+1) > 15, weight
// > 50 or a bad energy.  A bad energy is |E_L - e_orig| / |e_orig| > rel_cutoff
// (rel_cutoff <= 0: no energy rule); age is the count of consecutive
// rejected moves after this move.
struct WalkCtl {
  int *age;              // [Wc] consecutive rejected moves (null: no age rules)
  int old_walker;        // old_walker_acceptance_parameter (<= 0: off)
  int equilibrating;     // the equilibration rules apply at this reweighting
  double rel_cutoff;
  double e_orig;         // energy_estimated_original
  double branch_threshold;
};

__device__ inline double reweight_walker(double wt, double dW, int age, double el,
                                         const WalkCtl &c) {
  wt = (c.old_walker > 0 && age - c.old_walker > 0) ? wt * 0.9 : wt * dW;
  const bool bad_energy = c.rel_cutoff > 0.0 && fabs((el - c.e_orig) / c.e_orig) > c.rel_cutoff;
  if (c.equilibrating) {
    if (wt > 2.0 * c.branch_threshold || wt < 0.1 || !(dW <= 1.5 && dW >= 0.5) || bad_energy ||
        age >= 30)
      return 0.0;
    return wt;
  }
  const double growth = pow(dW, age + 1.0);
  if (growth > 15.0 || dW != dW || wt > 50.0 || bad_energy) return 0.0;
  return wt;
}

// ---- wavefront reductions (64 lanes) ---------------------------------------
// umrigar93_unequalelectrons drift factor of one electron with gradient g at
// x: nearest nucleus, its charge and the time step (QMCSCFJastrow.cpp:368-443).
// The nearest nucleus (first of equals) and the factor given it are split so
// that callers evaluating two gradients at one position search once.
__device__ inline int nearest_nucleus(const double x[3], int nn, const double *nuc_xyz) {
  int cn = 0;
  double best = 1e300;
  for (int A = 0; A < nn; A++) {
    const double dx = x[0] - nuc_xyz[3 * A], dy = x[1] - nuc_xyz[3 * A + 1],
                 dz = x[2] - nuc_xyz[3 * A + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r < best) { best = r; cn = A; }
  }
  return cn;
}

__device__ inline double drift_factor_near(const double g[3], const double x[3], int cn,
                                           const double *nuc_xyz, const double *nuc_z,
                                           double dt) {
  double nv[3], dsq = 0.0;
  for (int k = 0; k < 3; k++) { nv[k] = x[k] - nuc_xyz[3 * cn + k]; dsq += nv[k] * nv[k]; }
  const double inr = 1.0 / sqrt(dsq);
  const double v2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
  if (fabs(v2) < 1e-10) return 1.0;
  const double iv = 1.0 / sqrt(v2);
  const double dotp = (nv[0] * g[0] + nv[1] * g[1] + nv[2] * g[2]) * inr * iv;
  const double Z = nuc_z[cn];
  const double t = Z * Z * dsq;
  const double a = 0.5 * (1.0 + dotp) + t / (10.0 * (4.0 + t));
  return (-1.0 + sqrt(1.0 + 2.0 * a * v2 * dt)) / (a * v2 * dt);
}

__device__ inline double drift_factor_unequal(const double g[3], const double x[3], int nn,
                                              const double *nuc_xyz, const double *nuc_z,
                                              double dt) {
  return drift_factor_near(g, x, nearest_nucleus(x, nn, nuc_xyz), nuc_xyz, nuc_z, dt);
}

// Cross-lane moves through DPP (data-parallel primitives: the VALU reads a
// neighbouring lane's register in the same instruction) instead of the LDS
// permute path: a few cycles per step instead of an LDS round trip.  Every
// pattern used here has a valid source lane for every lane.
//   0xB1 quad_perm [1,0,3,2]: lane ^ 1      0x4E quad_perm [2,3,0,1]: lane ^ 2
//   0xA0 quad_perm [0,0,2,2]: even lane of each pair
//   0x141 row_half_mirror: i <-> 7 - i in each group of 8
//   0x140 row_mirror: i <-> 15 - i in each row of 16
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = dpp_i32<CTRL>((int)b), hi = dpp_i32<CTRL>((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the 64 lanes of a wave (all lanes active); every lane gets it.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);  // each lane: the sum of its row of 16
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Largest value over each row of 16 lanes with the lowest index among
// equals (the row's lanes active); every lane of the row gets the winner.
__device__ __forceinline__ void row16_argmax(double &best, int &bi) {
  auto take = [&](double ob, int oi) {
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
  };
  take(dpp_f64<0xB1>(best), dpp_i32<0xB1>(bi));
  take(dpp_f64<0x4E>(best), dpp_i32<0x4E>(bi));
  take(dpp_f64<0x141>(best), dpp_i32<0x141>(bi));
  take(dpp_f64<0x140>(best), dpp_i32<0x140>(bi));
}

// ---- Philox4x32-10 (Salmon et al., SC'11) counter-based RNG ---------------
struct u4 { uint32_t x, y, z, w; };

__host__ __device__ inline u4 philox4x32(u4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int i = 0; i < 10; i++) {
    uint64_t p0 = (uint64_t)M0 * ctr.x, p1 = (uint64_t)M1 * ctr.z;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    u4 n = {hi1 ^ ctr.y ^ k0, lo1, hi0 ^ ctr.w ^ k1, lo0};
    ctr = n;
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}

// uniform in (0,1] from 32 random bits pair (53-bit mantissa)
__host__ __device__ inline double u01(uint32_t a, uint32_t b) {
  uint64_t v = ((uint64_t)a << 21) ^ (uint64_t)b;
  v &= ((1ULL << 53) - 1);
  return ((double)v + 1.0) * (1.0 / 9007199254740992.0);
}

}  // namespace qmcb
