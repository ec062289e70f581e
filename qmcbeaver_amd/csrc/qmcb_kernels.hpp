// qmcb_kernels.hpp -- kernel launch interfaces (host side) and the POD
// argument records shared by the kernels.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "qmcb_common.hpp"

namespace qmcb {

// Derivative-matrix storage in BF16: the upper 16 bits of an FP32 (8-bit
// mantissa, the FP32 exponent range), rounded to nearest even on the way in
// and widened exactly on the way out; arithmetic is never done in it.
struct bf16s {
  uint16_t b;
  bf16s() = default;
  __host__ __device__ bf16s(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) {
      b = (uint16_t)((u >> 16) | 0x40u);  // NaN stays a quiet NaN
    } else {
      u += 0x7fffu + ((u >> 16) & 1u);
      b = (uint16_t)(u >> 16);
    }
  }
  __host__ __device__ bf16s(double d) : bf16s((float)d) {}
  __host__ __device__ operator double() const {
    return (double)__builtin_bit_cast(float, (uint32_t)b << 16);
  }
};
static_assert(sizeof(bf16s) == 2, "bf16s is two bytes");

// A proposed derivative value as the storage type will hold it: with BF16
// storage the moved electron's new row enters the traces rounded, as every
// other electron's stored row does (so a resident state equals a fresh
// evaluate); the wider storage types take the GEMM output as it is.
template <typename TS>
__device__ __forceinline__ double as_stored(double x) { return x; }
template <>
__device__ __forceinline__ double as_stored<bf16s>(double x) { return (double)bf16s(x); }

// Offset of derivative component k (x, y, z, Laplacian) at ro = o * n + e in
// one walker's stored derivative matrices ([rows][n][4]: the four components
// of an element interleaved, so that a lane streaming a row reads them with
// one 8-byte (BF16), one 16-byte (FP32) or two 16-byte (FP64) loads; cs, the
// elements of one component, is kept for other layouts).
template <typename TS>
__host__ __device__ __forceinline__ long dm_at(int k, long ro, long cs) {
  (void)cs;
  return (ro << 2) + k;
}

// A batch of evaluation points: point p lives at
//   base + (p / per_walker) * walker_stride + (p % per_walker) * point_stride
//   (with walker_map: base + walker_map[p / per_walker] * walker_stride + ...)
struct PointSet {
  const double *base;
  long P;
  int per_walker;
  long walker_stride;
  int point_stride;
  const int *walker_map;
};

// Output addressing of the orbital GEMM: element (component k, orbital o,
// point p) is stored at
//   out + k*s_comp + o*s_orb + (p / per_walker)*s_walker + (p % per_walker)*s_elec
//   (the walker index goes through walker_map when it is non-null)
//   With out_val set, component 0 goes to out_val (FP64) at
//   o*s_orb_val + walker*s_walker_val + (p % per_walker)*s_elec_val, and
//   components 1..4 to out at (k-1)*s_comp.
struct GemmOut {
  void *out;
  long s_comp, s_orb, s_walker, s_elec;
  int per_walker;
  const int *walker_map;
  double *out_val;
  long s_walker_val, s_orb_val, s_elec_val;
};

// cusp replacement added to GEMM output rows (device tables in qmcb_cusp.hpp)
struct CuspDev;
template <typename TO>
hipError_t launch_cusp(const CuspDev &c, const PointSet &pts, const GemmOut &o, hipStream_t st);

// scatter batch rows of positions into resident slots: R[map[b]] = Rin[b]
hipError_t launch_scatter_rows(double *R, const double *Rin, const int *map, int num, int row_len,
                               hipStream_t st);
// zero rows [B, ldk) of the five components of X[5][ldk][ldp] (elem bytes each)
hipError_t launch_zero_pad_rows(void *X, size_t elem, int B, int ldk, long ldp, hipStream_t st);

// ---- kernel (a) ----
template <typename TC, typename TS>
hipError_t launch_basis(const DevSystem &sys, const PointSet &pts, TS *X, int ldk, long ldp,
                        hipStream_t st);

// ---- kernel (b): D^T[k][o][p] = sum_b Ct[b][o] X[k][b][p] on MFMA ----
// TI: operand type (double, float, bf16); TO: output type (double/float)
template <typename TI, typename TO>
hipError_t launch_orbital_gemm(const TI *Ct, int ldc, int n_orb, const TI *X, int ldk, long ldp,
                               long P, int kpad, const GemmOut &o, hipStream_t st,
                               int wpb = 4, int mt = 0);
// orbital tiles per wave (16 orbitals each) chosen by rule for P rows, and
// whether mt tiles keep every tile inside ldc padded orbitals
int gemm_orbital_tiles(int n_orb, long P);
bool gemm_orbital_tiles_valid(int mt, int n_orb, int ldc);

// ---- kernel (c): Slater inverse, SM updates, traces ----
struct SlaterArgs {
  int n;                 // electrons of this spin
  int N;                 // all electrons
  int spin_start;        // first electron index of this spin
  double *inv;           // [Wc][n][n]
  double *dval;          // [Wc][n][n] FP64 value matrix, (o, e) = D(e, o)
  void *dmat;            // [Wc][n][n][4] Tstore: d/dx, d/dy, d/dz, Laplacian (dm_at)
  double *logdet, *sgn;  // [Wc]
  double *gradPR;        // [Wc][n][3]
  double *lapPR;         // [Wc]
  int *singular;         // [Wc]
  // delayed Sherman-Morrison updates: the current inverse is
  //   A^-1 = inv - sum_{a < dcount} ubuf[a] zbuf[a]^T
  int *dcount;           // [Wc]
  double *ubuf, *zbuf;   // [Wc][kDelay][n]
  // several determinants: the per-determinant arrays above (inv, logdet, sgn,
  // gradPR, lapPR, singular, dcount) are indexed by w * ndet + d, while dval /
  // dmat hold all nact active orbitals of walker w ([Wc][nact][n],
  // [Wc][nact][n][4]); determinant d takes its column c from active orbital
  // occ[d * n + c].  ndet <= 1, occ = null: one determinant, nact = n.
  int ndet, nact;
  const int *occ;
};
__device__ __forceinline__ int det_count(const SlaterArgs &A) { return A.ndet > 1 ? A.ndet : 1; }
__device__ __forceinline__ int act_count(const SlaterArgs &A) { return A.nact > 0 ? A.nact : A.n; }
__device__ __forceinline__ int occ_row(const SlaterArgs &A, int d, int c) {
  return A.occ ? A.occ[d * A.n + c] : c;
}
constexpr int kDelay = 8;
constexpr int kDer = 4;  // stored derivative components

template <typename TS>
hipError_t launch_lu_inverse(const SlaterArgs &a, const int *walkers, int num, hipStream_t st);
template <typename TS>
hipError_t launch_traces_all(const SlaterArgs &a, const int *walkers, int num, hipStream_t st);

// Single-electron proposal: everything the evaluate contract needs at the
// proposed position, plus optional in-kernel Metropolis and commit.
struct ProposeArgs {
  int N, moved, n_up;
  const double *R;
  const double *gradPR[2];
  const double *cgrad;   // several determinants: [Wc][N][3] grad ln D of the sum (else null)
  const double *Ux;
  double tau;
  // drift of the moved electron: gradPsiRatio (QMCB_DRIFT_NONE) or its
  // umrigar93_unequalelectrons modification with time step tau
  int drift_mod;
  int n_nuc;
  const double *nuc_xyz, *nuc_z;
  uint64_t seed;
  long step;
  double *propR, *gauss, *unif;
  // batch row of walker w: the launch's own index, or with row_from_walker
  // w - row_base (a launch over a subset of a step's walkers)
  int row_from_walker, row_base;
};
hipError_t launch_propose(const ProposeArgs &a, const int *walkers, int num, hipStream_t st);

// The importance-sampled proposal of electron A.moved of walker w into batch
// row `row` (QMCWalker.cpp:343-398): x' = x + tau F + sqrt(tau) g, with F the
// electron's drift v (gradPsiRatio) or its umrigar93_unequalelectrons
// modification, g ~ N(0,1)^3 by Box-Muller from Philox keyed by (seed, step,
// walker), and the uniform of the Metropolis test from the same draw.
__device__ inline void propose_row(const ProposeArgs &A, long row, int w, const double v_in[3],
                                   const double x[3]) {
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  double u[8];
  for (int c = 0; c < 2; c++) {
    u4 ctr = {(uint32_t)A.step, (uint32_t)((uint64_t)A.step >> 32), (uint32_t)w, (uint32_t)c};
    u4 r = philox4x32(ctr, k0, k1);
    u[4 * c + 0] = u01(r.x, r.y);
    u[4 * c + 1] = u01(r.y, r.z);
    u[4 * c + 2] = u01(r.z, r.w);
    u[4 * c + 3] = u01(r.w, r.x);
  }
  // Box-Muller on (u0,u1), (u2,u3)
  const double two_pi = 6.283185307179586;
  const double r1 = sqrt(-2.0 * log(u[0])), r2 = sqrt(-2.0 * log(u[2]));
  const double g[3] = {r1 * cos(two_pi * u[1]), r1 * sin(two_pi * u[1]), r2 * cos(two_pi * u[3])};
  const double st = sqrt(A.tau);
  double v[3] = {v_in[0], v_in[1], v_in[2]};
  if (A.drift_mod == QMCB_DRIFT_UMRIGAR93_UNEQUAL) {
    const double f = drift_factor_unequal(v, x, A.n_nuc, A.nuc_xyz, A.nuc_z, A.tau);
    for (int k = 0; k < 3; k++) v[k] *= f;
  }
  for (int k = 0; k < 3; k++) {
    const double chi = st * g[k];
    A.gauss[3L * row + k] = chi;
    A.propR[3L * row + k] = x[k] + A.tau * v[k] + chi;
  }
  A.unif[row] = u[4];
}

struct MoveArgs {
  SlaterArgs s;          // moved spin
  SlaterArgs other;      // the other spin (read only)
  int moved;             // moved electron (global index)
  int N;
  const void *vrow;      // new rows [batch][5][npad] (GEMM output type TO)
  long v_s_comp, v_s_walker;
  int n_nuc;
  const double *nuc_xyz, *nuc_z;
  double *R;             // [Wc][N][3] current positions
  const double *propR;   // [Wc][3] proposed position of `moved`
  double *U, *Ux, *Uxx, *Vee, *Ven;           // current Jastrow/potential
  const double *pU, *pUx, *pUxx, *pVee, *pVen;  // at the proposal
  double v_nn;
  // jastrow_fused: sm_move forms the Jastrow factor and potential at the
  // proposal itself (the moved electron's pairs; needs N <= the tile's
  // threads, sm_move_fuses_jastrow) instead of reading pU ... pVen
  int jastrow_fused;
  const qmcb_pade *ee;   // [3] up-dn, up-up, dn-dn
  const qmcb_pade *en;   // [n_nuc] of the moved spin
  int n_up;
  int drift_mod;
  double dt, umrigar_a;
  // outputs at the proposal (device, any may be null)
  qmcb_results out;
  // staging of the proposal (staged mode)
  double *st_inv, *st_grad, *st_lap, *st_logdet, *st_sgn;
  int *st_singular;
  // fused Metropolis (vmc mode)
  int mode;              // 0 = stage, 1 = vmc accept in kernel
  double tau;
  const double *gauss;   // [Wc][3] the Gaussian displacement used (vmc)
  const double *unif;    // [Wc] uniform for the accept test (vmc)
  double *acc_el, *acc_el2, *acc_n, *acc_prop_el, *cur_el;  // [Wc]
  // diffusion Monte Carlo (mode 2): fixed-node rejection, weights
  double *weight;        // [Wc]
  double *acc_w;         // [Wc] sum of weights over the steps
  double e_trial, tau_eff;
  int reweight;          // reweight this step (electron 0 moved)
  int energy_mod;        // 0: energy_modification_type none, 1: umrigar93
  double e_est;          // energy_estimated of the umrigar93 modification
  WalkCtl walk;          // walker age, old-walker acceptance, deletion rules
  // observable sample of this step (null except once per sweep):
  // obs[3 b + {0,1,2}] = {p E_L(R') + (1 - p) E_L(R), weight, p}
  double *obs;
  // VMC: the next step's proposal (electron next.moved, step next.step) of
  // each walker from the state this move leaves, into the same batch row
  int propose_next;
  ProposeArgs next;
};

// One row {sum w E, sum w, sum w p, walkers} of the per-sweep observable
// series from obs[3 * num] (single workgroup, fixed summation order).
hipError_t launch_series_reduce(const double *obs, int num, double *row, hipStream_t st);

// several determinants (multidet.hip): per-spin Slater records with ndet /
// nact / occ set, staging per (batch row, determinant) at b * D + d
struct MdetArgs {
  SlaterArgs s[2];
  int D;                 // determinants
  int N;
  int moved;             // moved electron (global index); -1: all-electron assembly
  const double *ci;      // [D]
  const void *vrow;      // new orbital rows of the moved electron (GEMM output type)
  long v_s_comp, v_s_walker;
  double *st_inv, *st_grad, *st_lap, *st_logdet, *st_sgn;  // [num * D] records
  int *st_singular;      // [num * D]
  int *st_row_singular;  // [num]: the combined proposal failed
  double *R;
  const double *propR;
  double *U, *Ux, *Uxx, *Vee, *Ven;
  const double *pU, *pUx, *pUxx, *pVee, *pVen;
  double v_nn;
  int n_nuc;
  const double *nuc_xyz, *nuc_z;
  int drift_mod;
  double dt, umrigar_a;
  qmcb_results out;      // device outputs per batch row
  int mode;              // 0: evaluate / stage, 1: VMC, 2: DMC Metropolis on the device
  double tau;
  const double *gauss, *unif;
  uint8_t *mask;         // [num] accept decisions (VMC)
  double *acc_el, *acc_el2, *acc_n, *acc_prop_el, *cur_el;
  double *obs;           // per-sweep observable rows (VMC, electron 0 moves)
  double *cgrad;         // [Wc][N][3] grad ln D of the determinant sum (resident)
  double *st_cgrad;      // [num][N][3] the same at the proposal
  // diffusion Monte Carlo (mode 2): fixed-node rejection, weights
  double *weight, *acc_w;
  double e_trial, tau_eff;
  int reweight;
  int energy_mod;        // 0: energy_modification_type none, 1: umrigar93
  double e_est;
  WalkCtl walk;          // walker age, old-walker acceptance, deletion rules
};
template <typename TS, typename TO>
hipError_t launch_mdet_move(const MdetArgs &a, const int *walkers, int num, hipStream_t st);
hipError_t launch_mdet_assemble(const MdetArgs &a, const int *walkers, int num, hipStream_t st);
template <typename TS, typename TO>
hipError_t launch_mdet_commit(const MdetArgs &a, const int *walkers, const uint8_t *mask, int num,
                              hipStream_t st);

// copy whole walker states slot -> slot (DMC branching)
constexpr int kMaxCopyArrays = 40;
struct CopyArgs {
  int narr;
  char *base[kMaxCopyArrays];
  long bytes[kMaxCopyArrays];  // bytes per walker (multiple of 4)
};
hipError_t launch_copy_walkers(const CopyArgs &a, const int *pairs /*[2*np]*/, int np,
                               hipStream_t st);

template <typename TS, typename TO>
hipError_t launch_sm_move(const MoveArgs &a, const int *walkers, int num, hipStream_t st);
// whether sm_move can form the Jastrow factor at the proposal itself for N
// electrons when n electrons of the moved spin set its tile (one electron
// per thread)
bool sm_move_fuses_jastrow(int n, int N);

// ---- kernel (d): Jastrow + Coulomb, all-electron and single-electron ----
struct JastrowArgs {
  DevSystem sys;
  const double *R;       // [Wc][N][3]
  double *U, *Ux, *Uxx, *Vee, *Ven;   // outputs
  // single-electron
  int moved;
  const double *propR;   // [Wc][3]
};
hipError_t launch_jastrow_all(const JastrowArgs &a, const int *walkers, int num, hipStream_t st);
hipError_t launch_jastrow_move(const JastrowArgs &a, const double *U, const double *Ux,
                               const double *Uxx, const double *Vee, const double *Ven,
                               const int *walkers, int num, hipStream_t st);

// all-electron assembly: psi, gradient, Laplacian, E_L, drift
struct AssembleArgs {
  DevSystem sys;
  SlaterArgs s[2];
  const double *R, *U, *Ux, *Uxx, *Vee, *Ven;
  int drift_mod;
  double dt, umrigar_a;
  qmcb_results out;      // device arrays indexed by batch row
  double *cur_el;        // [Wc] current E_L (resident)
  int *age;              // [Wc] walker age, reset for a fresh walker (may be null)
};
hipError_t launch_assemble(const AssembleArgs &a, const int *walkers, int num, hipStream_t st);
// out.mod_ratio and out.r2 of rows [0, num) from out.grad_lnpsi / out.mod_drift
// and the configuration: resident R of walkers[b] with electron whichE (>= 0)
// at propR[b]
hipError_t launch_results_extra(const qmcb_results &out, const double *R, const int *walkers,
                                const double *propR, int whichE, int N, int num, hipStream_t st);

// proposal of a single-electron importance-sampled move (QMCWalker.cpp:343-398)


// all-electron moves (walk_ae.hip; QMCWalker.cpp:343-398, 756-780,
// 1578-1685): the walker record at the current positions, the proposal
// written over R (R saved), the evaluate's outputs at the proposal by batch
// row, Metropolis and restore
constexpr uint32_t kAeTag = 0x40000000u;  // Philox counter word of these draws
struct AeArgs {
  int N;
  double *R, *Rsave;     // [Wc][N][3]
  double *drift;         // [Wc][N][3] modified drift at R
  double *logpsi, *sgn, *mr, *fwd;  // [Wc]
  double *cur_el;        // [Wc]
  qmcb_results prop;     // evaluate outputs at the proposal, by batch row
  double tau;
  uint64_t seed;
  long step;
  double *unif;          // [Wc] by batch row
  int mode;              // 0: take `prop` as the record; 1: VMC; 2: DMC
  double *acc_el, *acc_el2, *acc_n, *acc_prop_el;
  double *weight, *acc_w;
  double e_trial, tau_eff, e_est;
  int reweight, energy_mod;
  WalkCtl walk;
  double *obs;           // [num][3] per-step observable sample (may be null)
};
hipError_t launch_ae_propose(const AeArgs &a, const int *walkers, int num, hipStream_t st);
hipError_t launch_ae_accept(const AeArgs &a, const int *walkers, int num, hipStream_t st);

// commit staged single-electron proposals
template <typename TS, typename TO>
hipError_t launch_commit(const MoveArgs &a, const int *walkers, const uint8_t *mask, int num,
                         hipStream_t st);

}  // namespace qmcb
