/*
 * qmcbeaver_amd.h -- C ABI of the MI355X walker-evaluation library.
 *
 * One context holds one molecule (basis, MO coefficients, Jastrow) plus the
 * device-resident state of up to `max_walkers` walkers: positions, the FP64
 * inverse Slater matrices, log-determinants, derivative matrices and Jastrow
 * sums.  All entry points take plain pointers and sizes.
 *
 * Reference interface replaced (QMcBeaver, paths relative to QMcBeaver/):
 *   class QMCFunctions            include/QMCFunctions.h:26-105
 *   QMCFunctionsFactory           src/QMCFunctionsFactory.cpp:15-37
 *   QMCSCFJastrow::initialize     src/QMCSCFJastrow.cpp:61-89
 *   QMCSCFJastrow::evaluate       src/QMCSCFJastrow.cpp:91-192
 *   QMCWalkerData outputs         include/QMCWalkerData.h:64-226
 */
#ifndef QMCBEAVER_AMD_H
#define QMCBEAVER_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QMCB_ABI_VERSION 3
#define QMCB_PADE_MAX 8

typedef enum qmcb_status {
  QMCB_OK = 0,
  QMCB_INVALID_ARG = 1,
  QMCB_OOM = 2,
  QMCB_DEVICE_ERROR = 3,
  QMCB_BAD_STATE = 4
} qmcb_status;

/* Arithmetic of the basis values and the orbital contraction.  The inverse
 * Slater matrices, log-determinants, traces, Jastrow and energies are FP64
 * in every mode. */
typedef enum qmcb_precision {
  QMCB_FP64 = 0,        /* everything FP64 (reference arithmetic)                 */
  QMCB_MIXED_FP32 = 1,  /* orbital GEMM on FP32 MFMA with FP32 operands and
                           accumulate (basis computed in FP64, rounded); FP32
                           derivative storage; FP64 re-inversion                  */
  QMCB_MIXED_BF16 = 2,  /* as MIXED_FP32 with BF16 GEMM operands (basis values
                           and coefficients rounded from FP64) on
                           v_mfma_f32_16x16x32_bf16, FP32 accumulate            */
  QMCB_MIXED_F32STORE = 3, /* FP64 basis and GEMM, derivative matrices stored
                           in FP32 (the HBM-bound part halves)                    */
  QMCB_MIXED_BF16STORE = 4 /* as MIXED_FP32 (FP32 MFMA GEMM), derivative matrices
                           stored in BF16 (8-bit mantissa): a single-electron
                           move streams 16 n^2 bytes instead of 24 n^2          */
} qmcb_precision;

/* Drift (quantum force) modification, QMCSCFJastrow.cpp:329-472. */
typedef enum qmcb_drift_mod {
  QMCB_DRIFT_NONE = 0,
  QMCB_DRIFT_UMRIGAR93_UNEQUAL = 1,
  QMCB_DRIFT_UMRIGAR93_EQUAL = 2
} qmcb_drift_mod;

/* Pade correlation function u(r) = (num[0] r + num[1] r^2 + ...) /
 * (1 + den[0] r + den[1] r^2 + ...).  Covers the reference's "Pade" and
 * "FixedCuspPade" types (PadeCorrelationFunction.cpp:15-83,
 * FixedCuspPadeCorrelationFunction.cpp:15-117): both are this rational form
 * once constants and parameters are placed. */
typedef struct qmcb_pade {
  int32_t n_num;
  int32_t n_den;
  double num[QMCB_PADE_MAX];
  double den[QMCB_PADE_MAX];
} qmcb_pade;

/* Flat, self-contained system description (copied at create time; no host
 * pointer is retained).  Replaces the parts of the global QMCInput the
 * evaluate path reads: QMCMolecule, QMCBasisFunction(Coefficients),
 * QMCWavefunction::getCoeff, QMCJastrowParameters. */
typedef struct qmcb_system {
  int32_t n_up, n_dn;       /* electrons per spin, n_dn <= n_up; alpha first */
  int32_t n_nuc;            /* nuclei                                          */
  int32_t n_basis;          /* basis functions B                               */
  int32_t n_prim;           /* total primitives                                */
  const double *nuc_xyz;    /* [n_nuc][3] bohr                                 */
  const double *nuc_z;      /* [n_nuc] (effective) charges                     */
  const int32_t *bf_atom;   /* [n_basis] centre of each basis function         */
  const int32_t *bf_pow;    /* [n_basis][3] Cartesian powers a,b,c (<= 2)      */
  const int32_t *bf_prim_off; /* [n_basis+1] primitive ranges                  */
  const double *prim_alpha; /* [n_prim] exponents                              */
  const double *prim_coef;  /* [n_prim] raw contraction coefficients           */
  const double *mo_up;      /* [n_up][n_basis] occupied orbitals, alpha        */
  const double *mo_dn;      /* [n_dn][n_basis] occupied orbitals, beta         */
  qmcb_pade ee_updn, ee_upup, ee_dndn;
  const qmcb_pade *en_up;   /* [n_nuc] e-n function seen by alpha electrons    */
  const qmcb_pade *en_dn;   /* [n_nuc] e-n function seen by beta electrons     */
  /* Several determinants, Psi = sum_d ci_coef[d] det_up(d) det_dn(d) e^U
   * (QMCSCFJastrow.cpp:281-327).  n_det <= 1 with NULL tables: one
   * determinant of the n_up / n_dn rows of mo_up / mo_dn.  Otherwise mo_up has
   * n_act_up rows (the active orbitals of the alpha spin, in orbital order) and
   * determinant d uses rows occ_up[d][0..n_up) as its columns
   * (QMCWavefunction.cpp:873-905); likewise for beta. */
  int32_t n_det;
  int32_t n_act_up, n_act_dn;
  const double *ci_coef;    /* [n_det]                                         */
  const int32_t *occ_up;    /* [n_det][n_up]                                   */
  const int32_t *occ_dn;    /* [n_det][n_dn]                                   */
} qmcb_system;

typedef struct qmcb_options {
  int32_t precision;        /* qmcb_precision                                  */
  int32_t max_walkers;      /* resident walker slots                           */
  int32_t drift_mod;        /* qmcb_drift_mod                                  */
  int32_t refresh_interval; /* device VMC/DMC: each walker's Slater matrices of
                               a spin are refreshed to FP64 accuracy once every
                               K single-electron moves of that spin, staggered
                               by slot: the residues s mod K form phases of S
                               consecutive residues, S = min(K, ceil(K * 1024
                               / walkers)), and phase q is refreshed after the
                               move c of the spin with c mod K = min(qS + S - 1,
                               K - 1), so every S steps refresh about S/K of
                               the walkers of the moved spin in one launch of
                               at least 1024 walkers; 0 = never (the reference
                               never re-inverts in single-electron mode)      */
  double dt;                /* time step used by the drift modification        */
  double umrigar_a;         /* umrigar93_equalelectrons_parameter              */
  uint64_t seed;            /* device RNG seed (vmc entry points)              */
  int32_t device;           /* HIP device ordinal                              */
  int32_t replace_cusps;    /* 1: electron-nucleus cusp replacement of the
                               orbitals (replace_electron_nucleus_cusps 1)     */
  int64_t basis_scratch_bytes; /* cap on the basis-value scratch; 0 = 2 GiB.
                               All-electron evaluates run in chunks that fit;
                               at least one point per walker is kept          */
  int32_t walker_groups;    /* qmcb_vmc_steps and qmcb_dmc_steps with one
                               determinant: split the walkers into this many
                               groups (at most 8, each of at least 256
                               walkers), each enqueued on a stream of its own
                               so that their kernels overlap; results are
                               bit-identical for every value.  0 or 1 = one
                               group; 3 is fastest at 8192 walkers of 256
                               electrons (12 % over one group).  -1:
                               qmcb_vmc_steps times windows of 16 steps (S
                               when the refresh phase S above is 17-32) of
                               each of 1, 2 and 3 groups (an untimed round,
                               then two timed rounds, over as many calls as
                               needed, one group until then) and keeps the
                               fastest, again when the
                               batch changes by more than 2x; qmcb_dmc_steps
                               uses that choice                              */
  int32_t old_walker_acceptance; /* old_walker_acceptance_parameter of the
                               reference (QMCFlags.cpp:59, default 50): after
                               more than this many consecutive rejected moves a
                               walker's acceptance probability is multiplied by
                               1.1^(age - K) (QMCWalker.cpp:1578-1600) and, in
                               DMC, its weight decays by 0.9 per sweep instead
                               of being reweighted (:1181-1184).  <= 0: off    */
  int32_t all_electron_moves; /* 1: qmcb_vmc_steps moves every electron of a
                               walker at each step (the reference's
                               one_e_per_iter 0, its default; QMCWalker.cpp:
                               179-198, 343-398, 756-780): the whole evaluate
                               runs at the proposal and a rejected walker gets
                               its saved record back (:259-262); one observable
                               row per step.  0: one electron per step, moved
                               in order (one_e_per_iter 1)                   */
  int32_t graphs;           /* 1: with all_electron_moves, the evaluate at the
                               proposal (basis, orbital GEMM, inverse, traces,
                               Jastrow, assembly: no step-dependent argument)
                               is captured in a HIP graph once two steps in a
                               row share walker count, walker map and time
                               step, and replayed while they do; 0: launched
                               kernel by kernel.  Same results bit for bit;
                               kernel timing launches kernel by kernel       */
} qmcb_options;

/* Caller-owned outputs, one row per evaluated walker.  Any pointer may be
 * NULL (not written).  N = n_up + n_dn. */
typedef struct qmcb_results {
  double *log_psi;          /* [num] ln|Psi| = ln|D_up D_dn| + U                */
  int8_t *sign_psi;         /* [num] sign of Psi                               */
  double *grad_lnpsi;       /* [num][N][3] gradPsiRatio = D_x + U_x            */
  double *mod_drift;        /* [num][N][3] modifiedGradPsiRatio                */
  double *lap_ratio;        /* [num] Laplacian(Psi)/Psi                        */
  double *e_local;          /* [num] local energy                              */
  double *kinetic;          /* [num] -1/2 Laplacian(Psi)/Psi                    */
  double *potential;        /* [num] V_nn + V_en + V_ee                        */
  double *v_ne;             /* [num]                                           */
  double *v_ee;             /* [num]                                           */
  double *det_ratio;        /* [num] new/old determinant of the moved spin
                               (single-electron evaluate); 1 for all-electron  */
  double *jastrow_u;        /* [num] U                                         */
  int8_t *singular;         /* [num] numerical failure flag                    */
  double *mod_ratio;        /* [num] modificationRatio: |modified drift| /
                               |gradPsiRatio| over all electrons (1 without a
                               drift modification; QMCSCFJastrow.cpp:450-458).
                               The device-buffer entry points need grad_lnpsi
                               and mod_drift with it                          */
  double *r2;               /* [num][3] x2, y2, z2: sums over the electrons of
                               the squared coordinates of the evaluated
                               configuration (QMCSCFJastrow.cpp:164-169)      */
} qmcb_results;

typedef struct qmcb_ctx qmcb_ctx;

/* Library/ABI version; no device work. */
int32_t qmcb_abi_version(void);

/* Build a context: copies `sys` to the device and allocates the resident
 * walker state.  Replaces QMCFunctionsFactory::functionsFactory +
 * QMCSCFJastrow::initialize (QMCSCFJastrow.cpp:61-89).  The environment
 * variable QMCB_GEMM_TILE="wpb,mt" (wpb 2 or 4 waves per workgroup, mt 1, 2,
 * 4 or 8 orbital tiles of 16 per wave), read here, fixes the single-electron
 * orbital GEMM's tile instead of the timing autotuner; every tile gives the
 * same bits.  QMCB_SEPARATE_PROPOSE=1 makes the device VMC launch its
 * proposal kernel every step instead of drawing the next proposal inside the
 * move kernel (same bits). */
qmcb_status qmcb_create(const qmcb_system *sys, const qmcb_options *opt,
                        qmcb_ctx **out);

void qmcb_destroy(qmcb_ctx *ctx);

/* Last error message of this context (or of the last failed create when
 * ctx is NULL). */
const char *qmcb_last_error(const qmcb_ctx *ctx);

/* Evaluate a batch of walkers: QMCFunctions::evaluate(walkerData, xData, num)
 * (QMCFunctions.h:67-70, QMCSCFJastrow.cpp:91-181).
 *   walker_ids [num]   resident slots
 *   R          [num][N][3] host positions, alpha electrons first
 *   whichE     -1: all-electron evaluate; (re)initialises the resident state
 *              of each slot from R.
 *              k>=0: single-electron proposal; only row k of R is read, the
 *              other electrons are the slot's resident positions.  The
 *              proposal is staged until qmcb_accept.
 * Numerical failure sets results->singular, never an error status. */
qmcb_status qmcb_evaluate(qmcb_ctx *ctx, int32_t num, const int32_t *walker_ids,
                          const double *R, int32_t whichE, qmcb_results *out);

/* Same as qmcb_evaluate with device-resident R and device-resident result
 * arrays (every non-NULL pointer in `out` is a device pointer).  Asynchronous
 * on the context's stream. */
qmcb_status qmcb_evaluate_device(qmcb_ctx *ctx, int32_t num,
                                 const int32_t *walker_ids_dev,
                                 const double *R_dev, int32_t whichE,
                                 qmcb_results *out_dev);

/* qmcb_evaluate_device with whichE = -1 (all-electron) that allocates
 * nothing: no hipMalloc, hipFree or allocating library call.  No counterpart
 * in the reference; it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The basis values of each chunk of points go to a static
 * device array of 64 MiB per GPU (held until the process ends) instead of
 * the context's scratch; results equal qmcb_evaluate_device's.  Limit: one
 * walker's electrons of a spin must fit, 5 * ceil32(n_basis) * (8 FP64 and
 * fp32store, 4 FP32 and bf16store, 2 BF16) bytes each, else QMCB_INVALID_ARG.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams or threads, any context) race and must be serialised by the
 * caller.  Kernel timing must be off (QMCB_BAD_STATE otherwise).
 * Asynchronous on the context's stream. */
qmcb_status qmcb_evaluate_device_noalloc(qmcb_ctx *ctx, int32_t num,
                                         const int32_t *walker_ids_dev,
                                         const double *R_dev, qmcb_results *out_dev);

/* Commit (mask[i]=1) or discard (mask[i]=0) the staged single-electron
 * proposal of walker_ids[i].  Replaces the reference's re-evaluation at the
 * old position after a rejected move (QMCWalker.cpp:254-266). */
qmcb_status qmcb_accept(qmcb_ctx *ctx, int32_t num, const int32_t *walker_ids,
                        const uint8_t *mask);

/* FP64 re-inversion of the resident Slater matrices of the given slots
 * (no reference counterpart in single-electron mode, which never re-inverts;
 * the all-electron path is QMCSlater.cpp:675-683). */
qmcb_status qmcb_refresh(qmcb_ctx *ctx, int32_t num, const int32_t *walker_ids);

/* Copy resident positions of slots [0, num) to host R[num][N][3]. */
qmcb_status qmcb_get_positions(qmcb_ctx *ctx, int32_t num, double *R);

/* ---- device-resident VMC (the caller loop of the path) ------------------
 * QMCRun::propagateWalkers + QMCWalker::initializePropagation /
 * processPropagation with importance sampling (QMCRun.cpp:28-87,
 * QMCWalker.cpp:179-282, 343-398, 756-780, 1578-1685): one step moves
 * electron (step mod N) of every walker in slots [0, num), evaluates the
 * full contract at the proposal, and accepts or rejects on the device with a
 * counter-based RNG keyed by (seed, slot, step).  The drift is gradPsiRatio
 * (qmcb_options.drift_mod QMCB_DRIFT_NONE) or its
 * umrigar93_unequalelectrons modification (QMCB_DRIFT_UMRIGAR93_UNEQUAL, the
 * reference's default QF_modification_type; QMCSCFJastrow.cpp:368-443) with
 * the step's tau as its time step, in the proposal and the Green's function
 * ratio alike.  Same for qmcb_dmc_steps. */
typedef struct qmcb_vmc_stats {
  double sum_e_local;       /* sum over walkers and steps of E_L (current cfg) */
  double sum_e_local_sq;
  double sum_accept;        /* accepted moves                                  */
  double sum_prop_e_local;  /* sum of E_L at the proposals                     */
  int64_t n_samples;
} qmcb_vmc_stats;

qmcb_status qmcb_vmc_steps(qmcb_ctx *ctx, int32_t num, int64_t first_step,
                           int32_t n_steps, double tau, qmcb_vmc_stats *stats);

/* ---- device-resident DMC ------------------------------------------------
 * Fixed-node importance-sampled DMC with single-electron moves: Metropolis
 * with node-crossing rejection (QMCWalker.cpp:1578-1654), the
 * umrigar93_probability_weighted reweighting once per electron sweep
 * (QMCWalker.cpp:1060-1185), non-unit-weight branching after each sweep
 * (QMCRun.cpp:740-814, 1179-1206) and the trial-energy feedback
 * E_T = E_est - pop_control * ln(W_total / target)
 * (QMCManager.cpp:2263-2310).  Walker states (with several determinants,
 * every determinant's record) are copied on the device when a walker splits;
 * the population lives in the context's slots. */
typedef struct qmcb_dmc_params {
  double tau;               /* time step                                       */
  double e_trial;           /* initial trial energy (used when first_step == 0
                               or no estimate exists yet)                      */
  double pop_control;       /* population_control_parameter (reference 1.0)   */
  double branch_threshold;  /* split walkers above this weight (2.0)          */
  double fusion_threshold;  /* fuse walkers below this weight (0.45)          */
  int32_t target_walkers;   /* population the E_T feedback aims at            */
  int32_t reweight_after;   /* weights stay 1 before this step (100)          */
  int32_t energy_modification; /* 0: energy_modification_type none,
                               1: umrigar93 -- S = (E_T - E_est) + (E_est - E_L)
                               |modified drift| / |drift| (QMCWalker.cpp:1121-
                               1132); needs qmcb_options.drift_mod
                               QMCB_DRIFT_UMRIGAR93_UNEQUAL to differ from 0   */
  double rel_cutoff;        /* the reference's rel_cutoff (QMCFlags.cpp:61,
                               default 100): at a reweighting, a walker with
                               |E_L - E_0| / |E_0| > rel_cutoff is deleted
                               (weight 0, QMCWalker.cpp:1270-1290, 1386-1393);
                               E_0 = the e_trial given at the first call after
                               qmcb_dmc_init.  <= 0: no such deletion          */
  int32_t equilibration_steps; /* steps [0, equilibration_steps) are the
                               reference's equilibration (iteration < 0):
                               E_T = E_est - ln(W_total / target) / tau
                               (QMCManager.cpp:2277-2291), E_est stays E_0, and
                               while more than 100 of them remain the stricter
                               deletion rules hold (heavy, light, dW outside
                               [0.5, 1.5], age >= 30; QMCWalker.cpp:1207-1300).
                               Production (the rest): a walker is deleted when
                               dW^(age+1) > 15, its weight exceeds 50 or its
                               energy is bad (:1312-1393), and E_est is the
                               weighted mean of the production sweeps' samples
                               once there are at least 100 of them
                               (QMCManager.cpp:2247-2261), E_0 before          */
} qmcb_dmc_params;

typedef struct qmcb_dmc_stats {
  double sum_w_e_local;     /* sum over steps and walkers of w E_L             */
  double sum_w_e_local_sq;
  double sum_weights;
  double sum_accept;
  int64_t n_moves;
  int32_t population;       /* walkers alive after the last step               */
  int32_t n_split, n_killed, n_fused;
  double total_weight;
  double e_trial;           /* trial energy after the last update              */
  int32_t n_split_skipped;  /* walkers above branch_threshold left unsplit
                               because every slot was in use (they keep their
                               weight; the reference always splits,
                               QMCRun.cpp:748-757): raise max_walkers          */
  int32_t n_deleted;        /* walkers removed by the deletion rules (weight
                               set to 0 at a reweighting; part of n_killed)    */
  double e_estimated;       /* E_est after the last update                     */
} qmcb_dmc_stats;

/* Slots [0, num) (all-electron evaluated) become the DMC population, weight 1. */
qmcb_status qmcb_dmc_init(qmcb_ctx *ctx, int32_t num);
qmcb_status qmcb_dmc_steps(qmcb_ctx *ctx, int64_t first_step, int32_t n_steps,
                           const qmcb_dmc_params *params, qmcb_dmc_stats *stats);
/* Current population: up to cap slot ids and weights; *count = population. */
qmcb_status qmcb_dmc_population(qmcb_ctx *ctx, int32_t cap, int32_t *slots,
                                double *weights, int32_t *count);

qmcb_status qmcb_sync(qmcb_ctx *ctx);

/* Change qmcb_options.walker_groups of a live context (-1..8); takes effect
 * at the next qmcb_vmc_steps.  No reference counterpart (a scheduling knob). */
qmcb_status qmcb_set_walker_groups(qmcb_ctx *ctx, int32_t groups);

/* The walker-group count the last qmcb_vmc_steps / qmcb_dmc_steps call ran
 * its steps with (after the clamps to at least 256 walkers per group and one
 * group with several determinants; while a walker_groups = -1 choice is being
 * timed, the count of its last step); 0 before any such call.  No reference
 * counterpart (a scheduling knob). */
qmcb_status qmcb_get_walker_groups(const qmcb_ctx *ctx, int32_t *groups);

/* ---- device memory ------------------------------------------------------
 * qmcb_create makes one device allocation (an arena) holding every array of
 * the context, including the staging arrays the first single-electron
 * qmcb_evaluate (or, with several determinants, the first move) uses.
 * qmcb_plan_device_bytes: host-only projection of that arena for this system
 * and these options (cusp tables at their upper bound), for admission control
 * of many contexts on one device.  qmcb_device_bytes: what the context holds,
 * the arena plus the per-sweep series buffer (32 bytes per recorded sweep)
 * and, once walker groups are used, their slot map (4 bytes per walker
 * slot).
 * No reference counterpart: the reference sizes host memory per walker
 * (QMCWalker.cpp) and runs one molecule per process. */
qmcb_status qmcb_device_bytes(const qmcb_ctx *ctx, int64_t *bytes);
qmcb_status qmcb_plan_device_bytes(const qmcb_system *sys, const qmcb_options *opt,
                                   int64_t *bytes);

/* ---- electron-nucleus cusp replacement ---------------------------------
 * With qmcb_options.replace_cusps = 1 the s-type part of every orbital near
 * every nucleus is replaced inside a radius rc by C + sgn0 exp(alpha(r)), the
 * reference's replace_electron_nucleus_cusps 1 (QMCElectronNucleusCusp.cpp:
 * 46-177 fit, :587-640 replaceCusps; QMCSlater.cpp:88-93, 864-871).
 * qmcb_cusp_fit returns the fitted parameters of one spin's orbitals (host
 * only, no device needed): out[n_nuc][n_orbitals of the spin]; rc = 0 where
 * an orbital has no s-type part on a nucleus. */
typedef struct qmcb_cusp_orbital {
  double rc;        /* radius of correction (0: no replacement)               */
  double C;         /* constant shift                                          */
  double sgn0;      /* sign of the exponential                                 */
  double alpha[5];  /* alpha(r) = sum alpha[i] r^i                             */
  double phi0;      /* replacement value at the nucleus                        */
  double n0;        /* the orbital at the nucleus from the other atoms         */
  double zeff;      /* effective charge of the replacement's local energy      */
  double sigma_sq;  /* largest squared deviation from the ideal curve          */
  double ideal0;    /* constant term of the ideal curve                        */
} qmcb_cusp_orbital;

qmcb_status qmcb_cusp_fit(const qmcb_system *sys, int32_t spin, qmcb_cusp_orbital *out);

/* ---- per-sweep observable series ----------------------------------------
 * The reference collects statistics once per electron sweep, at the move of
 * electron 0 (QMCWalker.cpp:2302-2320), from the estimator
 * p E_L(R') + (1 - p) E_L(R) of every walker with its weight
 * (QMCWalker.cpp:2096-2123, 2340-2346), and feeds the walker average into its
 * accumulated properties with the total weight (QMCRun.cpp:544-576).  The
 * device steps record one row per such sweep:
 *   {sum_w w E, sum_w w, sum_w w p, walkers}
 * so that sample = row[0] / row[1] and weight = row[1].  Rows of the last
 * qmcb_vmc_steps / qmcb_dmc_steps call; *count = rows recorded (may exceed
 * cap, then only cap rows are copied). */
qmcb_status qmcb_sample_series(qmcb_ctx *ctx, int32_t cap, double *rows, int32_t *count);

/* ---- DDDA statistics ----------------------------------------------------
 * Weighted binary-blocking estimator of the reference (QMCProperty.cpp,
 * QMCStatistic.cpp): replaces QMCProperty::newSample (:427-478), operator+
 * (:511-598), reWeight (:492-503), getAverage / getStandardDeviation /
 * getDecorrDepth / getCorrelationLength (:127-237) and getBlockVariance
 * (:830-872).  Host-only; needs no device. */
#define QMCB_DDDA_LEVELS 30
typedef struct qmcb_property qmcb_property;

typedef struct qmcb_property_summary {
  int64_t n_samples;
  double average;
  double std_dev;            /* block std at the decorrelation depth; 1e300
                                below 100 samples, 99 without a plateau      */
  double std_dev_std;        /* its standard error (0 when depth < 0)        */
  double correlation_length; /* block variance ratio at the depth (or depth) */
  double serial_variance;    /* variance of the raw samples                  */
  double skewness, kurtosis; /* at the depth                                 */
  int32_t decorr_depth;      /* -2: < 100 samples, -1: no plateau yet        */
  int32_t n_levels;          /* levels holding samples                       */
} qmcb_property_summary;

qmcb_property *qmcb_property_create(void);
void qmcb_property_destroy(qmcb_property *p);
void qmcb_property_zero(qmcb_property *p);
void qmcb_property_add(qmcb_property *p, double sample, double weight);
/* weights may be NULL (unit weights) */
qmcb_status qmcb_property_add_series(qmcb_property *p, int64_t n, const double *samples,
                                     const double *weights);
/* dst = dst + src, the reference's operator+ (merge of two accumulations) */
qmcb_status qmcb_property_merge(qmcb_property *dst, const qmcb_property *src);
void qmcb_property_reweight(qmcb_property *p, double w);
qmcb_status qmcb_property_summary_get(const qmcb_property *p, qmcb_property_summary *out);
/* Blocking level i (block size 2^i): samples, summed weights, average and
 * block variance (variance / (n - 1)). */
qmcb_status qmcb_property_level(const qmcb_property *p, int32_t level, int64_t *n_samples,
                                double *sum_weights, double *average, double *block_variance);

/* Per-kernel timing of the last qmcb_vmc_steps call, measured with HIP
 * events on the context's stream.  names/ms arrays of length cap; returns
 * the number of entries. */
int32_t qmcb_kernel_times(qmcb_ctx *ctx, int32_t cap, const char **names,
                          double *ms, int64_t *launches);
void qmcb_enable_kernel_timing(qmcb_ctx *ctx, int32_t on);

#ifdef __cplusplus
}
#endif
#endif /* QMCBEAVER_AMD_H */
