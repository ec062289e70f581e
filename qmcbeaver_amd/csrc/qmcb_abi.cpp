// qmcb_abi.cpp -- the C ABI (include/qmcbeaver_amd.h) over the HIP kernels.
//
// A context owns one molecule and the device-resident state of up to
// max_walkers walkers.  Every entry point runs on the context's stream; the
// host-buffer entry points copy in, launch, copy out and synchronise.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "qmcb_cusp.hpp"
#include "qmcb_kernels.hpp"

using namespace qmcb;


namespace {

thread_local std::string g_create_error;

struct Timing {
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::string> names;
  std::map<std::string, int> index;
  std::vector<double> ms;
  std::vector<long long> launches;
};

}  // namespace

// Scratch of qmcb_evaluate_device_noalloc: the basis values of a chunk of
// points, in a file-scope device array of kScratchBytes, one copy on every
// GPU the library runs on, held until the process ends.  In host code the
// array's name is only a stand-in: its device address (different on each
// device) comes from hipGetSymbolAddress, looked up once per device at
// qmcb_create with that device current and kept in g_scratch_dev; kernels
// receive that address, never the name.  Every call on a GPU shares that
// GPU's array, so calls that overlap on one GPU (two streams, two threads)
// race; the entry point zeroes what must start at zero (the padded basis
// rows) on every call.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices] = {};

struct qmcb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // walker groups of the device VMC: one stream per group (default priority)
  // and the identity slot map the groups index into
  static constexpr int kMaxGroups = 8;
  hipStream_t gstream[kMaxGroups] = {};
  hipEvent_t gevent[kMaxGroups] = {};
  hipEvent_t gfront[kMaxGroups] = {};  // a group's orbital rows of the current step are done
  hipEvent_t main_event = nullptr;
  int *d_iota = nullptr;
  qmcb_options opt{};
  int N = 0, n[2] = {0, 0}, start[2] = {0, 0}, nn = 0, B = 0, kpad = 0, nshell = 0, nprim = 0;
  int npad[2] = {16, 16};
  int Wc = 0;
  DevSystem dsys{};
  std::vector<void *> allocs;
  size_t dev_bytes = 0;  // device memory this context holds (arena + own allocations)
  char *arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  // single-electron orbital GEMM: (waves per workgroup, orbital tiles per
  // wave) chosen by tune_gemm, per (spin, batch rows / 16); gemm_force: the
  // pair set by QMCB_GEMM_TILE="wpb,mt" at creation ({0, 0}: tuned)
  std::map<std::pair<int, int>, std::pair<int, int>> gemm_tile;
  std::pair<int, int> gemm_force{0, 0};
  bool separate_propose = false;  // QMCB_SEPARATE_PROPOSE set at creation
  size_t ts = 8, tx = 8, to = 8;  // bytes: derivative storage, GEMM operand, GEMM output

  // resident walker state
  double *R = nullptr;
  double *inv[2] = {nullptr, nullptr};
  double *dval[2] = {nullptr, nullptr};
  void *dmat[2] = {nullptr, nullptr};
  double *logdet[2] = {nullptr, nullptr}, *sgn[2] = {nullptr, nullptr};
  double *gradPR[2] = {nullptr, nullptr}, *lapPR[2] = {nullptr, nullptr};
  int *singular[2] = {nullptr, nullptr};
  int *dcount[2] = {nullptr, nullptr};
  double *ubuf[2] = {nullptr, nullptr}, *zbuf[2] = {nullptr, nullptr};
  double *U = nullptr, *Ux = nullptr, *Uxx = nullptr, *Vee = nullptr, *Ven = nullptr;
  double *cur_el = nullptr;
  int *age = nullptr;  // [Wc] consecutive rejected moves (QMCWalker::age)
  // all-electron moves (opt.all_electron_moves): the walker record at R --
  // modified drift [Wc][N][3], ln|Psi|, sign, |modified drift| / |drift| and
  // the move's forward Green's-function sum [Wc]; Rin holds R during a move
  double *ae_drift = nullptr, *ae_logpsi = nullptr, *ae_sgn = nullptr, *ae_mr = nullptr,
         *ae_fwd = nullptr;
  // staggered refresh: the walker list sorted by slot residue mod K (see
  // refresh_plan), valid for the VMC identity list of rplan_num walkers, or
  // built from the DMC active list (rplan_num = -1)
  int *d_rmap = nullptr;
  std::vector<int> rmap, rpos, roff;
  int rplan_num = 0, rplan_K = 0, rplan_S = 1;
  long moves_of_spin[2] = {0, 0};  // single-electron moves of each spin so far
  double *Rin = nullptr;
  double *propR = nullptr, *gauss = nullptr, *unif = nullptr;
  double *pU = nullptr, *pUx = nullptr, *pUxx = nullptr, *pVee = nullptr, *pVen = nullptr;
  void *vrow = nullptr;
  long ldv = 0;
  double *st_inv = nullptr, *st_grad = nullptr, *st_lap = nullptr, *st_logdet = nullptr,
         *st_sgn = nullptr;
  int *st_singular = nullptr;
  // basis scratch
  void *X = nullptr;
  long ldp = 0;      // point capacity (multiple of 64)
  // coefficients per spin in the GEMM operand type: transposed [kpad][npad]
  // (FP64, FP32), orbital-major [npad][kpad] (BF16)
  void *Ct[2] = {nullptr, nullptr};
  // device output arrays (batch rows) for the host-buffer entry points
  qmcb_results dout{};
  int *d_ids = nullptr;
  uint8_t *d_mask = nullptr;
  // vmc accumulators
  double *acc_el = nullptr, *acc_el2 = nullptr, *acc_n = nullptr, *acc_prop_el = nullptr;
  // several determinants: per-determinant records (inverse, log-det, traces)
  // at w * ndet + d; active orbitals nact[sp] >= n[sp] share one GEMM
  int ndet = 1;
  bool multi = false;
  int nact[2] = {0, 0};
  double *d_ci = nullptr;
  const qmcb_pade *d_ee = nullptr;  // DevSystem::ee on the device
  int *d_occ[2] = {nullptr, nullptr};
  int *st_rowsing = nullptr;  // [Wc] combined proposal failed
  double *cgrad = nullptr;     // [Wc][N][3] grad ln D of the determinant sum
  double *st_cgrad = nullptr;  // [Wc][N][3] the same at a staged proposal
  // electron-nucleus cusp replacement (per spin)
  bool cusp = false;
  CuspDev cuspd[2] = {};
  // per-sweep observable series (qmcb_sample_series)
  double *d_obs = nullptr;     // [Wc][3]
  double *d_series = nullptr;  // [series_cap][4]
  int series_cap = 0, series_len = 0;
  // dmc population
  double *weight = nullptr, *acc_w = nullptr;
  int *d_active = nullptr, *d_pairs = nullptr;
  std::vector<int> active, free_slots;
  bool dmc_ready = false;
  bool have_e_trial = false;
  double e_trial = 0.0;
  double e_est = 0.0;  // weighted mean local energy the trial energy is fed back from
  double e_orig = 0.0;  // energy_estimated_original: the first e_trial after dmc_init
  // production samples (one per sweep) the estimate averages, across calls
  double est_we = 0.0, est_w = 0.0;
  long est_n = 0;
  // staged single-electron pass
  int staged_moved = -1;
  int staged_num = 0;
  std::vector<int> staged_ids;
  std::map<int, int> staged_row;
  std::vector<int> initialized;  // per slot: has a resident state
  int auto_groups = 0;           // walker_groups < 0: the count chosen by timing (0: not yet)
  int auto_groups_num = 0;       // ... for this many walkers
  int tune_num = 0, tune_q = 0;  // timing in progress: batch size, windows run so far
  float tune_ms[3] = {0, 0, 0};  // best window per candidate
  int last_groups = 0;           // groups the last vmc/dmc steps ran with
  void *scratch_dev = nullptr;   // this device's address of `scratch`
  // timing
  bool timing = false;
  // all-electron moves: the evaluate at the proposal captured in a graph for
  // this walker count, walker map and time step; ae_last_* is the previous
  // step's, since a graph pays off only when the next step repeats it
  hipGraphExec_t ae_graph = nullptr;
  int ae_graph_num = -1, ae_last_num = -1;
  const int *ae_graph_map = nullptr, *ae_last_map = nullptr;
  double ae_graph_tau = 0.0, ae_last_tau = 0.0;
  Timing tm;
  std::string err;
};

namespace {

#define QCHK(expr)                                                              \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(_e);             \
      return QMCB_DEVICE_ERROR;                                                 \
    }                                                                           \
  } while (0)

// Device memory.  qmcb_create makes one allocation, the arena, of the size
// plan_device_bytes projects for the context, zero-fills it on the context's
// own stream (a non-blocking stream: work on the legacy null stream would not
// be ordered before its kernels) and carves every array out of it, each
// rounded up to 256 bytes; qmcb_create synchronises the stream before
// returning.  What the projection leaves out (the per-sweep series buffer,
// the walker-group slot map of a context created without groups) gets an
// allocation of its own, zero-filled the same way.
constexpr size_t kArenaAlign = 256;
inline size_t arena_round(size_t bytes) {
  return (std::max<size_t>(bytes, 16) + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
}

hipError_t dalloc_bytes(qmcb_ctx *ctx, void **p, size_t bytes, bool own = false) {
  const size_t r = arena_round(bytes);
  if (!own && ctx->arena && ctx->arena_used + r <= ctx->arena_cap) {
    *p = ctx->arena + ctx->arena_used;
    ctx->arena_used += r;
    return hipSuccess;
  }
  bytes = std::max<size_t>(bytes, 16);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return e;
  ctx->allocs.push_back(*p);
  ctx->dev_bytes += bytes;
  return hipMemsetAsync(*p, 0, bytes, ctx->stream);
}

template <typename T>
hipError_t dalloc(qmcb_ctx *ctx, T **p, size_t count, bool own = false) {
  void *q = nullptr;
  hipError_t e = dalloc_bytes(ctx, &q, count * sizeof(T), own);
  *p = (T *)q;
  return e;
}

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet)
uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

int pad_orb(int n) {
  if (n > 64) return ((n + 127) / 128) * 128;
  if (n > 32) return 64;
  if (n > 16) return 32;
  return 16;
}

// ---- timing helpers --------------------------------------------------------
void tm_begin(qmcb_ctx *ctx, const char *name, hipEvent_t *e0) {
  if (!ctx->timing) return;
  Timing &t = ctx->tm;
  if (t.used + 2 > t.pool.size()) {
    for (int i = 0; i < 64; i++) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      t.pool.push_back(e);
    }
  }
  *e0 = t.pool[t.used];
  (void)hipEventRecord(*e0, ctx->stream);
  (void)name;
}
void tm_end(qmcb_ctx *ctx, const char *name, hipEvent_t e0) {
  if (!ctx->timing) return;
  Timing &t = ctx->tm;
  hipEvent_t e1 = t.pool[t.used + 1];
  (void)hipEventRecord(e1, ctx->stream);
  auto it = t.index.find(name);
  int id;
  if (it == t.index.end()) {
    id = (int)t.names.size();
    t.index[name] = id;
    t.names.push_back(name);
    t.ms.push_back(0.0);
    t.launches.push_back(0);
  } else {
    id = it->second;
  }
  t.marks.push_back({id, {e0, e1}});
  t.used += 2;
}
void tm_collect(qmcb_ctx *ctx) {
  if (!ctx->timing) return;
  Timing &t = ctx->tm;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &m : t.marks) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, m.second.first, m.second.second);
    t.ms[m.first] += ms;
    t.launches[m.first] += 1;
  }
  t.marks.clear();
  t.used = 0;
}

#define TIMED(name, call)               \
  do {                                  \
    hipEvent_t _e0 = nullptr;           \
    tm_begin(ctx, name, &_e0);          \
    hipError_t _r = (call);             \
    tm_end(ctx, name, _e0);             \
    if (_r != hipSuccess) {             \
      ctx->err = std::string(name) + ": " + hipGetErrorString(_r); \
      return QMCB_DEVICE_ERROR;         \
    }                                   \
  } while (0)


SlaterArgs slater_args(qmcb_ctx *ctx, int sp) {
  SlaterArgs a{};
  a.n = ctx->n[sp];
  a.N = ctx->N;
  a.spin_start = ctx->start[sp];
  a.inv = ctx->inv[sp];
  a.dval = ctx->dval[sp];
  a.dmat = ctx->dmat[sp];
  a.logdet = ctx->logdet[sp];
  a.sgn = ctx->sgn[sp];
  a.gradPR = ctx->gradPR[sp];
  a.lapPR = ctx->lapPR[sp];
  a.singular = ctx->singular[sp];
  a.dcount = ctx->dcount[sp];
  a.ubuf = ctx->ubuf[sp];
  a.zbuf = ctx->zbuf[sp];
  a.ndet = ctx->ndet;
  a.nact = ctx->nact[sp];
  a.occ = ctx->multi ? ctx->d_occ[sp] : nullptr;
  return a;
}

// several determinants: arguments of the staging / assembly / commit kernels
MdetArgs mdet_args(qmcb_ctx *ctx, int m, int mode, const qmcb_results &out) {
  MdetArgs a{};
  a.s[0] = slater_args(ctx, 0);
  a.s[1] = slater_args(ctx, 1);
  a.D = ctx->ndet;
  a.N = ctx->N;
  a.moved = m;
  a.ci = ctx->d_ci;
  const int sp = (m >= 0 && m >= ctx->n[0]) ? 1 : 0;
  a.vrow = ctx->vrow;
  a.v_s_comp = ctx->npad[sp];
  a.v_s_walker = (long)kComp * ctx->npad[sp];
  a.st_inv = ctx->st_inv; a.st_grad = ctx->st_grad; a.st_lap = ctx->st_lap;
  a.st_logdet = ctx->st_logdet; a.st_sgn = ctx->st_sgn; a.st_singular = ctx->st_singular;
  a.st_row_singular = ctx->st_rowsing;
  a.R = ctx->R;
  a.propR = ctx->propR;
  a.U = ctx->U; a.Ux = ctx->Ux; a.Uxx = ctx->Uxx; a.Vee = ctx->Vee; a.Ven = ctx->Ven;
  a.pU = ctx->pU; a.pUx = ctx->pUx; a.pUxx = ctx->pUxx; a.pVee = ctx->pVee; a.pVen = ctx->pVen;
  a.v_nn = ctx->dsys.v_nn;
  a.n_nuc = ctx->nn;
  a.nuc_xyz = ctx->dsys.nuc_xyz;
  a.nuc_z = ctx->dsys.nuc_z;
  a.drift_mod = ctx->opt.drift_mod;
  a.dt = ctx->opt.dt;
  a.umrigar_a = ctx->opt.umrigar_a;
  a.out = out;
  a.mode = mode;
  a.gauss = ctx->gauss;
  a.unif = ctx->unif;
  a.mask = ctx->d_mask;
  a.acc_el = ctx->acc_el; a.acc_el2 = ctx->acc_el2; a.acc_n = ctx->acc_n;
  a.acc_prop_el = ctx->acc_prop_el; a.cur_el = ctx->cur_el;
  a.cgrad = ctx->cgrad;
  a.st_cgrad = ctx->st_cgrad;
  a.weight = ctx->weight;
  a.acc_w = ctx->acc_w;
  a.walk.age = ctx->age;
  a.walk.old_walker = ctx->opt.old_walker_acceptance;
  return a;
}

// ---- tile autotuning -----------------------------------------------------------
// The single-electron orbital GEMM has one row per walker, so its grid is
// small (num / 64 workgroups per component) and how evenly it spreads over
// the CUs decides its speed.  The first move of a spin at a batch size (per
// 16 rows; walker groups run at their own sizes) times the tiles on the real
// operands, three launches each, and keeps the fastest: workgroups of 4 and
// 2 waves, each with the rule's orbital tile per wave (gemm_orbital_tiles)
// and the half and double of it where they fit the padded orbitals.
// Batches below 4096 rows (where the timing runs would cost more than they
// can save) take 4 waves and the rule's tile.  Every candidate produces
// identical bits.
template <typename F>
qmcb_status tune_gemm(qmcb_ctx *ctx, int sp, int num, F &&gemm) {
  const std::pair<int, int> key(sp, num >> 4);
  const int n_orb = ctx->nact[sp], ldc = ctx->npad[sp];
  const int rule = gemm_orbital_tiles(n_orb, num);
  if (ctx->gemm_force.first) {
    if (!gemm_orbital_tiles_valid(ctx->gemm_force.second, n_orb, ldc)) {
      ctx->err = "QMCB_GEMM_TILE: orbital tiles do not fit this spin's orbitals";
      return QMCB_INVALID_ARG;
    }
    ctx->gemm_tile[key] = ctx->gemm_force;
    return QMCB_OK;
  }
  if (num < 4096) {  // small batches: the rule's tile, no timing runs
    ctx->gemm_tile[key] = {4, rule};
    return QMCB_OK;
  }
  std::vector<std::pair<int, int>> cand;
  for (int wpb : {4, 2})
    for (int mt : {rule, rule / 2, rule * 2})
      if (mt >= 1 && gemm_orbital_tiles_valid(mt, n_orb, ldc)) cand.push_back({wpb, mt});
  hipEvent_t e0, e1;
  QCHK(hipEventCreate(&e0));
  QCHK(hipEventCreate(&e1));
  std::vector<float> best_ms(cand.size(), 1e30f);
  for (size_t c = 0; c < cand.size(); c++)
    for (int rep = 0; rep < 3; rep++) {
      QCHK(hipEventRecord(e0, ctx->stream));
      QCHK(gemm(cand[c]));
      QCHK(hipEventRecord(e1, ctx->stream));
      QCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      QCHK(hipEventElapsedTime(&ms, e0, e1));
      best_ms[c] = std::min(best_ms[c], ms);
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  size_t best = 0;  // ties keep the earlier candidate (4 waves, the rule's tile first)
  for (size_t c = 1; c < cand.size(); c++)
    if (best_ms[c] < best_ms[best]) best = c;
  ctx->gemm_tile[key] = cand[best];
  return QMCB_OK;
}

// Whether the device walk's single-electron move of electron m lets sm_move
// form the Jastrow factor at the proposal (one determinant, N no more than the
// move tile's threads) instead of a jastrow_move launch before it.
static bool jastrow_fused(const qmcb_ctx *ctx, int m) {
  const int sp = m < ctx->n[0] ? 0 : 1;
  return !ctx->multi && sm_move_fuses_jastrow(ctx->n[sp], ctx->N);
}

// ---- precision dispatch ------------------------------------------------------
// TS: derivative storage, TX: basis/GEMM operand, TO: GEMM output
template <typename TS, typename TX, typename TC, typename TO>
struct Engine {
  // all-electron: slots given by d_map (device) for batch rows [0, num); the
  // basis values go to X with a point capacity of ldp (the context's own
  // scratch unless the caller brings one)
  static qmcb_status all_electron(qmcb_ctx *ctx, const int *d_map, int num, void *Xs = nullptr,
                                  long ldp = 0) {
    if (!Xs) {
      Xs = ctx->X;
      ldp = ctx->ldp;
    }
    for (int sp = 0; sp < 2; sp++) {
      const int n = ctx->n[sp];
      if (n == 0) continue;
      const int chunk = std::max<long>(1, std::min<long>(num, ldp / n));
      if (!d_map && chunk < num) {
        ctx->err = "all-electron evaluate in chunks needs a walker map";
        return QMCB_INVALID_ARG;
      }
      for (int c0 = 0; c0 < num; c0 += chunk) {
        const int cw = std::min(chunk, num - c0);
        PointSet pts{ctx->R + 3L * ctx->start[sp], (long)cw * n, n, 3L * ctx->N, 3,
                     d_map + c0};
        TIMED("basis", (launch_basis<TC, TX>(ctx->dsys, pts, (TX *)Xs, ctx->kpad, ldp,
                                            ctx->stream)));
        const long na = ctx->nact[sp];
        // derivative components at dm_at<TS>(k, o * n + e, na * n); values FP64 (o, e)
        const long e_s = dm_at<TS>(0, 1, 0), o_s = (long)n * e_s, k_s = dm_at<TS>(1, 0, na * n);
        GemmOut go{ctx->dmat[sp], k_s, o_s, (long)kDer * na * n, e_s, n, d_map + c0,
                   ctx->dval[sp], na * n, (long)n, 1};
        TIMED("orbital_gemm", (launch_orbital_gemm<TX, TS>((const TX *)ctx->Ct[sp], ctx->npad[sp],
                                                          (int)na,
                                                          (const TX *)Xs, ctx->kpad, ldp,
                                                          (long)cw * n, ctx->kpad, go,
                                                          ctx->stream)));
        if (ctx->cusp)
          TIMED("cusp", (launch_cusp<TS>(ctx->cuspd[sp], pts, go, ctx->stream)));
      }
      SlaterArgs sa = slater_args(ctx, sp);
      TIMED("lu_inverse", launch_lu_inverse<TS>(sa, d_map, num, ctx->stream));
      TIMED("traces_all", launch_traces_all<TS>(sa, d_map, num, ctx->stream));
    }
    JastrowArgs ja{};
    ja.sys = ctx->dsys;
    ja.R = ctx->R;
    ja.U = ctx->U; ja.Ux = ctx->Ux; ja.Uxx = ctx->Uxx; ja.Vee = ctx->Vee; ja.Ven = ctx->Ven;
    TIMED("jastrow_all", launch_jastrow_all(ja, d_map, num, ctx->stream));
    return QMCB_OK;
  }

  // what: bit 0 writes each walker's current E_L (cur_el), bit 1 resets its
  // age (a fresh walker); dt < 0: the context's time step for the drift
  // modification
  static qmcb_status assemble(qmcb_ctx *ctx, const int *d_map, int num, const qmcb_results &out,
                              double dt = -1.0, int what = 3) {
    if (ctx->multi) {
      MdetArgs ma = mdet_args(ctx, -1, 0, out);
      TIMED("mdet_assemble", launch_mdet_assemble(ma, d_map, num, ctx->stream));
      return QMCB_OK;
    }
    AssembleArgs aa{};
    aa.sys = ctx->dsys;
    aa.s[0] = slater_args(ctx, 0);
    aa.s[1] = slater_args(ctx, 1);
    aa.R = ctx->R; aa.U = ctx->U; aa.Ux = ctx->Ux; aa.Uxx = ctx->Uxx;
    aa.Vee = ctx->Vee; aa.Ven = ctx->Ven;
    aa.drift_mod = ctx->opt.drift_mod;
    aa.dt = dt < 0 ? ctx->opt.dt : dt;
    aa.umrigar_a = ctx->opt.umrigar_a;
    aa.out = out;
    aa.cur_el = (what & 1) ? ctx->cur_el : nullptr;
    aa.age = (what & 2) ? ctx->age : nullptr;
    TIMED("assemble", launch_assemble(aa, d_map, num, ctx->stream));
    return QMCB_OK;
  }

  // proposal rows: propR[b] set; Jastrow at the proposal (unless sm_move
  // forms it, jastrow = false), basis + GEMM row
  static qmcb_status move_front(qmcb_ctx *ctx, const int *d_map, int num, int m,
                                bool jastrow = true) {
    const int sp = m < ctx->n[0] ? 0 : 1;
    JastrowArgs ja{};
    ja.sys = ctx->dsys;
    ja.R = ctx->R;
    ja.U = ctx->pU; ja.Ux = ctx->pUx; ja.Uxx = ctx->pUxx; ja.Vee = ctx->pVee; ja.Ven = ctx->pVen;
    ja.moved = m;
    ja.propR = ctx->propR;
    if (jastrow)
      TIMED("jastrow_move", launch_jastrow_move(ja, ctx->U, ctx->Ux, ctx->Uxx, ctx->Vee,
                                                ctx->Ven, d_map, num, ctx->stream));
    PointSet pts{ctx->propR, (long)num, 1, 3, 0, nullptr};
    TIMED("basis", (launch_basis<TC, TX>(ctx->dsys, pts, (TX *)ctx->X, ctx->kpad, ctx->ldp,
                                        ctx->stream)));
    const int n = ctx->n[sp];
    GemmOut go{ctx->vrow, (long)ctx->npad[sp], 1, (long)kComp * ctx->npad[sp], 0, 1, nullptr,
               nullptr, 0, 0, 0};
    auto gemm = [&](std::pair<int, int> tile) {
      return launch_orbital_gemm<TX, TO>((const TX *)ctx->Ct[sp], ctx->npad[sp], ctx->nact[sp],
                                         (const TX *)ctx->X, ctx->kpad, ctx->ldp, (long)num,
                                         ctx->kpad, go, ctx->stream, tile.first, tile.second);
    };
    auto it = ctx->gemm_tile.find(std::make_pair(sp, num >> 4));
    if (it == ctx->gemm_tile.end()) {
      qmcb_status ts = tune_gemm(ctx, sp, num, gemm);
      if (ts != QMCB_OK) return ts;
      it = ctx->gemm_tile.find(std::make_pair(sp, num >> 4));
    }
    TIMED("orbital_gemm", gemm(it->second));
    (void)n;
    if (ctx->cusp) TIMED("cusp", (launch_cusp<TO>(ctx->cuspd[sp], pts, go, ctx->stream)));
    return QMCB_OK;
  }

  static MoveArgs move_args(qmcb_ctx *ctx, int m, int mode, const qmcb_results &out) {
    const int sp = m < ctx->n[0] ? 0 : 1;
    MoveArgs a{};
    a.s = slater_args(ctx, sp);
    a.other = slater_args(ctx, 1 - sp);
    a.moved = m;
    a.N = ctx->N;
    a.vrow = ctx->vrow;
    a.v_s_comp = ctx->npad[sp];
    a.v_s_walker = (long)kComp * ctx->npad[sp];
    a.n_nuc = ctx->nn;
    a.nuc_xyz = ctx->dsys.nuc_xyz;
    a.nuc_z = ctx->dsys.nuc_z;
    a.R = ctx->R;
    a.propR = ctx->propR;
    a.U = ctx->U; a.Ux = ctx->Ux; a.Uxx = ctx->Uxx; a.Vee = ctx->Vee; a.Ven = ctx->Ven;
    a.pU = ctx->pU; a.pUx = ctx->pUx; a.pUxx = ctx->pUxx; a.pVee = ctx->pVee; a.pVen = ctx->pVen;
    a.v_nn = ctx->dsys.v_nn;
    a.ee = ctx->d_ee;
    a.en = ctx->dsys.en[sp];
    a.n_up = ctx->n[0];
    a.drift_mod = ctx->opt.drift_mod;
    a.dt = ctx->opt.dt;
    a.umrigar_a = ctx->opt.umrigar_a;
    a.out = out;
    a.st_inv = ctx->st_inv; a.st_grad = ctx->st_grad; a.st_lap = ctx->st_lap;
    a.st_logdet = ctx->st_logdet; a.st_sgn = ctx->st_sgn; a.st_singular = ctx->st_singular;
    a.mode = mode;
    a.gauss = ctx->gauss;
    a.unif = ctx->unif;
    a.acc_el = ctx->acc_el; a.acc_el2 = ctx->acc_el2; a.acc_n = ctx->acc_n;
    a.acc_prop_el = ctx->acc_prop_el; a.cur_el = ctx->cur_el;
    a.weight = ctx->weight;
    a.acc_w = ctx->acc_w;
    a.walk.age = ctx->age;
    a.walk.old_walker = ctx->opt.old_walker_acceptance;
    return a;
  }

  static qmcb_status sm(qmcb_ctx *ctx, const MoveArgs &a, const int *d_map, int num) {
    TIMED("sm_move", (launch_sm_move<TS, TO>(a, d_map, num, ctx->stream)));
    return QMCB_OK;
  }

  static qmcb_status commit(qmcb_ctx *ctx, const MoveArgs &a, const int *d_map,
                            const uint8_t *d_mask, int num) {
    TIMED("commit", (launch_commit<TS, TO>(a, d_map, d_mask, num, ctx->stream)));
    return QMCB_OK;
  }

  // several determinants: stage the proposal of every determinant, combine
  // (and, in VMC mode, accept on the device), commit accepted rows
  static qmcb_status mdet_stage(qmcb_ctx *ctx, const int *d_map, int num, int m, int mode,
                                const qmcb_results &out, double tau, double *obs) {
    MdetArgs ma = mdet_args(ctx, m, mode, out);
    ma.tau = tau;
    ma.obs = obs;
    TIMED("mdet_move", (launch_mdet_move<TS, TO>(ma, d_map, num, ctx->stream)));
    TIMED("mdet_assemble", launch_mdet_assemble(ma, d_map, num, ctx->stream));
    return QMCB_OK;
  }

  // the same with arguments prepared by the caller (device DMC)
  static qmcb_status mdet_stage_args(qmcb_ctx *ctx, const MdetArgs &ma, const int *d_map,
                                     int num) {
    TIMED("mdet_move", (launch_mdet_move<TS, TO>(ma, d_map, num, ctx->stream)));
    TIMED("mdet_assemble", launch_mdet_assemble(ma, d_map, num, ctx->stream));
    return QMCB_OK;
  }

  static qmcb_status mdet_commit(qmcb_ctx *ctx, const int *d_map, const uint8_t *d_mask, int num,
                                 int m) {
    MdetArgs ma = mdet_args(ctx, m, 0, qmcb_results{});
    TIMED("mdet_commit", (launch_mdet_commit<TS, TO>(ma, d_map, d_mask, num, ctx->stream)));
    return QMCB_OK;
  }

  static qmcb_status refresh(qmcb_ctx *ctx, const int *d_map, int num, int spin_mask = 3) {
    for (int sp = 0; sp < 2; sp++) {
      if (ctx->n[sp] == 0 || !(spin_mask & (1 << sp))) continue;
      SlaterArgs sa = slater_args(ctx, sp);
      TIMED("lu_inverse", launch_lu_inverse<TS>(sa, d_map, num, ctx->stream));
      TIMED("traces_all", launch_traces_all<TS>(sa, d_map, num, ctx->stream));
    }
    return QMCB_OK;
  }
};

// Engine<TS derivative storage, TX GEMM operand, TC basis arithmetic, TO GEMM output>
typedef Engine<double, double, double, double> EngineF64;
typedef Engine<float, float, double, float> EngineF32;       // FP64 basis, FP32 MFMA
typedef Engine<float, double, double, double> EngineF32Store;  // FP64 GEMM, FP32 storage
typedef Engine<float, __hip_bfloat16, double, float> EngineBF16;  // FP64 basis, BF16 MFMA
typedef Engine<bf16s, float, double, float> EngineBF16Store;  // FP32 MFMA, BF16 storage

#define DISPATCH(ctx, call)                                                          \
  ((ctx)->opt.precision == QMCB_FP64                                                 \
       ? EngineF64::call                                                             \
       : ((ctx)->opt.precision == QMCB_MIXED_F32STORE                                \
              ? EngineF32Store::call                                                 \
              : ((ctx)->opt.precision == QMCB_MIXED_BF16                             \
                     ? EngineBF16::call                                              \
                     : ((ctx)->opt.precision == QMCB_MIXED_BF16STORE ? EngineBF16Store::call \
                                                                    : EngineF32::call))))

qmcb_status check_ids(qmcb_ctx *ctx, int num, const int32_t *ids) {
  if (num < 0 || num > ctx->Wc) {
    ctx->err = "num out of range";
    return QMCB_INVALID_ARG;
  }
  for (int i = 0; i < num; i++)
    if (ids[i] < 0 || ids[i] >= ctx->Wc) {
      ctx->err = "walker id out of range";
      return QMCB_INVALID_ARG;
    }
  return QMCB_OK;
}

// group consecutive basis functions of one atom with identical primitive lists
struct Shells {
  std::vector<int> atom, prim_off, nprim, bf0, ncart;
};

Shells make_shells(const qmcb_system *s) {
  Shells sh;
  int j = 0;
  while (j < s->n_basis) {
    const int a = s->bf_atom[j];
    const int p0 = s->bf_prim_off[j], np = s->bf_prim_off[j + 1] - p0;
    int k = j + 1;
    while (k < s->n_basis && s->bf_atom[k] == a && k - j < 6) {
      const int q0 = s->bf_prim_off[k], nq = s->bf_prim_off[k + 1] - q0;
      if (nq != np) break;
      bool same = true;
      for (int i = 0; i < np; i++)
        if (s->prim_alpha[q0 + i] != s->prim_alpha[p0 + i] ||
            s->prim_coef[q0 + i] != s->prim_coef[p0 + i]) {
          same = false;
          break;
        }
      if (!same) break;
      const int L0 = s->bf_pow[3 * j] + s->bf_pow[3 * j + 1] + s->bf_pow[3 * j + 2];
      const int L1 = s->bf_pow[3 * k] + s->bf_pow[3 * k + 1] + s->bf_pow[3 * k + 2];
      if (L0 != L1) break;
      k++;
    }
    sh.atom.push_back(a);
    sh.prim_off.push_back(p0);
    sh.nprim.push_back(np);
    sh.bf0.push_back(j);
    sh.ncart.push_back(k - j);
    j = k;
  }
  return sh;
}

template <typename T>
hipError_t upload(qmcb_ctx *ctx, const T **dst, const T *src, size_t count) {
  T *d = nullptr;
  hipError_t e = dalloc(ctx, &d, count);
  if (e != hipSuccess) return e;
  if (count) e = hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // src is pageable host memory
  *dst = d;
  return e;
}

size_t plan_device_bytes(const qmcb_system *s, const qmcb_options *o);

qmcb_status create_impl(const qmcb_system *s, const qmcb_options *o, qmcb_ctx *ctx) {
  ctx->opt = *o;
  if (const char *gt = std::getenv("QMCB_GEMM_TILE")) {
    int wpb = 0, mt = 0;
    if (std::sscanf(gt, "%d,%d", &wpb, &mt) != 2 || (wpb != 2 && wpb != 4) ||
        (mt != 1 && mt != 2 && mt != 4 && mt != 8)) {
      ctx->err = "QMCB_GEMM_TILE must be \"wpb,mt\": wpb 2 or 4, mt 1, 2, 4 or 8";
      return QMCB_INVALID_ARG;
    }
    ctx->gemm_force = {wpb, mt};
  }
  if (const char *sp = std::getenv("QMCB_SEPARATE_PROPOSE")) ctx->separate_propose = sp[0] == '1';
  if (ctx->opt.dt <= 0) ctx->opt.dt = 0.1;
  if (ctx->opt.umrigar_a <= 0) ctx->opt.umrigar_a = 0.5;
  ctx->N = s->n_up + s->n_dn;
  ctx->n[0] = s->n_up;
  ctx->n[1] = s->n_dn;
  ctx->start[0] = 0;
  ctx->start[1] = s->n_up;
  ctx->nn = s->n_nuc;
  ctx->B = s->n_basis;
  ctx->nprim = s->n_prim;
  ctx->kpad = ((s->n_basis + 31) / 32) * 32;  // GEMM k-chunks of 32 (zero padded)
  ctx->Wc = o->max_walkers;
  const bool f64 = o->precision == QMCB_FP64;
  const bool store32 = o->precision == QMCB_MIXED_F32STORE;
  ctx->ts = f64 ? 8 : (o->precision == QMCB_MIXED_BF16STORE ? 2 : 4);
  ctx->tx = ctx->to = (f64 || store32) ? 8 : 4;
  if (o->precision == QMCB_MIXED_BF16) ctx->tx = 2;  // BF16 GEMM operands, FP32 output
  QCHK(hipSetDevice(o->device));
  ctx->device = o->device;
  QCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  {
    // the device address of the no-allocation scratch, once per device
    if (o->device < 0 || o->device >= kMaxDevices) {
      ctx->err = "device ordinal out of range";
      return QMCB_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[o->device]) {
      void *ptr = nullptr;
      QCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      g_scratch_dev[o->device] = ptr;
    }
    ctx->scratch_dev = g_scratch_dev[o->device];
  }
  {
    const size_t cap = plan_device_bytes(s, o);
    QCHK(hipMalloc((void **)&ctx->arena, cap));
    ctx->allocs.push_back(ctx->arena);
    ctx->arena_cap = cap;
    ctx->dev_bytes += cap;
    QCHK(hipMemsetAsync(ctx->arena, 0, cap, ctx->stream));
  }


  // ---- system tables
  DevSystem &d = ctx->dsys;
  d.N = ctx->N;
  d.n[0] = ctx->n[0]; d.n[1] = ctx->n[1];
  d.start[0] = 0; d.start[1] = ctx->n[0];
  d.nn = ctx->nn;
  d.B = ctx->B;
  d.nprim = ctx->nprim;
  Shells sh = make_shells(s);
  ctx->nshell = d.nshell = (int)sh.atom.size();
  QCHK(upload(ctx, &d.nuc_xyz, s->nuc_xyz, 3 * (size_t)ctx->nn));
  QCHK(upload(ctx, &d.nuc_z, s->nuc_z, (size_t)ctx->nn));
  QCHK(upload(ctx, &d.sh_atom, sh.atom.data(), sh.atom.size()));
  QCHK(upload(ctx, &d.sh_prim_off, sh.prim_off.data(), sh.prim_off.size()));
  QCHK(upload(ctx, &d.sh_nprim, sh.nprim.data(), sh.nprim.size()));
  QCHK(upload(ctx, &d.sh_bf0, sh.bf0.data(), sh.bf0.size()));
  QCHK(upload(ctx, &d.sh_ncart, sh.ncart.data(), sh.ncart.size()));
  QCHK(upload(ctx, &d.bf_pow, (const int *)s->bf_pow, 3 * (size_t)ctx->B));
  {
    // cost-balanced shell chunks for the basis kernel's second grid dimension
    const int nch = std::max(1, std::min(32, ctx->nshell / 4));
    std::vector<double> cost(ctx->nshell);
    double tot = 0.0;
    for (int i = 0; i < ctx->nshell; i++) {
      cost[i] = sh.nprim[i] + 2.0 * sh.ncart[i];
      tot += cost[i];
    }
    std::vector<int> beg(1, 0);
    double accum = 0.0;
    for (int i = 0; i < ctx->nshell; i++) {
      accum += cost[i];
      if ((int)beg.size() < nch && accum >= tot * beg.size() / nch && i + 1 < ctx->nshell)
        beg.push_back(i + 1);
    }
    beg.push_back(ctx->nshell);
    d.nchunk = (int)beg.size() - 1;
    QCHK(upload(ctx, &d.chunk_beg, beg.data(), beg.size()));
  }
  QCHK(upload(ctx, &d.prim_alpha, s->prim_alpha, (size_t)ctx->nprim));
  QCHK(upload(ctx, &d.prim_coef, s->prim_coef, (size_t)ctx->nprim));
  d.ee[0] = s->ee_updn;
  d.ee[1] = s->ee_upup;
  d.ee[2] = s->ee_dndn;
  QCHK(upload(ctx, &d.en[0], s->en_up, (size_t)ctx->nn));
  QCHK(upload(ctx, &d.en[1], s->en_dn, (size_t)ctx->nn));
  QCHK(upload(ctx, &ctx->d_ee, d.ee, 3));  // for kernels that take no DevSystem
  double vnn = 0.0;
  for (int i = 0; i < ctx->nn; i++)
    for (int j = i + 1; j < ctx->nn; j++) {
      double dx = s->nuc_xyz[3 * i] - s->nuc_xyz[3 * j];
      double dy = s->nuc_xyz[3 * i + 1] - s->nuc_xyz[3 * j + 1];
      double dz = s->nuc_xyz[3 * i + 2] - s->nuc_xyz[3 * j + 2];
      vnn += s->nuc_z[i] * s->nuc_z[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  d.v_nn = vnn;

  // ---- determinants: active orbitals and occupation maps
  ctx->multi = s->n_det > 1 || s->occ_up || s->occ_dn;
  ctx->ndet = std::max(1, (int)s->n_det);
  if (ctx->ndet > 64) {
    ctx->err = "at most 64 determinants";
    return QMCB_INVALID_ARG;
  }
  for (int sp = 0; sp < 2; sp++) {
    const int n = ctx->n[sp];
    const int na = sp == 0 ? s->n_act_up : s->n_act_dn;
    ctx->nact[sp] = (ctx->multi && na > 0) ? na : n;
    if (ctx->nact[sp] < n) {
      ctx->err = "fewer active orbitals than electrons";
      return QMCB_INVALID_ARG;
    }
    if (!ctx->multi || n == 0) continue;
    const int32_t *occ = sp == 0 ? s->occ_up : s->occ_dn;
    if (!occ || !s->ci_coef) {
      ctx->err = "several determinants need ci_coef, occ_up and occ_dn";
      return QMCB_INVALID_ARG;
    }
    for (long i = 0; i < (long)ctx->ndet * n; i++)
      if (occ[i] < 0 || occ[i] >= ctx->nact[sp]) {
        ctx->err = "occupation index outside the active orbitals";
        return QMCB_INVALID_ARG;
      }
    QCHK(upload(ctx, (const int **)&ctx->d_occ[sp], occ, (size_t)ctx->ndet * n));
  }
  if (ctx->multi) {
    if (std::max(ctx->n[0], ctx->n[1]) > 128) {
      ctx->err = "several determinants: at most 128 electrons per spin";
      return QMCB_INVALID_ARG;
    }
    QCHK(upload(ctx, (const double **)&ctx->d_ci, s->ci_coef, (size_t)ctx->ndet));
  }

  // ---- transposed, padded MO coefficients (active orbitals) in the operand type
  for (int sp = 0; sp < 2; sp++) {
    const int n = ctx->nact[sp];
    ctx->npad[sp] = pad_orb(std::max(n, 1));
    const double *C = sp == 0 ? s->mo_up : s->mo_dn;
    const size_t cnt = (size_t)ctx->kpad * ctx->npad[sp];
    std::vector<double> ct(cnt, 0.0);
    for (int o2 = 0; o2 < n; o2++)
      for (int b = 0; b < ctx->B; b++) ct[(size_t)b * ctx->npad[sp] + o2] = C[(size_t)o2 * ctx->B + b];
    QCHK(dalloc_bytes(ctx, &ctx->Ct[sp], cnt * ctx->tx));
    if (ctx->tx == 8) {
      QCHK(hipMemcpyAsync(ctx->Ct[sp], ct.data(), cnt * 8, hipMemcpyHostToDevice, ctx->stream));
      QCHK(hipStreamSynchronize(ctx->stream));
    } else if (ctx->tx == 2) {
      // BF16 operands orbital-major, [npad][kpad] (the GEMM's 16x16x32 MFMA
      // takes eight consecutive basis functions of one orbital per lane)
      std::vector<uint16_t> cb(cnt);
      const long np = ctx->npad[sp], kp = ctx->kpad;
      for (long o2 = 0; o2 < np; o2++)
        for (long b = 0; b < kp; b++) cb[o2 * kp + b] = bf16_bits((float)ct[b * np + o2]);
      QCHK(hipMemcpyAsync(ctx->Ct[sp], cb.data(), cnt * 2, hipMemcpyHostToDevice, ctx->stream));
      QCHK(hipStreamSynchronize(ctx->stream));
    } else {
      std::vector<float> cf(cnt);
      for (size_t i = 0; i < cnt; i++) cf[i] = (float)ct[i];
      QCHK(hipMemcpyAsync(ctx->Ct[sp], cf.data(), cnt * 4, hipMemcpyHostToDevice, ctx->stream));
      QCHK(hipStreamSynchronize(ctx->stream));
    }
  }

  // ---- cusp replacement tables (host fit, device copy)
  if (o->replace_cusps) {
    ctx->cusp = true;
    for (int sp = 0; sp < 2; sp++) {
      const int n = ctx->n[sp];
      if (n == 0) continue;
      CuspTable t;
      std::string err;
      if (!fit_cusps(s, sp == 0 ? s->mo_up : s->mo_dn, ctx->nact[sp], t, err)) {
        ctx->err = err;
        return QMCB_INVALID_ARG;
      }
      std::vector<double> orb8((size_t)t.nn * ctx->nact[sp] * 8);
      for (size_t i = 0; i < t.orb.size(); i++) {
        const qmcb_cusp_orbital &q = t.orb[i];
        double *f = &orb8[i * 8];
        f[0] = q.rc; f[1] = q.C; f[2] = q.sgn0;
        for (int k = 0; k < 5; k++) f[3 + k] = q.alpha[k];
      }
      CuspDev &c = ctx->cuspd[sp];
      c.nn = t.nn;
      c.norb = ctx->nact[sp];
      c.ns = (int)t.s_alpha.size();
      c.nuc_xyz = d.nuc_xyz;
      QCHK(upload(ctx, &c.rc_max, t.rc_max.data(), t.rc_max.size()));
      QCHK(upload(ctx, &c.s_off, t.s_off.data(), t.s_off.size()));
      QCHK(upload(ctx, &c.s_alpha, t.s_alpha.data(), t.s_alpha.size()));
      QCHK(upload(ctx, &c.s_coef, t.s_coef.data(), t.s_coef.size()));
      QCHK(upload(ctx, &c.orb8, orb8.data(), orb8.size()));
    }
  }

  // ---- resident walker state
  const long W = ctx->Wc, N = ctx->N;
  QCHK(dalloc(ctx, &ctx->R, W * N * 3));
  const long WD = W * ctx->ndet;  // per-determinant records
  for (int sp = 0; sp < 2; sp++) {
    const long n = ctx->n[sp], na = ctx->nact[sp];
    QCHK(dalloc(ctx, &ctx->inv[sp], WD * n * n));
    QCHK(dalloc(ctx, &ctx->dval[sp], W * na * n));
    QCHK(dalloc_bytes(ctx, &ctx->dmat[sp], (size_t)W * kDer * na * n * ctx->ts));
    QCHK(dalloc(ctx, &ctx->logdet[sp], WD));
    QCHK(dalloc(ctx, &ctx->sgn[sp], WD));
    QCHK(dalloc(ctx, &ctx->gradPR[sp], WD * std::max(n, 1L) * 3));
    QCHK(dalloc(ctx, &ctx->lapPR[sp], WD));
    QCHK(dalloc(ctx, &ctx->singular[sp], WD));
    QCHK(dalloc(ctx, &ctx->dcount[sp], WD));
    QCHK(dalloc(ctx, &ctx->ubuf[sp], W * kDelay * std::max(n, 1L)));
    QCHK(dalloc(ctx, &ctx->zbuf[sp], W * kDelay * std::max(n, 1L)));
  }
  QCHK(dalloc(ctx, &ctx->U, W));
  QCHK(dalloc(ctx, &ctx->Ux, W * N * 3));
  if (ctx->multi) QCHK(dalloc(ctx, &ctx->cgrad, W * N * 3));
  QCHK(dalloc(ctx, &ctx->Uxx, W));
  QCHK(dalloc(ctx, &ctx->Vee, W));
  QCHK(dalloc(ctx, &ctx->Ven, W));
  QCHK(dalloc(ctx, &ctx->cur_el, W));
  QCHK(dalloc(ctx, &ctx->age, W));
  QCHK(dalloc(ctx, &ctx->d_rmap, W));
  QCHK(dalloc(ctx, &ctx->Rin, W * N * 3));
  QCHK(dalloc(ctx, &ctx->propR, W * 3));
  QCHK(dalloc(ctx, &ctx->gauss, W * 3));
  QCHK(dalloc(ctx, &ctx->unif, W));
  QCHK(dalloc(ctx, &ctx->pU, W));
  QCHK(dalloc(ctx, &ctx->pUx, W * N * 3));
  QCHK(dalloc(ctx, &ctx->pUxx, W));
  QCHK(dalloc(ctx, &ctx->pVee, W));
  QCHK(dalloc(ctx, &ctx->pVen, W));
  ctx->ldv = ((W + 63) / 64) * 64;
  const int npmax = std::max(ctx->npad[0], ctx->npad[1]);
  QCHK(dalloc_bytes(ctx, &ctx->vrow, (size_t)kComp * npmax * ctx->ldv * ctx->to));
  QCHK(dalloc(ctx, &ctx->acc_el, W));
  QCHK(dalloc(ctx, &ctx->acc_el2, W));
  QCHK(dalloc(ctx, &ctx->acc_n, W));
  QCHK(dalloc(ctx, &ctx->acc_prop_el, W));
  QCHK(dalloc(ctx, &ctx->d_obs, 3 * W));
  QCHK(dalloc(ctx, &ctx->weight, W));
  QCHK(dalloc(ctx, &ctx->acc_w, W));
  QCHK(dalloc(ctx, &ctx->d_active, W));
  QCHK(dalloc(ctx, &ctx->d_pairs, 2 * W));
  QCHK(dalloc(ctx, &ctx->d_ids, W));
  QCHK(dalloc(ctx, &ctx->d_mask, W));
  if (o->all_electron_moves) {
    QCHK(dalloc(ctx, &ctx->ae_drift, W * N * 3));
    QCHK(dalloc(ctx, &ctx->ae_logpsi, W));
    QCHK(dalloc(ctx, &ctx->ae_sgn, W));
    QCHK(dalloc(ctx, &ctx->ae_mr, W));
    QCHK(dalloc(ctx, &ctx->ae_fwd, W));
  }
  // basis scratch: at least one point per walker; all-electron chunks reuse it
  const size_t per_point = (size_t)kComp * ctx->kpad * ctx->tx;
  const size_t budget = o->basis_scratch_bytes > 0 ? (size_t)o->basis_scratch_bytes
                                                   : (size_t)1 << 31;  // basis values per chunk
  long cap = std::max<long>(W, std::min<long>(W * std::max(ctx->n[0], 1), (long)(budget / per_point)));
  cap = std::max<long>(cap, std::max(ctx->n[0], 1));
  ctx->ldp = ((cap + 63) / 64) * 64;
  QCHK(dalloc_bytes(ctx, &ctx->X, per_point * ctx->ldp));
  // device outputs (batch rows)
  qmcb_results &r = ctx->dout;
  QCHK(dalloc(ctx, &r.log_psi, W));
  QCHK(dalloc(ctx, &r.sign_psi, W));
  QCHK(dalloc(ctx, &r.grad_lnpsi, W * N * 3));
  QCHK(dalloc(ctx, &r.mod_drift, W * N * 3));
  QCHK(dalloc(ctx, &r.lap_ratio, W));
  QCHK(dalloc(ctx, &r.e_local, W));
  QCHK(dalloc(ctx, &r.kinetic, W));
  QCHK(dalloc(ctx, &r.potential, W));
  QCHK(dalloc(ctx, &r.v_ne, W));
  QCHK(dalloc(ctx, &r.v_ee, W));
  QCHK(dalloc(ctx, &r.det_ratio, W));
  QCHK(dalloc(ctx, &r.jastrow_u, W));
  QCHK(dalloc(ctx, &r.singular, W));
  QCHK(dalloc(ctx, &r.mod_ratio, W));
  QCHK(dalloc(ctx, &r.r2, W * 3));
  ctx->initialized.assign(ctx->Wc, 0);
  QCHK(hipStreamSynchronize(ctx->stream));
  return QMCB_OK;
}

qmcb_status ensure_staging(qmcb_ctx *ctx) {
  if (ctx->st_inv) return QMCB_OK;
  const long W = ctx->Wc * ctx->ndet;  // one record per (batch row, determinant)
  const long nm = std::max(ctx->n[0], ctx->n[1]);
  QCHK(dalloc(ctx, &ctx->st_inv, W * nm * nm));
  QCHK(dalloc(ctx, &ctx->st_grad, W * nm * 3));
  QCHK(dalloc(ctx, &ctx->st_lap, W));
  QCHK(dalloc(ctx, &ctx->st_logdet, W));
  QCHK(dalloc(ctx, &ctx->st_sgn, W));
  QCHK(dalloc(ctx, &ctx->st_singular, W));
  QCHK(dalloc(ctx, &ctx->st_rowsing, ctx->Wc));
  if (ctx->multi) QCHK(dalloc(ctx, &ctx->st_cgrad, (long)ctx->Wc * ctx->N * 3));
  QCHK(hipStreamSynchronize(ctx->stream));
  return QMCB_OK;
}

template <typename T>
hipError_t d2h(qmcb_ctx *ctx, T *host, const T *dev, size_t count) {
  if (!host || !dev || count == 0) return hipSuccess;
  return hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream);
}

qmcb_status copy_out(qmcb_ctx *ctx, int num, qmcb_results *out) {
  if (!out) return QMCB_OK;
  const size_t N3 = (size_t)ctx->N * 3;
  const qmcb_results &d = ctx->dout;
  QCHK(d2h(ctx, out->log_psi, d.log_psi, num));
  QCHK(d2h(ctx, out->sign_psi, d.sign_psi, num));
  QCHK(d2h(ctx, out->grad_lnpsi, d.grad_lnpsi, num * N3));
  QCHK(d2h(ctx, out->mod_drift, d.mod_drift, num * N3));
  QCHK(d2h(ctx, out->lap_ratio, d.lap_ratio, num));
  QCHK(d2h(ctx, out->e_local, d.e_local, num));
  QCHK(d2h(ctx, out->kinetic, d.kinetic, num));
  QCHK(d2h(ctx, out->potential, d.potential, num));
  QCHK(d2h(ctx, out->v_ne, d.v_ne, num));
  QCHK(d2h(ctx, out->v_ee, d.v_ee, num));
  QCHK(d2h(ctx, out->det_ratio, d.det_ratio, num));
  QCHK(d2h(ctx, out->jastrow_u, d.jastrow_u, num));
  QCHK(d2h(ctx, out->singular, d.singular, num));
  QCHK(d2h(ctx, out->mod_ratio, d.mod_ratio, num));
  QCHK(d2h(ctx, out->r2, d.r2, num * 3));
  QCHK(hipStreamSynchronize(ctx->stream));
  return QMCB_OK;
}

// evaluate with R already on the device (batch rows) and ids on the device
// modificationRatio and x2/y2/z2 once the drift outputs are written
static qmcb_status results_extra(qmcb_ctx *ctx, int num, const int *d_ids, int whichE,
                                 const qmcb_results &dout) {
  if (dout.mod_ratio && (!dout.grad_lnpsi || !dout.mod_drift)) {
    ctx->err = "mod_ratio needs the grad_lnpsi and mod_drift outputs";
    return QMCB_INVALID_ARG;
  }
  QCHK(launch_results_extra(dout, ctx->R, d_ids, ctx->propR, whichE, ctx->N, num, ctx->stream));
  return QMCB_OK;
}

qmcb_status evaluate_core(qmcb_ctx *ctx, int num, const int *d_ids, const double *d_R,
                          int whichE, const qmcb_results &dout) {
  if (whichE == -1) {
    QCHK(launch_scatter_rows(ctx->R, d_R, d_ids, num, ctx->N * 3, ctx->stream));
    qmcb_status s = DISPATCH(ctx, all_electron(ctx, d_ids, num));
    if (s != QMCB_OK) return s;
    s = DISPATCH(ctx, assemble(ctx, d_ids, num, dout));
    if (s != QMCB_OK) return s;
    return results_extra(ctx, num, d_ids, -1, dout);
  }
  qmcb_status s = ensure_staging(ctx);
  if (s != QMCB_OK) return s;
  s = DISPATCH(ctx, move_front(ctx, d_ids, num, whichE));
  if (s != QMCB_OK) return s;
  if (ctx->multi) {
    s = DISPATCH(ctx, mdet_stage(ctx, d_ids, num, whichE, 0, dout, 0.0, nullptr));
  } else {
    MoveArgs a = DISPATCH(ctx, move_args(ctx, whichE, 0, dout));
    s = DISPATCH(ctx, sm(ctx, a, d_ids, num));
  }
  if (s != QMCB_OK) return s;
  return results_extra(ctx, num, d_ids, whichE, dout);
}

// The device memory a context of this system and these options holds in its
// arena: the arrays create_impl allocates plus the staging arrays
// (ensure_staging: first single-electron evaluate, or first move with
// several determinants), each rounded as the arena rounds them, with the cusp tables at their upper bound (every primitive an s
// Gaussian).  The per-sweep series buffer (32 bytes per recorded sweep) is
// not included.
size_t plan_device_bytes(const qmcb_system *s, const qmcb_options *o) {
  size_t total = 0;
  auto add = [&](size_t count, size_t elem) { total += arena_round(count * elem); };
  const bool f64 = o->precision == QMCB_FP64;
  const bool store32 = o->precision == QMCB_MIXED_F32STORE;
  const size_t ts = f64 ? 8 : (o->precision == QMCB_MIXED_BF16STORE ? 2 : 4);
  const size_t to = (f64 || store32) ? 8 : 4;
  const size_t tx = o->precision == QMCB_MIXED_BF16 ? 2 : to;
  const long W = o->max_walkers, N = s->n_up + s->n_dn, nn = s->n_nuc, B = s->n_basis;
  const long kpad = ((B + 31) / 32) * 32;
  const long n[2] = {s->n_up, s->n_dn};
  // system tables
  const Shells sh = make_shells(s);
  const long nshell = (long)sh.atom.size();
  add(3 * nn, 8);
  add(nn, 8);
  for (int i = 0; i < 5; i++) add(nshell, 4);
  add(3 * B, 4);
  add(std::max(1L, std::min(32L, nshell / 4)) + 1, 4);
  add(s->n_prim, 8);
  add(s->n_prim, 8);
  add(nn, sizeof(qmcb_pade));
  add(nn, sizeof(qmcb_pade));
  add(3, sizeof(qmcb_pade));
  const bool multi = s->n_det > 1 || s->occ_up || s->occ_dn;
  const long ndet = std::max(1, (int)s->n_det);
  long nact[2];
  for (int sp = 0; sp < 2; sp++) {
    const int na = sp == 0 ? s->n_act_up : s->n_act_dn;
    nact[sp] = (multi && na > 0) ? na : n[sp];
    if (multi && n[sp] > 0) add(ndet * n[sp], 4);
  }
  if (multi) add(ndet, 8);
  long npad[2];
  for (int sp = 0; sp < 2; sp++) {
    npad[sp] = pad_orb((int)std::max(nact[sp], 1L));
    add(kpad * npad[sp], tx);
  }
  if (o->replace_cusps)
    for (int sp = 0; sp < 2; sp++) {
      if (n[sp] == 0) continue;
      add(nn, 8);
      add(nn + 1, 4);
      add(s->n_prim, 8);
      add(nact[sp] * s->n_prim, 8);
      add(nn * nact[sp] * 8, 8);
    }
  // resident walker state
  const long WD = W * ndet;
  add(W * N * 3, 8);
  for (int sp = 0; sp < 2; sp++) {
    add(WD * n[sp] * n[sp], 8);
    add(W * nact[sp] * n[sp], 8);
    add((size_t)W * kDer * nact[sp] * n[sp], ts);
    add(WD, 8);
    add(WD, 8);
    add(WD * std::max(n[sp], 1L) * 3, 8);
    add(WD, 8);
    add(WD, 4);
    add(WD, 4);
    add(W * kDelay * std::max(n[sp], 1L), 8);
    add(W * kDelay * std::max(n[sp], 1L), 8);
  }
  add(W, 8);
  add(W * N * 3, 8);
  if (multi) add(W * N * 3, 8);
  for (int i = 0; i < 4; i++) add(W, 8);   // Uxx, Vee, Ven, cur_el
  add(W, 4);                                 // age
  add(W, 4);                                 // refresh map
  add(W * N * 3, 8);                         // Rin
  add(W * 3, 8);
  add(W * 3, 8);
  add(W, 8);
  add(W, 8);
  add(W * N * 3, 8);
  for (int i = 0; i < 3; i++) add(W, 8);   // pUxx, pVee, pVen
  const long ldv = ((W + 63) / 64) * 64;
  add((size_t)kComp * std::max(npad[0], npad[1]) * ldv, to);
  for (int i = 0; i < 4; i++) add(W, 8);   // energy accumulators
  add(3 * W, 8);
  add(W, 8);
  add(W, 8);
  add(W, 4);
  add(2 * W, 4);
  add(W, 4);
  add(W, 1);
  if (o->all_electron_moves) {
    add(W * N * 3, 8);                       // all-electron record: drift,
    for (int i = 0; i < 4; i++) add(W, 8);  // ln|Psi|, sign, ratio, forward sum
  }
  const size_t per_point = (size_t)kComp * kpad * tx;
  const size_t budget = o->basis_scratch_bytes > 0 ? (size_t)o->basis_scratch_bytes
                                                   : (size_t)1 << 31;
  long cap = std::max<long>(W, std::min<long>(W * std::max(n[0], 1L), (long)(budget / per_point)));
  cap = std::max<long>(cap, std::max(n[0], 1L));
  add(per_point * (size_t)(((cap + 63) / 64) * 64), 1);
  add(W, 8);                                 // log_psi
  add(W, 1);                                 // sign_psi
  add(W * N * 3, 8);
  add(W * N * 3, 8);
  for (int i = 0; i < 8; i++) add(W, 8);   // lap_ratio .. jastrow_u
  add(W, 1);                                 // singular
  add(W, 8);                                 // mod_ratio
  add(W * 3, 8);                             // r2
  // staging (first single-electron evaluate or move)
  const long nm = std::max(n[0], n[1]);
  add(WD * nm * nm, 8);
  add(WD * nm * 3, 8);
  add(WD, 8);
  add(WD, 8);
  add(WD, 8);
  add(WD, 4);
  add(W, 4);
  if (multi) add(W * N * 3, 8);
  return total;
}

}  // namespace

extern "C" {

int32_t qmcb_abi_version(void) { return QMCB_ABI_VERSION; }

const char *qmcb_last_error(const qmcb_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

qmcb_status qmcb_create(const qmcb_system *sys, const qmcb_options *opt, qmcb_ctx **out) {
  g_create_error.clear();
  if (!sys || !opt || !out) {
    g_create_error = "null argument";
    return QMCB_INVALID_ARG;
  }
  *out = nullptr;
  if (sys->n_up < 0 || sys->n_dn < 0 || sys->n_dn > sys->n_up || sys->n_up + sys->n_dn < 1 ||
      sys->n_up > 128 || sys->n_nuc < 1 || sys->n_basis < 1 || sys->n_prim < 1 ||
      opt->max_walkers < 1 || sys->n_up + sys->n_dn > 512) {
    g_create_error = "system sizes out of range (1 <= n_dn <= n_up <= 128 electrons per spin)";
    return QMCB_INVALID_ARG;
  }
  if (opt->precision != QMCB_FP64 && opt->precision != QMCB_MIXED_FP32 &&
      opt->precision != QMCB_MIXED_F32STORE && opt->precision != QMCB_MIXED_BF16 &&
      opt->precision != QMCB_MIXED_BF16STORE) {
    g_create_error = "precision not supported by this library";
    return QMCB_INVALID_ARG;
  }
  for (int j = 0; j < sys->n_basis; j++) {
    const int *p = sys->bf_pow + 3 * j;
    if (p[0] < 0 || p[1] < 0 || p[2] < 0 || p[0] > 2 || p[1] > 2 || p[2] > 2 ||
        p[0] + p[1] + p[2] > 2 || sys->bf_atom[j] < 0 || sys->bf_atom[j] >= sys->n_nuc) {
      g_create_error = "basis function powers must be s/p/d (L <= 2)";
      return QMCB_INVALID_ARG;
    }
  }
  qmcb_ctx *ctx = new (std::nothrow) qmcb_ctx();
  if (!ctx) return QMCB_OOM;
  qmcb_status st = create_impl(sys, opt, ctx);
  if (st != QMCB_OK) {
    g_create_error = ctx->err;
    qmcb_destroy(ctx);
    return st;
  }
  *out = ctx;
  return QMCB_OK;
}

void qmcb_destroy(qmcb_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (void *p : ctx->allocs) (void)hipFree(p);
  for (hipEvent_t e : ctx->tm.pool) (void)hipEventDestroy(e);
  if (ctx->ae_graph) (void)hipGraphExecDestroy(ctx->ae_graph);

  for (int g = 0; g < qmcb_ctx::kMaxGroups; g++) {
    if (ctx->gstream[g]) {
      (void)hipStreamSynchronize(ctx->gstream[g]);
      (void)hipStreamDestroy(ctx->gstream[g]);
    }
    if (ctx->gevent[g]) (void)hipEventDestroy(ctx->gevent[g]);
    if (ctx->gfront[g]) (void)hipEventDestroy(ctx->gfront[g]);
  }
  if (ctx->main_event) (void)hipEventDestroy(ctx->main_event);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

qmcb_status qmcb_evaluate(qmcb_ctx *ctx, int32_t num, const int32_t *ids, const double *R,
                          int32_t whichE, qmcb_results *out) {
  if (!ctx || (num > 0 && (!ids || !R))) return QMCB_INVALID_ARG;
  QCHK(hipSetDevice(ctx->device));
  qmcb_status s = check_ids(ctx, num, ids);
  if (s != QMCB_OK) return s;
  if (whichE < -1 || whichE >= ctx->N) {
    ctx->err = "whichE out of range";
    return QMCB_INVALID_ARG;
  }
  if (num == 0) return QMCB_OK;
  if (whichE >= 0) {
    for (int i = 0; i < num; i++)
      if (!ctx->initialized[ids[i]]) {
        ctx->err = "single-electron evaluate of a walker without an all-electron evaluate";
        return QMCB_BAD_STATE;
      }
  }
  QCHK(hipMemcpyAsync(ctx->d_ids, ids, sizeof(int) * num, hipMemcpyHostToDevice, ctx->stream));
  std::vector<double> pr;  // must outlive the asynchronous copy below
  if (whichE == -1) {
    QCHK(hipMemcpyAsync(ctx->Rin, R, sizeof(double) * num * ctx->N * 3, hipMemcpyHostToDevice,
                        ctx->stream));
  } else {
    pr.resize((size_t)num * 3);
    for (int i = 0; i < num; i++)
      for (int k = 0; k < 3; k++) pr[3 * i + k] = R[((size_t)i * ctx->N + whichE) * 3 + k];
    QCHK(hipMemcpyAsync(ctx->propR, pr.data(), sizeof(double) * num * 3, hipMemcpyHostToDevice,
                        ctx->stream));
  }
  s = evaluate_core(ctx, num, ctx->d_ids, ctx->Rin, whichE, ctx->dout);
  if (s != QMCB_OK) return s;
  s = copy_out(ctx, num, out);
  if (s != QMCB_OK) return s;
  if (whichE == -1) {
    for (int i = 0; i < num; i++) ctx->initialized[ids[i]] = 1;
    ctx->staged_moved = -1;
  } else {
    ctx->staged_moved = whichE;
    ctx->staged_num = num;
    ctx->staged_ids.assign(ids, ids + num);
    ctx->staged_row.clear();
    for (int i = 0; i < num; i++) ctx->staged_row[ids[i]] = i;
  }
  return QMCB_OK;
}

qmcb_status qmcb_evaluate_device(qmcb_ctx *ctx, int32_t num, const int32_t *ids_dev,
                                 const double *R_dev, int32_t whichE, qmcb_results *out_dev) {
  if (!ctx || !out_dev || (num > 0 && (!ids_dev || !R_dev))) return QMCB_INVALID_ARG;
  if (num < 0 || num > ctx->Wc || whichE < -1 || whichE >= ctx->N) return QMCB_INVALID_ARG;
  QCHK(hipSetDevice(ctx->device));
  if (whichE >= 0) {
    // the proposal row of each walker: R_dev[b][whichE]
    std::vector<double> dummy;
    (void)dummy;
    QCHK(hipMemcpy2DAsync(ctx->propR, 3 * sizeof(double), R_dev + 3L * whichE,
                          (size_t)ctx->N * 3 * sizeof(double), 3 * sizeof(double), num,
                          hipMemcpyDeviceToDevice, ctx->stream));
  }
  return evaluate_core(ctx, num, ids_dev, R_dev, whichE, *out_dev);
}

qmcb_status qmcb_evaluate_device_noalloc(qmcb_ctx *ctx, int32_t num, const int32_t *ids_dev,
                                         const double *R_dev, qmcb_results *out_dev) {
  if (!ctx || !out_dev || (num > 0 && (!ids_dev || !R_dev))) return QMCB_INVALID_ARG;
  if (num < 0 || num > ctx->Wc) {
    ctx->err = "num out of range";
    return QMCB_INVALID_ARG;
  }
  if (ctx->timing) {
    ctx->err = "qmcb_evaluate_device_noalloc: kernel timing creates events; switch it off";
    return QMCB_BAD_STATE;
  }
  if (!ctx->scratch_dev) {
    ctx->err = "no device address for the scratch array";
    return QMCB_BAD_STATE;
  }
  // point capacity of the scratch for this system's basis (a multiple of 64,
  // as the basis and GEMM kernels index whole 64-point tiles); one walker's
  // electrons of a spin must fit
  const size_t per_point = (size_t)kComp * ctx->kpad * ctx->tx;
  const long cap = (long)(kScratchBytes / per_point) / 64 * 64;
  const int nmax = std::max(ctx->n[0], ctx->n[1]);
  if (cap < nmax) {
    ctx->err = "qmcb_evaluate_device_noalloc: one walker's basis values (" +
               std::to_string(per_point * nmax) + " bytes) exceed the " +
               std::to_string(kScratchBytes) + "-byte scratch";
    return QMCB_INVALID_ARG;
  }
  if (num == 0) return QMCB_OK;
  QCHK(hipSetDevice(ctx->device));
  QCHK(launch_scatter_rows(ctx->R, R_dev, ids_dev, num, ctx->N * 3, ctx->stream));
  QCHK(launch_zero_pad_rows(ctx->scratch_dev, ctx->tx, ctx->B, ctx->kpad, cap, ctx->stream));
  qmcb_status s = DISPATCH(ctx, all_electron(ctx, ids_dev, num, ctx->scratch_dev, cap));
  if (s != QMCB_OK) return s;
  s = DISPATCH(ctx, assemble(ctx, ids_dev, num, *out_dev));
  if (s != QMCB_OK) return s;
  return results_extra(ctx, num, ids_dev, -1, *out_dev);
}

qmcb_status qmcb_accept(qmcb_ctx *ctx, int32_t num, const int32_t *ids, const uint8_t *mask) {
  if (!ctx || (num > 0 && (!ids || !mask))) return QMCB_INVALID_ARG;
  if (ctx->staged_moved < 0) {
    ctx->err = "no staged single-electron proposal";
    return QMCB_BAD_STATE;
  }
  QCHK(hipSetDevice(ctx->device));
  std::vector<uint8_t> rowmask(ctx->staged_num, 0);
  for (int i = 0; i < num; i++) {
    auto it = ctx->staged_row.find(ids[i]);
    if (it == ctx->staged_row.end()) {
      ctx->err = "walker id was not part of the staged proposal";
      return QMCB_INVALID_ARG;
    }
    rowmask[it->second] = mask[i] ? 1 : 0;
  }
  QCHK(hipMemcpyAsync(ctx->d_ids, ctx->staged_ids.data(), sizeof(int) * ctx->staged_num,
                      hipMemcpyHostToDevice, ctx->stream));
  QCHK(hipMemcpyAsync(ctx->d_mask, rowmask.data(), ctx->staged_num, hipMemcpyHostToDevice,
                      ctx->stream));
  qmcb_status s;
  if (ctx->multi) {
    s = DISPATCH(ctx, mdet_commit(ctx, ctx->d_ids, ctx->d_mask, ctx->staged_num,
                                  ctx->staged_moved));
  } else {
    MoveArgs a = DISPATCH(ctx, move_args(ctx, ctx->staged_moved, 0, ctx->dout));
    s = DISPATCH(ctx, commit(ctx, a, ctx->d_ids, ctx->d_mask, ctx->staged_num));
  }
  if (s != QMCB_OK) return s;
  QCHK(hipStreamSynchronize(ctx->stream));
  ctx->staged_moved = -1;
  return QMCB_OK;
}

qmcb_status qmcb_refresh(qmcb_ctx *ctx, int32_t num, const int32_t *ids) {
  if (!ctx) return QMCB_INVALID_ARG;
  QCHK(hipSetDevice(ctx->device));
  qmcb_status s = check_ids(ctx, num, ids);
  if (s != QMCB_OK) return s;
  QCHK(hipMemcpyAsync(ctx->d_ids, ids, sizeof(int) * num, hipMemcpyHostToDevice, ctx->stream));
  s = DISPATCH(ctx, refresh(ctx, ctx->d_ids, num));
  if (s != QMCB_OK) return s;
  QCHK(hipStreamSynchronize(ctx->stream));
  return QMCB_OK;
}

qmcb_status qmcb_get_positions(qmcb_ctx *ctx, int32_t num, double *R) {
  if (!ctx || num < 0 || num > ctx->Wc || (num > 0 && !R)) return QMCB_INVALID_ARG;
  QCHK(hipSetDevice(ctx->device));
  QCHK(hipMemcpyAsync(R, ctx->R, sizeof(double) * num * ctx->N * 3, hipMemcpyDeviceToHost,
                      ctx->stream));
  QCHK(hipStreamSynchronize(ctx->stream));
  return QMCB_OK;
}

// Rows of the observable series a step range records: one per step that
// moves electron 0.  Grows the device buffer before the step loop.
// Drift of the device walk: gradPsiRatio (QF_modification_type none) or its
// umrigar93_unequalelectrons modification.
static bool walk_drift_ok(qmcb_ctx *ctx) {
  const int dm = ctx->opt.drift_mod;
  if (dm == QMCB_DRIFT_NONE || dm == QMCB_DRIFT_UMRIGAR93_UNEQUAL) return true;
  ctx->err = "device VMC/DMC drift: QF_modification_type none or umrigar93_unequalelectrons";
  return false;
}

static void walk_propose_args(qmcb_ctx *ctx, ProposeArgs &pa) {
  pa.drift_mod = ctx->opt.drift_mod;
  pa.n_nuc = ctx->nn;
  pa.nuc_xyz = ctx->dsys.nuc_xyz;
  pa.nuc_z = ctx->dsys.nuc_z;
}

// Walker groups.  The device VMC can split its walkers into G contiguous
// groups and enqueue every group's step on a stream of its own, so that one
// group's memory-bound Sherman-Morrison pass overlaps another group's
// compute-bound Jastrow, basis and GEMM kernels.  A group touches only its
// own slots (through the identity map d_iota + first) and its own batch rows
// (the staging pointers below, offset by `first` rows); random numbers are
// keyed by (seed, step, slot), so results are bit-identical for every G.
// GroupView points ctx's stream and batch-row staging at one group while the
// group's kernels are enqueued, and restores them when it goes out of scope.
struct GroupView {
  qmcb_ctx *ctx;
  hipStream_t stream;
  double *propR, *gauss, *unif, *pU, *pUx, *pUxx, *pVee, *pVen, *d_obs;
  void *X, *vrow;
  GroupView(qmcb_ctx *c, int g, long first, int sp) : ctx(c) {
    stream = c->stream;
    propR = c->propR; gauss = c->gauss; unif = c->unif;
    pU = c->pU; pUx = c->pUx; pUxx = c->pUxx; pVee = c->pVee; pVen = c->pVen;
    d_obs = c->d_obs; X = c->X; vrow = c->vrow;
    c->stream = c->gstream[g];
    c->propR += 3 * first; c->gauss += 3 * first; c->unif += first;
    c->pU += first; c->pUx += first * c->N * 3; c->pUxx += first; c->pVee += first;
    c->pVen += first; c->d_obs += 3 * first;
    c->X = (char *)c->X + first * c->tx;                                  // point columns
    c->vrow = (char *)c->vrow + first * (long)kComp * c->npad[sp] * c->to;  // [row][5][npad]
  }
  ~GroupView() {
    ctx->stream = stream;
    ctx->propR = propR; ctx->gauss = gauss; ctx->unif = unif;
    ctx->pU = pU; ctx->pUx = pUx; ctx->pUxx = pUxx; ctx->pVee = pVee; ctx->pVen = pVen;
    ctx->d_obs = d_obs; ctx->X = X; ctx->vrow = vrow;
  }
};

static qmcb_status ensure_groups(qmcb_ctx *ctx, int G) {
  if (!ctx->d_iota) {
    std::vector<int> iota(ctx->Wc);
    for (int i = 0; i < ctx->Wc; i++) iota[i] = i;
    QCHK(dalloc(ctx, &ctx->d_iota, iota.size(), true));
    QCHK(hipMemcpyAsync(ctx->d_iota, iota.data(), iota.size() * sizeof(int),
                        hipMemcpyHostToDevice, ctx->stream));
    QCHK(hipStreamSynchronize(ctx->stream));  // iota is pageable host memory
  }
  if (!ctx->main_event) QCHK(hipEventCreateWithFlags(&ctx->main_event, hipEventDisableTiming));
  for (int g = 0; g < G; g++) {
    if (!ctx->gstream[g]) QCHK(hipStreamCreateWithFlags(&ctx->gstream[g], hipStreamNonBlocking));
    if (!ctx->gevent[g]) QCHK(hipEventCreateWithFlags(&ctx->gevent[g], hipEventDisableTiming));
    if (!ctx->gfront[g]) QCHK(hipEventCreateWithFlags(&ctx->gfront[g], hipEventDisableTiming));
  }
  return QMCB_OK;
}

// every group stream waits for the work queued so far on the main stream
static qmcb_status groups_wait_main(qmcb_ctx *ctx, int G) {
  QCHK(hipEventRecord(ctx->main_event, ctx->stream));
  for (int g = 0; g < G; g++) QCHK(hipStreamWaitEvent(ctx->gstream[g], ctx->main_event, 0));
  return QMCB_OK;
}

// the main stream waits for every group's queued work
static qmcb_status main_wait_groups(qmcb_ctx *ctx, int G) {
  for (int g = 0; g < G; g++) {
    QCHK(hipEventRecord(ctx->gevent[g], ctx->gstream[g]));
    QCHK(hipStreamWaitEvent(ctx->stream, ctx->gevent[g], 0));
  }
  return QMCB_OK;
}

static qmcb_status series_reserve(qmcb_ctx *ctx, int64_t first_step, int32_t n_steps,
                                  int64_t N = 0) {
  if (N <= 0) N = ctx->N;
  const int64_t first = ((first_step + N - 1) / N) * N;  // first multiple of N >= first_step
  const int64_t last = first_step + n_steps;             // exclusive
  const int64_t rows = last > first ? (last - first + N - 1) / N : 0;
  ctx->series_len = 0;
  if (rows > INT32_MAX / 8) {
    ctx->err = "too many sweeps in one call for the observable series";
    return QMCB_INVALID_ARG;
  }
  if (rows > ctx->series_cap) {
    const int cap = (int)std::max<int64_t>(rows, 2 * (int64_t)ctx->series_cap);
    QCHK(dalloc(ctx, &ctx->d_series, 4 * (size_t)cap, true));
    ctx->series_cap = cap;
  }
  return QMCB_OK;
}

// Staggered refresh.  Slot s of a spin is refreshed (FP64 re-inversion and
// traces) once every K single-electron moves of that spin (counted over the
// context's life), on the stream of the walker group that moved it.  The K
// residues s mod K are grouped into phases of S consecutive residues; the
// slots of phase q = (s mod K) / S are refreshed right after the move c with
// c mod K = min(q S + S - 1, K - 1), i.e. at the last move of their phase.
// S is the smallest phase length whose refresh launches carry at least
// kRefreshBatch walkers (one re-inversion launch costs about the same for a
// few matrices as for a full wave of workgroups), so a window of S or more
// steps carries its share of the refresh work and every walker still
// re-inverts exactly once per K moves of a spin.  The plan sorts the walker
// list the steps run over (slots = null: the identity list of the VMC; else
// the DMC active list) by phase, then list position: rmap[roff[q] ..
// roff[q + 1]) are the slots of phase q and rpos their list positions, so the
// share of a walker group (a contiguous range of list positions) is a
// contiguous piece of rmap.  rmap is copied to d_rmap; the caller
// synchronises the stream before rmap changes again.
constexpr int kRefreshBatch = 1024;
static int refresh_phase(int K, int num) {
  return std::max(1, std::min(K, (int)(((long)K * kRefreshBatch + std::max(num, 1) - 1) /
                                       std::max(num, 1))));
}
static qmcb_status refresh_plan(qmcb_ctx *ctx, const int *slots, int num) {
  const int K = ctx->opt.refresh_interval;
  const int S = refresh_phase(K, num);
  const int P = (K + S - 1) / S;
  ctx->rmap.assign(std::max(num, 1), 0);
  ctx->rpos.assign(std::max(num, 1), 0);
  ctx->roff.assign((size_t)P + 1, 0);
  for (int i = 0; i < num; i++) ctx->roff[((slots ? slots[i] : i) % K) / S + 1]++;
  for (int q = 0; q < P; q++) ctx->roff[q + 1] += ctx->roff[q];
  std::vector<int> fill(ctx->roff.begin(), ctx->roff.end() - 1);
  for (int i = 0; i < num; i++) {
    const int sl = slots ? slots[i] : i, at = fill[(sl % K) / S]++;
    ctx->rmap[at] = sl;
    ctx->rpos[at] = i;
  }
  if (num > 0)
    QCHK(hipMemcpyAsync(ctx->d_rmap, ctx->rmap.data(), sizeof(int) * num, hipMemcpyHostToDevice,
                        ctx->stream));
  ctx->rplan_num = slots ? -1 : num;
  ctx->rplan_K = K;
  ctx->rplan_S = S;
  return QMCB_OK;
}

// The refresh share, after a move of spin sp, of the walkers at list
// positions [first, last): the phase whose last move this is, if any.
// (*map, *cnt): the device map and count of the walkers re-inverted (0 if none)
static qmcb_status staggered_refresh(qmcb_ctx *ctx, long first, long last, int sp,
                                     const int **map = nullptr, int *cnt = nullptr) {
  if (cnt) *cnt = 0;
  const int K = ctx->rplan_K, S = ctx->rplan_S;
  if (K <= 0) return QMCB_OK;
  const int c = (int)(ctx->moves_of_spin[sp] % K), q = c / S;
  if (c != std::min(q * S + S - 1, K - 1)) return QMCB_OK;
  const int *p0 = ctx->rpos.data() + ctx->roff[q], *p1 = ctx->rpos.data() + ctx->roff[q + 1];
  const int *b = std::lower_bound(p0, p1, (int)first), *e = std::lower_bound(b, p1, (int)last);
  if (e == b) return QMCB_OK;
  const int *d = ctx->d_rmap + (b - ctx->rpos.data());
  if (map) *map = d;
  if (cnt) *cnt = (int)(e - b);
  return DISPATCH(ctx, refresh(ctx, d, (int)(e - b), 1 << sp));
}

// All-electron moves (opt.all_electron_moves).  ae_args fills the record and
// accumulator pointers of `a`, whose mode (1 VMC, 2 DMC) and DMC fields the
// caller sets.  ae_begin takes each walker's record at its current
// positions from the resident state (which describes R between calls);
// ae_step makes one move of the walkers d_map[0, num): the proposal written
// over R (R saved in Rin), the ordinary all-electron evaluate and assembly
// at it, the Metropolis decision with restore and the observable sample in
// d_obs; ae_end evaluates every walker at R again, since a rejected
// walker's Slater and Jastrow state describes its proposal.
static void ae_args(qmcb_ctx *ctx, AeArgs &a, double tau) {
  a.N = ctx->N;
  a.R = ctx->R;
  a.Rsave = ctx->Rin;
  a.drift = ctx->ae_drift;
  a.logpsi = ctx->ae_logpsi;
  a.sgn = ctx->ae_sgn;
  a.mr = ctx->ae_mr;
  a.fwd = ctx->ae_fwd;
  a.cur_el = ctx->cur_el;
  a.prop = ctx->dout;
  a.tau = tau;
  a.seed = ctx->opt.seed;
  a.unif = ctx->unif;
  a.acc_el = ctx->acc_el; a.acc_el2 = ctx->acc_el2; a.acc_n = ctx->acc_n;
  a.acc_prop_el = ctx->acc_prop_el;
  a.walk.age = ctx->age;
  a.walk.old_walker = ctx->opt.old_walker_acceptance;
}

static qmcb_status ae_begin(qmcb_ctx *ctx, const int *d_map, int num, const AeArgs &a) {
  qmcb_status s = DISPATCH(ctx, assemble(ctx, d_map, num, ctx->dout, a.tau, 0));
  if (s != QMCB_OK) return s;
  AeArgs a0 = a;
  a0.mode = 0;
  TIMED("ae_accept", launch_ae_accept(a0, d_map, num, ctx->stream));
  return QMCB_OK;
}

// The evaluate at the proposal: its launches (basis, orbital GEMM chunks,
// cusp, inverse, traces, Jastrow, assembly) take no step-dependent argument,
// so they are captured once per (num, walker map, tau) and replayed.  A step
// whose key differs from the previous step's launches eagerly (a DMC
// population that changes every step would otherwise capture every step);
// so do kernel timing and graphs = 0 (the default: on this runtime a replay
// measured no faster than the launches it holds).
static qmcb_status ae_evaluate(qmcb_ctx *ctx, const int *d_map, int num, double tau) {
  auto eager = [&]() -> qmcb_status {
    qmcb_status s = DISPATCH(ctx, all_electron(ctx, d_map, num));
    if (s != QMCB_OK) return s;
    return DISPATCH(ctx, assemble(ctx, d_map, num, ctx->dout, tau, 0));
  };
  if (ctx->timing || !ctx->opt.graphs) return eager();
  if (!(ctx->ae_graph && ctx->ae_graph_num == num && ctx->ae_graph_map == d_map &&
        ctx->ae_graph_tau == tau)) {
    const bool repeat = ctx->ae_last_num == num && ctx->ae_last_map == d_map &&
                        ctx->ae_last_tau == tau;
    ctx->ae_last_num = num;
    ctx->ae_last_map = d_map;
    ctx->ae_last_tau = tau;
    if (!repeat) return eager();
    if (ctx->ae_graph) {
      (void)hipGraphExecDestroy(ctx->ae_graph);
      ctx->ae_graph = nullptr;
    }
    QCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const qmcb_status s = eager();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (s != QMCB_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (s != QMCB_OK) return s;
      QCHK(e);
    }
    const hipError_t ie = hipGraphInstantiate(&ctx->ae_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      ctx->ae_graph = nullptr;
      QCHK(ie);
    }
    ctx->ae_graph_num = num;
    ctx->ae_graph_map = d_map;
    ctx->ae_graph_tau = tau;
  }
  QCHK(hipGraphLaunch(ctx->ae_graph, ctx->stream));
  return QMCB_OK;
}

static qmcb_status ae_step(qmcb_ctx *ctx, const int *d_map, int num, const AeArgs &a) {
  TIMED("ae_propose", launch_ae_propose(a, d_map, num, ctx->stream));
  const qmcb_status s = ae_evaluate(ctx, d_map, num, a.tau);
  if (s != QMCB_OK) return s;
  TIMED("ae_accept", launch_ae_accept(a, d_map, num, ctx->stream));
  return QMCB_OK;
}

static qmcb_status ae_end(qmcb_ctx *ctx, const int *d_map, int num) {
  qmcb_status s = DISPATCH(ctx, all_electron(ctx, d_map, num));
  if (s != QMCB_OK) return s;
  return DISPATCH(ctx, assemble(ctx, d_map, num, ctx->dout, -1.0, 1));
}

qmcb_status qmcb_vmc_steps(qmcb_ctx *ctx, int32_t num, int64_t first_step, int32_t n_steps,
                           double tau, qmcb_vmc_stats *stats) {
  if (!ctx || num < 1 || num > ctx->Wc || n_steps < 0 || first_step < 0 || !(tau > 0))
    return QMCB_INVALID_ARG;
  if (!walk_drift_ok(ctx)) return QMCB_INVALID_ARG;
  for (int i = 0; i < num; i++)
    if (!ctx->initialized[i]) {
      ctx->err = "vmc slots must be initialised by an all-electron evaluate";
      return QMCB_BAD_STATE;
    }
  QCHK(hipSetDevice(ctx->device));
  QCHK(hipMemsetAsync(ctx->acc_el, 0, sizeof(double) * num, ctx->stream));
  QCHK(hipMemsetAsync(ctx->acc_el2, 0, sizeof(double) * num, ctx->stream));
  QCHK(hipMemsetAsync(ctx->acc_n, 0, sizeof(double) * num, ctx->stream));
  QCHK(hipMemsetAsync(ctx->acc_prop_el, 0, sizeof(double) * num, ctx->stream));
  const bool ae = ctx->opt.all_electron_moves != 0;
  {
    qmcb_status rs = series_reserve(ctx, first_step, n_steps, ae ? 1 : 0);
    if (rs != QMCB_OK) return rs;
    if (ctx->multi && !ae) {
      rs = ensure_staging(ctx);
      if (rs != QMCB_OK) return rs;
    }
  }
  const int K = ctx->opt.refresh_interval;
  qmcb_results none{};
  auto propose_args = [&](long step, int m) {
    ProposeArgs pa{};
    pa.N = ctx->N;
    pa.moved = m;
    pa.n_up = ctx->n[0];
    pa.R = ctx->R;
    pa.gradPR[0] = ctx->gradPR[0];
    pa.gradPR[1] = ctx->gradPR[1];
    pa.Ux = ctx->Ux;
    pa.cgrad = ctx->multi ? ctx->cgrad : nullptr;
    pa.tau = tau;
    walk_propose_args(ctx, pa);
    pa.seed = ctx->opt.seed;
    pa.step = step;
    pa.propR = ctx->propR;
    pa.gauss = ctx->gauss;
    pa.unif = ctx->unif;
    return pa;
  };
  auto propose = [&](long step, int m, const int *d_map, int cnt) -> qmcb_status {
    const ProposeArgs pa = propose_args(step, m);
    TIMED("propose", launch_propose(pa, d_map, cnt, ctx->stream));
    return QMCB_OK;
  };
  // One determinant: a step's sm_move also draws the next step's proposal
  // (propose_next), so only a call's first step launches propose_kernel;
  // walkers a re-inversion touched between the two steps are proposed again
  // from their refreshed gradients (the same draws), which leaves every bit as
  // the separate launch would.  Not with BF16 derivative storage, whose
  // shorter move launch measured slower with the draw's serial tail;
  // QMCB_SEPARATE_PROPOSE (read at creation) keeps one launch per step too.
  const bool fuse_propose = !ctx->multi && !ctx->separate_propose &&
                            ctx->opt.precision != QMCB_MIXED_BF16STORE;
  auto next_proposal = [&](MoveArgs &a, long step, int s, int count) {
    if (!fuse_propose || s + 1 >= count) return;
    a.propose_next = 1;
    a.next = propose_args(step + 1, (int)((step + 1) % ctx->N));
  };
  auto repropose = [&](long step, int s, int count, long first, const int *map, int cnt)
      -> qmcb_status {
    if (!fuse_propose || s + 1 >= count || cnt <= 0) return QMCB_OK;
    ProposeArgs pa = propose_args(step + 1, (int)((step + 1) % ctx->N));
    pa.row_from_walker = 1;
    pa.row_base = (int)first;
    TIMED("propose", launch_propose(pa, map, cnt, ctx->stream));
    return QMCB_OK;
  };
  // `count` steps from step0 with the walkers in G groups
  auto run = [&](long step0, int count, int G) -> qmcb_status {
    if (K > 0 && !(ctx->rplan_num == num && ctx->rplan_K == K)) {
      qmcb_status ps = refresh_plan(ctx, nullptr, num);
      if (ps != QMCB_OK) return ps;
      QCHK(hipStreamSynchronize(ctx->stream));  // rmap is pageable host memory
    }
    ctx->last_groups = G;
    if (G > 1) {
      qmcb_status gs = ensure_groups(ctx, G);
      if (gs != QMCB_OK) return gs;
      gs = groups_wait_main(ctx, G);
      if (gs != QMCB_OK) return gs;
    }
    for (int s = 0; s < count; s++) {
      const long step = step0 + s;
      const int m = (int)(step % ctx->N);
      const int sp = m < ctx->n[0] ? 0 : 1;
      qmcb_status st = QMCB_OK;
      if (G > 1) {
        for (int g = 0; g < G; g++) {
          const long first = (long)num * g / G, cnt = (long)num * (g + 1) / G - first;
          const int *d_map = ctx->d_iota + first;
          GroupView view(ctx, g, first, sp);
          // stagger: group g's front half (proposal, Jastrow, basis, GEMM) of a
          // step starts after group g-1's front half of the same step, so each
          // group trails the one before it and its front overlaps that group's
          // Sherman-Morrison pass; group 0 runs free
          if (g > 0) QCHK(hipStreamWaitEvent(ctx->stream, ctx->gfront[g - 1], 0));
          if (s == 0 || !fuse_propose) st = propose(step, m, d_map, (int)cnt);
          if (st == QMCB_OK)
            st = DISPATCH(ctx, move_front(ctx, d_map, (int)cnt, m, !jastrow_fused(ctx, m)));
          if (st == QMCB_OK) QCHK(hipEventRecord(ctx->gfront[g], ctx->stream));
          if (st == QMCB_OK) {
            MoveArgs a = DISPATCH(ctx, move_args(ctx, m, 1, none));
            a.tau = tau;
            a.jastrow_fused = jastrow_fused(ctx, m);
            if (m == 0) a.obs = ctx->d_obs;
            next_proposal(a, step, s, count);
            st = DISPATCH(ctx, sm(ctx, a, d_map, (int)cnt));
          }
          const int *rmap = nullptr;
          int rcnt = 0;
          if (st == QMCB_OK) st = staggered_refresh(ctx, first, first + cnt, sp, &rmap, &rcnt);
          if (st == QMCB_OK) st = repropose(step, s, count, first, rmap, rcnt);
          if (st != QMCB_OK) return st;
        }
        ctx->moves_of_spin[sp]++;
        if (m == 0) {
          // the sweep's observable row sums every walker in one fixed order
          st = main_wait_groups(ctx, G);
          if (st != QMCB_OK) return st;
          TIMED("series", launch_series_reduce(ctx->d_obs, num,
                                               ctx->d_series + 4L * ctx->series_len++,
                                               ctx->stream));
          st = groups_wait_main(ctx, G);
          if (st != QMCB_OK) return st;
        }
        continue;
      }
      if (s == 0 || !fuse_propose) st = propose(step, m, nullptr, num);
      if (st != QMCB_OK) return st;
      st = DISPATCH(ctx, move_front(ctx, nullptr, num, m, !jastrow_fused(ctx, m)));
      if (st != QMCB_OK) return st;
      if (ctx->multi) {
        // several determinants: stage every determinant, combine and accept on
        // the device, then commit the accepted rows
        st = DISPATCH(ctx, mdet_stage(ctx, nullptr, num, m, 1, none, tau,
                                      m == 0 ? ctx->d_obs : nullptr));
        if (st != QMCB_OK) return st;
        st = DISPATCH(ctx, mdet_commit(ctx, nullptr, ctx->d_mask, num, m));
      } else {
        MoveArgs a = DISPATCH(ctx, move_args(ctx, m, 1, none));
        a.tau = tau;
        a.jastrow_fused = jastrow_fused(ctx, m);
        if (m == 0) a.obs = ctx->d_obs;
        next_proposal(a, step, s, count);
        st = DISPATCH(ctx, sm(ctx, a, nullptr, num));
      }
      if (st != QMCB_OK) return st;
      if (m == 0)
        TIMED("series", launch_series_reduce(ctx->d_obs, num,
                                             ctx->d_series + 4L * ctx->series_len++, ctx->stream));
      const int *rmap = nullptr;
      int rcnt = 0;
      st = staggered_refresh(ctx, 0, num, sp, &rmap, &rcnt);
      if (st == QMCB_OK && !ctx->multi) st = repropose(step, s, count, 0, rmap, rcnt);
      if (st != QMCB_OK) return st;
      ctx->moves_of_spin[sp]++;
    }
    if (G > 1) {
      qmcb_status gs = main_wait_groups(ctx, G);
      if (gs != QMCB_OK) return gs;
    }
    return QMCB_OK;
  };
  // walker groups: single determinant only; at least 256 walkers per group
  auto clamp_groups = [&](int g) {
    g = ctx->multi ? 1 : std::max(1, std::min<int>(g, qmcb_ctx::kMaxGroups));
    while (g > 1 && num / g < 256) g--;
    return g;
  };
  int done = 0;
  int G = clamp_groups(ctx->opt.walker_groups);
  if (ae) {
    // all-electron moves: one group, one observable row per step
    qmcb_status as = ensure_groups(ctx, 1);
    if (as != QMCB_OK) return as;
    AeArgs a{};
    ae_args(ctx, a, tau);
    a.mode = 1;
    a.obs = ctx->d_obs;
    as = ae_begin(ctx, ctx->d_iota, num, a);
    for (int32_t i = 0; as == QMCB_OK && i < n_steps; i++) {
      a.step = first_step + i;
      as = ae_step(ctx, ctx->d_iota, num, a);
      if (as == QMCB_OK) {
        TIMED("series", launch_series_reduce(ctx->d_obs, num,
                                             ctx->d_series + 4L * ctx->series_len++, ctx->stream));
      }
    }
    if (as == QMCB_OK) as = ae_end(ctx, ctx->d_iota, num);
    if (as != QMCB_OK) return as;
    ctx->last_groups = 1;
    done = n_steps;
  } else if (ctx->opt.walker_groups < 0) {
    // chosen once per context (again when the batch changes by more than 2x)
    // by timing windows of kTune steps of each candidate count on the real
    // walkers, in this call and, if it is too short, the next ones (one group
    // until the choice is made).  One untimed round of windows comes first
    // (it absorbs the GEMM tuning of each candidate's group batch and the
    // first launches), then two timed rounds, interleaved, keeping each
    // candidate's faster window.  The count changes no result bit, so the
    // windows are ordinary steps.  A window is 16 steps, or S steps when the
    // refresh phase S (one staggered re-inversion launch every S moves of a
    // spin) is 17-32: windows whose length is a multiple of S hold equally
    // many launches, others may differ by one.
    const int S_ref = ctx->opt.refresh_interval > 0
                          ? refresh_phase(ctx->opt.refresh_interval, num) : 1;
    const int kTune = S_ref > 16 && S_ref <= 32 ? S_ref : 16;
    auto valid = [&]() {
      return ctx->auto_groups > 0 && num <= 2 * ctx->auto_groups_num &&
             2 * num >= ctx->auto_groups_num;
    };
    std::vector<int> cand;
    for (int g = 1; g <= 3; g++) {
      const int c = clamp_groups(g);
      if (std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
    }
    const int nc = (int)cand.size(), windows = 3 * nc;
    if (!valid() && nc == 1) {
      ctx->auto_groups = cand[0];
      ctx->auto_groups_num = num;
    } else if (!valid()) {
      if (ctx->tune_num != num) {
        ctx->tune_num = num;
        ctx->tune_q = 0;
      }
      qmcb_status ts = ensure_groups(ctx, cand.back());
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (ts == QMCB_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        ts = QMCB_DEVICE_ERROR;
      while (ts == QMCB_OK && ctx->tune_q < windows && done + kTune <= n_steps) {
        const int c = ctx->tune_q % nc, round = ctx->tune_q / nc;
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) { ts = QMCB_DEVICE_ERROR; break; }
        ts = run(first_step + done, kTune, cand[c]);
        done += kTune;
        float ms = 0.0f;
        if (ts == QMCB_OK && (hipEventRecord(e1, ctx->stream) != hipSuccess ||
                              hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
          ts = QMCB_DEVICE_ERROR;
        if (ts != QMCB_OK) break;
        if (round == 1) ctx->tune_ms[c] = ms;
        if (round == 2) ctx->tune_ms[c] = std::min(ctx->tune_ms[c], ms);
        ctx->tune_q++;
      }
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (ts != QMCB_OK) {
        if (ts == QMCB_DEVICE_ERROR && ctx->err.empty()) ctx->err = "walker-group tuning failed";
        return ts;
      }
      if (ctx->tune_q == windows) {  // every candidate timed twice: keep the fastest
        int best = 0;
        for (int q = 1; q < nc; q++)
          if (ctx->tune_ms[q] < ctx->tune_ms[best]) best = q;
        ctx->auto_groups = cand[best];
        ctx->auto_groups_num = num;
        ctx->tune_num = 0;
      }
    }
    G = valid() ? clamp_groups(ctx->auto_groups) : 1;
  }
  if (done < n_steps) {
    qmcb_status rs = run(first_step + done, n_steps - done, G);
    if (rs != QMCB_OK) return rs;
  }
  if (stats) {
    std::vector<double> a(num), b(num), c(num), d(num);
    QCHK(hipMemcpyAsync(a.data(), ctx->acc_el, sizeof(double) * num, hipMemcpyDeviceToHost, ctx->stream));
    QCHK(hipMemcpyAsync(b.data(), ctx->acc_el2, sizeof(double) * num, hipMemcpyDeviceToHost, ctx->stream));
    QCHK(hipMemcpyAsync(c.data(), ctx->acc_n, sizeof(double) * num, hipMemcpyDeviceToHost, ctx->stream));
    QCHK(hipMemcpyAsync(d.data(), ctx->acc_prop_el, sizeof(double) * num, hipMemcpyDeviceToHost, ctx->stream));
    QCHK(hipStreamSynchronize(ctx->stream));
    stats->sum_e_local = stats->sum_e_local_sq = stats->sum_accept = stats->sum_prop_e_local = 0;
    for (int i = 0; i < num; i++) {
      stats->sum_e_local += a[i];
      stats->sum_e_local_sq += b[i];
      stats->sum_accept += c[i];
      stats->sum_prop_e_local += d[i];
    }
    stats->n_samples = (int64_t)num * n_steps;
  }
  tm_collect(ctx);
  return QMCB_OK;
}

static CopyArgs walker_copy_args(qmcb_ctx *ctx) {
  CopyArgs c{};
  auto add = [&](void *p, long bytes) {
    if (bytes <= 0) return;
    c.base[c.narr] = (char *)p;
    c.bytes[c.narr] = bytes;
    c.narr++;
  };
  const long N = ctx->N, D = ctx->ndet;  // D per-determinant records per walker
  add(ctx->R, N * 3 * 8);
  for (int sp = 0; sp < 2; sp++) {
    const long n = ctx->n[sp], na = ctx->nact[sp];
    if (n == 0) continue;
    add(ctx->inv[sp], D * n * n * 8);
    add(ctx->dval[sp], na * n * 8);
    add(ctx->dmat[sp], kDer * na * n * (long)ctx->ts);
    add(ctx->ubuf[sp], kDelay * n * 8);
    add(ctx->zbuf[sp], kDelay * n * 8);
    add(ctx->dcount[sp], D * 4);
    add(ctx->logdet[sp], D * 8);
    add(ctx->sgn[sp], D * 8);
    add(ctx->gradPR[sp], D * n * 3 * 8);
    add(ctx->lapPR[sp], D * 8);
    add(ctx->singular[sp], D * 4);
  }
  if (ctx->multi) add(ctx->cgrad, N * 3 * 8);
  add(ctx->U, 8);
  add(ctx->Ux, N * 3 * 8);
  add(ctx->Uxx, 8);
  add(ctx->Vee, 8);
  add(ctx->Ven, 8);
  add(ctx->cur_el, 8);
  add(ctx->weight, 8);
  add(ctx->age, 4);
  if (ctx->ae_drift) {  // the all-electron record
    add(ctx->ae_drift, N * 3 * 8);
    add(ctx->ae_logpsi, 8);
    add(ctx->ae_sgn, 8);
    add(ctx->ae_mr, 8);
  }
  return c;
}

qmcb_status qmcb_dmc_init(qmcb_ctx *ctx, int32_t num) {
  if (!ctx || num < 1 || num > ctx->Wc) return QMCB_INVALID_ARG;
  for (int i = 0; i < num; i++)
    if (!ctx->initialized[i]) {
      ctx->err = "dmc slots must be initialised by an all-electron evaluate";
      return QMCB_BAD_STATE;
    }
  QCHK(hipSetDevice(ctx->device));
  ctx->active.resize(num);
  for (int i = 0; i < num; i++) ctx->active[i] = i;
  ctx->free_slots.clear();
  for (int i = ctx->Wc - 1; i >= num; i--) ctx->free_slots.push_back(i);
  std::vector<double> ones(ctx->Wc, 1.0);
  QCHK(hipMemcpyAsync(ctx->weight, ones.data(), sizeof(double) * ctx->Wc, hipMemcpyHostToDevice,
                      ctx->stream));
  QCHK(hipMemcpyAsync(ctx->d_active, ctx->active.data(), sizeof(int) * num,
                      hipMemcpyHostToDevice, ctx->stream));
  QCHK(hipStreamSynchronize(ctx->stream));
  ctx->dmc_ready = true;
  ctx->have_e_trial = false;
  ctx->est_we = ctx->est_w = 0.0;
  ctx->est_n = 0;
  return QMCB_OK;
}

qmcb_status qmcb_dmc_steps(qmcb_ctx *ctx, int64_t first_step, int32_t n_steps,
                           const qmcb_dmc_params *prm, qmcb_dmc_stats *stats) {
  if (!ctx || !prm || n_steps < 0 || first_step < 0 || !(prm->tau > 0)) return QMCB_INVALID_ARG;
  if (!ctx->dmc_ready) {
    ctx->err = "qmcb_dmc_init first";
    return QMCB_BAD_STATE;
  }
  if (!walk_drift_ok(ctx)) return QMCB_INVALID_ARG;
  if (prm->energy_modification != 0 && prm->energy_modification != 1) {
    ctx->err = "energy_modification: 0 (none) or 1 (umrigar93)";
    return QMCB_INVALID_ARG;
  }
  QCHK(hipSetDevice(ctx->device));
  const int W = ctx->Wc;
  QCHK(hipMemsetAsync(ctx->acc_el, 0, sizeof(double) * W, ctx->stream));
  QCHK(hipMemsetAsync(ctx->acc_el2, 0, sizeof(double) * W, ctx->stream));
  QCHK(hipMemsetAsync(ctx->acc_n, 0, sizeof(double) * W, ctx->stream));
  QCHK(hipMemsetAsync(ctx->acc_prop_el, 0, sizeof(double) * W, ctx->stream));
  QCHK(hipMemsetAsync(ctx->acc_w, 0, sizeof(double) * W, ctx->stream));
  if (!ctx->have_e_trial) {
    // energy_estimated and energy_estimated_original start at the trial
    // energy (QMCFlags.cpp:918-919)
    ctx->e_trial = prm->e_trial;
    ctx->e_est = prm->e_trial;
    ctx->e_orig = prm->e_trial;
    ctx->have_e_trial = true;
  }
  const bool ae = ctx->opt.all_electron_moves != 0;
  {
    qmcb_status rs = series_reserve(ctx, first_step, n_steps, ae ? 1 : 0);
    if (rs != QMCB_OK) return rs;
    if (ctx->multi && !ae) {
      rs = ensure_staging(ctx);
      if (rs != QMCB_OK) return rs;
    }
  }
  const int target = prm->target_walkers > 0 ? prm->target_walkers : (int)ctx->active.size();
  const int K = ctx->opt.refresh_interval;
  // the fusion draws are counter-based, keyed by (seed, step, fusion index),
  // so a run split over several calls draws the same numbers as one call
  auto fuse_uniform = [&](long step, int k) {
    const u4 ctr = {(uint32_t)step, (uint32_t)((uint64_t)step >> 32), 0xFFFFFFFFu, (uint32_t)k};
    const u4 r = philox4x32(ctr, (uint32_t)ctx->opt.seed, (uint32_t)(ctx->opt.seed >> 32));
    return u01(r.x, r.y);
  };
  qmcb_dmc_stats st{};
  std::vector<double> wts(W), row(4);
  const CopyArgs cargs = walker_copy_args(ctx);
  long moves = 0;
  // walker groups (single determinant): as in qmcb_vmc_steps, over the active
  // list; walker_groups < 0 uses the count a qmcb_vmc_steps call chose, one
  // group before one was chosen
  const int want = ctx->opt.walker_groups >= 0 ? ctx->opt.walker_groups : ctx->auto_groups;
  const int G = (ctx->multi || ae) ? 1 : std::max(1, std::min<int>(want, qmcb_ctx::kMaxGroups));
  if (G > 1) {
    qmcb_status gs = ensure_groups(ctx, G);
    if (gs != QMCB_OK) return gs;
  }
  if (K > 0 && !ae) {
    qmcb_status ps = refresh_plan(ctx, ctx->active.data(), (int)ctx->active.size());
    if (ps != QMCB_OK) return ps;
    QCHK(hipStreamSynchronize(ctx->stream));  // rmap is pageable host memory
  }
  if (G > 1) {
    qmcb_status gs = groups_wait_main(ctx, G);
    if (gs != QMCB_OK) return gs;
  }
  WalkCtl walk{};
  walk.age = ctx->age;
  walk.old_walker = ctx->opt.old_walker_acceptance;
  walk.rel_cutoff = prm->rel_cutoff;
  walk.e_orig = ctx->e_orig;
  walk.branch_threshold = prm->branch_threshold;
  // all-electron moves: every step moves every electron, reweights (from
  // reweight_after on; QMCWalker.cpp:226-228) and branches
  AeArgs ae_a{};
  if (ae) {
    ae_args(ctx, ae_a, prm->tau);
    ae_a.mode = 2;
    ae_a.weight = ctx->weight;
    ae_a.acc_w = ctx->acc_w;
    ae_a.energy_mod = prm->energy_modification;
    ae_a.tau_eff = prm->tau;
    ae_a.obs = ctx->d_obs;
    if (!ctx->active.empty()) {
      qmcb_status as = ae_begin(ctx, ctx->d_active, (int)ctx->active.size(), ae_a);
      if (as != QMCB_OK) return as;
    }
  }
  for (int s = 0; s < n_steps; s++) {
    const long step = first_step + s;
    const int m = ae ? 0 : (int)(step % ctx->N);
    const int pop = (int)ctx->active.size();
    if (pop == 0) {
      ctx->err = "DMC population died out";
      return QMCB_BAD_STATE;
    }
    const int sp = m < ctx->n[0] ? 0 : 1;
    // the reference's iteration = step - equilibration_steps; the stricter
    // deletion rules hold while iteration < -100 (QMCWalker.cpp:1207)
    walk.equilibrating = step < (long)prm->equilibration_steps - 100 ? 1 : 0;
    // one group's step over the active-list rows [first, first + cnt) (g < 0:
    // no groups)
    auto group_step = [&](long first, int cnt, int g) -> qmcb_status {
      const int *d_map = ctx->d_active + first;
      ProposeArgs pa{};
      pa.N = ctx->N;
      pa.moved = m;
      pa.n_up = ctx->n[0];
      pa.R = ctx->R;
      pa.gradPR[0] = ctx->gradPR[0];
      pa.gradPR[1] = ctx->gradPR[1];
      pa.Ux = ctx->Ux;
      pa.cgrad = ctx->multi ? ctx->cgrad : nullptr;
      pa.tau = prm->tau;
      walk_propose_args(ctx, pa);
      pa.seed = ctx->opt.seed;
      pa.step = step;
      pa.propR = ctx->propR;
      pa.gauss = ctx->gauss;
      pa.unif = ctx->unif;
      TIMED("propose", launch_propose(pa, d_map, cnt, ctx->stream));
      qmcb_status gst = DISPATCH(ctx, move_front(ctx, d_map, cnt, m, !jastrow_fused(ctx, m)));
      if (gst != QMCB_OK) return gst;
      if (g >= 0) QCHK(hipEventRecord(ctx->gfront[g], ctx->stream));
      qmcb_results none{};
      const int reweight = (m == 0 && step >= prm->reweight_after) ? 1 : 0;
      if (ctx->multi) {
        // several determinants: stage every determinant, combine, accept
        // and reweight on the device, commit the accepted rows
        MdetArgs ma = mdet_args(ctx, m, 2, none);
        ma.tau = prm->tau;
        ma.e_trial = ctx->e_trial;
        ma.e_est = ctx->e_est;
        ma.energy_mod = prm->energy_modification;
        ma.tau_eff = prm->tau;
        ma.reweight = reweight;
        ma.walk = walk;
        ma.obs = m == 0 ? ctx->d_obs : nullptr;
        gst = DISPATCH(ctx, mdet_stage_args(ctx, ma, d_map, cnt));
        if (gst != QMCB_OK) return gst;
        gst = DISPATCH(ctx, mdet_commit(ctx, d_map, ctx->d_mask, cnt, m));
      } else {
        MoveArgs a = DISPATCH(ctx, move_args(ctx, m, 2, none));
        a.tau = prm->tau;
        a.jastrow_fused = jastrow_fused(ctx, m);
        a.e_trial = ctx->e_trial;
        a.e_est = ctx->e_est;
        a.energy_mod = prm->energy_modification;
        a.tau_eff = prm->tau;
        a.reweight = reweight;
        a.walk = walk;
        if (m == 0) a.obs = ctx->d_obs;
        gst = DISPATCH(ctx, sm(ctx, a, d_map, cnt));
      }
      if (gst != QMCB_OK) return gst;
      return staggered_refresh(ctx, first, first + cnt, sp);
    };
    const int Geff = (G > 1 && pop / G >= 256) ? G : 1;
    ctx->last_groups = Geff;
    qmcb_status sst = QMCB_OK;
    if (ae) {
      ae_a.step = step;
      ae_a.reweight = step >= prm->reweight_after ? 1 : 0;
      ae_a.e_trial = ctx->e_trial;
      ae_a.e_est = ctx->e_est;
      ae_a.walk = walk;
      sst = ae_step(ctx, ctx->d_active, pop, ae_a);
      if (sst != QMCB_OK) return sst;
    } else if (Geff > 1) {
      for (int g = 0; g < G; g++) {
        const long first = (long)pop * g / G, cnt = (long)pop * (g + 1) / G - first;
        GroupView view(ctx, g, first, sp);
        if (g > 0) QCHK(hipStreamWaitEvent(ctx->stream, ctx->gfront[g - 1], 0));
        sst = group_step(first, (int)cnt, g);
        if (sst != QMCB_OK) return sst;
      }
      sst = main_wait_groups(ctx, G);
      if (sst != QMCB_OK) return sst;
    } else {
      sst = group_step(0, pop, -1);
      if (sst != QMCB_OK) return sst;
    }
    if (!ae) ctx->moves_of_spin[sp]++;
    if (m == 0)
      TIMED("series", launch_series_reduce(ctx->d_obs, pop, ctx->d_series + 4L * ctx->series_len++,
                                           ctx->stream));
    moves += pop;
    if (m != 0) {
      if (G > 1) {
        // the groups of the next step start after everything queued so far
        sst = groups_wait_main(ctx, G);
        if (sst != QMCB_OK) return sst;
      }
      continue;
    }
    // ---- branching once per electron sweep (after the electron-0 move,
    // QMCRun.cpp:1195-1206): the weights and this sweep's observable row
    QCHK(hipMemcpyAsync(wts.data(), ctx->weight, sizeof(double) * W, hipMemcpyDeviceToHost,
                        ctx->stream));
    QCHK(hipMemcpyAsync(row.data(), ctx->d_series + 4L * (ctx->series_len - 1),
                        sizeof(double) * 4, hipMemcpyDeviceToHost, ctx->stream));
    QCHK(hipStreamSynchronize(ctx->stream));
    // nonunitWeightBranching (QMCRun.cpp:739-811): split above the branching
    // threshold, delete below 0.01, fuse pairs below the fusion threshold in
    // list order
    std::vector<int> next, pairs, dead;
    next.reserve(pop * 2);
    int fuse_pending = -1, n_fuse = 0;
    for (int idx = 0; idx < pop; idx++) {
      const int w = ctx->active[idx];
      const double wt = wts[w];
      if (wt == 0.0) st.n_deleted++;
      if (wt > prm->branch_threshold) {
        if (ctx->free_slots.empty()) {
          // no slot left: the walker keeps its weight unsplit (counted)
          st.n_split_skipped++;
          next.push_back(w);
          continue;
        }
        wts[w] = wt / 2.0;
        const int d = ctx->free_slots.back();
        ctx->free_slots.pop_back();
        pairs.push_back(w);
        pairs.push_back(d);
        ctx->initialized[d] = 1;  // the copy carries a complete resident state
        wts[d] = wt / 2.0;
        next.push_back(w);
        next.push_back(d);
        st.n_split++;
      } else if (wt < 0.01) {
        dead.push_back(w);
        st.n_killed++;
      } else if (wt < prm->fusion_threshold) {
        if (fuse_pending < 0) {
          fuse_pending = w;
        } else {
          const double w1 = wts[fuse_pending], w3 = w1 + wt;
          if (fuse_uniform(step, n_fuse++) < w1 / w3) {
            wts[fuse_pending] = w3;
            next.push_back(fuse_pending);
            dead.push_back(w);
          } else {
            wts[w] = w3;
            next.push_back(w);
            dead.push_back(fuse_pending);
          }
          fuse_pending = -1;
          st.n_fused++;
        }
      } else {
        next.push_back(w);
      }
    }
    if (fuse_pending >= 0) next.push_back(fuse_pending);
    for (int w : dead) ctx->free_slots.push_back(w);
    ctx->active.swap(next);
    QCHK(hipMemcpyAsync(ctx->weight, wts.data(), sizeof(double) * W, hipMemcpyHostToDevice,
                        ctx->stream));
    if (!pairs.empty()) {
      QCHK(hipMemcpyAsync(ctx->d_pairs, pairs.data(), sizeof(int) * pairs.size(),
                          hipMemcpyHostToDevice, ctx->stream));
      TIMED("copy_walkers", launch_copy_walkers(cargs, ctx->d_pairs, (int)pairs.size() / 2,
                                                ctx->stream));
    }
    if (!ctx->active.empty())
      QCHK(hipMemcpyAsync(ctx->d_active, ctx->active.data(), sizeof(int) * ctx->active.size(),
                          hipMemcpyHostToDevice, ctx->stream));
    if (K > 0 && !ae) {
      qmcb_status ps = refresh_plan(ctx, ctx->active.data(), (int)ctx->active.size());
      if (ps != QMCB_OK) return ps;
    }
    QCHK(hipStreamSynchronize(ctx->stream));  // host vectors above must outlive the copies
    if (G > 1) {
      sst = groups_wait_main(ctx, G);
      if (sst != QMCB_OK) return sst;
    }
    // energy estimate and trial energy (QMCManager.cpp:2247-2310), once per
    // sweep: the sweep's sample row[0] / row[1] with weight row[1]
    double wtot = 0.0;
    for (int w : ctx->active) wtot += wts[w];
    const double ratio = wtot / target;
    const double lr = ratio > 1e-250 ? std::log(ratio) : 0.0;
    if (step < (long)prm->equilibration_steps) {
      // equilibration: E_est stays the input estimate, E_T follows the weight
      // change (UNR93 eq. 11)
      ctx->e_est = ctx->e_orig;
      ctx->e_trial = ctx->e_est - lr / prm->tau;
    } else {
      if (row[1] > 0.0) {
        ctx->est_we += row[0];
        ctx->est_w += row[1];
        ctx->est_n++;
      }
      ctx->e_est = (ctx->est_n >= 100 && ctx->est_w > 0.0) ? ctx->est_we / ctx->est_w : ctx->e_orig;
      ctx->e_trial = ctx->e_est - prm->pop_control * lr;
    }
  }
  if (ae && !ctx->active.empty()) {
    qmcb_status as = ae_end(ctx, ctx->d_active, (int)ctx->active.size());
    if (as != QMCB_OK) return as;
  }
  // statistics
  {
    std::vector<double> a(W), b2(W), c(W), d(W), wv(W);
    QCHK(hipMemcpyAsync(a.data(), ctx->acc_el, sizeof(double) * W, hipMemcpyDeviceToHost, ctx->stream));
    QCHK(hipMemcpyAsync(b2.data(), ctx->acc_el2, sizeof(double) * W, hipMemcpyDeviceToHost, ctx->stream));
    QCHK(hipMemcpyAsync(c.data(), ctx->acc_n, sizeof(double) * W, hipMemcpyDeviceToHost, ctx->stream));
    QCHK(hipMemcpyAsync(d.data(), ctx->acc_w, sizeof(double) * W, hipMemcpyDeviceToHost, ctx->stream));
    QCHK(hipMemcpyAsync(wv.data(), ctx->weight, sizeof(double) * W, hipMemcpyDeviceToHost, ctx->stream));
    QCHK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < W; i++) {
      st.sum_w_e_local += a[i];
      st.sum_w_e_local_sq += b2[i];
      st.sum_accept += c[i];
      st.sum_weights += d[i];
    }
    for (int w : ctx->active) st.total_weight += wv[w];
    st.n_moves = moves;
    st.population = (int)ctx->active.size();
    st.e_trial = ctx->e_trial;
    st.e_estimated = ctx->e_est;
  }
  if (stats) *stats = st;
  tm_collect(ctx);
  return QMCB_OK;
}

qmcb_status qmcb_dmc_population(qmcb_ctx *ctx, int32_t cap, int32_t *slots, double *weights,
                                int32_t *count) {
  if (!ctx || cap < 0) return QMCB_INVALID_ARG;
  QCHK(hipSetDevice(ctx->device));
  const int pop = (int)ctx->active.size();
  if (count) *count = pop;
  std::vector<double> wv(ctx->Wc);
  QCHK(hipMemcpyAsync(wv.data(), ctx->weight, sizeof(double) * ctx->Wc, hipMemcpyDeviceToHost,
                      ctx->stream));
  QCHK(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < pop && i < cap; i++) {
    if (slots) slots[i] = ctx->active[i];
    if (weights) weights[i] = wv[ctx->active[i]];
  }
  return QMCB_OK;
}

qmcb_status qmcb_sample_series(qmcb_ctx *ctx, int32_t cap, double *rows, int32_t *count) {
  if (!ctx || cap < 0 || (cap > 0 && !rows)) return QMCB_INVALID_ARG;
  QCHK(hipSetDevice(ctx->device));
  if (count) *count = ctx->series_len;
  const int n = std::min(cap, ctx->series_len);
  if (n > 0) {
    QCHK(hipMemcpyAsync(rows, ctx->d_series, sizeof(double) * 4 * n, hipMemcpyDeviceToHost,
                        ctx->stream));
  }
  QCHK(hipStreamSynchronize(ctx->stream));
  return QMCB_OK;
}

qmcb_status qmcb_sync(qmcb_ctx *ctx) {
  if (!ctx) return QMCB_INVALID_ARG;
  QCHK(hipSetDevice(ctx->device));
  QCHK(hipStreamSynchronize(ctx->stream));
  tm_collect(ctx);
  return QMCB_OK;
}

qmcb_status qmcb_set_walker_groups(qmcb_ctx *ctx, int32_t groups) {
  if (!ctx || groups < -1 || groups > qmcb_ctx::kMaxGroups) return QMCB_INVALID_ARG;
  ctx->opt.walker_groups = groups;
  return QMCB_OK;
}

qmcb_status qmcb_get_walker_groups(const qmcb_ctx *ctx, int32_t *groups) {
  if (!ctx || !groups) return QMCB_INVALID_ARG;
  *groups = ctx->last_groups;
  return QMCB_OK;
}

qmcb_status qmcb_device_bytes(const qmcb_ctx *ctx, int64_t *bytes) {
  if (!ctx || !bytes) return QMCB_INVALID_ARG;
  *bytes = (int64_t)ctx->dev_bytes;
  return QMCB_OK;
}

qmcb_status qmcb_plan_device_bytes(const qmcb_system *sys, const qmcb_options *opt,
                                   int64_t *bytes) {
  if (!sys || !opt || !bytes || opt->max_walkers < 1 || sys->n_basis < 1 || sys->n_nuc < 1)
    return QMCB_INVALID_ARG;
  *bytes = (int64_t)plan_device_bytes(sys, opt);
  return QMCB_OK;
}

void qmcb_enable_kernel_timing(qmcb_ctx *ctx, int32_t on) {
  if (!ctx) return;
  ctx->timing = on != 0;
  ctx->tm.names.clear();
  ctx->tm.index.clear();
  ctx->tm.ms.clear();
  ctx->tm.launches.clear();
  ctx->tm.marks.clear();
  ctx->tm.used = 0;
}

int32_t qmcb_kernel_times(qmcb_ctx *ctx, int32_t cap, const char **names, double *ms,
                          int64_t *launches) {
  if (!ctx) return 0;
  tm_collect(ctx);
  const int n = (int)ctx->tm.names.size();
  for (int i = 0; i < n && i < cap; i++) {
    if (names) names[i] = ctx->tm.names[i].c_str();
    if (ms) ms[i] = ctx->tm.ms[i];
    if (launches) launches[i] = ctx->tm.launches[i];
  }
  return n;
}

}  // extern "C"
