// slater.hip -- kernel (c): Slater inverse, determinant, Sherman-Morrison
// single-electron updates and derivative-ratio traces, all FP64.
//
// Reference: QMCSlater::update_Ds / calculate_DerivativeRatios
// (QMCSlater.cpp:390-764), Array2D::determinant_and_inverse (Array2D.h:
// 1521-1548, LU in ludcmp/lubksb :1204-1421) and inverseUpdateOneRow
// (Array2D.h:1566-1610).
//
// Storage (per walker, per spin): inv(o,e) row-major, dmat(k,o,e) =
// d_k D(e,o) (transposed), so the thread that owns electron column e of
// A^-1 also streams column e of every derivative matrix with coalesced loads.
#include "qmcb_kernels.hpp"

namespace qmcb {

template <typename TS>
__device__ __forceinline__ double ld(const TS *p) { return (double)*p; }

// ---------------------------------------------------------------------------
// Batched inverse by Gauss-Jordan elimination with partial (row) pivoting,
// register resident: 256 threads as a 16x16 grid, thread (tr, tc) owns the
// cyclic block a[tr + 16 i][tc + 16 j], i, j < R (n <= 16R).  Per pivot step
// only the pivot row and pivot column pass through LDS.  ln|det| and sign
// are accumulated from the pivots and row swaps; the LU of the reference
// (Crout, implicit scaling, long double sums) gives the same inverse and
// determinant up to rounding.  Singular (an all-zero pivot column) sets the
// walker's singular flag, as calcOK = false does (QMCSlater.cpp:675-683).
template <typename TS, int R>
__global__ void __launch_bounds__(256)
lu_inverse_kernel(SlaterArgs A, const int *walkers, int num) {
  const int b = blockIdx.x;
  if (b >= num) return;
  const int w = walkers ? walkers[b] : b;
  const int n = A.n;
  const int tid = threadIdx.x;
  const int tr = tid & 15, tc = tid >> 4;  // consecutive lanes -> consecutive rows
  __shared__ double prow[16 * R], krow[16 * R], fcol[16 * R];
  __shared__ double cmax[16];
  __shared__ int cidx[16];
  __shared__ int ipiv[16 * R];
  __shared__ int dest[16 * R];
  __shared__ int cur[16 * R];
  __shared__ int s_piv, s_sing;

  const TS *dv = (const TS *)A.dmat + (long)w * kComp * n * n;  // component 0
  double a[R][R];
#pragma unroll
  for (int i = 0; i < R; i++)
#pragma unroll
    for (int j = 0; j < R; j++) {
      const int gi = tr + 16 * i, gj = tc + 16 * j;  // gi: electron, gj: orbital
      double v = (gi == gj) ? 1.0 : 0.0;
      if (gi < n && gj < n) v = ld(dv + (long)gj * n + gi);
      a[i][j] = v;
    }
  double logdet = 0.0, sgn = 1.0;
  if (tid == 0) s_sing = 0;
  for (int k = 0; k < n; k++) {
    // ---- pivot search in column k, rows >= k
    const int kj = k >> 4, kt = k & 15;
    if (tc == kt) {
      double best = -1.0;
      int bi = -1;
#pragma unroll
      for (int i = 0; i < R; i++) {
        const int gi = tr + 16 * i;
        double aik = 0.0;
#pragma unroll
        for (int j = 0; j < R; j++) aik = (j == kj) ? a[i][j] : aik;
        if (gi >= k && gi < n) {
          const double v = fabs(aik);
          if (v > best) { best = v; bi = gi; }
        }
      }
      cmax[tr] = best;
      cidx[tr] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double best = -1.0;
      int bi = -1;
      for (int i = 0; i < 16; i++)
        if (cidx[i] >= 0 && (cmax[i] > best || (cmax[i] == best && cidx[i] < bi))) {
          best = cmax[i]; bi = cidx[i];
        }
      s_piv = bi;
      ipiv[k] = bi;
      if (!(best > 0.0)) s_sing = 1;
    }
    __syncthreads();
    if (s_sing) break;
    const int p = s_piv;
    // ---- rows p and k to LDS
    const int pi = p >> 4;
    if (tr == (p & 15)) {
#pragma unroll
      for (int i = 0; i < R; i++)
        if (i == pi) {
#pragma unroll
          for (int j = 0; j < R; j++) prow[tc + 16 * j] = a[i][j];
        }
    }
    if (tr == kt && p != k) {
#pragma unroll
      for (int i = 0; i < R; i++)
        if (i == kj) {
#pragma unroll
          for (int j = 0; j < R; j++) krow[tc + 16 * j] = a[i][j];
        }
    }
    __syncthreads();
    if (p != k) {
      if (tr == kt) {
#pragma unroll
        for (int i = 0; i < R; i++)
          if (i == kj) {
#pragma unroll
            for (int j = 0; j < R; j++) a[i][j] = prow[tc + 16 * j];
          }
      }
      if (tr == (p & 15)) {
#pragma unroll
        for (int i = 0; i < R; i++)
          if (i == pi) {
#pragma unroll
            for (int j = 0; j < R; j++) a[i][j] = krow[tc + 16 * j];
          }
      }
      sgn = -sgn;
    }
    const double piv = prow[k];
    logdet += log(fabs(piv));
    if (piv < 0) sgn = -sgn;
    const double ip = 1.0 / piv;
    if (tc == kt) {
#pragma unroll
      for (int i = 0; i < R; i++) {
        double aik = 0.0;
#pragma unroll
        for (int j = 0; j < R; j++) aik = (j == kj) ? a[i][j] : aik;
        fcol[tr + 16 * i] = aik;
      }
    }
    __syncthreads();
    // ---- elimination
#pragma unroll
    for (int i = 0; i < R; i++) {
      const int gi = tr + 16 * i;
      const double f = fcol[gi];
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int gj = tc + 16 * j;
        const double pr = (gj == k) ? ip : prow[gj] * ip;
        if (gi == k) a[i][j] = pr;
        else a[i][j] = (gj == k) ? -f * ip : a[i][j] - f * pr;
      }
    }
    __syncthreads();
  }
  const bool sing = s_sing != 0;
  if (tid == 0) {
    // undo the row interchanges as column interchanges, in reverse order
    for (int c = 0; c < n; c++) dest[c] = c;
    if (!sing) {
      for (int c = 0; c < n; c++) cur[c] = c;
      for (int k = n - 1; k >= 0; k--) {
        const int q = ipiv[k];
        if (q != k) { const int t = cur[k]; cur[k] = cur[q]; cur[q] = t; }
      }
      for (int c = 0; c < n; c++) dest[cur[c]] = c;
    }
  }
  __syncthreads();
  double *inv = A.inv + (long)w * n * n;
  if (!sing) {
#pragma unroll
    for (int i = 0; i < R; i++)
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int gi = tr + 16 * i, gj = tc + 16 * j;
        if (gi < n && gj < n) inv[(long)gi * n + dest[gj]] = a[i][j];
      }
  }
  if (tid == 0) {
    A.logdet[w] = sing ? 0.0 : logdet;
    A.sgn[w] = sing ? 0.0 : sgn;
    A.singular[w] = sing ? 1 : 0;
  }
}

template <typename TS>
hipError_t launch_lu_inverse(const SlaterArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0 || a.n <= 0) return hipSuccess;
  if (a.n > 256) return hipErrorInvalidValue;
  if (a.n <= 16)
    hipLaunchKernelGGL((lu_inverse_kernel<TS, 1>), dim3(num), dim3(256), 0, st, a, walkers, num);
  else if (a.n <= 32)
    hipLaunchKernelGGL((lu_inverse_kernel<TS, 2>), dim3(num), dim3(256), 0, st, a, walkers, num);
  else if (a.n <= 64)
    hipLaunchKernelGGL((lu_inverse_kernel<TS, 4>), dim3(num), dim3(256), 0, st, a, walkers, num);
  else if (a.n <= 128)
    hipLaunchKernelGGL((lu_inverse_kernel<TS, 8>), dim3(num), dim3(256), 0, st, a, walkers, num);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Traces after an all-electron evaluate (QMCSlater.cpp:731-762):
//   gradPR[e][k] = sum_o d_k D(e,o) A^-1(o,e),  lapPR = sum_e sum_o lap D(e,o) A^-1(o,e)
template <typename TS>
__global__ void __launch_bounds__(256)
traces_all_kernel(SlaterArgs A, const int *walkers, int num) {
  const int b = blockIdx.x;
  if (b >= num) return;
  const int w = walkers ? walkers[b] : b;
  const int n = A.n;
  __shared__ double part[256];
  const double *inv = A.inv + (long)w * n * n;
  const TS *dm = (const TS *)A.dmat + (long)w * kComp * n * n;
  double lap = 0.0;
  const bool sing = A.singular[w] != 0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    double g0 = 0, g1 = 0, g2 = 0, l = 0;
    for (int o = 0; o < n; o++) {
      const double iv = inv[(long)o * n + e];
      g0 += ld(dm + 1L * n * n + (long)o * n + e) * iv;
      g1 += ld(dm + 2L * n * n + (long)o * n + e) * iv;
      g2 += ld(dm + 3L * n * n + (long)o * n + e) * iv;
      l += ld(dm + 4L * n * n + (long)o * n + e) * iv;
    }
    if (!sing) {
      double *g = A.gradPR + ((long)w * n + e) * 3;
      g[0] = g0; g[1] = g1; g[2] = g2;
    }
    lap += l;
  }
  part[threadIdx.x] = lap;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < (int)blockDim.x; i++) s += part[i];
    if (!sing) A.lapPR[w] = s;
  }
}

template <typename TS>
hipError_t launch_traces_all(const SlaterArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0 || a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL((traces_all_kernel<TS>), dim3(num), dim3(256), 0, st, a, walkers, num);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Drift modification of one electron, umrigar93_unequalelectrons
// (QMCSCFJastrow.cpp:368-443).
__device__ inline double drift_factor_unequal(const double g[3], const double x[3], int nn,
                                              const double *nuc_xyz, const double *nuc_z,
                                              double dt) {
  int cn = 0;
  double best = 1e300;
  for (int A = 0; A < nn; A++) {
    const double dx = x[0] - nuc_xyz[3 * A], dy = x[1] - nuc_xyz[3 * A + 1],
                 dz = x[2] - nuc_xyz[3 * A + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r < best) { best = r; cn = A; }
  }
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

// ---------------------------------------------------------------------------
// Single-electron move of spin s for one walker per workgroup.
//
// NT threads = NMAX columns x SL row slices; thread (c, sl) keeps rows
// [sl*RPT, (sl+1)*RPT) of A^-1 column c in registers for the whole kernel:
//   1. w(c) = v . A^-1(:,c) (v = new orbital row), ratio = w(r)
//   2. A'^-1(:,c) = A^-1(:,c) - A^-1(:,r) w(c)/ratio (c != r), A^-1(:,r)/ratio
//      -- the same row-replacement formula as inverseUpdateOneRow
//   3. traces at the proposal: gradPR'(c) and lap, streaming dmat column c
//      (new row r taken from v)
//   4. assembly over all N electrons: grad ln Psi', Laplacian, E_L', drift
//   5. stage the proposal (mode 0) or Metropolis-accept it and commit A'^-1
//      in place (mode 1: device VMC).
// HBM traffic per move: A^-1 read once (+ write when accepted), four
// derivative matrices read once; everything else is O(N).
template <typename TS, typename TO, int NMAX, int NT>
__global__ void __launch_bounds__(NT)
sm_move_kernel(MoveArgs A, const int *walkers, int num) {
  constexpr int SL = NT / NMAX;
  constexpr int RPT = NMAX / SL;
  constexpr int NW = NT / 64;
  const int b = blockIdx.x;
  if (b >= num) return;
  const int w = walkers ? walkers[b] : b;
  const int t = threadIdx.x, c = t % NMAX, sl = t / NMAX;
  const int n = A.s.n, N = A.N, r = A.moved - A.s.spin_start;
  const long nn2 = (long)n * n;

  __shared__ double v[kComp][NMAX];
  __shared__ double red[4][SL][NMAX];
  __shared__ double colr[NMAX];
  __shared__ double wv[NMAX];
  __shared__ double gq[4][NMAX];
  __shared__ double wred[4][NW];
  __shared__ double s_bc[8];

  const TO *vr = (const TO *)A.vrow;
  for (int i = t; i < kComp * n; i += NT) {
    const int k = i / n, o = i % n;
    v[k][o] = (double)vr[(long)k * A.v_s_comp + (long)o * A.v_s_orb + b];
  }
  __syncthreads();

  // ---- phase 1: w(c)
  const double *inv = A.s.inv + (long)w * nn2;
  double a[RPT];
  double wp = 0.0;
#pragma unroll
  for (int i = 0; i < RPT; i++) {
    const int o = sl * RPT + i;
    a[i] = (o < n && c < n) ? inv[(long)o * n + c] : 0.0;
    if (o < n) wp += v[0][o] * a[i];
  }
  red[0][sl][c] = wp;
  if (c == r) {
#pragma unroll
    for (int i = 0; i < RPT; i++) {
      const int o = sl * RPT + i;
      if (o < n) colr[o] = a[i];
    }
  }
  __syncthreads();
  if (sl == 0 && c < n) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < SL; q++) s += red[0][q][c];
    wv[c] = s;
  }
  __syncthreads();
  const double ratio = wv[r];
  const bool ok = (ratio != 0.0) && isfinite(ratio);
  const double qinv = 1.0 / ratio;
  const double wc = c < n ? wv[c] * qinv : 0.0;

  // ---- phase 2: A'^-1 in registers
#pragma unroll
  for (int i = 0; i < RPT; i++) {
    const int o = sl * RPT + i;
    if (o < n) a[i] = (c == r) ? a[i] * qinv : a[i] - wc * colr[o];
  }

  // ---- phase 3: traces at the proposal
  const TS *dm = (const TS *)A.s.dmat + (long)w * kComp * nn2;
  double gp[4] = {0.0, 0.0, 0.0, 0.0};
  if (c < n) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const TS *col = dm + (long)(1 + k) * nn2 + c;
#pragma unroll
      for (int i = 0; i < RPT; i++) {
        const int o = sl * RPT + i;
        if (o < n) {
          const double d = (c == r) ? v[1 + k][o] : ld(col + (long)o * n);
          gp[k] += d * a[i];
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) red[k][sl][c] = gp[k];
  __syncthreads();
  if (sl == 0) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < SL; q++) s += red[k][q][c];
      gq[k][c] = (c < n) ? s : 0.0;
    }
  }
  __syncthreads();

  // ---- phase 4: assembly over all electrons
  const int so = A.other.spin_start, no = A.other.n, ss = A.s.spin_start;
  const double *pUx = A.pUx + (long)b * N * 3;
  const double *cUx = A.Ux + (long)w * N * 3;
  const double *Rw = A.R + (long)w * N * 3;
  const double *gpo = no > 0 ? A.other.gradPR + (long)w * no * 3 : nullptr;
  const double *gps = A.s.gradPR + (long)w * n * 3;
  const double tau = A.tau;
  double lsum = 0.0, fwd = 0.0, rev = 0.0, lapc = 0.0, gsq = 0.0;
  for (int e = t; e < N; e += NT) {
    double dx[3], dxc[3];
    if (e >= ss && e < ss + n) {
      for (int k = 0; k < 3; k++) { dx[k] = gq[k][e - ss]; dxc[k] = gps[3 * (e - ss) + k]; }
      lapc += gq[3][e - ss];
    } else {
      for (int k = 0; k < 3; k++) { dx[k] = gpo[3 * (e - so) + k]; dxc[k] = dx[k]; }
    }
    double g[3], gc[3];
    for (int k = 0; k < 3; k++) {
      const double ux = pUx[3 * e + k];
      g[k] = dx[k] + ux;
      gc[k] = dxc[k] + cUx[3 * e + k];
      lsum += ux * dx[k] * 2.0;
      lsum += ux * ux;
      gsq += g[k] * g[k];
    }
    if (A.out.grad_lnpsi) {
      double *o = A.out.grad_lnpsi + ((long)b * N + e) * 3;
      o[0] = g[0]; o[1] = g[1]; o[2] = g[2];
    }
    if (A.mode == 1) {
      // Green's function terms (QMCWalker.cpp:343-398, 756-780); drift =
      // gradPsiRatio (QF_modification_type none)
      if (e != A.moved) {
        for (int k = 0; k < 3; k++) {
          fwd += tau * gc[k] * tau * gc[k];
          rev += tau * g[k] * tau * g[k];
        }
      } else {
        for (int k = 0; k < 3; k++) {
          const double tt = Rw[3 * e + k] - A.propR[3L * b + k] - tau * g[k];
          rev += tt * tt;
          fwd += A.gauss[3L * b + k] * A.gauss[3L * b + k];
        }
      }
    }
  }
  // block reduction of lsum, fwd, rev, lapc (+gsq for the equal-electron drift)
  double vals[5] = {lsum, fwd, rev, lapc, gsq};
  __shared__ double wred5[5][NW];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const double s = wave_sum(vals[k]);
    if ((t & 63) == 0) wred5[k][t >> 6] = s;
  }
  __syncthreads();
  if (t == 0) {
    double tot[5];
    for (int k = 0; k < 5; k++) {
      double s = 0.0;
      for (int q = 0; q < NW; q++) s += wred5[k][q];
      tot[k] = s;
    }
    s_bc[0] = tot[0]; s_bc[1] = tot[1]; s_bc[2] = tot[2]; s_bc[3] = tot[3]; s_bc[4] = tot[4];
  }
  __syncthreads();
  const double lap_s = s_bc[3];
  const double lap_o = no > 0 ? A.other.lapPR[w] : 0.0;
  const double lapR = lap_s + lap_o + A.pUxx[b] + s_bc[0];
  const double V = A.v_nn + A.pVee[b] + A.pVen[b];
  const double EL = -0.5 * lapR + V;
  const double ld_s = A.s.logdet[w], ld_o = no > 0 ? A.other.logdet[w] : 0.0;
  const double logpsi_new = ld_s + log(fabs(ratio)) + ld_o + A.pU[b];
  const double sg_o = no > 0 ? A.other.sgn[w] : 1.0;
  const double sgn_s_new = A.s.sgn[w] * (ratio < 0 ? -1.0 : 1.0);
  bool sing = !ok || A.s.singular[w] || (no > 0 && A.other.singular[w]);
  if (!(s_bc[4] > 0.0) || !isfinite(s_bc[4])) sing = true;

  // modified drift output (staged/evaluate outputs)
  if (A.out.mod_drift) {
    double feq = 1.0;
    if (A.drift_mod == QMCB_DRIFT_UMRIGAR93_EQUAL) {
      const double m = s_bc[4];
      feq = (-1.0 + sqrt(1.0 + 2 * A.umrigar_a * m * A.dt)) / (A.umrigar_a * m * A.dt);
    }
    for (int e = t; e < N; e += NT) {
      double g[3];
      const double *src = A.out.grad_lnpsi + ((long)b * N + e) * 3;
      g[0] = src[0]; g[1] = src[1]; g[2] = src[2];
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
  if (t == 0) {
    const qmcb_results &o = A.out;
    if (o.log_psi) o.log_psi[b] = logpsi_new;
    if (o.sign_psi) o.sign_psi[b] = (int8_t)(sgn_s_new * sg_o < 0 ? -1 : 1);
    if (o.lap_ratio) o.lap_ratio[b] = lapR;
    if (o.e_local) o.e_local[b] = EL;
    if (o.kinetic) o.kinetic[b] = -0.5 * lapR;
    if (o.potential) o.potential[b] = V;
    if (o.v_ne) o.v_ne[b] = A.pVen[b];
    if (o.v_ee) o.v_ee[b] = A.pVee[b];
    if (o.det_ratio) o.det_ratio[b] = ratio;
    if (o.jastrow_u) o.jastrow_u[b] = A.pU[b];
    if (o.singular) o.singular[b] = sing ? 1 : 0;
  }

  if (A.mode == 0) {
    // ---- stage the proposal
    if (c < n) {
      double *st = A.st_inv + (long)b * nn2;
#pragma unroll
      for (int i = 0; i < RPT; i++) {
        const int o = sl * RPT + i;
        if (o < n) st[(long)o * n + c] = a[i];
      }
      if (sl == 0) {
        double *g = A.st_grad + ((long)b * n + c) * 3;
        g[0] = gq[0][c]; g[1] = gq[1][c]; g[2] = gq[2][c];
      }
    }
    if (t == 0) {
      A.st_lap[b] = lap_s;
      A.st_logdet[b] = ld_s + log(fabs(ratio));
      A.st_sgn[b] = sgn_s_new;
      A.st_singular[b] = sing ? 1 : 0;
    }
    return;
  }

  // ---- mode 1: Metropolis with the Green's function ratio
  // (QMCWalker.cpp:1578-1685); p = |Psi'/Psi|^2 G(R<-R')/G(R'<-R)
  const double logpsi_old = ld_s + ld_o + A.U[w];
  double p = sing ? 0.0 : exp(2.0 * (logpsi_new - logpsi_old) - (s_bc[2] - s_bc[1]) / (2.0 * tau));
  if (!(p == p)) p = 0.0;
  if (p > 1.0) p = 1.0;
  const bool accept = p > A.unif[b];
  __syncthreads();  // all reads of the current state are done
  if (accept) {
    double *invw = A.s.inv + (long)w * nn2;
    if (c < n) {
#pragma unroll
      for (int i = 0; i < RPT; i++) {
        const int o = sl * RPT + i;
        if (o < n) invw[(long)o * n + c] = a[i];
      }
      if (sl == 0) {
        double *g = A.s.gradPR + ((long)w * n + c) * 3;
        g[0] = gq[0][c]; g[1] = gq[1][c]; g[2] = gq[2][c];
      }
    }
    // new row r of D and its derivatives -> column r of the stored matrices
    TS *dmw = (TS *)A.s.dmat + (long)w * kComp * nn2;
    for (int i = t; i < kComp * n; i += NT) {
      const int k = i / n, o = i % n;
      dmw[(long)k * nn2 + (long)o * n + r] = (TS)v[k][o];
    }
    double *Uxw = A.Ux + (long)w * N * 3;
    for (int i = t; i < 3 * N; i += NT) Uxw[i] = pUx[i];
    if (t == 0) {
      A.s.lapPR[w] = lap_s;
      A.s.logdet[w] = ld_s + log(fabs(ratio));
      A.s.sgn[w] = sgn_s_new;
      double *Rm = A.R + ((long)w * N + A.moved) * 3;
      Rm[0] = A.propR[3L * b]; Rm[1] = A.propR[3L * b + 1]; Rm[2] = A.propR[3L * b + 2];
      A.U[w] = A.pU[b];
      A.Uxx[w] = A.pUxx[b];
      A.Vee[w] = A.pVee[b];
      A.Ven[w] = A.pVen[b];
      A.cur_el[w] = EL;
    }
  }
  if (t == 0) {
    const double el = accept ? EL : A.cur_el[w];
    A.acc_el[w] += el;
    A.acc_el2[w] += el * el;
    A.acc_n[w] += accept ? 1.0 : 0.0;
    A.acc_prop_el[w] += sing ? 0.0 : EL;
  }
}

template <typename TS, typename TO>
hipError_t launch_sm_move(const MoveArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  const int n = a.s.n;
  if (n <= 0) return hipErrorInvalidValue;
  if (n <= 16)
    hipLaunchKernelGGL((sm_move_kernel<TS, TO, 16, 64>), dim3(num), dim3(64), 0, st, a, walkers,
                       num);
  else if (n <= 32)
    hipLaunchKernelGGL((sm_move_kernel<TS, TO, 32, 128>), dim3(num), dim3(128), 0, st, a,
                       walkers, num);
  else if (n <= 64)
    hipLaunchKernelGGL((sm_move_kernel<TS, TO, 64, 256>), dim3(num), dim3(256), 0, st, a,
                       walkers, num);
  else if (n <= 128)
    hipLaunchKernelGGL((sm_move_kernel<TS, TO, 128, 512>), dim3(num), dim3(512), 0, st, a,
                       walkers, num);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Commit of staged proposals (qmcb_accept).  One workgroup per staged row.
template <typename TS, typename TO>
__global__ void __launch_bounds__(256)
commit_kernel(MoveArgs A, const int *walkers, const uint8_t *mask, int num) {
  const int b = blockIdx.x;
  if (b >= num || !mask[b]) return;
  const int w = walkers[b];
  const int n = A.s.n, N = A.N, r = A.moved - A.s.spin_start;
  const long nn2 = (long)n * n;
  const int t = threadIdx.x;
  if (A.st_singular[b]) return;
  double *invw = A.s.inv + (long)w * nn2;
  const double *st = A.st_inv + (long)b * nn2;
  for (long i = t; i < nn2; i += blockDim.x) invw[i] = st[i];
  double *g = A.s.gradPR + (long)w * n * 3;
  const double *sg = A.st_grad + (long)b * n * 3;
  for (int i = t; i < 3 * n; i += blockDim.x) g[i] = sg[i];
  const TO *vr = (const TO *)A.vrow;
  TS *dmw = (TS *)A.s.dmat + (long)w * kComp * nn2;
  for (int i = t; i < kComp * n; i += blockDim.x) {
    const int k = i / n, o = i % n;
    dmw[(long)k * nn2 + (long)o * n + r] = (TS)(double)vr[(long)k * A.v_s_comp + (long)o * A.v_s_orb + b];
  }
  double *Uxw = A.Ux + (long)w * N * 3;
  const double *pUx = A.pUx + (long)b * N * 3;
  for (int i = t; i < 3 * N; i += blockDim.x) Uxw[i] = pUx[i];
  if (t == 0) {
    A.s.lapPR[w] = A.st_lap[b];
    A.s.logdet[w] = A.st_logdet[b];
    A.s.sgn[w] = A.st_sgn[b];
    double *Rm = A.R + ((long)w * N + A.moved) * 3;
    Rm[0] = A.propR[3L * b]; Rm[1] = A.propR[3L * b + 1]; Rm[2] = A.propR[3L * b + 2];
    A.U[w] = A.pU[b];
    A.Uxx[w] = A.pUxx[b];
    A.Vee[w] = A.pVee[b];
    A.Ven[w] = A.pVen[b];
  }
}

template <typename TS, typename TO>
hipError_t launch_commit(const MoveArgs &a, const int *walkers, const uint8_t *mask, int num,
                         hipStream_t st) {
  if (num <= 0) return hipSuccess;
  hipLaunchKernelGGL((commit_kernel<TS, TO>), dim3(num), dim3(256), 0, st, a, walkers, mask, num);
  return hipGetLastError();
}

#define INST(TS, TO)                                                                          \
  template hipError_t launch_sm_move<TS, TO>(const MoveArgs &, const int *, int, hipStream_t); \
  template hipError_t launch_commit<TS, TO>(const MoveArgs &, const int *, const uint8_t *, int, \
                                            hipStream_t);
INST(double, double)
INST(float, float)
#undef INST
template hipError_t launch_lu_inverse<double>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_lu_inverse<float>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_traces_all<double>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_traces_all<float>(const SlaterArgs &, const int *, int, hipStream_t);

}  // namespace qmcb
