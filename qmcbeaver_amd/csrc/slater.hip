// slater.hip -- kernel (c): Slater inverse, determinant, Sherman-Morrison
// single-electron updates and derivative-ratio traces, all FP64.
//
// Reference: QMCSlater::update_Ds / calculate_DerivativeRatios
// (QMCSlater.cpp:390-764), Array2D::determinant_and_inverse (Array2D.h:
// 1521-1548, LU in ludcmp/lubksb :1204-1421) and inverseUpdateOneRow
// (Array2D.h:1566-1610).
//
// Storage (per walker, per spin): inv(o,e) row-major, dval(o,e) = D(e,o) in
// FP64, dmat(k,o,e) = d_k D(e,o) for k in {x, y, z, Laplacian} (transposed),
// so the thread that owns electron column e of A^-1 also streams column e of
// every derivative matrix with coalesced loads.
#include "qmcb_kernels.hpp"

namespace qmcb {

template <typename TS>
__device__ __forceinline__ double ld(const TS *p) { return (double)*p; }

// ---------------------------------------------------------------------------
// Batched inverse by Gauss-Jordan elimination with partial (row) pivoting,
// register resident: G*G threads as a GxG grid, thread (tr, tc) owns the
// cyclic block a[tr + G i][tc + G j], i, j < R (n <= G R).  Per pivot step
// the G owners of the pivot column find the pivot with wavefront shuffles,
// and only the pivot row and column pass through LDS (two barriers/step).  ln|det| and sign
// are accumulated from the pivots and row swaps; the LU of the reference
// (Crout, implicit scaling, long double sums) gives the same inverse and
// determinant up to rounding.  Singular (an all-zero pivot column) sets the
// walker's singular flag, as calcOK = false does (QMCSlater.cpp:675-683).
template <typename TS, int R, int G>
__global__ void __launch_bounds__(G * G)
lu_inverse_kernel(SlaterArgs A, const int *walkers, int num) {
  const int D = det_count(A), d = blockIdx.x % D;
  const int b = blockIdx.x / D;
  if (b >= num) return;
  const int w0 = walkers ? walkers[b] : b;
  const int w = w0 * D + d;  // this determinant's record
  const int n = A.n, nact = act_count(A);
  const int tid = threadIdx.x;
  // tid = tr + G tc: the G owners of one column are G consecutive lanes
  const int tr = tid % G, tc = tid / G;
  constexpr int M = G * R;
  __shared__ double prow[M], krow[M];
  __shared__ double fcol[2][M];
  __shared__ int s_piv[2];
  __shared__ double s_max[2];
  __shared__ int ipiv[M];
  __shared__ int dest[M];
  __shared__ int cur[M];

  const double *dv = A.dval + (long)w0 * nact * n;  // value matrix, (o, e)
  double a[R][R];
#pragma unroll
  for (int i = 0; i < R; i++)
#pragma unroll
    for (int j = 0; j < R; j++) {
      const int gi = tr + G * i, gj = tc + G * j;  // gi: electron, gj: orbital
      double v = (gi == gj) ? 1.0 : 0.0;
      if (gi < n && gj < n) v = ld(dv + (long)occ_row(A, d, gj) * n + gi);
      a[i][j] = v;
    }
  // |det| kept as mantissa * 2^expo to stay in range
  double mant = 1.0, sgn = 1.0;
  int expo = 0;
  bool sing = false;
  for (int k = 0; k < n; k++) {
    const int kj = k / G, kt = k % G, buf = k & 1;
    // ---- pivot search in column k (rows >= k) by its 16 owner lanes
    if (tc == kt) {
      double best = -1.0;
      int bi = 0x7fffffff;
#pragma unroll
      for (int i = 0; i < R; i++) {
        const int gi = tr + G * i;
        double aik = 0.0;
#pragma unroll
        for (int j = 0; j < R; j++) aik = (j == kj) ? a[i][j] : aik;
        fcol[buf][gi] = aik;
        if (gi >= k && gi < n) {
          const double v = fabs(aik);
          if (v > best) { best = v; bi = gi; }
        }
      }
      if constexpr (G == 16) {
        row16_argmax(best, bi);  // the owners are one DPP row of 16 lanes
      } else {
#pragma unroll
        for (int off = G / 2; off >= 1; off >>= 1) {
          const double ob = __shfl_xor(best, off, 64);
          const int oi = __shfl_xor(bi, off, 64);
          if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
      }
      if (tr == 0) { s_piv[buf] = bi; s_max[buf] = best; }
    }
    __syncthreads();
    if (!(s_max[buf] > 0.0)) { sing = true; break; }
    const int p = s_piv[buf], pi = p / G;
    if (tid == 0) ipiv[k] = p;
    if (tr == p % G) {
#pragma unroll
      for (int i = 0; i < R; i++)
        if (i == pi) {
#pragma unroll
          for (int j = 0; j < R; j++) prow[tc + G * j] = a[i][j];
        }
    }
    if (tr == kt && p != k) {
#pragma unroll
      for (int i = 0; i < R; i++)
        if (i == kj) {
#pragma unroll
          for (int j = 0; j < R; j++) krow[tc + G * j] = a[i][j];
        }
    }
    __syncthreads();
    if (p != k) {
      if (tr == kt) {
#pragma unroll
        for (int i = 0; i < R; i++)
          if (i == kj) {
#pragma unroll
            for (int j = 0; j < R; j++) a[i][j] = prow[tc + G * j];
          }
      }
      if (tr == p % G) {
#pragma unroll
        for (int i = 0; i < R; i++)
          if (i == pi) {
#pragma unroll
            for (int j = 0; j < R; j++) a[i][j] = krow[tc + G * j];
          }
      }
      sgn = -sgn;
    }
    const double piv = prow[k];
    if (piv < 0) sgn = -sgn;
    int e2;
    mant = frexp(mant * fabs(piv), &e2);
    expo += e2;
    const double ip = 1.0 / piv;
    // ---- elimination: row k scaled, other rows reduced (in-place Gauss-Jordan)
#pragma unroll
    for (int i = 0; i < R; i++) {
      const int gi = tr + G * i;
      // column-k entry of row gi after the interchange
      const double f = (gi == p) ? fcol[buf][k] : fcol[buf][gi];
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int gj = tc + G * j;
        const double pr = (gj == k) ? ip : prow[gj] * ip;
        if (gi == k) a[i][j] = pr;
        else a[i][j] = (gj == k) ? -f * ip : a[i][j] - f * pr;
      }
    }
  }
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
        const int gi = tr + G * i, gj = tc + G * j;
        if (gi < n && gj < n) inv[(long)gi * n + dest[gj]] = a[i][j];
      }
  }
  if (tid == 0) {
    A.logdet[w] = sing ? 0.0 : log(mant) + expo * 0.6931471805599453;
    A.sgn[w] = sing ? 0.0 : sgn;
    A.singular[w] = sing ? 1 : 0;
    A.dcount[w] = 0;
  }
}

// ---------------------------------------------------------------------------
// Row-resident Gauss-Jordan for NP in {64, 128}: row i of the matrix lives in
// the registers of the lane pair (2i, 2i+1), NP/2 doubles each (half h holds
// register positions h*NP/2 ...).  Pivot columns are processed in blocks of 16
// with compile-time register indices; after each block every row is rotated
// left by 16 positions (16 values cross between the partner lanes by
// shuffle), so the next block's pivot columns are again positions 0..15 of
// half 0 and no register array is indexed dynamically.  All rows share the
// rotation state, so LDS rows (pivot row, swapped row) are exchanged in
// position order.  Per pivot: a wavefront argmax of |a[k]| over rows >= k,
// one row interchange through LDS, NP/2 FMAs per lane.  The inverse is written
// through an LDS tile so that the column permutation of the interchanges
// lands in coalesced stores of rows of A^-1(o, e).
template <typename TS, int NP>
__global__ void __launch_bounds__(2 * NP)
gj_rows_kernel(SlaterArgs A, const int *walkers, int num) {
  constexpr int KB = 16, NB = NP / KB, H = NP / 2, NT = 2 * NP, NWV = NT / 64, TR = 32;
  const int D = det_count(A), d = blockIdx.x % D;
  const int b = blockIdx.x / D;
  if (b >= num) return;
  const int w0 = walkers ? walkers[b] : b;
  const int w = w0 * D + d;  // this determinant's record
  const int n = A.n, nact = act_count(A);
  const int t = threadIdx.x, i = t >> 1, h = t & 1;
  // pivot rows and arg-max slots alternate between two buffers by pivot
  // parity: pivot k+1 writes the buffer pivot k is not reading, and pivot k+2
  // rewrites pivot k's buffer only after every wave has passed both barriers
  // of pivot k+1, so one barrier per pivot is saved
  __shared__ double prow_b[2][NP], krow_b[2][NP];
  __shared__ double wmax_b[2][NWV];
  __shared__ int widx_b[2][NWV];
  __shared__ int ipiv[NP], dest[NP], cur[NP];
  __shared__ double tile[TR][NP + 1];

  const double *dv = A.dval + (long)w0 * nact * n;  // value matrix, (o, e)
  double a[H];
#pragma unroll
  for (int j = 0; j < H; j++) {
    const int col = h * H + j;
    a[j] = (i < n && col < n) ? ld(dv + (long)occ_row(A, d, col) * n + i)
                              : (i == col ? 1.0 : 0.0);
  }
  double mant = 1.0, sgn = 1.0;
  int expo = 0;
  bool sing = false;
  for (int kb = 0; kb < NB; kb++) {
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const int k = kb * KB + kk;
      if (k < n && !sing) {
        double *prow = prow_b[k & 1], *krow = krow_b[k & 1], *wmax = wmax_b[k & 1];
        int *widx = widx_b[k & 1];
        const double akk = dpp_f64<0xA0>(a[kk]);  // position kk of half 0
        double best = (h == 0 && i >= k && i < n) ? fabs(akk) : -1.0;
        int bi = i;
        wave_argmax(best, bi);
        if ((t & 63) == 0) { wmax[t >> 6] = best; widx[t >> 6] = bi; }
        __syncthreads();
        best = wmax[0];
        bi = widx[0];
#pragma unroll
        for (int q = 1; q < NWV; q++)
          if (wmax[q] > best || (wmax[q] == best && widx[q] < bi)) { best = wmax[q]; bi = widx[q]; }
        if (!(best > 0.0)) {
          sing = true;
        } else {
          const int p = bi;
          if (i == p) {
#pragma unroll
            for (int j = 0; j < H; j++) prow[h * H + j] = a[j];
          }
          if (i == k && p != k) {
#pragma unroll
            for (int j = 0; j < H; j++) krow[h * H + j] = a[j];
          }
          if (t == 0) ipiv[k] = p;
          __syncthreads();
          if (p != k) {
            if (i == k) {
#pragma unroll
              for (int j = 0; j < H; j++) a[j] = prow[h * H + j];
            } else if (i == p) {
#pragma unroll
              for (int j = 0; j < H; j++) a[j] = krow[h * H + j];
            }
            sgn = -sgn;
          }
          const double piv = prow[kk];
          if (piv < 0) sgn = -sgn;
          int e2;
          mant = frexp(mant * fabs(piv), &e2);
          expo += e2;
          const double ip = 1.0 / piv;
          const double f = dpp_f64<0xA0>(a[kk]);  // after the interchange
          if (i == k) {
#pragma unroll
            for (int j = 0; j < H; j++) a[j] = (h == 0 && j == kk) ? ip : prow[h * H + j] * ip;
          } else {
            const double fi = f * ip;
#pragma unroll
            for (int j = 0; j < H; j++)
              a[j] = (h == 0 && j == kk) ? -fi : a[j] - fi * prow[h * H + j];
          }
        }
      }
    }
    // rotate every row left by KB positions: half 0 takes its own [KB, H)
    // plus half 1's [0, KB); half 1 takes its own [KB, H) plus half 0's [0, KB)
    double tq[KB];
#pragma unroll
    for (int q = 0; q < KB; q++) tq[q] = dpp_f64<0xB1>(a[q]);  // partner lane (lane ^ 1)
#pragma unroll
    for (int j = 0; j < H - KB; j++) a[j] = a[j + KB];
#pragma unroll
    for (int q = 0; q < KB; q++) a[H - KB + q] = tq[q];
  }
  if (t == 0) {
    for (int c = 0; c < n; c++) dest[c] = c;
    if (!sing) {
      for (int c = 0; c < n; c++) cur[c] = c;
      for (int k = n - 1; k >= 0; k--) {
        const int q = ipiv[k];
        if (q != k) { const int tt = cur[k]; cur[k] = cur[q]; cur[q] = tt; }
      }
      for (int c = 0; c < n; c++) dest[cur[c]] = c;
    }
  }
  __syncthreads();
  // row o = i of A^-1; register position j of half h is column h*H + j,
  // which belongs to electron dest[h*H + j]
  double *inv = A.inv + (long)w * n * n;
  if (!sing) {
    for (int r0 = 0; r0 < n; r0 += TR) {
      if (i >= r0 && i < r0 + TR) {
#pragma unroll
        for (int j = 0; j < H; j++) {
          const int col = h * H + j;
          if (col < n) tile[i - r0][dest[col]] = a[j];
        }
      }
      __syncthreads();
      for (int idx = t; idx < TR * n; idx += NT) {
        const int r = idx / n, e = idx % n;
        if (r0 + r < n) inv[(long)(r0 + r) * n + e] = tile[r][e];
      }
      __syncthreads();
    }
  }
  if (t == 0) {
    A.logdet[w] = sing ? 0.0 : log(mant) + expo * 0.6931471805599453;
    A.sgn[w] = sing ? 0.0 : sgn;
    A.singular[w] = sing ? 1 : 0;
    A.dcount[w] = 0;
  }
}

template <typename TS>
hipError_t launch_lu_inverse(const SlaterArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0 || a.n <= 0) return hipSuccess;
  const unsigned g = (unsigned)num * (a.ndet > 1 ? a.ndet : 1);
  if (a.n <= 16)
    hipLaunchKernelGGL((lu_inverse_kernel<TS, 1, 16>), dim3(g), dim3(256), 0, st, a, walkers, num);
  else if (a.n <= 32)
    hipLaunchKernelGGL((lu_inverse_kernel<TS, 2, 16>), dim3(g), dim3(256), 0, st, a, walkers, num);
  else if (a.n <= 64)
    hipLaunchKernelGGL((gj_rows_kernel<TS, 64>), dim3(g), dim3(128), 0, st, a, walkers, num);
  else if (a.n <= 128)
    hipLaunchKernelGGL((gj_rows_kernel<TS, 128>), dim3(g), dim3(256), 0, st, a, walkers, num);
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
  const int D = det_count(A), d = blockIdx.x % D;
  const int b = blockIdx.x / D;
  if (b >= num) return;
  const int w0 = walkers ? walkers[b] : b;
  const int w = w0 * D + d;  // this determinant's record
  const int n = A.n, nact = act_count(A);
  const long cs = (long)nact * n;  // component stride of the derivative matrices
  __shared__ double part[256];
  const double *inv = A.inv + (long)w * n * n;
  const TS *dm = (const TS *)A.dmat + (long)w0 * kDer * cs;
  double lap = 0.0;
  const bool sing = A.singular[w] != 0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    double g0 = 0, g1 = 0, g2 = 0, l = 0;
    for (int o = 0; o < n; o++) {
      const double iv = inv[(long)o * n + e];
      const long ro = (long)occ_row(A, d, o) * n + e;
      g0 += ld(dm + ro) * iv;
      g1 += ld(dm + cs + ro) * iv;
      g2 += ld(dm + 2 * cs + ro) * iv;
      l += ld(dm + 3 * cs + ro) * iv;
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
  const unsigned g = (unsigned)num * (a.ndet > 1 ? a.ndet : 1);
  hipLaunchKernelGGL((traces_all_kernel<TS>), dim3(g), dim3(256), 0, st, a, walkers, num);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Single-electron move of spin s, one walker per workgroup, one streaming pass.
//
// With w(c) = v . A^-1(:,c) (v = the moved electron's new orbital row) and
// ratio = w(r), the row-replacement update (inverseUpdateOneRow) is
//   A'^-1 = A^-1 - A^-1(:,r) z^T,   z(c) = (w(c) - delta_cr) / ratio,
// so every trace at the proposal splits into two sums over the current
// inverse:
//   sum_o d_k D(c,o) A'^-1(o,c) = T1_k(c) - (w(c)/ratio) T2_k(c),
//   T1_k(c) = sum_o d_k D(c,o) A^-1(o,c),  T2_k(c) = sum_o d_k D(c,o) A^-1(o,r).
//
// Delayed updates: the resident inverse is kept as A_0 - sum_a u_a z_a^T with
// up to kDelay pending rank-1 terms (u_a = column r_a of the inverse at update
// a).  An accepted move only appends (u, z) -- 2n doubles -- instead of
// rewriting the n x n inverse; the pending terms are folded into A_0 during
// the streaming pass of the move that finds the buffer full.
//
// Pass 1 streams A_0 and the four derivative matrices once, forms the current
// inverse element on the fly and accumulates w, T1, T2 per column (9 registers).
// Thread mapping: c = t % NC (electron column), sl = t / NC (row slice).
// Rows per software-pipeline stage of the read-only stream (2U rows in
// flight per lane).  For the 128-column tile the deeper pipelines raise the
// register count above 128 (FP32 derivatives: 131 VGPRs, 3 waves per SIMD;
// FP64: 213, 2 waves per SIMD) with the streaming rate unchanged, which frees
// wave slots for the basis, Jastrow and GEMM kernels of the other walker
// groups running beside sm_move (+7 % / +8 % walker-moves/s with three groups
// at 256 electrons, equal with one group).
#ifndef QMCB_SM_UNROLL_F32
#define QMCB_SM_UNROLL_F32 4
#endif
#ifndef QMCB_SM_UNROLL_F64
#define QMCB_SM_UNROLL_F64 6
#endif
#ifndef QMCB_SM_UNROLL_SMALL
#define QMCB_SM_UNROLL_SMALL 2
#endif
template <typename TS, int NC>
__host__ __device__ constexpr int sm_unroll() {
  return NC < 128 ? QMCB_SM_UNROLL_SMALL : (sizeof(TS) == 4 ? QMCB_SM_UNROLL_F32 : QMCB_SM_UNROLL_F64);
}
#ifndef QMCB_SM_MINWAVES
#define QMCB_SM_MINWAVES 1
#endif
#ifndef QMCB_SM_NT
#define QMCB_SM_NT 1
#endif


// threads per walker for the 32- and 64-column tiles: one and two waves
// (more walkers resident per CU; 5 % faster on the batched many-molecule
// workload than two and four waves)
#ifndef QMCB_SM_NT32
#define QMCB_SM_NT32 64
#endif
#ifndef QMCB_SM_NT64
#define QMCB_SM_NT64 128
#endif
// a 96-column tile (three waves) for 65-96 electrons per spin, where the
// 128-column tile leaves 25-49 % of its columns idle (about 5 % more
// walker-moves/s on the batched many-molecule workload)
#ifndef QMCB_SM_TILE96
#define QMCB_SM_TILE96 1
#endif

// streamed-once operands: optionally non-temporal loads
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
#if QMCB_SM_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

template <typename TS, typename TO, int NC, int NT>
__global__ void __launch_bounds__(NT, QMCB_SM_MINWAVES)
sm_move_kernel(MoveArgs A, const int *walkers, int num) {
  constexpr int SL = NT / NC;
  constexpr int NW = NT / 64;
  constexpr int KD = kDelay;
  const int b = blockIdx.x;
  if (b >= num) return;
  const int w = walkers ? walkers[b] : b;
  const int t = threadIdx.x, c = t % NC, sl = t / NC;
  const int n = A.s.n, N = A.N, r = A.moved - A.s.spin_start;
  const long nn2 = (long)n * n;
  const int rows = (n + SL - 1) / SL;
  const int o_beg = sl * rows, o_end = min(n, o_beg + rows);
  const int dc = A.s.dcount[w];
  const bool flush = (A.mode >= 1) && dc >= KD;

  __shared__ double v[kComp][NC];
  __shared__ double colr[NC];
  __shared__ double ub[KD][NC];
  __shared__ double zr[KD];
  __shared__ double red[9][SL > 1 ? SL - 1 : 1][NC];
  __shared__ double wv[NC];
  __shared__ double gq[4][NC];
  __shared__ double wred5[8][NW];
  __shared__ double s_bc[8];

  const TO *vr = (const TO *)A.vrow + (long)b * A.v_s_walker;
  double *inv = A.s.inv + (long)w * nn2;
  const double *ubw = A.s.ubuf + (long)w * KD * n;
  const double *zbw = A.s.zbuf + (long)w * KD * n;
  for (int i = t; i < kComp * n; i += NT) {
    const int k = i / n, o = i % n;
    v[k][o] = (double)vr[(long)k * A.v_s_comp + o];
  }
  for (int i = t; i < KD * NC; i += NT) {
    const int a = i / NC, o = i % NC;
    ub[a][o] = (a < dc && o < n) ? ubw[(long)a * n + o] : 0.0;
  }
  if (t < KD) zr[t] = t < dc ? zbw[(long)t * n + r] : 0.0;
  double zc[KD];
#pragma unroll
  for (int a = 0; a < KD; a++) zc[a] = (a < dc && c < n) ? zbw[(long)a * n + c] : 0.0;
  // column r of A_0 (one element per thread: NT >= NC >= n), loaded with the
  // first batch instead of after its barrier
  static_assert(NT >= NC, "one column element per thread");
  const double xr = t < n ? inv[(long)t * n + r] : 0.0;
  __syncthreads();

  // ---- pass 1 set-up: the read-only stream's first U rows are requested
  // before column r is formed, so their latency overlaps it and the barrier
  const TS *dm = (const TS *)A.s.dmat + (long)w * kDer * nn2;
  const TS *__restrict__ d1 = dm + c;
  const TS *__restrict__ d2 = dm + 1 * nn2 + c;
  const TS *__restrict__ d3 = dm + 2 * nn2 + c;
  const TS *__restrict__ d4 = dm + 3 * nn2 + c;
  // software pipeline in registers: the loads of the next U rows are issued
  // before the current U rows are consumed, so every wave keeps 2U rows
  // (U x 24/40 bytes per lane) in flight
  constexpr int U = sm_unroll<TS, NC>();
  const double *__restrict__ ip_s = inv + c;
  double ca[U], na[U];
  TS cx[U], cy[U], cz[U], cl[U], nx[U], ny[U], nz[U], nl[U];
  auto load = [&](int o0, double *a, TS *x, TS *y, TS *z, TS *l) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int o = min(o0 + u, o_end - 1);  // clamped: in bounds, unused past o_end
      const long off = (long)o * n;
      a[u] = ld_stream(ip_s + off);
      x[u] = ld_stream(d1 + off);
      y[u] = ld_stream(d2 + off);
      z[u] = ld_stream(d3 + off);
      l[u] = ld_stream(d4 + off);
    }
  };
  const bool stream = !flush && c < n;
  if (stream && o_beg < o_end) load(o_beg, ca, cx, cy, cz, cl);
  if (t < n) {
    double x = xr;
#pragma unroll
    for (int a = 0; a < KD; a++) x -= ub[a][t] * zr[a];
    colr[t] = x;
  }
  __syncthreads();

  // ---- pass 1
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // w, T1[4], T2[4]
  if (c < n) {
    // one row of the pass: the current inverse element from A_0 and the
    // pending terms, then the nine accumulators
    auto row = [&](int o, double a, double x, double y, double z, double l) {
#pragma unroll
      for (int q = 0; q < KD; q++) a -= ub[q][o] * zc[q];
      const double cr = colr[o];
      acc[0] += v[0][o] * a;
      acc[1] += x * a; acc[2] += y * a; acc[3] += z * a; acc[4] += l * a;
      acc[5] += x * cr; acc[6] += y * cr; acc[7] += z * cr; acc[8] += l * cr;
      return a;
    };
    if (stream) {
      // read-only stream (its first U rows are already requested)
      for (int o0 = o_beg; o0 < o_end; o0 += U) {
        if (o0 + U < o_end) load(o0 + U, na, nx, ny, nz, nl);
#pragma unroll
        for (int u = 0; u < U; u++)
          if (o0 + u < o_end)
            row(o0 + u, ca[u], (double)cx[u], (double)cy[u], (double)cz[u], (double)cl[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
          ca[u] = na[u]; cx[u] = nx[u]; cy[u] = ny[u]; cz[u] = nz[u]; cl[u] = nl[u];
        }
      }
    } else {
      // the pending terms are folded into the stored inverse on the way
      double *ip = inv + c;
      for (int o = o_beg; o < o_end; o++) {
        const long off = (long)o * n;
        ip[off] = row(o, ld_stream(ip + off), (double)ld_stream(d1 + off),
                      (double)ld_stream(d2 + off), (double)ld_stream(d3 + off),
                      (double)ld_stream(d4 + off));
      }
    }
  }
  if (sl > 0) {
#pragma unroll
    for (int k = 0; k < 9; k++) red[k][sl - 1][c] = acc[k];
  }
  __syncthreads();
  if (sl == 0) {
#pragma unroll
    for (int q = 1; q < SL; q++)
#pragma unroll
      for (int k = 0; k < 9; k++) acc[k] += red[k][q - 1][c];
    wv[c] = c < n ? acc[0] : 0.0;
  }
  // new derivative row of electron r times the new column r: v_k . colr / ratio
  if (t < 64) {
    double dv[4] = {0, 0, 0, 0};
    for (int o = t; o < n; o += 64)
#pragma unroll
      for (int k = 0; k < 4; k++) dv[k] += v[1 + k][o] * colr[o];
#pragma unroll
    for (int k = 0; k < 4; k++) dv[k] = wave_sum(dv[k]);
    if (t == 0)
#pragma unroll
      for (int k = 0; k < 4; k++) s_bc[4 + k] = dv[k];
  }
  __syncthreads();
  const double ratio = wv[r];
  const bool ok = (ratio != 0.0) && isfinite(ratio);
  const double qinv = 1.0 / ratio;
  if (sl == 0) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      double g = 0.0;
      if (c < n) g = (c == r) ? s_bc[4 + k] * qinv : acc[1 + k] - wv[c] * qinv * acc[5 + k];
      gq[k][c] = g;
    }
  }
  __syncthreads();

  // ---- assembly over all electrons
  const int so = A.other.spin_start, no = A.other.n, ss = A.s.spin_start;
  const double *pUx = A.pUx + (long)b * N * 3;
  const double *cUx = A.Ux + (long)w * N * 3;
  const double *Rw = A.R + (long)w * N * 3;
  const double *gpo = no > 0 ? A.other.gradPR + (long)w * no * 3 : nullptr;
  const double *gps = A.s.gradPR + (long)w * n * 3;
  const double tau = A.tau;
  const bool qf = A.mode >= 1 && A.drift_mod == QMCB_DRIFT_UMRIGAR93_UNEQUAL;
  double lsum = 0.0, fwd = 0.0, rev = 0.0, lapc = 0.0, gsq = 0.0;
  double vbt = 0.0, vbo = 0.0, vo = 0.0;  // |modified drift|^2 (trial, current), |v|^2 current
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
    if (A.mode >= 1) {
      // Green's function terms (QMCWalker.cpp:343-398, 756-780) with the
      // drift gradPsiRatio (QF_modification_type none) or its per-electron
      // umrigar93_unequalelectrons modification (QMCSCFJastrow.cpp:368-443)
      // at the trial (ft) and current (fo) states
      double ft = 1.0, fo = 1.0;
      if (qf) {
        const double xo[3] = {Rw[3 * e], Rw[3 * e + 1], Rw[3 * e + 2]};
        const int co = nearest_nucleus(xo, A.n_nuc, A.nuc_xyz);
        fo = drift_factor_near(gc, xo, co, A.nuc_xyz, A.nuc_z, tau);
        if (e != A.moved) {
          ft = drift_factor_near(g, xo, co, A.nuc_xyz, A.nuc_z, tau);
        } else {
          const double xt[3] = {A.propR[3L * b], A.propR[3L * b + 1], A.propR[3L * b + 2]};
          ft = drift_factor_unequal(g, xt, A.n_nuc, A.nuc_xyz, A.nuc_z, tau);
        }
        double g2 = 0.0, gc2 = 0.0;
        for (int k = 0; k < 3; k++) { g2 += g[k] * g[k]; gc2 += gc[k] * gc[k]; }
        vbt += ft * ft * g2;
        vbo += fo * fo * gc2;
        vo += gc2;
      }
      if (e != A.moved) {
        for (int k = 0; k < 3; k++) {
          fwd += tau * fo * gc[k] * tau * fo * gc[k];
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
  double vals[8] = {lsum, fwd, rev, lapc, gsq, vbt, vbo, vo};
  const int nred = qf ? 8 : 5;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (k >= nred) break;
    const double s = wave_sum(vals[k]);
    if ((t & 63) == 0) wred5[k][t >> 6] = s;
  }
  __syncthreads();
  if (t == 0) {
    for (int k = 0; k < nred; k++) {
      double s = 0.0;
      for (int q = 0; q < NW; q++) s += wred5[k][q];
      s_bc[k] = s;
    }
  }
  __syncthreads();
  // |modified drift| / |drift| of the trial and current states, for the
  // umrigar93 energy modification of the DMC weights
  const double mr_t = (qf && s_bc[4] > 0.0) ? sqrt(s_bc[5]) / sqrt(s_bc[4]) : 1.0;
  const double mr_o = (qf && s_bc[7] > 0.0) ? sqrt(s_bc[6]) / sqrt(s_bc[7]) : 1.0;
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
    // ---- stage the proposal: A'^-1 = A^-1 - colr z^T written in full
    if (sl == 0 && c < n) {
      double *g = A.st_grad + ((long)b * n + c) * 3;
      g[0] = gq[0][c]; g[1] = gq[1][c]; g[2] = gq[2][c];
    }
    if (t == 0) {
      A.st_lap[b] = lap_s;
      A.st_logdet[b] = ld_s + log(fabs(ratio));
      A.st_sgn[b] = sgn_s_new;
      A.st_singular[b] = sing ? 1 : 0;
    }
    if (ok && c < n) {
      const double zcur = (wv[c] - (c == r ? 1.0 : 0.0)) * qinv;
      const double *ip = inv + c;
      double *op = A.st_inv + (long)b * nn2 + c;
      for (int o = o_beg; o < o_end; o++) {
        double a = ip[(long)o * n];
#pragma unroll
        for (int q = 0; q < KD; q++) a -= ub[q][o] * zc[q];
        op[(long)o * n] = a - colr[o] * zcur;
      }
    }
    return;
  }

  // ---- modes 1 (VMC) and 2 (DMC): Metropolis with the Green's function
  // ratio (QMCWalker.cpp:1578-1685): p = |Psi'/Psi|^2 G(R<-R')/G(R'<-R);
  // DMC additionally rejects node crossings (fixed node, :1641-1646)
  const double logpsi_old = ld_s + ld_o + A.U[w];
  const int age_old = A.walk.age ? A.walk.age[w] : 0;
  double p = sing ? 0.0
                  : exp(2.0 * (logpsi_new - logpsi_old) - (s_bc[2] - s_bc[1]) / (2.0 * tau));
  if (!(p == p)) p = 0.0;
  else if (!sing) p = old_walker_boost(p, age_old, A.walk.old_walker);
  if (A.mode == 2 && ratio < 0) p = 0.0;
  if (p > 1.0) p = 1.0;
  const bool accept = p > A.unif[b];
  const int age_new = accept ? 0 : age_old + 1;  // QMCWalker.cpp:1656-1672
  const double el_old = A.cur_el[w];
  __syncthreads();  // every read of the current state is done
  const int base = flush ? 0 : dc;
  if (accept) {
    // append the rank-1 term of this update to the delay buffer
    double *ubo = A.s.ubuf + ((long)w * KD + base) * n;
    double *zbo = A.s.zbuf + ((long)w * KD + base) * n;
    for (int o = t; o < n; o += NT) ubo[o] = colr[o];
    if (sl == 0 && c < n) {
      zbo[c] = (wv[c] - (c == r ? 1.0 : 0.0)) * qinv;
      double *g = A.s.gradPR + ((long)w * n + c) * 3;
      g[0] = gq[0][c]; g[1] = gq[1][c]; g[2] = gq[2][c];
    }
    TS *dmw = (TS *)A.s.dmat + (long)w * kDer * nn2;
    double *dvw = A.s.dval + (long)w * nn2;
    for (int i = t; i < kComp * n; i += NT) {
      const int k = i / n, o = i % n;
      if (k == 0) dvw[(long)o * n + r] = v[0][o];
      else dmw[(long)(k - 1) * nn2 + (long)o * n + r] = (TS)v[k][o];
    }
    double *Uxw = A.Ux + (long)w * N * 3;
    for (int i = t; i < 3 * N; i += NT) Uxw[i] = pUx[i];
    if (t == 0) {
      A.s.dcount[w] = base + 1;
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
  } else if (flush && t == 0) {
    A.s.dcount[w] = 0;
  }
  if (t == 0) {
    const double el = accept ? EL : el_old;
    if (A.walk.age) A.walk.age[w] = age_new;
    double wt = 1.0;
    if (A.mode == 2) {
      // umrigar93_probability_weighted reweighting once per electron sweep
      // (QMCWalker.cpp:1060-1185, UNR93 eq. 25): dW = exp(tau_eff *
      // (p (S'+S)/2 + (1-p) S)), S = E_T - E_L
      wt = A.weight[w];
      if (A.reweight) {
        // energy_modification_type none: S = E_T - E_L; umrigar93 (UNR93
        // eq. 33-34, QMCWalker.cpp:1121-1132): S = (E_T - E_est) +
        // (E_est - E_L) |modified drift| / |drift|
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
}

// ---------------------------------------------------------------------------
// Per-sweep observable row: one workgroup sums the walkers in a fixed order
// (strided per thread, then a tree over the waves), so rows are reproducible.
__global__ void __launch_bounds__(1024)
series_reduce_kernel(const double *obs, int num, double *row) {
  __shared__ double red[3][16];
  const int t = threadIdx.x;
  double se = 0.0, sw = 0.0, sp = 0.0;
  for (int b = t; b < num; b += 1024) {
    const double e = obs[3L * b], w = obs[3L * b + 1], p = obs[3L * b + 2];
    se += w * e;
    sw += w;
    sp += w * p;
  }
  se = wave_sum(se);
  sw = wave_sum(sw);
  sp = wave_sum(sp);
  if ((t & 63) == 0) {
    red[0][t >> 6] = se;
    red[1][t >> 6] = sw;
    red[2][t >> 6] = sp;
  }
  __syncthreads();
  if (t < 3) {
    double s = 0.0;
    for (int q = 0; q < 16; q++) s += red[t][q];
    row[t] = s;
  }
  if (t == 3) row[3] = (double)num;
}

hipError_t launch_series_reduce(const double *obs, int num, double *row, hipStream_t st) {
  hipLaunchKernelGGL(series_reduce_kernel, dim3(1), dim3(1024), 0, st, obs, num, row);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Walker-state copies for branching: one workgroup per (src, dst) pair.
__global__ void __launch_bounds__(256)
copy_walkers_kernel(CopyArgs A, const int *pairs, int np) {
  const int b = blockIdx.x;
  if (b >= np) return;
  const long src = pairs[2 * b], dst = pairs[2 * b + 1];
  for (int k = 0; k < A.narr; k++) {
    const long nw = A.bytes[k] / 4;
    const uint32_t *s = (const uint32_t *)(A.base[k] + src * A.bytes[k]);
    uint32_t *d = (uint32_t *)(A.base[k] + dst * A.bytes[k]);
    for (long i = threadIdx.x; i < nw; i += blockDim.x) d[i] = s[i];
  }
}

hipError_t launch_copy_walkers(const CopyArgs &a, const int *pairs, int np, hipStream_t st) {
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL(copy_walkers_kernel, dim3(np), dim3(256), 0, st, a, pairs, np);
  return hipGetLastError();
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
    hipLaunchKernelGGL((sm_move_kernel<TS, TO, 32, QMCB_SM_NT32>), dim3(num), dim3(QMCB_SM_NT32),
                       0, st, a, walkers, num);
  else if (n <= 64)
    hipLaunchKernelGGL((sm_move_kernel<TS, TO, 64, QMCB_SM_NT64>), dim3(num), dim3(QMCB_SM_NT64),
                       0, st, a, walkers, num);
  else if (n <= 96 && QMCB_SM_TILE96)
    hipLaunchKernelGGL((sm_move_kernel<TS, TO, 96, 192>), dim3(num), dim3(192), 0, st, a,
                       walkers, num);
  else if (n <= 128)
    hipLaunchKernelGGL((sm_move_kernel<TS, TO, 128, 256>), dim3(num), dim3(256), 0, st, a,
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
  TS *dmw = (TS *)A.s.dmat + (long)w * kDer * nn2;
  double *dvw = A.s.dval + (long)w * nn2;
  for (int i = t; i < kComp * n; i += blockDim.x) {
    const int k = i / n, o = i % n;
    const double x = (double)vr[(long)b * A.v_s_walker + (long)k * A.v_s_comp + o];
    if (k == 0) dvw[(long)o * n + r] = x;
    else dmw[(long)(k - 1) * nn2 + (long)o * n + r] = (TS)x;
  }
  double *Uxw = A.Ux + (long)w * N * 3;
  const double *pUx = A.pUx + (long)b * N * 3;
  for (int i = t; i < 3 * N; i += blockDim.x) Uxw[i] = pUx[i];
  if (t == 0) {
    A.s.dcount[w] = 0;
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
INST(float, double)
#undef INST
template hipError_t launch_lu_inverse<double>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_lu_inverse<float>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_traces_all<double>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_traces_all<float>(const SlaterArgs &, const int *, int, hipStream_t);

}  // namespace qmcb
