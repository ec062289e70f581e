// slater.hip -- kernel (c): Slater inverse, determinant, Sherman-Morrison
// single-electron updates and derivative-ratio traces, all FP64.
//
// Reference: QMCSlater::update_Ds / calculate_DerivativeRatios
// (QMCSlater.cpp:390-764), Array2D::determinant_and_inverse (Array2D.h:
// 1521-1548, LU in ludcmp/lubksb :1204-1421) and inverseUpdateOneRow
// (Array2D.h:1566-1610).
//
// Storage (per walker, per spin): inv(o,e) row-major, dval(o,e) = D(e,o) in
// FP64, dmat(o,e,k) = d_k D(e,o) for k in {x, y, z, Laplacian} (transposed,
// the four components of an element together), so the thread that owns
// electron column e of A^-1 also streams column e of the derivative matrices
// with coalesced loads, one per element.
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
// Blocked Gauss-Jordan inverse on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64), 32 < n <= NP, NP in {64, 128}.  It inverts
// B = A^T, whose rows are the stored rows (o, e) of dval, so the loads are
// coalesced (inv(B) = inv(A)^T, det B = det A).  B, padded to NP with the
// identity, stays in registers in the accumulator layout of the instruction
// (element i of lane l of a 16x16 tile: row 4i + l/16, column l%16); wave v
// of the four owns the 16-column tiles v, v + 4, ...  Columns are eliminated
// in panels of 16 with partial pivoting and no row interchanges:
//   1. the wave that owns the panel moves its NP x 16 block to one row per
//      lane (two for NP = 128) and runs the 16 Gauss-Jordan steps of the
//      panel alone: pivot = lowest row index among the largest |column| over
//      rows not yet pivoted (wavefront max, then a ballot), the pivot row read
//      across lanes, every other row reduced.  Pivot rows stay unscaled until
//      the panel ends (a row is linear in its own scale, so only the pivot
//      column entry, 1, and a pending factor 1/pivot differ), which keeps the
//      steps free of branches.  Padded columns (>= n) are eliminated like the
//      others: their pivot is their own identity row and changes nothing.
//      This leaves XK, the panel columns of the final matrix;
//   2. every wave copies the panel's pivot rows P (as they were before the
//      panel) out of its other columns and zeroes them there;
//   3. every other column becomes  B[:, j] <- B[:, j] + XK B_P[:, j]  (four
//      MFMA per 16x16 tile), which is what the 16 sequential steps do to it.
// Two workgroup barriers per panel.  With pivot row p_k for column k the
// final X holds inv(B)[k][p_c] = X[p_k][c], so
//   inv(o, e) = inv(B)[e][o] = X[p_e][c(o)],   c(o): the column pivoted on row o,
// and det B = sign(k -> p_k) * prod_k pivot_k; the sign of the permutation is
// the parity of its inversions, counted per pivot from the mask of rows
// pivoted before it.  The operands of step 3 pass through LDS already in the
// lane order of the instruction (A operand lane l: XK[16 rt + l%16][4 s + l/16];
// B operand lane l: B_P[4 s + l/16][16 ct + l%16]).
typedef double d4v __attribute__((ext_vector_type(4)));

// Largest value over the 64 lanes (all active); every lane gets it.
template <int CTRL>
__device__ __forceinline__ uint32_t max_dpp(uint32_t v) {
  // old value 0 (the identity of max): foldable into one DPP max instruction
  return max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true));
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max_dpp<0xB1>(v);
  v = max_dpp<0x4E>(v);
  v = max_dpp<0x141>(v);
  v = max_dpp<0x140>(v);
  const uint32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
  const uint32_t c = __builtin_amdgcn_readlane(v, 32), e = __builtin_amdgcn_readlane(v, 48);
  return max(max(a, b), max(c, e));
}

// Orders LDS accesses of one wave across its lanes (a compiler barrier; the
// LDS itself serves one wave's accesses in order).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename TS, int NP>
__global__ void __launch_bounds__(256, 2)
gj_mfma_kernel(SlaterArgs A, const int *walkers, int num) {
  constexpr int NT = NP / 16, NWV = 4, CT = NT / NWV, RPL = NP / 64, TILE = 256, MW = NP / 32;
  const int D = det_count(A), d = blockIdx.x % D;
  const int b = blockIdx.x / D;
  if (b >= num) return;
  const int w0 = walkers ? walkers[b] : b;
  const int w = w0 * D + d;  // this determinant's record
  const int n = A.n, nact = act_count(A);
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);  // wave-uniform: scalar branches
  const int lr = lane & 15, lq = lane >> 4;

  // [0, 2) XK fragments by panel parity, [2, 3) pivot-row fragments, [3, 4)
  // dummy slots; after the last panel, each wave's staging area for its part
  // of the output
  __shared__ double lds[4 * NT * TILE];
  __shared__ int pivrow[NP];        // pivot row of column k
  __shared__ int colofp[NP];        // column pivoted on row 16rt + 4i + q, at 16rt + 4q + i
  __shared__ uint32_t s_pm[MW];     // rows pivoted so far
  __shared__ double s_mant[NWV], s_sgn[NWV];
  __shared__ int s_expo[NWV], s_par[NWV];
  __shared__ int s_sing;
  // fragment offset of XK[row][col] (col < 16); pivot row slot k, column j
  // sits at fidx(j, k)
  auto fidx = [](int row, int col) {
    return (row >> 4) * TILE + (col >> 2) * 64 + (col & 3) * 16 + (row & 15);
  };
  double *fp = lds + 2 * NT * TILE;

  if (t < MW) s_pm[t] = 0u;
  if (t == 0) s_sing = 0;
  const double *dv = A.dval + (long)w0 * nact * n;
  d4v acc[NT][CT];
#pragma unroll
  for (int rt = 0; rt < NT; rt++) {
#pragma unroll
    for (int c = 0; c < CT; c++) {
      const int col = 16 * (wv + NWV * c) + lr;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int r = 16 * rt + 4 * i + lq;
        const double v = ld(dv + (long)occ_row(A, d, min(r, n - 1)) * n + min(col, n - 1));
        acc[rt][c][i] = (r < n && col < n) ? v : (r == col ? 1.0 : 0.0);
      }
    }
  }
  __syncthreads();

  double mant = 1.0, sgn = 1.0;  // this wave's pivots: prod |pivot| = mant 2^expo
  int expo = 0, par = 0;         // par: parity of this wave's inversion count
  uint32_t pm[MW];               // rows pivoted before the current panel (uniform)
#pragma unroll
  for (int q = 0; q < MW; q++) pm[q] = 0u;
  bool sing = false;
  // panel K = NWV c + v is column tile K: wave v, its slot c (compile time)
#pragma unroll
  for (int c = 0; c < CT; c++) {
    for (int v = 0; v < NWV; v++) {
      const int K = NWV * c + v, K0 = 16 * K;
      if (sing || K0 >= n) break;
      double *fx = lds + (K & 1) * NT * TILE;
      if (wv == v) {
        // ---- 1. the panel's 16 elimination steps, one or two rows per lane
#pragma unroll
        for (int rt = 0; rt < NT; rt++)
#pragma unroll
          for (int i = 0; i < 4; i++) fx[fidx(16 * rt + 4 * i + lq, lr)] = acc[rt][c][i];
        wave_lds_sync();
        double x[RPL][16], dsc[RPL];
        bool live[RPL];  // row may still become a pivot
#pragma unroll
        for (int s = 0; s < RPL; s++) {
          const int r = lane + 64 * s;
#pragma unroll
          for (int j = 0; j < 16; j++) x[s][j] = fx[fidx(r, j)];
          const uint32_t m = (lane >> 5) ? pm[2 * s + 1] : pm[2 * s];
          live[s] = !((m >> (lane & 31)) & 1u);
          dsc[s] = 1.0;
        }
        wave_lds_sync();
        bool psing = false;
        int pv = 0;  // lane kk: pivot row of column K0 + kk
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
          // pivot: the lowest row index among the live rows whose |x[.][kk]|
          // has the largest upper word (sign, exponent, 20 mantissa bits):
          // within 2^-20 of the largest |x|; an upper word of 0 (below
          // 2^-1022) counts as zero, i.e. singular
          uint32_t key[RPL], best = 0u;
#pragma unroll
          for (int s = 0; s < RPL; s++) {
            const uint32_t hi = (uint32_t)(__builtin_bit_cast(uint64_t, x[s][kk]) >> 32) & 0x7fffffffu;
            key[s] = live[s] ? hi + 1u : 0u;
            best = max(best, key[s]);
          }
          best = wave_max_u32(best);
          psing = psing || best <= 1u;
          int p = 0;
#pragma unroll
          for (int s = RPL - 1; s >= 0; s--) {
            const unsigned long long hit = __ballot(key[s] == best);
            p = hit ? 64 * s + (int)__builtin_ctzll(hit) : p;
          }
          // the pivot row into scalar registers by readlane; a scalar branch
          // picks the row half (the empty asm keeps the compiler from reading
          // both halves and selecting)
          const int pl = p & 63;
          double prow[16];
          if (RPL > 1 && p >= 64) {
            asm volatile("");
#pragma unroll
            for (int j = 0; j < 16; j++) prow[j] = readlane_f64(x[RPL - 1][j], pl);
          } else {
            asm volatile("");
#pragma unroll
            for (int j = 0; j < 16; j++) prow[j] = readlane_f64(x[0][j], pl);
          }
          const double piv = prow[kk];
          sgn = piv < 0 ? -sgn : sgn;
          {
            int e2;
            mant = frexp(mant * fabs(piv), &e2);
            expo += e2;
          }
          // 1/piv: reciprocal estimate and two Newton steps
          double ip = __builtin_amdgcn_rcp(piv);
          ip = fma(ip, fma(-piv, ip, 1.0), ip);
          ip = fma(ip, fma(-piv, ip, 1.0), ip);
          double fi[RPL];
#pragma unroll
          for (int s = 0; s < RPL; s++) fi[s] = lane + 64 * s == p ? 0.0 : x[s][kk] * ip;
#pragma unroll
          for (int j = 0; j < 16; j++) {
            if (j == kk) continue;
#pragma unroll
            for (int s = 0; s < RPL; s++) x[s][j] -= fi[s] * prow[j];
          }
#pragma unroll
          for (int s = 0; s < RPL; s++) {
            const bool isp = lane + 64 * s == p;
            x[s][kk] = isp ? 1.0 : -fi[s];
            dsc[s] = isp ? ip : dsc[s];
            live[s] = live[s] && !isp;
          }
          pv = lane == kk ? p : pv;
        }
        // pending scale of the pivot rows, then XK: the MFMA A operand and
        // this wave's panel tile
#pragma unroll
        for (int s = 0; s < RPL; s++)
#pragma unroll
          for (int kk = 0; kk < 16; kk++) fx[fidx(lane + 64 * s, kk)] = x[s][kk] * dsc[s];
        // inversions of k -> p_k: rows pivoted earlier that lie above p_k
        {
          const int pw = pv >> 5, pb = pv & 31;
          int cnt = 0;
#pragma unroll
          for (int q = 0; q < MW; q++)
            cnt += q > pw ? __builtin_popcount(pm[q]) : (q == pw ? __builtin_popcount(pm[q] >> pb) : 0);
          uint32_t nm[MW];
#pragma unroll
          for (int q = 0; q < MW; q++) nm[q] = pm[q];
#pragma unroll
          for (int k2 = 0; k2 < 16; k2++) {
            const int p2 = __builtin_amdgcn_readlane(pv, k2);
            cnt += (k2 < lane && p2 > pv) ? 1 : 0;
#pragma unroll
            for (int q = 0; q < MW; q++) nm[q] |= (p2 >> 5) == q ? 1u << (p2 & 31) : 0u;
          }
          par ^= __popcll(__ballot(lane < 16 && (cnt & 1))) & 1;
          if (lane < MW) {
            uint32_t m = nm[0];
#pragma unroll
            for (int q = 1; q < MW; q++) m = lane == q ? nm[q] : m;
            s_pm[lane] = m;
          }
        }
        if (lane < 16) {
          pivrow[K0 + lane] = pv;
          colofp[(pv & ~15) | ((pv & 3) << 2) | ((pv >> 2) & 3)] = K0 + lane;
        }
        if (lane == 0 && psing) s_sing = 1;
        wave_lds_sync();
#pragma unroll
        for (int rt = 0; rt < NT; rt++)
#pragma unroll
          for (int i = 0; i < 4; i++) acc[rt][c][i] = fx[fidx(16 * rt + 4 * i + lq, lr)];
      }
      __syncthreads();
      if (s_sing) { sing = true; break; }
      // ---- 2. pivot rows out of the other columns (zeroed there)
      uint32_t nm[MW];  // this panel's pivot rows
#pragma unroll
      for (int q = 0; q < MW; q++) {
        const uint32_t m = __builtin_amdgcn_readfirstlane(s_pm[q]);
        nm[q] = m & ~pm[q];
        pm[q] = m;
      }
      // LDS writes only for the groups of four rows that hold a pivot (lanes
      // on other rows of such a group write to a dummy slot of their own);
      // the zeroing is a select over every element, outside any branch, so
      // that the accumulator tiles are never modified under control flow
#pragma unroll
      for (int rt = 0; rt < NT; rt++) {
        const uint32_t half = (nm[rt >> 1] >> (16 * (rt & 1))) & 0xFFFFu;
        if (!half) continue;
        const int4 cof = *reinterpret_cast<const int4 *>(colofp + 16 * rt + 4 * lq);
        const int cv[4] = {cof.x, cof.y, cof.z, cof.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const uint32_t nib = (half >> (4 * i)) & 0xFu;
          if (!nib) continue;
          const bool mine = (nib >> lq) & 1u;
          const int k = mine ? cv[i] - K0 : 0;  // colofp is valid on the pivot rows only
#pragma unroll
          for (int c2 = 0; c2 < CT; c2++) {
            const int ct = wv + NWV * c2;
            if (ct == K || 16 * ct >= n) continue;
            lds[mine ? 2 * NT * TILE + fidx(16 * ct + lr, k) : 3 * NT * TILE + lane] = acc[rt][c2][i];
          }
        }
      }
#pragma unroll
      for (int c2 = 0; c2 < CT; c2++) {
        const bool other = wv + NWV * c2 != K;
#pragma unroll
        for (int rt = 0; rt < NT; rt++) {
          const uint32_t bits = (nm[rt >> 1] >> (16 * (rt & 1) + lq)) & 0x1111u;
#pragma unroll
          for (int i = 0; i < 4; i++)
            acc[rt][c2][i] = (other && ((bits >> (4 * i)) & 1u)) ? 0.0 : acc[rt][c2][i];
        }
      }
      __syncthreads();
      // ---- 3. rank-16 update of every other column
      double bf[CT][4];
#pragma unroll
      for (int c2 = 0; c2 < CT; c2++)
#pragma unroll
        for (int s = 0; s < 4; s++) bf[c2][s] = fp[(wv + NWV * c2) * TILE + s * 64 + lane];
#pragma unroll
      for (int s = 0; s < 4; s++) {
#pragma unroll
        for (int rt = 0; rt < NT; rt++) {
          if (16 * rt >= n) continue;
          const double af = fx[rt * TILE + s * 64 + lane];
#pragma unroll
          for (int c2 = 0; c2 < CT; c2++) {
            const int ct = wv + NWV * c2;
            if (ct != K && 16 * ct < n)
              acc[rt][c2] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf[c2][s], acc[rt][c2], 0, 0, 0);
          }
        }
      }
    }
  }
  // ---- determinant: the waves' pivot products and inversion parities
  if (lane == 0) { s_mant[wv] = mant; s_expo[wv] = expo; s_sgn[wv] = sgn; s_par[wv] = par; }
  __syncthreads();
  if (!sing) {
    // ---- inverse out: inv(o, e) = X[p_e][c(o)]; for column c of X, o = p_c.
    // Each wave stages its 16-column tiles transposed (rows swizzled) in its
    // own LDS area and writes the rows o of A^-1, coalesced.
    int pe[RPL];
#pragma unroll
    for (int s = 0; s < RPL; s++) pe[s] = pivrow[min(lane + 64 * s, n - 1)];
    double *inv = A.inv + (long)w * n * n;
    double *st = lds + wv * 16 * NP;
#pragma unroll
    for (int q = 0; q < CT; q++) {
      const int c0 = 16 * (wv + NWV * q);
      if (c0 >= n) break;
      if (q > 0) wave_lds_sync();
#pragma unroll
      for (int rt = 0; rt < NT; rt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int r = 16 * rt + 4 * i + lq;
          st[lr * NP + (r ^ (lr << 2))] = acc[rt][q][i];
        }
      wave_lds_sync();
      const int4 *pv4 = reinterpret_cast<const int4 *>(pivrow + c0);
      int ov[16];
#pragma unroll
      for (int m4 = 0; m4 < 4; m4++) {
        const int4 v = pv4[m4];
        ov[4 * m4] = v.x; ov[4 * m4 + 1] = v.y; ov[4 * m4 + 2] = v.z; ov[4 * m4 + 3] = v.w;
      }
#pragma unroll
      for (int m = 0; m < 16; m++) {
        if (c0 + m < n) {
#pragma unroll
          for (int s = 0; s < RPL; s++) {
            const int e = lane + 64 * s;
            if (e < n) inv[(long)ov[m] * n + e] = st[m * NP + (pe[s] ^ (m << 2))];
          }
        }
      }
    }
  }
  if (t == 0) {
    double m = 1.0, sg = 1.0;
    int ex = 0, pr = 0;
    for (int q = 0; q < NWV; q++) {
      int e2;
      m = frexp(m * s_mant[q], &e2);
      ex += e2 + s_expo[q];
      sg *= s_sgn[q];
      pr ^= s_par[q];
    }
    A.logdet[w] = sing ? 0.0 : log(m) + ex * 0.6931471805599453;
    A.sgn[w] = sing ? 0.0 : (pr ? -sg : sg);
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
    hipLaunchKernelGGL((gj_mfma_kernel<TS, 64>), dim3(g), dim3(256), 0, st, a, walkers, num);
  else if (a.n <= 128)
    hipLaunchKernelGGL((gj_mfma_kernel<TS, 128>), dim3(g), dim3(256), 0, st, a, walkers, num);
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
      g0 += ld(dm + dm_at<TS>(0, ro, cs)) * iv;
      g1 += ld(dm + dm_at<TS>(1, ro, cs)) * iv;
      g2 += ld(dm + dm_at<TS>(2, ro, cs)) * iv;
      l += ld(dm + dm_at<TS>(3, ro, cs)) * iv;
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
// flight per lane).  For the FP64 128-column tile the deeper pipeline raises
// the register count to about 190 (2 waves per SIMD) with the streaming rate
// unchanged, which leaves wave slots to the basis and GEMM kernels of the
// other walker groups running beside sm_move (+8 % walker-moves/s with three
// groups at 256 electrons, equal with one group).
#ifndef QMCB_SM_UNROLL_F32
#define QMCB_SM_UNROLL_F32 4
#endif
#ifndef QMCB_SM_UNROLL_F64
#define QMCB_SM_UNROLL_F64 6
#endif
#ifndef QMCB_SM_UNROLL_SMALL
#define QMCB_SM_UNROLL_SMALL 2
#endif
// BF16 derivatives (16 bytes per row and lane in two 8-byte loads): the FP32
// depth measured best (9.41 M walker-moves/s at 256 electrons against 9.21 /
// 8.97 M at 6 / 8 rows)
#ifndef QMCB_SM_UNROLL_BF16
#define QMCB_SM_UNROLL_BF16 4
#endif
template <typename TS, int NC>
__host__ __device__ constexpr int sm_unroll() {
  return NC < 128 ? QMCB_SM_UNROLL_SMALL
                  : (sizeof(TS) == 4 ? QMCB_SM_UNROLL_F32
                                     : (sizeof(TS) == 2 ? QMCB_SM_UNROLL_BF16 : QMCB_SM_UNROLL_F64));
}
// waves per SIMD the register allocation must allow: four everywhere (at
// most 128 VGPRs, no spills), two for the FP64 128-column tile, whose deeper
// unroll holds about 200
// (the FP32 128-column tile at four waves and 4 rows per stage measured
// 8.42-8.44 M walker-moves/s with three groups at 256 electrons, against
// 8.32-8.45 M for three or two waves with 4, 6 or 8 rows)
#ifndef QMCB_SM_WAVES_F32
#define QMCB_SM_WAVES_F32 4
#endif
template <typename TS, int NC>
__host__ __device__ constexpr int sm_waves() {
  return NC == 128 ? (sizeof(TS) == 8 ? 2 : (sizeof(TS) == 4 ? QMCB_SM_WAVES_F32 : 4)) : 4;
}
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
template <>
__device__ __forceinline__ bf16s ld_stream<bf16s>(const bf16s *p) {
  bf16s v;
#if QMCB_SM_NT
  v.b = __builtin_nontemporal_load(reinterpret_cast<const uint16_t *>(p));
#else
  v.b = p->b;
#endif
  return v;
}

// the four derivative components of one element, adjacent in the
// interleaved layout, with one store (two for FP64); FP32 and BF16 ones
// non-temporal, since a column write touches one line per row and nothing
// reads it back before the line is evicted (measured faster; for the two
// FP64 stores per element, slower)
__device__ __forceinline__ void st4(float *p, double x, double y, double z, double l) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  __builtin_nontemporal_store(f4{(float)x, (float)y, (float)z, (float)l},
                              reinterpret_cast<f4 *>(p));
}
__device__ __forceinline__ void st4(double *p, double x, double y, double z, double l) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  reinterpret_cast<d2v *>(p)[0] = d2v{x, y};
  reinterpret_cast<d2v *>(p)[1] = d2v{z, l};
}
__device__ __forceinline__ void st4(bf16s *p, double x, double y, double z, double l) {
  const uint64_t q = (uint64_t)bf16s(x).b | ((uint64_t)bf16s(y).b << 16) |
                     ((uint64_t)bf16s(z).b << 32) | ((uint64_t)bf16s(l).b << 48);
  __builtin_nontemporal_store(q, reinterpret_cast<uint64_t *>(p));
}

template <typename TS, typename TO, int NC, int NT>
__global__ void
__attribute__((amdgpu_flat_work_group_size(NT, NT), amdgpu_waves_per_eu(sm_waves<TS, NC>())))
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
  __shared__ double s_ux[NT][3];  // jastrow_fused: electron t's Jastrow gradient at the proposal
  __shared__ double wj[7][NW];    // jastrow_fused: per-wave sums

  const TO *vr = (const TO *)A.vrow + (long)b * A.v_s_walker;
  double *inv = A.s.inv + (long)w * nn2;
  const double *ubw = A.s.ubuf + (long)w * KD * n;
  const double *zbw = A.s.zbuf + (long)w * KD * n;
  for (int i = t; i < kComp * n; i += NT) {
    const int k = i / n, o = i % n;
    const double x = (double)vr[(long)k * A.v_s_comp + o];
    v[k][o] = k == 0 ? x : as_stored<TS>(x);
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
  if (A.jastrow_fused) {
    // the Jastrow factor and potential at the proposal, the terms of
    // jastrow_move_kernel (N <= NT): thread t takes electron t's pair with
    // the moved electron at its old and new position, and nucleus t; the
    // sums dU, dUxx, dVee, dVen, grad_m U go to wj per wave, electron t's
    // new gradient to s_ux (before the stream, whose in-flight rows would
    // otherwise hold registers across it)
    const int m = A.moved;
    const double *Rw = A.R + (long)w * N * 3;
    const double xo[3] = {Rw[3 * m], Rw[3 * m + 1], Rw[3 * m + 2]};
    const double xn[3] = {A.propR[3L * b], A.propR[3L * b + 1], A.propR[3L * b + 2]};
    double js[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (t < N && t != m) {
      double dn[3], dold[3], rn2 = 0.0, ro2 = 0.0;
      for (int k = 0; k < 3; k++) {
        dn[k] = xn[k] - Rw[3 * t + k]; rn2 += dn[k] * dn[k];
        dold[k] = xo[k] - Rw[3 * t + k]; ro2 += dold[k] * dold[k];
      }
      const double rn = sqrt(rn2), ro = sqrt(ro2), irn = 1.0 / rn, iro = 1.0 / ro;
      const qmcb_pade &f = A.ee[ee_type(m, t, A.n_up)];
      double un, dun, d2un, uo, duo, d2uo;
      pade_eval(f, rn, un, dun, d2un);
      pade_eval(f, ro, uo, duo, d2uo);
      js[0] = un - uo;
      js[1] = 2.0 * (2.0 * irn * dun + d2un) - 2.0 * (2.0 * iro * duo + d2uo);
      js[2] = irn - iro;
      const double gn = dun * irn, go = duo * iro;
      const double *uxw = A.Ux + (long)w * N * 3;
      for (int k = 0; k < 3; k++) {
        s_ux[t][k] = uxw[3 * t + k] - gn * dn[k] + go * dold[k];
        js[4 + k] = gn * dn[k];
      }
    }
    for (int a_ = t; a_ < A.n_nuc; a_ += NT) {
      double dn[3], dold[3], rn2 = 0.0, ro2 = 0.0;
      for (int k = 0; k < 3; k++) {
        dn[k] = xn[k] - A.nuc_xyz[3 * a_ + k]; rn2 += dn[k] * dn[k];
        dold[k] = xo[k] - A.nuc_xyz[3 * a_ + k]; ro2 += dold[k] * dold[k];
      }
      const double rn = sqrt(rn2), ro = sqrt(ro2), irn = 1.0 / rn, iro = 1.0 / ro;
      double un, dun, d2un, uo, duo, d2uo;
      pade_eval(A.en[a_], rn, un, dun, d2un);
      pade_eval(A.en[a_], ro, uo, duo, d2uo);
      js[0] += un - uo;
      js[1] += (2.0 * irn * dun + d2un) - (2.0 * iro * duo + d2uo);
      js[3] += A.nuc_z[a_] * (iro - irn);
      const double gn = dun * irn;
      for (int k = 0; k < 3; k++) js[4 + k] += gn * dn[k];
    }
#pragma unroll
    for (int k = 0; k < 7; k++) {
      const double s = wave_sum(js[k]);
      if ((t & 63) == 0) wj[k][t >> 6] = s;
    }
  }
  __syncthreads();

  // ---- pass 1 set-up: the read-only stream's first U rows are requested
  // before column r is formed, so their latency overlaps it and the barrier
  const TS *dm = (const TS *)A.s.dmat + (long)w * kDer * nn2;
  const TS *__restrict__ d1 = dm + dm_at<TS>(0, c, nn2);
  const TS *__restrict__ d2 = dm + dm_at<TS>(1, c, nn2);
  const TS *__restrict__ d3 = dm + dm_at<TS>(2, c, nn2);
  const TS *__restrict__ d4 = dm + dm_at<TS>(3, c, nn2);
  // element (o, c) of the four components: 4 consecutive values at d1 + 4 o n
  (void)d2; (void)d3; (void)d4;
  auto ld4 = [&](long off, TS &x, TS &y, TS &z, TS &l) {
    if constexpr (std::is_same<TS, bf16s>::value) {
      const uint64_t q = ld_stream(reinterpret_cast<const uint64_t *>(d1 + 4 * off));
      x.b = (uint16_t)q;
      y.b = (uint16_t)(q >> 16);
      z.b = (uint16_t)(q >> 32);
      l.b = (uint16_t)(q >> 48);
    } else if constexpr (std::is_same<TS, float>::value) {
      typedef float f4 __attribute__((ext_vector_type(4)));
      const f4 q = ld_stream(reinterpret_cast<const f4 *>(d1 + 4 * off));
      x = q[0]; y = q[1]; z = q[2]; l = q[3];
    } else {
      typedef double d2v __attribute__((ext_vector_type(2)));
      const d2v q0 = ld_stream(reinterpret_cast<const d2v *>(d1 + 4 * off));
      const d2v q1 = ld_stream(reinterpret_cast<const d2v *>(d1 + 4 * off + 2));
      x = q0[0]; y = q0[1]; z = q1[0]; l = q1[1];
    }
  };
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
      ld4(off, x[u], y[u], z[u], l[u]);
    }
  };
  // every move streams the rows through the same register pipeline; a flush
  // (8 pending terms) also writes each folded row of A^-1 back
  const bool stream = c < n;
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
    // the stream (its first U rows are already requested); the loads of the
    // next U rows are issued before this batch's rows are folded and, on a
    // flush, stored, so the stores never wait on them
    double *ip = inv + c;
    for (int o0 = o_beg; o0 < o_end; o0 += U) {
      if (o0 + U < o_end) load(o0 + U, na, nx, ny, nz, nl);
#pragma unroll
      for (int u = 0; u < U; u++)
        if (o0 + u < o_end) {
          const double a =
              row(o0 + u, ca[u], (double)cx[u], (double)cy[u], (double)cz[u], (double)cl[u]);
          if (flush) ip[(long)(o0 + u) * n] = a;
        }
#pragma unroll
      for (int u = 0; u < U; u++) {
        ca[u] = na[u]; cx[u] = nx[u]; cy[u] = ny[u]; cz[u] = nz[u]; cl[u] = nl[u];
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
  // jastrow_fused: sum k of the Jastrow terms at the proposal (dU, dUxx,
  // dVee, dVen, grad_m U), read where used so none stays live in registers
  auto jsum = [&](int k) {
    double s = 0.0;
    for (int q = 0; q < NW; q++) s += wj[k][q];
    return s;
  };
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
      const double ux = !A.jastrow_fused ? pUx[3 * e + k]
                                         : (e == A.moved ? jsum(4 + k) : s_ux[e][k]);
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
  const double pU_b = A.jastrow_fused ? A.U[w] + jsum(0) : A.pU[b];
  const double pUxx_b = A.jastrow_fused ? A.Uxx[w] + jsum(1) : A.pUxx[b];
  const double pVee_b = A.jastrow_fused ? A.Vee[w] + jsum(2) : A.pVee[b];
  const double pVen_b = A.jastrow_fused ? A.Ven[w] + jsum(3) : A.pVen[b];
  const double lapR = lap_s + lap_o + pUxx_b + s_bc[0];
  const double V = A.v_nn + pVee_b + pVen_b;
  const double EL = -0.5 * lapR + V;
  const double ld_s = A.s.logdet[w], ld_o = no > 0 ? A.other.logdet[w] : 0.0;
  const double logpsi_new = ld_s + log(fabs(ratio)) + ld_o + pU_b;
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
    if (o.v_ne) o.v_ne[b] = pVen_b;
    if (o.v_ee) o.v_ee[b] = pVee_b;
    if (o.det_ratio) o.det_ratio[b] = ratio;
    if (o.jastrow_u) o.jastrow_u[b] = pU_b;
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
    // column r of the value and derivative matrices: one thread per row o
    // writes the element's four derivative components with one vector store
    // (they are adjacent), not four threads with a scalar store each
    TS *dmw = (TS *)A.s.dmat + (long)w * kDer * nn2;
    double *dvw = A.s.dval + (long)w * nn2;
    for (int o = t; o < n; o += NT) {
      __builtin_nontemporal_store(v[0][o], dvw + (long)o * n + r);
      st4(dmw + dm_at<TS>(0, (long)o * n + r, nn2), v[1][o], v[2][o], v[3][o], v[4][o]);
    }
    double *Uxw = A.Ux + (long)w * N * 3;
    if (A.jastrow_fused) {
      if (t < N)
        for (int k = 0; k < 3; k++) Uxw[3 * t + k] = t == A.moved ? jsum(4 + k) : s_ux[t][k];
    } else {
      for (int i = t; i < 3 * N; i += NT) Uxw[i] = pUx[i];
    }
    if (t == 0) {
      A.s.dcount[w] = base + 1;
      A.s.lapPR[w] = lap_s;
      A.s.logdet[w] = ld_s + log(fabs(ratio));
      A.s.sgn[w] = sgn_s_new;
      double *Rm = A.R + ((long)w * N + A.moved) * 3;
      Rm[0] = A.propR[3L * b]; Rm[1] = A.propR[3L * b + 1]; Rm[2] = A.propR[3L * b + 2];
      A.U[w] = pU_b;
      A.Uxx[w] = pUxx_b;
      A.Vee[w] = pVee_b;
      A.Ven[w] = pVen_b;
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
    if (A.propose_next) {
      // the next step's proposal from the values this move leaves behind,
      // the ones the next step's propose_kernel would read: electron m2's
      // position, its ln D gradient (this spin's gq if accepted) and its
      // Jastrow gradient (the proposal's if accepted)
      const ProposeArgs &P = A.next;
      const int m2 = P.moved;
      const bool moved2 = m2 == A.moved;
      double x[3], gd[3], ux[3], vd[3];
      for (int k = 0; k < 3; k++) x[k] = (moved2 && accept) ? A.propR[3L * b + k] : Rw[3 * m2 + k];
      if (m2 >= ss && m2 < ss + n) {
        for (int k = 0; k < 3; k++) gd[k] = accept ? gq[k][m2 - ss] : gps[3 * (m2 - ss) + k];
      } else {
        for (int k = 0; k < 3; k++) gd[k] = gpo[3 * (m2 - so) + k];
      }
      for (int k = 0; k < 3; k++)
        ux[k] = !accept ? cUx[3 * m2 + k]
                        : (A.jastrow_fused ? (moved2 ? jsum(4 + k) : s_ux[m2][k]) : pUx[3 * m2 + k]);
      for (int k = 0; k < 3; k++) vd[k] = gd[k] + ux[k];
      propose_row(P, b, w, vd, x);
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

// threads of the tile that serves n electrons of the moved spin
static int sm_move_threads(int n) {
  if (n <= 16) return 64;
  if (n <= 32) return QMCB_SM_NT32;
  if (n <= 64) return QMCB_SM_NT64;
  if (n <= 96 && QMCB_SM_TILE96) return 192;
  return 256;
}

bool sm_move_fuses_jastrow(int n, int N) {
  return n > 0 && n <= 128 && N <= sm_move_threads(n);
}

template <typename TS, typename TO>
hipError_t launch_sm_move(const MoveArgs &a, const int *walkers, int num, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  const int n = a.s.n;
  if (n <= 0) return hipErrorInvalidValue;
  if (a.jastrow_fused && !sm_move_fuses_jastrow(n, a.N)) return hipErrorInvalidValue;
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
    else dmw[dm_at<TS>(k - 1, (long)o * n + r, nn2)] = (TS)x;
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
INST(bf16s, float)
#undef INST
template hipError_t launch_lu_inverse<double>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_lu_inverse<float>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_traces_all<double>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_traces_all<float>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_lu_inverse<bf16s>(const SlaterArgs &, const int *, int, hipStream_t);
template hipError_t launch_traces_all<bf16s>(const SlaterArgs &, const int *, int, hipStream_t);

}  // namespace qmcb
