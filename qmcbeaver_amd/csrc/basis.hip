// basis.hip -- kernel (a): contracted Cartesian Gaussians, value, gradient
// and Laplacian, for a batch of points.
//
// Reference: QMCBasisFunction::evaluateBasisFunctions
// (QMCBasisFunction.cpp:287-403).  Same quantities, different evaluation
// order: primitives are summed once per *shell* (functions of one centre
// sharing exponents and coefficients, e.g. px/py/pz) into the radial sums
//   s0 = sum c e^{-a r^2},  s1 = sum c a e^{-a r^2},  s2 = sum c a^2 e^{-a r^2}
// and each Cartesian component P = x^a y^b z^c is then formed by the product
// rule:
//   chi       = P s0
//   d chi/dx  = (dP/dx) s0 - 2 x P s1
//   lap chi   = (lap P) s0 - (4L+6) P s1 + 4 r^2 P s2
// which equals the reference's per-primitive sum of
// e (a/x - 2 alpha x) and e (a(a-1)/x^2 + ... - alpha(4L+6) + 4 alpha^2 r^2)
// without dividing by coordinates.
//
// Mapping: one lane per point (64 points per wavefront), all lanes walk the
// shell list in lockstep, so shell/primitive parameters are wave-uniform
// scalar loads and every store X[k][bf][p] is a coalesced 64-wide row.
// grid.y splits the shell list into cost-balanced chunks so that a batch of
// a few thousand points (one moved electron per walker) still fills all CUs.
#include "qmcb_kernels.hpp"

namespace qmcb {

template <typename TC>
__device__ __forceinline__ TC dev_exp(TC x);
template <>
__device__ __forceinline__ double dev_exp<double>(double x) { return exp(x); }
template <>
__device__ __forceinline__ float dev_exp<float>(float x) { return __expf(x); }

template <typename TC, typename TS>
__device__ __forceinline__ TS to_store(TC v) { return (TS)v; }
template <>
__device__ __forceinline__ __hip_bfloat16 to_store<float, __hip_bfloat16>(float v) {
  return __float2bfloat16(v);
}
// FP64 basis values rounded to FP32, then to BF16 (both round-to-nearest-
// even), as the host rounds the coefficients
template <>
__device__ __forceinline__ __hip_bfloat16 to_store<double, __hip_bfloat16>(double v) {
  return __float2bfloat16((float)v);
}

// X layout: [5][ldk][ldp]: component, basis function (padded rows zero),
// point (padded columns untouched).
template <typename TC, typename TS>
__global__ void __launch_bounds__(256)
basis_kernel(DevSystem sys, PointSet pts, TS *__restrict__ X, int ldk, long ldp) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = p < pts.P;
  double r[3] = {0.0, 0.0, 0.0};
  if (live) {
    const long row = p / pts.per_walker;
    const long wk = pts.walker_map ? pts.walker_map[row] : row;
    const double *q = pts.base + wk * pts.walker_stride + (p % pts.per_walker) * pts.point_stride;
    r[0] = q[0]; r[1] = q[1]; r[2] = q[2];
  }
  const long cs = (long)ldk * ldp;
  int prev_atom = -1;
  TC d[3] = {0, 0, 0}, rsq = 0;
  const int s_beg = sys.chunk_beg[blockIdx.y], s_end = sys.chunk_beg[blockIdx.y + 1];
  for (int s = s_beg; s < s_end; s++) {
    const int atom = sys.sh_atom[s];
    if (atom != prev_atom) {
      prev_atom = atom;
      // displacement formed in FP64, then rounded: keeps the small
      // electron-nucleus offsets accurate in the FP32 path
      for (int k = 0; k < 3; k++) d[k] = (TC)(r[k] - sys.nuc_xyz[3 * atom + k]);
      rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    }
    const int p0 = sys.sh_prim_off[s], np = sys.sh_nprim[s];
    TC s0 = 0, s1 = 0, s2 = 0;
    for (int i = 0; i < np; i++) {
      const TC al = (TC)sys.prim_alpha[p0 + i];
      const TC e = (TC)sys.prim_coef[p0 + i] * dev_exp<TC>(-al * rsq);
      s0 += e;
      s1 += al * e;
      s2 += al * al * e;
    }
    const int bf0 = sys.sh_bf0[s], nc = sys.sh_ncart[s];
    for (int j = 0; j < nc; j++) {
      const int bf = bf0 + j;
      const int a = sys.bf_pow[3 * bf], b = sys.bf_pow[3 * bf + 1], c = sys.bf_pow[3 * bf + 2];
      // powers and their derivatives for exponents 0..2
      TC px[3], dpx[3], lpx[3];
      const int pw[3] = {a, b, c};
      for (int k = 0; k < 3; k++) {
        const int n = pw[k];
        px[k] = n == 0 ? (TC)1 : (n == 1 ? d[k] : d[k] * d[k]);
        dpx[k] = n == 0 ? (TC)0 : (n == 1 ? (TC)1 : (TC)2 * d[k]);
        lpx[k] = n == 2 ? (TC)2 : (TC)0;
      }
      const TC P = px[0] * px[1] * px[2];
      const TC L = (TC)(a + b + c);
      const TC val = P * s0;
      const TC gx = dpx[0] * px[1] * px[2] * s0 - (TC)2 * d[0] * P * s1;
      const TC gy = px[0] * dpx[1] * px[2] * s0 - (TC)2 * d[1] * P * s1;
      const TC gz = px[0] * px[1] * dpx[2] * s0 - (TC)2 * d[2] * P * s1;
      const TC lapP = lpx[0] * px[1] * px[2] + px[0] * lpx[1] * px[2] + px[0] * px[1] * lpx[2];
      const TC lap = lapP * s0 - ((TC)4 * L + (TC)6) * P * s1 + (TC)4 * rsq * P * s2;
      if (live) {
        TS *o = X + (long)bf * ldp + p;
        o[0] = to_store<TC, TS>(val);
        o[cs] = to_store<TC, TS>(gx);
        o[2 * cs] = to_store<TC, TS>(gy);
        o[3 * cs] = to_store<TC, TS>(gz);
        o[4 * cs] = to_store<TC, TS>(lap);
      }
    }
  }
}

__global__ void scatter_rows_kernel(double *R, const double *Rin, const int *map, int num,
                                   int row_len) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)num * row_len) return;
  const long b = i / row_len, k = i % row_len;
  R[(long)map[b] * row_len + k] = Rin[i];
}

hipError_t launch_scatter_rows(double *R, const double *Rin, const int *map, int num, int row_len,
                               hipStream_t st) {
  const long tot = (long)num * row_len;
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(scatter_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     R, Rin, map, num, row_len);
  return hipGetLastError();
}

// Zero rows [B, ldk) of the five components of an X buffer whose point
// columns are [0, ldp): the padded basis rows the orbital GEMM reads as k
// values beyond B.  A buffer shared by contexts of different systems (the
// no-allocation evaluate's scratch) holds another layout's values there.
// Rows [B, ldk) of one component are contiguous, so the fill is five byte
// ranges of (ldk - B) * ldp * elem bytes, written as 8-byte words.
__global__ void zero_pad_rows_kernel(unsigned long long *X, long comp_words, long first_word,
                                     long pad_words) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 5 * pad_words) return;
  const long c = i / pad_words, k = i % pad_words;
  X[c * comp_words + first_word + k] = 0ull;
}

hipError_t launch_zero_pad_rows(void *X, size_t elem, int B, int ldk, long ldp, hipStream_t st) {
  const size_t comp_bytes = (size_t)ldk * ldp * elem, first = (size_t)B * ldp * elem;
  const size_t pad = (size_t)(ldk - B) * ldp * elem;
  if (pad == 0) return hipSuccess;
  if (comp_bytes % 8 || first % 8 || pad % 8) return hipErrorInvalidValue;
  const long words = (long)(pad / 8), tot = 5 * words;
  hipLaunchKernelGGL(zero_pad_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     (unsigned long long *)X, (long)(comp_bytes / 8), (long)(first / 8), words);
  return hipGetLastError();
}

template <typename TC, typename TS>
hipError_t launch_basis(const DevSystem &sys, const PointSet &pts, TS *X, int ldk, long ldp,
                        hipStream_t st) {
  if (pts.P <= 0) return hipSuccess;
  dim3 grid((unsigned)((pts.P + 255) / 256), (unsigned)sys.nchunk);
  hipLaunchKernelGGL((basis_kernel<TC, TS>), grid, dim3(256), 0, st, sys, pts, X, ldk, ldp);
  return hipGetLastError();
}

template hipError_t launch_basis<double, double>(const DevSystem &, const PointSet &, double *,
                                                 int, long, hipStream_t);
template hipError_t launch_basis<double, float>(const DevSystem &, const PointSet &, float *, int,
                                                long, hipStream_t);
template hipError_t launch_basis<float, float>(const DevSystem &, const PointSet &, float *, int,
                                               long, hipStream_t);
template hipError_t launch_basis<float, __hip_bfloat16>(const DevSystem &, const PointSet &,
                                                        __hip_bfloat16 *, int, long,
                                                        hipStream_t);
template hipError_t launch_basis<double, __hip_bfloat16>(const DevSystem &, const PointSet &,
                                                         __hip_bfloat16 *, int, long,
                                                         hipStream_t);

}  // namespace qmcb
