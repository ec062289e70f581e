// cusp.hip -- electron-nucleus cusp replacement applied to the orbital rows.
//
// After the orbital GEMM, every point (electron position) closer to nucleus A
// than the fitted radius rc(A, m) gets, for orbital m, the replacement minus
// the original s-type part added to its value, gradient and Laplacian
// (QMCElectronNucleusCusp.cpp:587-640 replaceCusps;
// QMCElectronNucleusCuspParameters.cpp:104-253 replaceOrbitalValues,
// evaluateOriginalOrbital, evaluateReplacementOrbital).  One workgroup per
// point, one thread per orbital: the nucleus loop is uniform over the
// workgroup and, outside every rc, costs one distance per nucleus.
#include "qmcb_cusp.hpp"
#include "qmcb_kernels.hpp"

namespace qmcb {

template <typename TO>
__global__ void __launch_bounds__(256)
cusp_kernel(CuspDev c, PointSet pts, GemmOut o) {
  const long p = blockIdx.x;
  if (p >= pts.P) return;
  const long row = p / pts.per_walker;
  const long pw = pts.walker_map ? pts.walker_map[row] : row;
  const double *q = pts.base + pw * pts.walker_stride + (p % pts.per_walker) * pts.point_stride;
  const double x = q[0], y = q[1], z = q[2];
  const long wk = o.walker_map ? o.walker_map[row] : row;
  const long pe = (p % o.per_walker) * o.s_elec;
  for (int m = threadIdx.x; m < c.norb; m += blockDim.x) {
    double dv = 0.0, dgx = 0.0, dgy = 0.0, dgz = 0.0, dl = 0.0;
    bool hit = false;
    for (int a = 0; a < c.nn; a++) {
      const double dx = x - c.nuc_xyz[3 * a], dy = y - c.nuc_xyz[3 * a + 1],
                   dz = z - c.nuc_xyz[3 * a + 2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double rcm = c.rc_max[a];
      if (!(r2 < rcm * rcm)) continue;
      const double r = sqrt(r2);
      const double *f = c.orb8 + ((long)a * c.norb + m) * 8;
      if (!(r < f[0])) continue;
      hit = true;
      // original s-type part on this nucleus
      const double *cf = c.s_coef + (long)m * c.ns;
      for (int g = c.s_off[a]; g < c.s_off[a + 1]; g++) {
        const double a0 = c.s_alpha[g];
        const double t = cf[g] * exp(-a0 * r2);
        const double tg = -2 * a0 * t;
        dv -= t;
        dgx -= dx * tg;
        dgy -= dy * tg;
        dgz -= dz * tg;
        dl -= 2 * a0 * (2 * a0 * r2 - 3) * t;
      }
      // replacement C + sgn0 exp(alpha(r)); alpha and two derivatives by Horner
      double pf = f[7], pd = 0.0, pd2 = 0.0;
      for (int i = 3; i >= 0; i--) {
        pd2 = pd2 * r + pd;
        pd = pd * r + pf;
        pf = pf * r + f[3 + i];
      }
      pd2 *= 2;
      const double e = f[2] * exp(pf);
      const double tg = e * pd / r;
      dv += f[1] + e;
      dgx += tg * dx;
      dgy += tg * dy;
      dgz += tg * dz;
      dl += e * (pd2 + pd * (2.0 / r + pd));
    }
    if (!hit) continue;
    const double d[kComp] = {dv, dgx, dgy, dgz, dl};
#pragma unroll
    for (int k = 0; k < kComp; k++) {
      if (k == 0 && o.out_val) {
        double *ov = o.out_val + wk * o.s_walker_val + (p % o.per_walker) * o.s_elec_val +
                     (long)m * o.s_orb_val;
        *ov += d[0];
        continue;
      }
      const int kc = o.out_val ? k - 1 : k;
      TO *dst = (TO *)o.out + (long)kc * o.s_comp + wk * o.s_walker + pe + (long)m * o.s_orb;
      *dst = (TO)((double)*dst + d[k]);
    }
  }
}

template <typename TO>
hipError_t launch_cusp(const CuspDev &c, const PointSet &pts, const GemmOut &o, hipStream_t st) {
  if (pts.P <= 0 || c.norb <= 0 || c.nn <= 0) return hipSuccess;
  const int threads = c.norb > 128 ? 256 : (c.norb > 64 ? 128 : 64);
  hipLaunchKernelGGL((cusp_kernel<TO>), dim3((unsigned)pts.P), dim3(threads), 0, st, c, pts, o);
  return hipGetLastError();
}

template hipError_t launch_cusp<double>(const CuspDev &, const PointSet &, const GemmOut &,
                                        hipStream_t);
template hipError_t launch_cusp<float>(const CuspDev &, const PointSet &, const GemmOut &,
                                       hipStream_t);
template hipError_t launch_cusp<bf16s>(const CuspDev &, const PointSet &, const GemmOut &,
                                       hipStream_t);

}  // namespace qmcb
