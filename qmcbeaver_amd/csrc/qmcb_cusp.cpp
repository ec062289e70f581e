// qmcb_cusp.cpp -- electron-nucleus cusp replacement: host-side fit of the
// replacement orbitals (set-up time, once per context and spin).
//
// Near nucleus A, the part of molecular orbital m that comes from the
// s-type Gaussians centred on A, phi_s(r) = sum_g c_g C(m, bf_g) e^{-a_g r^2},
// is replaced inside r < rc by  C + sgn0 * exp(p(r)),  p a quartic fitted so
// that value, first and second derivative match phi_s at rc, the slope at
// r = 0 satisfies the nuclear cusp, and the value at the nucleus phi0 is the
// one (searched by golden section) that minimises the largest squared
// deviation of the replacement's one-electron local energy from an "ideal"
// curve.  The device kernel then adds (replacement - phi_s) and its
// derivatives to the Slater rows of electrons inside rc.
//
// Reference: QMCElectronNucleusCusp.cpp:46-177 (fitReplacementOrbitals: the
// s-type list, n0 from the other atoms' functions at the nucleus),
// :179-418 (determineRc: ideal curve, node bracketing and bisection, scan
// for the first deviation > Z^2/50), :420-585 (fitOrbitalParameters: golden
// section over phi0); QMCElectronNucleusCuspParameters.cpp:334-454
// (fitReplacementOrbital: sign, shift C, Zeff, the five matching
// conditions), :456-515 (calculateSigmaSq, excluding +-rc/5 around a node of
// the replacement), :517-570 (findMaxDeviation), :255-300
// (calculateLocalEnergy); Polynomial.cpp:127-149 (Horner for value and two
// derivatives).  The root finder for the replacement's node is this file's
// own (Aberth iteration on the quartic), selecting, like the reference, the
// smallest real root inside (0, rc).
#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "qmcb_cusp.hpp"

namespace qmcb {
namespace {

// value, first and second derivative of sum_i c[i] x^i (Horner, highest first)
struct Poly {
  std::vector<double> c;
  double f = 0, df = 0, d2f = 0;
  void eval(double x) {
    const int n = (int)c.size() - 1;
    if (n < 0) { f = df = d2f = 0; return; }
    f = c[n];
    df = 0.0;
    d2f = 0.0;
    for (int i = n - 1; i >= 0; i--) {
      d2f = d2f * x + df;
      df = df * x + f;
      f = f * x + c[i];
    }
    d2f *= 2;
  }
  double value(double x) { eval(x); return f; }
};

// The s-type Gaussians of one nucleus, weighted by one orbital's coefficients.
struct SPart {
  std::vector<double> a, c;  // exponent, coefficient x MO coefficient
  double value(double r) const {
    const double r2 = r * r;
    double v = 0.0;
    for (size_t i = 0; i < a.size(); i++) v += c[i] * std::exp(-a[i] * r2);
    return v;
  }
  double gradient(double r) const {
    const double r2 = r * r;
    double g = 0.0;
    for (size_t i = 0; i < a.size(); i++) g += -a[i] * 2 * r * c[i] * std::exp(-a[i] * r2);
    return g;
  }
  // one-electron local energy of phi_s in the field of a bare nucleus Z
  double local_energy(double r, int Z) const {
    const double r2 = r * r;
    double v = 0.0, lap = 0.0;
    for (size_t i = 0; i < a.size(); i++) {
      const double t = c[i] * std::exp(-a[i] * r2);
      v += t;
      lap += 2 * a[i] * (2 * a[i] * r2 - 3) * t;
    }
    lap = lap / v;
    return -0.5 * lap - Z / r;
  }
};

struct Fit {
  // inputs
  const SPart *s = nullptr;
  double n0 = 0, rc = 0;
  int Z = 0;
  Poly ideal;
  // result of fit(phi0)
  double phi0 = 0, C = 0, zeff = 0, sigma_sq = -1;
  int sgn0 = 0;
  bool noC = true;
  Poly alpha;

  double rep_local_energy(double r, bool at_zero) {
    if (at_zero) {
      alpha.eval(0.0);
      if (noC) return -0.5 * (alpha.d2f + alpha.df * alpha.df);
      const double e = sgn0 * std::exp(alpha.f);
      return (-0.5 * e / (C + e)) * (alpha.d2f + alpha.df * alpha.df);
    }
    alpha.eval(r);
    const double ke = -0.5 * (alpha.d2f + alpha.df * (2.0 / r + alpha.df));
    if (noC) return ke - zeff / r;
    const double e = sgn0 * std::exp(alpha.f);
    return ke * e / (C + e) - zeff / r;
  }

  double sq_dev(double r, bool at_zero) {
    const double rep = rep_local_energy(r, at_zero);
    const double id = Z * Z * ideal.value(r);
    return (rep - id) * (rep - id);
  }

  // golden-section search for the largest squared deviation on [lo, hi]
  double max_deviation(double lower, double upper, bool lower_zero) {
    const double G = 0.61803399;
    double lowR = lower, lowV = sq_dev(lowR, lower_zero);
    double highR = upper, highV = sq_dev(highR, false);
    double midR = G * (upper - lower) + lower, midV = sq_dev(midR, false);
    for (int it = 0; highR - lowR > 1e-4 && it < 10000; it++) {
      if (highV > lowV) {
        lowR = midR;
        lowV = midV;
        midR = highV > midV ? lowR + G * (highR - lowR) : highR - G * (highR - lowR);
      } else if (highV < lowV) {
        highR = midR;
        highV = midV;
        midR = lowV > midV ? highR - G * (highR - lowR) : lowR + G * (highR - lowR);
      } else {
        break;  // equal end values: the interval would not shrink
      }
      midV = sq_dev(midR, false);
    }
    if (highV > midV) return highV > lowV ? highV : lowV;
    return midV > lowV ? midV : lowV;
  }

  void fit(double p0) {
    phi0 = p0;
    const double rc2 = rc * rc;
    double v = 0.0, d1 = 0.0, d2 = 0.0;
    for (size_t i = 0; i < s->a.size(); i++) {
      const double a = s->a[i], t = s->c[i] * std::exp(-a * rc2);
      v += t;
      d1 += -2 * a * rc * t;
      d2 += 2 * a * (2 * a * rc2 - 1) * t;
    }
    sgn0 = phi0 - v > 0 ? 1 : -1;
    noC = true;
    C = 0.0;
    if (v * phi0 < 0) {
      // the replacement changes sign inside rc: shift it by C
      noC = false;
      if (v > 0.0 && v < 0.1) C = v + 0.1;
      else if (v < 0.0 && v > -0.1) C = v - 0.1;
      else C = 2 * v;
    } else if (v > 0.0 && v < 0.1) {
      noC = false;
      C = v - 0.1;
    } else if (v < 0.0 && v > -0.1) {
      noC = false;
      C = v + 0.1;
    }
    if (sgn0 == 1 && phi0 < 0.0) {
      noC = false;
      C = v > -0.1 ? v - 0.1 : 2 * v;
    } else if (sgn0 == -1 && phi0 > 0.0) {
      noC = false;
      C = v < 0.1 ? v + 0.1 : 2 * v;
    }
    zeff = Z * std::fabs(1 + (n0 / phi0));
    const double x1 = std::log(std::fabs(v - C));  // value at rc
    const double x2 = d1 / (v - C);                 // slope at rc
    const double x3 = d2 / (v - C);                 // curvature at rc
    double x4 = n0 / phi0 >= -1.0 ? -Z * (phi0 + n0) / (phi0 - C) : Z * (phi0 + n0) / (phi0 - C);
    if ((sgn0 == 1 && phi0 < 0.0) || (sgn0 == -1 && phi0 > 0.0)) x4 *= -1;
    const double x5 = std::log(std::fabs(phi0 - C));  // value at the nucleus
    const double ri = 1 / rc;
    alpha.c.assign(5, 0.0);
    alpha.c[0] = x5;
    alpha.c[1] = x4;
    alpha.c[2] = (6 * (x1 - x5) * ri - 3 * (x2 + x4)) * ri + 0.5 * (x3 - x2 * x2);
    alpha.c[3] = ((8 * (x5 - x1) * ri + 5 * x2 + 3 * x4) * ri + x2 * x2 - x3) * ri;
    alpha.c[4] = ((3 * (x1 - x5) * ri - 2 * x2 - x4) * ri + 0.5 * (x3 - x2 * x2)) * ri * ri;
    sigma();
  }

  void sigma() {
    if (noC) {
      sigma_sq = max_deviation(0.0, rc, true);
      return;
    }
    // node of the replacement: p(r) = ln|C|
    std::vector<double> q = alpha.c;
    q[0] -= std::log(std::fabs(C));
    double root = -2;
    for (double re : real_roots(q))
      if (re > 0 && re < rc) { root = re; break; }
    if (root < 0) {
      sigma_sq = max_deviation(0.0, rc, true);
      return;
    }
    double m1 = 0.0, m2 = 0.0;
    if (root - rc / 5 > 0.0) m1 = max_deviation(0.0, root - rc / 5, true);
    if (root + rc / 5 < rc) m2 = max_deviation(root + rc / 5, rc, false);
    sigma_sq = m1 > m2 ? m1 : m2;
  }

  // Real roots (|Im| < 1e-8) of sum q[i] x^i in increasing order: Aberth
  // simultaneous iteration on the monic polynomial, then Newton polishing.
  static std::vector<double> real_roots(std::vector<double> q) {
    while (!q.empty() && std::fabs(q.back()) < 1.0e-50) q.pop_back();
    const int n = (int)q.size() - 1;
    std::vector<double> out;
    if (n < 1) return out;
    typedef std::complex<double> cd;
    auto P = [&](cd x, cd &dp) {
      cd p = q[n];
      dp = 0.0;
      for (int i = n - 1; i >= 0; i--) {
        dp = dp * x + p;
        p = p * x + q[i];
      }
      return p;
    };
    double rad = 0.0;
    for (int i = 0; i < n; i++) rad = std::max(rad, std::fabs(q[i] / q[n]));
    rad = 1.0 + rad;
    std::vector<cd> z(n);
    for (int k = 0; k < n; k++) z[k] = std::polar(0.5 * rad, 2 * M_PI * (k + 0.25) / n);
    for (int it = 0; it < 500; it++) {
      double move = 0.0;
      for (int k = 0; k < n; k++) {
        cd dp;
        const cd p = P(z[k], dp);
        if (std::abs(p) == 0.0) continue;
        const cd ratio = p / dp;
        cd s = 0.0;
        for (int j = 0; j < n; j++)
          if (j != k) s += 1.0 / (z[k] - z[j]);
        const cd w = ratio / (1.0 - ratio * s);
        z[k] -= w;
        move = std::max(move, std::abs(w) / std::max(1.0, std::abs(z[k])));
      }
      if (move < 1e-15) break;
    }
    for (int k = 0; k < n; k++) {
      for (int it = 0; it < 3; it++) {
        cd dp;
        const cd p = P(z[k], dp);
        if (std::abs(dp) == 0.0) break;
        z[k] -= p / dp;
      }
      if (std::fabs(z[k].imag()) < 1e-8) out.push_back(z[k].real());
    }
    std::sort(out.begin(), out.end());
    return out;
  }
};

constexpr double kIdealCurve[9] = {0.0, 0.0, 3.25819, -15.0126, 33.7308,
                                   -42.8705, 31.2276, -12.1316, 1.94692};

// The ideal local-energy curve, with its constant set so that it passes
// through the original orbital's local energy at r (scaled by 1/Z^2 for Z > 1).
Poly ideal_curve_through(const SPart &s, int Z, double r) {
  Poly p;
  if (Z > 1) {
    p.c.assign(kIdealCurve, kIdealCurve + 9);
    p.c[0] = s.local_energy(r, Z) / (Z * Z) - p.value(r);
  } else {
    p.c.assign(1, s.local_energy(r, Z));
  }
  return p;
}

// Radius of correction: scan inwards from 1/Z for the first point (away from
// the orbital's nodes) where the local energy leaves the ideal curve by more
// than Z^2/50.  Returns false where the reference would stop with an error.
bool determine_rc(const SPart &s, int Z, double &rc, Poly &ideal, std::string &err) {
  const double rcmax = 1.0 / Z;
  ideal = ideal_curve_through(s, Z, rcmax);
  const double dr = rcmax / 100;
  std::vector<double> roots;
  {
    double old_v = s.value(rcmax);
    std::vector<int> idx;
    for (int i = 0; i < 101; i++) {
      const double nv = s.value(rcmax - i * dr);
      if (nv * old_v < 0) idx.push_back(i);
      old_v = nv;
    }
    const double tol = 1e-8;
    for (int i : idx) {
      double rl = rcmax - i * dr, fl = s.value(rl);
      double rh = rl + dr, fh = s.value(rh);
      if (fl * fh > 0) {
        err = "cusp fit: orbital node bracket lost";
        return false;
      }
      if (std::fabs(fl) < tol) { roots.push_back(rl); continue; }
      if (std::fabs(fh) < tol) { roots.push_back(rh); continue; }
      bool found = false;
      for (int it = 0; it < 9999; it++) {
        const double rm = (rl + rh) / 2, fm = s.value(rm);
        if (std::fabs(fm) < tol) { roots.push_back(rm); found = true; break; }
        if (fm * fl < 0) { rh = rm; fh = fm; }
        else { rl = rm; fl = fm; }
      }
      if (!found) {
        err = "cusp fit: orbital node bisection did not converge";
        return false;
      }
    }
  }
  const double g_near = s.gradient(rcmax / 1000);
  const double v0 = s.value(0.0);
  rc = rcmax;
  double r = rcmax, r_bound = 0.0;
  int ri = -1;
  const double ddr = dr / 10;
  const double maxdev = Z * Z / 50.0;
  const int nroots = (int)roots.size();
  if (nroots > 0) {
    ri = nroots - 1;
    r_bound = roots[ri] + rcmax / 5;
  }
  auto dev = [&](double x) { return std::fabs(Z * Z * ideal.value(x) - s.local_energy(x, Z)); };
  auto same_side = [&](double x) { return g_near * (s.value(x) - v0) > 0.0; };
  while (r > 0.0) {
    if (same_side(r)) {
      if (r > r_bound) {
        if (dev(r) > maxdev) {
          r += dr;
          while (r > 0.0) {
            r -= ddr;
            if (same_side(r) && dev(r) > maxdev) { rc = r; break; }
          }
          break;
        }
      } else {
        if (same_side(r) && dev(r_bound) > maxdev) {
          while (r > 0.0) {
            r -= ddr;
            if (dev(r) > maxdev) { rc = r; break; }
          }
          break;
        }
        r = roots[ri] - rcmax / 5 + dr;
        if (ri > 0) {
          ri--;
          r_bound = roots[ri] + rcmax / 5;
        } else if (ri == 0) {
          r_bound = 0.0;
        }
      }
    }
    r -= dr;
  }
  // the ideal curve passes through the original local energy at rc
  if (Z > 1) {
    Poly p;
    p.c.assign(kIdealCurve, kIdealCurve + 9);
    p.c[0] = 0.0;
    p.c[0] = s.local_energy(rc, Z) / (Z * Z) - p.value(rc);
    ideal = p;
  } else {
    ideal.c.assign(1, s.local_energy(rc, Z));
  }
  return true;
}

// Golden-section search over the value at the nucleus, within 10% of the
// original orbital's change between 0 and rc.
Fit fit_orbital(const SPart &s, double n0, double rc, const Poly &ideal, int Z) {
  Fit lo, mid, hi;
  for (Fit *f : {&lo, &mid, &hi}) {
    f->s = &s;
    f->n0 = n0;
    f->rc = rc;
    f->ideal = ideal;
    f->Z = Z;
  }
  const double o0 = s.value(0.0), orc = s.value(rc);
  double lo_phi, mid_phi, hi_phi;
  if (o0 - orc < 0.0) {
    hi_phi = o0;
    mid_phi = o0 - 0.05 * std::fabs(o0 - orc);
    lo_phi = o0 - 0.1 * std::fabs(o0 - orc);
  } else {
    lo_phi = o0;
    mid_phi = o0 + 0.05 * std::fabs(o0 - orc);
    hi_phi = o0 + 0.1 * std::fabs(o0 - orc);
  }
  lo.fit(lo_phi);
  mid.fit(mid_phi);
  hi.fit(hi_phi);
  double ls = lo.sigma_sq, ms = mid.sigma_sq, hs = hi.sigma_sq;
  const double G = 0.618034;
  while (hi_phi - lo_phi > std::fabs(mid_phi / 10000)) {
    bool move_hi;  // replace the high end by mid (else the low end)
    bool from_lo;  // new mid measured from the low end (else from the high end)
    if (ls < ms) {
      if (ms < hs || ls < hs) { move_hi = true; from_lo = true; }
      else { move_hi = false; from_lo = false; }
    } else {
      if (hs < ms) { move_hi = false; from_lo = false; }
      else if (ls < hs) { move_hi = true; from_lo = false; }
      else { move_hi = false; from_lo = true; }
    }
    if (move_hi) {
      hi_phi = mid_phi;
      hs = ms;
      hi = mid;
    } else {
      lo_phi = mid_phi;
      ls = ms;
      lo = mid;
    }
    mid_phi = from_lo ? lo_phi + G * (hi_phi - lo_phi) : hi_phi - G * (hi_phi - lo_phi);
    mid.fit(mid_phi);
    ms = mid.sigma_sq;
  }
  return mid;
}

}  // namespace

bool fit_cusps(const qmcb_system *sys, const double *mo, int norb, CuspTable &t,
               std::string &err) {
  const int nn = sys->n_nuc, B = sys->n_basis;
  t.nn = nn;
  t.norb = norb;
  t.s_off.assign(nn + 1, 0);
  t.s_alpha.clear();
  t.s_coef.clear();
  t.orb.assign((size_t)nn * norb, qmcb_cusp_orbital{});
  t.rc_max.assign(nn, 0.0);
  // basis-function ranges per atom (functions are grouped by atom in order)
  std::vector<int> bf_beg(nn + 1, B);
  for (int j = B - 1; j >= 0; j--) bf_beg[sys->bf_atom[j]] = j;
  for (int a = nn - 1; a >= 0; a--)
    if (bf_beg[a] > bf_beg[a + 1]) bf_beg[a] = bf_beg[a + 1];
  auto bf_end = [&](int a) {
    int e = bf_beg[a];
    while (e < B && sys->bf_atom[e] == a) e++;
    return e;
  };
  // s-type Gaussians of each nucleus (exponents; coefficients per orbital below)
  std::vector<std::vector<int>> s_bf(nn), s_prim(nn);
  for (int a = 0; a < nn; a++) {
    t.s_off[a] = (int)t.s_alpha.size();
    for (int j = bf_beg[a]; j < bf_end(a); j++) {
      if (sys->bf_pow[3 * j] || sys->bf_pow[3 * j + 1] || sys->bf_pow[3 * j + 2]) continue;
      for (int p = sys->bf_prim_off[j]; p < sys->bf_prim_off[j + 1]; p++) {
        s_bf[a].push_back(j);
        s_prim[a].push_back(p);
        t.s_alpha.push_back(sys->prim_alpha[p]);
      }
    }
  }
  t.s_off[nn] = (int)t.s_alpha.size();
  const int ns = t.s_off[nn];
  t.s_coef.assign((size_t)norb * ns, 0.0);
  for (int m = 0; m < norb; m++)
    for (int a = 0; a < nn; a++)
      for (size_t g = 0; g < s_bf[a].size(); g++)
        t.s_coef[(size_t)m * ns + t.s_off[a] + g] =
            sys->prim_coef[s_prim[a][g]] * mo[(size_t)m * B + s_bf[a][g]];

  for (int a = 0; a < nn; a++) {
    if (s_bf[a].empty()) continue;  // no s functions: no replacement
    const int Z = (int)std::lround(sys->nuc_z[a]);
    if (Z < 1) {
      err = "cusp replacement needs nuclear charges >= 1";
      return false;
    }
    for (int m = 0; m < norb; m++) {
      SPart s;
      bool any = false;
      for (size_t g = 0; g < s_bf[a].size(); g++) {
        s.a.push_back(t.s_alpha[t.s_off[a] + g]);
        const double c = t.s_coef[(size_t)m * ns + t.s_off[a] + g];
        s.c.push_back(c);
        if (std::fabs(c) > 1e-15) any = true;
      }
      if (!any) continue;
      double rc = 0.0;
      Poly ideal;
      if (!determine_rc(s, Z, rc, ideal, err)) return false;
      // orbital m at nucleus a from the other atoms' functions
      double n0 = 0.0;
      const double *xa = sys->nuc_xyz + 3 * a;
      for (int b = 0; b < nn; b++) {
        if (b == a) continue;
        const double *xb = sys->nuc_xyz + 3 * b;
        const double dx = xa[0] - xb[0], dy = xa[1] - xb[1], dz = xa[2] - xb[2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        for (int j = bf_beg[b]; j < bf_end(b); j++) {
          const double xyz = std::pow(dx, sys->bf_pow[3 * j]) * std::pow(dy, sys->bf_pow[3 * j + 1]) *
                             std::pow(dz, sys->bf_pow[3 * j + 2]);
          double rad = 0.0;
          for (int p = sys->bf_prim_off[j]; p < sys->bf_prim_off[j + 1]; p++)
            rad += sys->prim_coef[p] * std::exp(-sys->prim_alpha[p] * r2);
          n0 += rad * xyz * mo[(size_t)m * B + j];
        }
      }
      Fit f = fit_orbital(s, n0, rc, ideal, Z);
      qmcb_cusp_orbital &o = t.orb[(size_t)a * norb + m];
      o.rc = rc;
      o.C = f.C;
      o.sgn0 = f.sgn0;
      for (int k = 0; k < 5; k++) o.alpha[k] = f.alpha.c[k];
      o.phi0 = f.phi0;
      o.n0 = n0;
      o.zeff = f.zeff;
      o.sigma_sq = f.sigma_sq;
      o.ideal0 = ideal.c[0];
      if (rc > t.rc_max[a]) t.rc_max[a] = rc;
    }
  }
  return true;
}

}  // namespace qmcb

extern "C" qmcb_status qmcb_cusp_fit(const qmcb_system *sys, int32_t spin,
                                     qmcb_cusp_orbital *out) {
  if (!sys || (spin != 0 && spin != 1) || !out) return QMCB_INVALID_ARG;
  const int norb = spin == 0 ? sys->n_up : sys->n_dn;
  const double *mo = spin == 0 ? sys->mo_up : sys->mo_dn;
  if (norb == 0) return QMCB_OK;
  qmcb::CuspTable t;
  std::string err;
  if (!qmcb::fit_cusps(sys, mo, norb, t, err)) return QMCB_INVALID_ARG;
  for (size_t i = 0; i < t.orb.size(); i++) out[i] = t.orb[i];
  return QMCB_OK;
}
