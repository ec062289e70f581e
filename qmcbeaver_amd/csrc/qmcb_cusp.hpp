// qmcb_cusp.hpp -- cusp-replacement tables (host fit, device correction).
#pragma once
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/qmcbeaver_amd.h"

namespace qmcb {

// Fitted replacement orbitals of one spin: per (nucleus, orbital) the fit,
// per nucleus its s-type exponents, per (orbital, s-Gaussian) coefficient
// times MO coefficient.
struct CuspTable {
  int nn = 0, norb = 0;
  std::vector<int> s_off;         // [nn + 1] slice of each nucleus's s Gaussians
  std::vector<double> s_alpha;    // [ns]
  std::vector<double> s_coef;     // [norb][ns]
  std::vector<qmcb_cusp_orbital> orb;  // [nn][norb], rc = 0: no replacement
  std::vector<double> rc_max;     // [nn]
};

bool fit_cusps(const qmcb_system *sys, const double *mo, int norb, CuspTable &t,
               std::string &err);

// Device copy of a CuspTable.  orb8[(a * norb + m) * 8 + {0..7}] =
// {rc, C, sgn0, alpha0..alpha4}.
struct CuspDev {
  int nn, norb, ns;
  const double *nuc_xyz;
  const double *rc_max;   // [nn]
  const int *s_off;       // [nn + 1]
  const double *s_alpha;  // [ns]
  const double *s_coef;   // [norb][ns]
  const double *orb8;     // [nn][norb][8]
};

}  // namespace qmcb
