// dropin_vmc.cpp -- a host-driven VMC pass through the QMCFunctions-shaped
// C++ interface (include/qmcb_functions.hpp), the way the reference's
// QMCRun::propagateWalkers (QMCRun.cpp:28-87) drives QMCFunctions::evaluate:
// all-electron evaluate of every walker, then single-electron proposals,
// Metropolis on the host, commit.  Reads a flat system file written by
// qmcbeaver_amd.sysio.write_flat and prints one JSON line.
//
//   dropin_vmc <system.bin> <walkers.bin> <n_moves>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../include/qmcb_functions.hpp"

template <typename T>
static std::vector<T> rd(std::FILE *f, size_t n) {
  std::vector<T> v(n);
  if (n && std::fread(v.data(), sizeof(T), n, f) != n) { std::fprintf(stderr, "short read\n"); std::exit(3); }
  return v;
}

int main(int argc, char **argv) {
  if (argc < 4) { std::fprintf(stderr, "usage: %s system.bin walkers.bin n_moves\n", argv[0]); return 2; }
  std::FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  auto hdr = rd<int32_t>(f, 5);
  const int nu = hdr[0], nd = hdr[1], nn = hdr[2], B = hdr[3], np = hdr[4];
  auto xyz = rd<double>(f, 3 * nn), z = rd<double>(f, nn);
  auto bfa = rd<int32_t>(f, B), bfp = rd<int32_t>(f, 3 * B), off = rd<int32_t>(f, B + 1);
  auto al = rd<double>(f, np), co = rd<double>(f, np);
  auto mu = rd<double>(f, (size_t)nu * B), md = rd<double>(f, (size_t)nd * B);
  auto pades = rd<qmcb_pade>(f, 3 + 2 * nn);
  std::fclose(f);
  qmcb_system sys{};
  sys.n_up = nu; sys.n_dn = nd; sys.n_nuc = nn; sys.n_basis = B; sys.n_prim = np;
  sys.nuc_xyz = xyz.data(); sys.nuc_z = z.data(); sys.bf_atom = bfa.data();
  sys.bf_pow = bfp.data(); sys.bf_prim_off = off.data(); sys.prim_alpha = al.data();
  sys.prim_coef = co.data(); sys.mo_up = mu.data(); sys.mo_dn = md.data();
  sys.ee_updn = pades[0]; sys.ee_upup = pades[1]; sys.ee_dndn = pades[2];
  sys.en_up = pades.data() + 3; sys.en_dn = pades.data() + 3 + nn;

  std::FILE *g = std::fopen(argv[2], "rb");
  if (!g) return 2;
  auto wh = rd<int32_t>(g, 2);
  const int W = wh[0], N = wh[1];
  auto R = rd<double>(g, (size_t)W * N * 3);
  std::fclose(g);
  const int n_moves = std::atoi(argv[3]);
  const double tau = 0.02;

  qmcb_options opt{};
  opt.precision = QMCB_FP64;
  opt.max_walkers = W;
  opt.dt = tau;
  qmcb::QMCFunctions qmf(sys, opt);
  std::vector<qmcb::WalkerData> wd(W);
  std::vector<qmcb::WalkerData *> wp(W);
  std::vector<const double *> xp(W);
  for (int i = 0; i < W; i++) {
    wd[i].slot = i;
    wd[i].whichE = -1;
    wp[i] = &wd[i];
    xp[i] = R.data() + (size_t)i * N * 3;
  }
  qmf.evaluate(wp, xp, W);
  const double e0 = wd[0].localEnergy;
  std::vector<double> lp(W), F((size_t)W * N * 3);
  for (int i = 0; i < W; i++) {
    lp[i] = wd[i].logPsi;
    std::copy(wd[i].modifiedGradPsiRatio.begin(), wd[i].modifiedGradPsiRatio.end(), F.begin() + (size_t)i * N * 3);
  }
  std::mt19937_64 rng(7);
  std::normal_distribution<double> gauss(0.0, 1.0);
  std::uniform_real_distribution<double> unif(0.0, 1.0);
  std::vector<double> Rp(R);
  double acc = 0, esum = 0;
  for (int m = 0; m < n_moves; m++) {
    const int e = m % N;
    std::vector<double> fwd(W, 0.0), old((size_t)W * 3);
    for (int i = 0; i < W; i++) {
      wd[i].whichE = e;
      double *Fi = F.data() + (size_t)i * N * 3;
      for (int j = 0; j < N; j++)
        if (j != e) for (int k = 0; k < 3; k++) fwd[i] += tau * Fi[3 * j + k] * tau * Fi[3 * j + k];
      for (int k = 0; k < 3; k++) {
        const double chi = std::sqrt(tau) * gauss(rng);
        old[3 * i + k] = Rp[((size_t)i * N + e) * 3 + k];
        Rp[((size_t)i * N + e) * 3 + k] += tau * Fi[3 * e + k] + chi;
        fwd[i] += chi * chi;
      }
      xp[i] = Rp.data() + (size_t)i * N * 3;
    }
    qmf.evaluate(wp, xp, W);
    std::vector<bool> ok(W);
    for (int i = 0; i < W; i++) {
      const auto &mg = wd[i].modifiedGradPsiRatio;
      double rev = 0;
      for (int j = 0; j < N; j++)
        for (int k = 0; k < 3; k++) {
          const double oc = j == e ? old[3 * i + k] : Rp[((size_t)i * N + j) * 3 + k];
          const double t = oc - Rp[((size_t)i * N + j) * 3 + k] - tau * mg[3 * j + k];
          rev += t * t;
        }
      double p = wd[i].singular ? 0.0 : std::exp(2 * (wd[i].logPsi - lp[i]) - (rev - fwd[i]) / (2 * tau));
      ok[i] = std::min(p, 1.0) > unif(rng);
      if (ok[i]) {
        lp[i] = wd[i].logPsi;
        std::copy(mg.begin(), mg.end(), F.begin() + (size_t)i * N * 3);
        acc += 1;
      } else {
        for (int k = 0; k < 3; k++) Rp[((size_t)i * N + e) * 3 + k] = old[3 * i + k];
      }
      esum += wd[i].localEnergy;
    }
    qmf.accept(wp, ok);
  }
  std::printf("{\"e_local_0\": %.12f, \"acceptance\": %.6f, \"mean_prop_e_local\": %.8f}\n", e0,
              acc / ((double)W * n_moves), esum / ((double)W * std::max(n_moves, 1)));
  return 0;
}
