// ref_driver.cpp -- this project's driver around the reference's own
// evaluate path, compiled from the reference sources by oracle/ref.mk into
// oracle/_ref/ (git-ignored).  Test infrastructure: it writes the golden
// vectors under tests/golden/ and times the reference CPU path.
//
// It calls, unchanged, QMCInput::read (QMCInput.cpp:26-53),
// QMCSCFJastrow::initialize/evaluate (QMCSCFJastrow.cpp:61-192) and
// QMCWalkerData::initialize/updateDistances (QMCWalkerData.cpp:56-161,
// 302-346).  Walker records are freshly initialised for every all-electron
// evaluation (zero Jastrow caches).
//
//   qmcb_ref_driver eval <input.ckmf> <configs.bin> <out.bin>
//   qmcb_ref_driver seq  <input.ckmf> <moves.bin>   <out.bin>
//   qmcb_ref_driver vmc  <input.ckmf> <configs.bin> <n_moves> <tau> <seed>
//   qmcb_ref_driver vmc_ae <input.ckmf> <configs.bin> <n_eq> <n_steps> <tau> <seed>
//   qmcb_ref_driver ddda <series.bin> <split>
//
// configs.bin: int32 nconf, int32 N, double R[nconf][N][3]
// moves.bin:   int32 nwalk, int32 N, int32 nmoves, then per walker
//              double R[N][3] followed by nmoves x double (e, x, y, z)
// out.bin:     per evaluation one record of 12 + 6N doubles:
//              detA detB U E_L KE PE V_ne V_ee singular modRatio lnPsi signPsi,
//              gradPsiRatio[N][3], modifiedGradPsiRatio[N][3]
//              (detA/detB only with one_e_per_iter 1; lnPsi and its sign come
//              from the QMCDouble psi = k a^b e^c as ln|k| + b ln a + c)
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <sstream>
#include <string>
#include <vector>
#include <sys/time.h>

#include "QMCInput.h"
#include "QMCProperty.h"
#include "QMCSCFJastrow.h"
#include "QMCWalkerData.h"

QMCInput globalInput;

// ln|psi| and sign from the printed k ( a ^ b ) Exp[ c ] form of a QMCDouble
static void log_qmcdouble(const QMCDouble &q, double &lnv, double &sgn) {
  std::ostringstream os;
  os.precision(17);
  os << q;
  std::string t = os.str();
  for (size_t i = 0; i < t.size(); i++)
    if (t[i] == '(' || t[i] == ')' || t[i] == '^' || t[i] == '[' || t[i] == ']') t[i] = ' ';
  std::istringstream is(t);
  double k = 0, a = 1, b = 0, c = 0;
  std::string ex;
  is >> k >> a >> b >> ex >> c;
  lnv = std::log(std::fabs(k)) + (b != 0.0 ? b * std::log(a) : 0.0) + c;
  sgn = k < 0 ? -1.0 : 1.0;
}

static void record(std::FILE *f, QMCWalkerData &wd, int N) {
  int na = globalInput.WF.getNumberElectrons(true);
  int nb = globalInput.WF.getNumberElectrons(false);
  const bool onee = globalInput.flags.one_e_per_iter != 0;
  double detA = (onee && na > 0) ? wd.DcA(0) : 1.0;
  double detB = (onee && nb > 0) ? wd.DcB(0) : 1.0;
  double lnpsi, sgn;
  log_qmcdouble(wd.psi, lnpsi, sgn);
  std::vector<double> r;
  r.push_back(detA);
  r.push_back(detB);
  r.push_back(wd.U);
  r.push_back(wd.localEnergy);
  r.push_back(wd.kineticEnergy);
  r.push_back(wd.potentialEnergy);
  r.push_back(wd.neEnergy);
  r.push_back(wd.eeEnergy);
  r.push_back(wd.singular ? 1.0 : 0.0);
  r.push_back(wd.modificationRatio);
  r.push_back(lnpsi);
  r.push_back(sgn);
  for (int e = 0; e < N; e++)
    for (int k = 0; k < 3; k++) r.push_back(wd.gradPsiRatio(e, k));
  for (int e = 0; e < N; e++)
    for (int k = 0; k < 3; k++) r.push_back(wd.modifiedGradPsiRatio(e, k));
  std::fwrite(&r[0], sizeof(double), r.size(), f);
}

static double now() {
  struct timeval tv;
  gettimeofday(&tv, 0);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

// splitmix64 + Box-Muller: this program's own move generator (the reference's
// ran1/gasdev streams are not part of the evaluate path).
static unsigned long long g_state;
static double unif() {
  unsigned long long z = (g_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (z >> 11) * (1.0 / 9007199254740992.0);
}
static double gauss() {
  double u1 = unif(), u2 = unif();
  if (u1 < 1e-300) u1 = 1e-300;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// The reference's DDDA estimator (QMCProperty.cpp:187-237, 427-478, 511-598,
// 830-872; QMCStatistic.cpp) over a weighted series.
//   qmcb_ref_driver ddda <series.bin> <split>
// series.bin: int32 n, then n x (double sample, double weight).  The samples
// [0, split) and [split, n) go into two properties that are merged with the
// reference's operator+; split = n means one property.  Prints one JSON line
// with the public getters, including getBlockVariance for every level
// (non-finite values as null).
static void jnum(double v) {
  if (std::isfinite(v)) std::printf("%.17g", v);
  else std::printf("null");
}

static void print_prop(QMCProperty &p) {
  std::printf("{\"n\": %lu, \"average\": ", p.getNumberSamples());
  jnum(p.getAverage());
  std::printf(", \"std\": ");
  jnum(p.getStandardDeviation());
  std::printf(", \"corr_len\": ");
  jnum(p.getCorrelationLength());
  std::printf(", \"serial_var\": ");
  jnum(p.getSeriallyCorrelatedVariance());
  std::printf(", \"std_std\": ");
  jnum(p.getStandardDeviationStandardDeviation());
  std::printf(", \"block_var\": [");
  for (int i = 0; i < DCL; i++) {
    if (i) std::printf(", ");
    jnum(p.getBlockVariance(i));
  }
  std::printf("]}\n");
}

static int ddda_main(int argc, char **argv) {
  if (argc < 4) return 2;
  std::FILE *f = std::fopen(argv[2], "rb");
  if (!f) return 1;
  int n = 0;
  if (std::fread(&n, sizeof(int), 1, f) != 1) return 3;
  std::vector<double> sw(2 * (size_t)n);
  if (n > 0 && std::fread(&sw[0], sizeof(double), sw.size(), f) != sw.size()) return 4;
  std::fclose(f);
  int split = std::atoi(argv[3]);
  if (split < 0 || split > n) return 5;
  // static: QMCStatistic leaves its sums uninitialised until the first
  // sample, and operator+ adds them for empty levels
  static QMCProperty a, b;
  for (int i = 0; i < split; i++) a.newSample(sw[2 * i], sw[2 * i + 1]);
  for (int i = split; i < n; i++) b.newSample(sw[2 * i], sw[2 * i + 1]);
  QMCProperty m = a + b;
  print_prop(m);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s eval|seq|vmc input.ckmf in out|n_moves [tau seed]\n", argv[0]);
    return 2;
  }
  std::string mode = argv[1];
  if (mode == "ddda") return ddda_main(argc, argv);
  // full precision for the cusp-replacement parameters the reference prints
  // with print_replacement_orbitals 1 (QMCElectronNucleusCusp.cpp:163-168)
  std::clog.precision(17);
  globalInput.read(std::string(argv[2]));
  QMCSCFJastrow qmf;
  qmf.initialize(&globalInput, 0);
  int N = globalInput.WF.getNumberElectrons();

  std::FILE *fin = std::fopen(argv[3], "rb");
  if (!fin) { std::perror("input"); return 1; }
  if (mode == "eval") {
    int hdr[2];
    if (std::fread(hdr, sizeof(int), 2, fin) != 2 || hdr[1] != N) return 3;
    std::FILE *fo = std::fopen(argv[4], "wb");
    for (int c = 0; c < hdr[0]; c++) {
      Array2D<double> R(N, 3);
      for (int e = 0; e < N; e++)
        for (int k = 0; k < 3; k++) {
          double v;
          if (std::fread(&v, sizeof(double), 1, fin) != 1) return 4;
          R(e, k) = v;
        }
      QMCWalkerData wd;
      wd.initialize(3, 0, 0);
      wd.whichE = -1;
      wd.updateDistances(R);
      qmf.evaluate(R, wd);
      record(fo, wd, N);
    }
    std::fclose(fo);
  } else if (mode == "seq") {
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, fin) != 3 || hdr[1] != N) return 3;
    std::FILE *fo = std::fopen(argv[4], "wb");
    for (int w = 0; w < hdr[0]; w++) {
      Array2D<double> R(N, 3);
      for (int e = 0; e < N; e++)
        for (int k = 0; k < 3; k++) {
          double v;
          if (std::fread(&v, sizeof(double), 1, fin) != 1) return 4;
          R(e, k) = v;
        }
      QMCWalkerData wd;
      wd.initialize(3, 0, 0);
      wd.whichE = -1;
      wd.updateDistances(R);
      qmf.evaluate(R, wd);
      record(fo, wd, N);
      for (int m = 0; m < hdr[2]; m++) {
        double mv[4];
        if (std::fread(mv, sizeof(double), 4, fin) != 4) return 4;
        int e = (int)mv[0];
        for (int k = 0; k < 3; k++) R(e, k) = mv[1 + k];
        wd.whichE = e;
        wd.updateDistances(R);
        qmf.evaluate(R, wd);
        record(fo, wd, N);
      }
    }
    std::fclose(fo);
  } else if (mode == "vmc") {
    // Importance-sampled single-electron VMC with the reference evaluate:
    // propose (QMCWalker.cpp:343-398), evaluate at the proposal, Metropolis
    // with the Green's function ratio (:756-780, 1578-1685), and on rejection
    // re-evaluate at the old position (:254-266).
    int hdr[2];
    if (std::fread(hdr, sizeof(int), 2, fin) != 2 || hdr[1] != N) return 3;
    int nw = hdr[0];
    int n_moves = std::atoi(argv[4]);
    double tau = argc > 5 ? std::atof(argv[5]) : 0.01;
    g_state = argc > 6 ? std::strtoull(argv[6], 0, 10) : 1;
    std::vector<Array2D<double> > R(nw);
    std::vector<QMCWalkerData *> wd(nw);
    std::vector<double> lp(nw), el(nw);
    std::vector<Array2D<double> > F(nw);
    for (int w = 0; w < nw; w++) {
      R[w].allocate(N, 3);
      for (int e = 0; e < N; e++)
        for (int k = 0; k < 3; k++) {
          double v;
          if (std::fread(&v, sizeof(double), 1, fin) != 1) return 4;
          R[w](e, k) = v;
        }
      wd[w] = new QMCWalkerData();
      wd[w]->initialize(3, 0, 0);
      wd[w]->whichE = -1;
      wd[w]->updateDistances(R[w]);
      qmf.evaluate(R[w], *wd[w]);
      lp[w] = std::log(std::fabs(wd[w]->DcA(0))) + std::log(std::fabs(wd[w]->DcB(0))) + wd[w]->U;
      el[w] = wd[w]->localEnergy;
      F[w] = wd[w]->modifiedGradPsiRatio;
    }
    double t0 = now(), esum = 0.0, acc = 0.0, sweep_e = 0.0;
    static QMCProperty energy;  // static: zero-filled sums (see ddda_main)
    double sig = std::sqrt(tau);
    for (int m = 0; m < n_moves; m++) {
      int e = m % N;
      for (int w = 0; w < nw; w++) {
        double old[3], fwd = 0.0;
        for (int i = 0; i < N; i++)
          if (i != e)
            for (int k = 0; k < 3; k++) fwd += tau * F[w](i, k) * tau * F[w](i, k);
        for (int k = 0; k < 3; k++) {
          old[k] = R[w](e, k);
          double g = sig * gauss();
          R[w](e, k) += tau * F[w](e, k) + g;
          fwd += g * g;
        }
        wd[w]->whichE = e;
        wd[w]->updateDistances(R[w]);
        qmf.evaluate(R[w], *wd[w]);
        double lpn = std::log(std::fabs(wd[w]->DcA(0))) + std::log(std::fabs(wd[w]->DcB(0))) + wd[w]->U;
        double rev = 0.0;
        for (int i = 0; i < N; i++)
          for (int k = 0; k < 3; k++) {
            double oc = i == e ? old[k] : R[w](i, k);
            double t = oc - R[w](i, k) - tau * wd[w]->modifiedGradPsiRatio(i, k);
            rev += t * t;
          }
        double p = wd[w]->singular ? 0.0 : std::exp(2.0 * (lpn - lp[w]) - (rev - fwd) / (2.0 * tau));
        if (!(p == p)) p = 0.0;
        if (e == 0) {
          // statistics once per sweep, at the electron-0 move, from
          // p E_L(R') + (1 - p) E_L(R) (QMCWalker.cpp:2096-2104, 2302-2346)
          const double pc = p > 1.0 ? 1.0 : p;
          sweep_e += pc > 0.0 ? pc * wd[w]->localEnergy + (1.0 - pc) * el[w] : el[w];
        }
        if (p > unif()) {
          lp[w] = lpn; el[w] = wd[w]->localEnergy; F[w] = wd[w]->modifiedGradPsiRatio;
          acc += 1.0;
        } else {
          for (int k = 0; k < 3; k++) R[w](e, k) = old[k];
          wd[w]->updateDistances(R[w]);
          qmf.evaluate(R[w], *wd[w]);
        }
        esum += el[w];
      }
      if (e == 0) {
        // walker average with the total weight (QMCRun.cpp:544-576)
        energy.newSample(sweep_e / nw, (double)nw);
        sweep_e = 0.0;
      }
    }
    double dt = now() - t0;
    std::printf("{\"seconds\": %.6f, \"walker_moves\": %lld, \"mean_e_local\": %.10f, "
                "\"acceptance\": %.6f, \"ddda_samples\": %lu, \"ddda_energy\": %.12f, "
                "\"ddda_std\": %.12g}\n",
                dt, (long long)nw * n_moves, esum / ((double)nw * n_moves), acc / ((double)nw * n_moves),
                energy.getNumberSamples(), energy.getAverage(), energy.getStandardDeviation());
  } else if (mode == "vmc_ae") {
    // Importance-sampled all-electron VMC (one_e_per_iter 0) with the
    // reference evaluate: every electron moves, R' = R + tau F(R) +
    // sqrt(tau) xi (QMCWalker.cpp:343-398); the reverse Green's function
    // over all electrons (:756-780); p = min(1, |Psi'/Psi|^2 G(R<-R') /
    // G(R'<-R)) (:1578-1653); a rejected walker gets its saved record back
    // (:259-262).  Statistics per step after n_eq steps: the energy sample
    // sum_w (p E_L(R') + (1 - p) E_L(R)) / W through QMCProperty
    // (QMCWalker.cpp:2096-2104, QMCRun.cpp:544-576), the mean acceptance
    // probability p over walkers and steps, and the fraction of accepted
    // moves.
    int hdr[2];
    if (std::fread(hdr, sizeof(int), 2, fin) != 2 || hdr[1] != N) return 3;
    const int nw = hdr[0];
    const int n_eq = std::atoi(argv[4]), n_steps = argc > 5 ? std::atoi(argv[5]) : 100;
    const double tau = argc > 6 ? std::atof(argv[6]) : 0.1;
    g_state = argc > 7 ? std::strtoull(argv[7], 0, 10) : 1;
    std::vector<Array2D<double> > R(nw);
    std::vector<QMCWalkerData *> wd(nw);
    for (int w = 0; w < nw; w++) {
      R[w].allocate(N, 3);
      for (int e = 0; e < N; e++)
        for (int k = 0; k < 3; k++) {
          double v;
          if (std::fread(&v, sizeof(double), 1, fin) != 1) return 4;
          R[w](e, k) = v;
        }
      wd[w] = new QMCWalkerData();
      wd[w]->initialize(3, 0, 0);
      wd[w]->whichE = -1;
      wd[w]->updateDistances(R[w]);
      qmf.evaluate(R[w], *wd[w]);
    }
    QMCWalkerData trial;
    trial.initialize(3, 0, 0);
    trial.whichE = -1;
    static QMCProperty energy, accp;  // static: zero-filled sums (see ddda_main)
    double psum = 0.0, acc = 0.0, t0 = now();
    const double sig = std::sqrt(tau);
    Array2D<double> Rp(N, 3);
    for (int it = 0; it < n_eq + n_steps; it++) {
      double step_e = 0.0, step_p = 0.0;
      for (int w = 0; w < nw; w++) {
        double fwd = 0.0;
        for (int e = 0; e < N; e++)
          for (int k = 0; k < 3; k++) {
            const double g = sig * gauss();
            Rp(e, k) = R[w](e, k) + tau * wd[w]->modifiedGradPsiRatio(e, k) + g;
            fwd += g * g;
          }
        trial = *wd[w];  // the walker record, as createChildWalkers copies it
        trial.whichE = -1;
        trial.updateDistances(Rp);
        qmf.evaluate(Rp, trial);
        double rev = 0.0;
        for (int e = 0; e < N; e++)
          for (int k = 0; k < 3; k++) {
            const double t = R[w](e, k) - Rp(e, k) - tau * trial.modifiedGradPsiRatio(e, k);
            rev += t * t;
          }
        double lo, so, ln, sn;
        log_qmcdouble(wd[w]->psi, lo, so);
        log_qmcdouble(trial.psi, ln, sn);
        double p = trial.singular ? 0.0 : std::exp(2.0 * (ln - lo) - (rev - fwd) / (2.0 * tau));
        if (!(p == p)) p = 0.0;
        if (p > 1.0) p = 1.0;
        step_e += p * trial.localEnergy + (1.0 - p) * wd[w]->localEnergy;
        step_p += p;
        if (p > unif()) {
          *wd[w] = trial;
          R[w] = Rp;
          if (it >= n_eq) acc += 1.0;
        }
      }
      if (it >= n_eq) {
        energy.newSample(step_e / nw, (double)nw);
        accp.newSample(step_p / nw, (double)nw);
        psum += step_p;
      }
    }
    const double dt = now() - t0, moves = (double)nw * n_steps;
    std::printf("{\"seconds\": %.6f, \"walker_moves\": %.0f, \"acceptance_probability\": %.8f, "
                "\"accepted_fraction\": %.8f, \"ddda_samples\": %lu, \"ddda_energy\": %.12f, "
                "\"ddda_std\": %.12g, \"acceptance_probability_std\": %.12g}\n",
                dt, moves, psum / moves, acc / moves, energy.getNumberSamples(),
                energy.getAverage(), energy.getStandardDeviation(), accp.getStandardDeviation());
  } else {
    return 2;
  }
  std::fclose(fin);
  return 0;
}
