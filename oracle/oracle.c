/*
 * oracle.c -- CPU restatement (plain C99) of QMcBeaver's FP64 evaluate path.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the HIP path is tested against.  Never linked into
 * the product library; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it (as liboracle.so via ctypes).
 *
 * Each function cites the reference code it restates (paths relative to
 * QMcBeaver/).  Parity pins (tests/test_oracle.py):
 *   orc_basis         golden vectors from the reference build (evaluate level),
 *                     closed form of an s Gaussian, finite differences of the
 *                     value for gradient and Laplacian.
 *   orc_pade          closed form a r/(1+b r) and its derivatives, FD.
 *   orc_det_inverse   numpy slogdet/inv, A A^-1 = I.
 *   orc_inverse_update_row  fresh inverse of the row-replaced matrix; det ratio.
 *   orc_evaluate      golden vectors written by oracle/_ref (the reference's own
 *                     QMCSCFJastrow::evaluate compiled from its sources), all-
 *                     electron and single-electron sequences; FD of ln|Psi|
 *                     for the gradient and Laplacian; hydrogen-like closed form.
 *                     Both umrigar93 drift modifications against reference
 *                     evaluates with those flags.
 *   orc_vmc_time      timing loop only (no parity claim beyond orc_evaluate).
 *
 * Scope: one determinant without cusp replacement.  Several determinants and
 * the cusp replacement are pinned on the device directly against golden
 * vectors of the reference build (tests/golden/make_golden_multidet.py,
 * make_golden_cusp.py), and the cusp fit on the host (tests/test_cusp.py).
 *
 * Deliberate difference: the reference's exponent cache is not reset between
 * atoms (QMCBasisFunction.cpp:310,372); here it is reset per atom.  The input
 * generator never produces the triggering case (equal exponents across an
 * atom boundary), so both agree on every generated system.
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

struct orc_walker {
  int N, n[2], start[2], nn, B;
  double *R;                      /* [N][3]                          */
  double *rij, *rij_uvec;         /* [N][N], [N][N][3]  (i > j)     */
  double *riA, *riA_uvec;         /* [N][nn], [N][nn][3]            */
  double *Uij, *Uij_x, *Uij_xx;   /* pair caches                      */
  double *UiA, *UiA_x, *UiA_xx;
  double U, *U_x, U_xx;
  double *inv[2];                 /* [n][n], inv(o,e)                 */
  double det[2];
  double *Dx[2][3], *Dxx[2];      /* [n][n], D(e,o)                   */
  double *gradPR[2];              /* [n][3]                           */
  double lapPR[2];
  int singular[2];
};

static void *zalloc(size_t n) {
  void *p = calloc(n ? n : 1, 1);
  return p;
}

orc_walker *orc_walker_new(const qmcb_system *s) {
  orc_walker *w = (orc_walker *)zalloc(sizeof(orc_walker));
  int N = s->n_up + s->n_dn, nn = s->n_nuc;
  w->N = N; w->n[0] = s->n_up; w->n[1] = s->n_dn;
  w->start[0] = 0; w->start[1] = s->n_up;
  w->nn = nn; w->B = s->n_basis;
  w->R = zalloc(sizeof(double) * N * 3);
  w->rij = zalloc(sizeof(double) * N * N);
  w->rij_uvec = zalloc(sizeof(double) * N * N * 3);
  w->riA = zalloc(sizeof(double) * N * nn);
  w->riA_uvec = zalloc(sizeof(double) * N * nn * 3);
  w->Uij = zalloc(sizeof(double) * N * N);
  w->Uij_x = zalloc(sizeof(double) * N * N * 3);
  w->Uij_xx = zalloc(sizeof(double) * N * N);
  w->UiA = zalloc(sizeof(double) * N * nn);
  w->UiA_x = zalloc(sizeof(double) * N * nn * 3);
  w->UiA_xx = zalloc(sizeof(double) * N * nn);
  w->U_x = zalloc(sizeof(double) * N * 3);
  for (int sp = 0; sp < 2; sp++) {
    int n = w->n[sp];
    w->inv[sp] = zalloc(sizeof(double) * n * n);
    for (int k = 0; k < 3; k++) w->Dx[sp][k] = zalloc(sizeof(double) * n * n);
    w->Dxx[sp] = zalloc(sizeof(double) * n * n);
    w->gradPR[sp] = zalloc(sizeof(double) * n * 3);
  }
  return w;
}

void orc_walker_free(orc_walker *w) {
  if (!w) return;
  free(w->R); free(w->rij); free(w->rij_uvec); free(w->riA); free(w->riA_uvec);
  free(w->Uij); free(w->Uij_x); free(w->Uij_xx);
  free(w->UiA); free(w->UiA_x); free(w->UiA_xx); free(w->U_x);
  for (int sp = 0; sp < 2; sp++) {
    free(w->inv[sp]);
    for (int k = 0; k < 3; k++) free(w->Dx[sp][k]);
    free(w->Dxx[sp]); free(w->gradPR[sp]);
  }
  free(w);
}

#define CP(f, cnt) memcpy(dst->f, src->f, sizeof(double) * (size_t)(cnt))
void orc_walker_copy(orc_walker *dst, const orc_walker *src) {
  int N = src->N, nn = src->nn;
  CP(R, N * 3); CP(rij, N * N); CP(rij_uvec, N * N * 3); CP(riA, N * nn);
  CP(riA_uvec, N * nn * 3); CP(Uij, N * N); CP(Uij_x, N * N * 3);
  CP(Uij_xx, N * N); CP(UiA, N * nn); CP(UiA_x, N * nn * 3); CP(UiA_xx, N * nn);
  CP(U_x, N * 3);
  dst->U = src->U; dst->U_xx = src->U_xx;
  for (int sp = 0; sp < 2; sp++) {
    int n = src->n[sp];
    CP(inv[sp], n * n);
    for (int k = 0; k < 3; k++) CP(Dx[sp][k], n * n);
    CP(Dxx[sp], n * n); CP(gradPR[sp], n * 3);
    dst->det[sp] = src->det[sp]; dst->lapPR[sp] = src->lapPR[sp];
    dst->singular[sp] = src->singular[sp];
  }
}
#undef CP

/* ---------------------------------------------------------------- basis -- */
/* QMCBasisFunction::evaluateBasisFunctions (QMCBasisFunction.cpp:287-403):
 * chi = x^a y^b z^c sum_i c_i exp(-alpha_i r^2),
 * d chi/dx = sum_i e_i (a/x - 2 alpha_i x),
 * lap chi = sum_i e_i (a(a-1)/x^2 + b(b-1)/y^2 + c(c-1)/z^2
 *                      - alpha_i (4L+6) + 4 r^2 alpha_i^2),
 * with e_i = c_i exp(-alpha_i r^2) x^a y^b z^c; exp reused while consecutive
 * primitives share the exponent (:372-379). */
void orc_basis(const qmcb_system *s, const double *r, double *val, double *gx,
               double *gy, double *gz, double *lap) {
  int prev_atom = -1;
  double d[3] = {0, 0, 0}, rsq = 0;
  double pw[3][3], inv1[3][3], inv2[3][3];
  double p0 = 0, t1 = 0, t2 = 0, t3[3] = {0, 0, 0};
  for (int j = 0; j < s->n_basis; j++) {
    int A = s->bf_atom[j];
    if (A != prev_atom) {
      prev_atom = A;
      for (int k = 0; k < 3; k++) {
        d[k] = r[k] - s->nuc_xyz[3 * A + k];
        pw[k][0] = 1.0; pw[k][1] = d[k]; pw[k][2] = d[k] * d[k];
        inv1[k][0] = 0.0; inv2[k][0] = 0.0;
        inv1[k][1] = 1.0 / d[k]; inv2[k][1] = 0.0 / pw[k][2];
        inv1[k][2] = 2.0 / d[k]; inv2[k][2] = 2.0 / pw[k][2];
      }
      rsq = pw[0][2] + pw[1][2] + pw[2][2];
      p0 = 0.0; /* cache reset per atom (see file header) */
    }
    int a = s->bf_pow[3 * j], b = s->bf_pow[3 * j + 1], c = s->bf_pow[3 * j + 2];
    double poly = pw[0][a] * pw[1][b] * pw[2][c];
    double lcoef = 4.0 * (a + b + c) + 6.0;
    double xv = 0, xx = 0, xy = 0, xz = 0, xl = 0;
    for (int p = s->bf_prim_off[j]; p < s->bf_prim_off[j + 1]; p++) {
      double al = s->prim_alpha[p];
      if (fabs(p0 - al) > 1e-10) {
        p0 = al;
        t1 = exp(-p0 * rsq);
        t2 = 4.0 * rsq * p0 * p0;
        for (int k = 0; k < 3; k++) t3[k] = -2.0 * p0 * d[k];
      }
      double e = t1 * s->prim_coef[p] * poly;
      xv += e;
      xx += e * (inv1[0][a] + t3[0]);
      xy += e * (inv1[1][b] + t3[1]);
      xz += e * (inv1[2][c] + t3[2]);
      xl += e * (inv2[0][a] + inv2[1][b] + inv2[2][c] - p0 * lcoef + t2);
    }
    val[j] = xv; gx[j] = xx; gy[j] = xy; gz[j] = xz; lap[j] = xl;
  }
}

/* ----------------------------------------------------------------- Pade -- */
/* Polynomial::evaluateAll (Polynomial.cpp:127-149), Horner for f, f', f''. */
static void poly_all(const double *coef, int n, double x, double *f, double *df,
                     double *d2f) {
  /* coef[0..n] */
  double F = coef[n], D = 0.0, D2 = 0.0;
  for (int i = n - 1; i >= 0; i--) {
    D2 = D2 * x + D;
    D = D * x + F;
    F = F * x + coef[i];
  }
  *f = F; *df = D; *d2f = 2.0 * D2;
}

/* PadeCorrelationFunction::evaluate (PadeCorrelationFunction.cpp:64-83). */
void orc_pade(const qmcb_pade *fn, double r, double *u, double *du, double *d2u) {
  double num[QMCB_PADE_MAX + 1], den[QMCB_PADE_MAX + 1];
  num[0] = 0.0; den[0] = 1.0;
  for (int i = 0; i < fn->n_num; i++) num[i + 1] = fn->num[i];
  for (int i = 0; i < fn->n_den; i++) den[i + 1] = fn->den[i];
  double a, ap, app, b, bp, bpp;
  poly_all(num, fn->n_num, r, &a, &ap, &app);
  poly_all(den, fn->n_den, r, &b, &bp, &bpp);
  *u = a / b;
  *du = ap / b - a * bp / b / b;
  *d2u = app / b - 2 * ap * bp / b / b + 2 * a * bp * bp / b / b / b - a * bpp / b / b;
}

/* ------------------------------------------------------------ LU / SM ---- */
/* Array2D::ludcmp (Array2D.h:1204-1340): Crout with implicit-scaling partial
 * pivoting and long double accumulators; zero pivot -> 1e-300. */
static int ludcmp(int n, double *a, int *indx, double *dsign) {
  long double *vv = (long double *)malloc(sizeof(long double) * (n ? n : 1));
  int imax = -1;
  *dsign = 1.0;
  for (int i = 0; i < n; i++) {
    long double big = 0.0L, t;
    for (int j = 0; j < n; j++)
      if ((t = fabs(a[i * n + j])) > big) big = t;
    if (big == 0.0L) { free(vv); return 0; }
    vv[i] = (double)(1.0L / big);
  }
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < j; i++) {
      long double sum = -1.0 * a[i * n + j];
      for (int k = 0; k < i; k++) sum += (long double)a[i * n + k] * a[k * n + j];
      a[i * n + j] = (double)(-1.0L * sum);
    }
    long double big = 0.0L, dum;
    for (int i = j; i < n; i++) {
      long double sum = -1.0 * a[i * n + j];
      for (int k = 0; k < j; k++) sum += (long double)a[i * n + k] * a[k * n + j];
      a[i * n + j] = (double)(-1.0L * sum);
      if ((dum = vv[i] * (double)fabs((double)sum)) >= big) { big = dum; imax = i; }
    }
    if (j != imax) {
      for (int k = 0; k < n; k++) {
        double t = a[imax * n + k];
        a[imax * n + k] = a[j * n + k];
        a[j * n + k] = t;
      }
      *dsign = -*dsign;
      vv[imax] = vv[j];
    }
    indx[j] = imax;
    if (a[j * n + j] == 0.0) a[j * n + j] = 1e-300;
    if (j != n) {
      long double d1 = 1.0L / a[j * n + j];
      for (int i = j + 1; i < n; i++) a[i * n + j] *= (double)d1;
    }
  }
  free(vv);
  return 1;
}

/* Array2D::lubksb (Array2D.h:1352-1421). */
static void lubksb(int n, const double *a, const int *indx, double *b) {
  int ii = -1;
  for (int i = 0; i < n; i++) {
    int ip = indx[i];
    long double sum = -1.0 * b[ip];
    b[ip] = b[i];
    if (ii >= 0) {
      for (int j = ii; j <= i - 1; j++) sum += (long double)a[i * n + j] * b[j];
    } else if (sum != 0.0L) {
      ii = i;
    }
    b[i] = (double)(-1.0L * sum);
  }
  for (int i = n - 1; i >= 0; i--) {
    long double sum = -1.0 * b[i];
    for (int j = i + 1; j < n; j++) sum += (long double)a[i * n + j] * b[j];
    b[i] = (double)(-1.0L * sum / a[i * n + i]);
  }
}

/* Array2D::determinant_and_inverse (Array2D.h:1521-1548). */
int orc_det_inverse(int n, double *a, double *inv, double *det) {
  int *indx = (int *)malloc(sizeof(int) * (n ? n : 1));
  double *col = (double *)malloc(sizeof(double) * (n ? n : 1));
  double d;
  int ok = ludcmp(n, a, indx, &d);
  if (ok) {
    for (int j = 0; j < n; j++) {
      for (int i = 0; i < n; i++) col[i] = 0.0;
      col[j] = 1.0;
      lubksb(n, a, indx, col);
      for (int i = 0; i < n; i++) inv[i * n + j] = col[i];
    }
  }
  for (int i = 0; i < n; i++) d *= a[i * n + i];
  *det = d;
  free(indx); free(col);
  return ok;
}

/* Array2D::inverseUpdateOneRow (Array2D.h:1566-1610). */
double orc_inverse_update_row(int n, double *inv, int row, const double *nr) {
  double ratio = 0.0;
  for (int r = 0; r < n; r++) ratio += nr[r] * inv[r * n + row];
  double q = 1.0 / ratio;
  for (int c = 0; c < n; c++) {
    if (c == row) continue;
    double v = 0.0;
    for (int l = 0; l < n; l++) v += nr[l] * inv[l * n + c];
    v *= q;
    for (int r = 0; r < n; r++) inv[r * n + c] -= v * inv[r * n + row];
  }
  for (int r = 0; r < n; r++) inv[r * n + row] *= q;
  return ratio;
}

/* ------------------------------------------------------------ distances -- */
/* QMCWalkerData::updateDistances (QMCWalkerData.cpp:302-346). */
static void update_distances(const qmcb_system *s, orc_walker *w, int whichE) {
  int N = w->N, nn = w->nn;
  const double *R = w->R;
  for (int i = 0; i < N; i++) {
    if (whichE != -1 && i != whichE) continue;
    if (whichE == -1) {
      for (int j = 0; j < i; j++) {
        double dx = R[3 * i] - R[3 * j], dy = R[3 * i + 1] - R[3 * j + 1],
               dz = R[3 * i + 2] - R[3 * j + 2];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        w->rij[i * N + j] = r;
        for (int k = 0; k < 3; k++)
          w->rij_uvec[(i * N + j) * 3 + k] = (R[3 * i + k] - R[3 * j + k]) / r;
      }
    } else {
      for (int j = 0; j < N; j++) {
        if (j == whichE) continue;
        int e1 = j > whichE ? j : whichE, e2 = j > whichE ? whichE : j;
        double dx = R[3 * j] - R[3 * whichE], dy = R[3 * j + 1] - R[3 * whichE + 1],
               dz = R[3 * j + 2] - R[3 * whichE + 2];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        w->rij[e1 * N + e2] = r;
        for (int k = 0; k < 3; k++)
          w->rij_uvec[(e1 * N + e2) * 3 + k] = (R[3 * e1 + k] - R[3 * e2 + k]) / r;
      }
    }
    for (int A = 0; A < nn; A++) {
      double dx = R[3 * i] - s->nuc_xyz[3 * A], dy = R[3 * i + 1] - s->nuc_xyz[3 * A + 1],
             dz = R[3 * i + 2] - s->nuc_xyz[3 * A + 2];
      double r = sqrt(dx * dx + dy * dy + dz * dz);
      w->riA[i * nn + A] = r;
      for (int k = 0; k < 3; k++)
        w->riA_uvec[(i * nn + A) * 3 + k] = (R[3 * i + k] - s->nuc_xyz[3 * A + k]) / r;
    }
  }
}

/* -------------------------------------------------------------- Jastrow -- */
/* QMCJastrowElectronElectron::collectForPair (QMCJastrowElectronElectron.cpp:
 * 229-266). */
static void collect_pair(orc_walker *w, int E1, int E2, const qmcb_pade *f) {
  int N = w->N;
  double r = w->rij[E1 * N + E2], u, du, d2u;
  orc_pade(f, r, &u, &du, &d2u);
  double uxx = 2.0 * (2.0 / r * du + d2u);
  w->U -= w->Uij[E1 * N + E2];
  w->U += u;
  w->Uij[E1 * N + E2] = u;
  for (int i = 0; i < 3; i++) {
    double t = w->Uij_x[(E1 * N + E2) * 3 + i];
    w->U_x[E1 * 3 + i] -= t;
    w->U_x[E2 * 3 + i] += t;
    t = du * w->rij_uvec[(E1 * N + E2) * 3 + i];
    w->Uij_x[(E1 * N + E2) * 3 + i] = t;
    w->U_x[E1 * 3 + i] += t;
    w->U_x[E2 * 3 + i] -= t;
  }
  w->U_xx -= w->Uij_xx[E1 * N + E2];
  w->U_xx += uxx;
  w->Uij_xx[E1 * N + E2] = uxx;
}

/* updateAll / updateOne (QMCJastrowElectronElectron.cpp:68-227). */
static void jastrow_ee(const qmcb_system *s, orc_walker *w, int E) {
  int na = w->n[0], nb = w->n[1], N = w->N;
  if (E == -1) {
    if (na > 0 && nb > 0)
      for (int e1 = 0; e1 < na; e1++)
        for (int e2 = na; e2 < N; e2++) collect_pair(w, e2, e1, &s->ee_updn);
    if (na > 1)
      for (int e1 = 0; e1 < na; e1++)
        for (int e2 = 0; e2 < e1; e2++) collect_pair(w, e1, e2, &s->ee_upup);
    if (nb > 1)
      for (int e1 = na; e1 < N; e1++)
        for (int e2 = na; e2 < e1; e2++) collect_pair(w, e1, e2, &s->ee_dndn);
    return;
  }
  int isA = E < na;
  if (na > 0 && nb > 0) {
    if (isA) for (int eb = na; eb < N; eb++) collect_pair(w, eb, E, &s->ee_updn);
    else for (int ea = 0; ea < na; ea++) collect_pair(w, E, ea, &s->ee_updn);
  }
  if (na > 1 && isA) {
    for (int ea = 0; ea < E; ea++) collect_pair(w, E, ea, &s->ee_upup);
    for (int ea = E + 1; ea < na; ea++) collect_pair(w, ea, E, &s->ee_upup);
  }
  if (nb > 1 && !isA) {
    for (int eb = na; eb < E; eb++) collect_pair(w, E, eb, &s->ee_dndn);
    for (int eb = E + 1; eb < N; eb++) collect_pair(w, eb, E, &s->ee_dndn);
  }
}

/* QMCJastrowElectronNuclear::evaluate (QMCJastrowElectronNuclear.cpp:54-187). */
static void jastrow_en(const qmcb_system *s, orc_walker *w, int E) {
  int nn = w->nn, na = w->n[0];
  for (int A = 0; A < nn; A++) {
    for (int e = 0; e < w->N; e++) {
      if (E != e && E != -1) continue;
      const qmcb_pade *f = e < na ? &s->en_up[A] : &s->en_dn[A];
      double r = w->riA[e * nn + A], u, du, d2u;
      orc_pade(f, r, &u, &du, &d2u);
      double uxx = 2.0 / r * du + d2u;
      w->U -= w->UiA[e * nn + A];
      w->U += u;
      w->UiA[e * nn + A] = u;
      w->U_xx -= w->UiA_xx[e * nn + A];
      w->U_xx += uxx;
      w->UiA_xx[e * nn + A] = uxx;
      for (int i = 0; i < 3; i++) {
        w->U_x[e * 3 + i] -= w->UiA_x[(e * nn + A) * 3 + i];
        double t = du * w->riA_uvec[(e * nn + A) * 3 + i];
        w->UiA_x[(e * nn + A) * 3 + i] = t;
        w->U_x[e * 3 + i] += t;
      }
    }
  }
}

/* --------------------------------------------------------------- Slater -- */
/* QMCSlater::evaluate + update_Ds + calculate_DerivativeRatios
 * (QMCSlater.cpp:390-764, 828-884) for one spin; returns singular flag. */
static int slater(const qmcb_system *s, orc_walker *w, int sp, int whichE) {
  int n = w->n[sp], B = s->n_basis, start = w->start[sp];
  if (n == 0) return 0;
  if (whichE != -1 && (whichE < start || whichE >= start + n)) return w->singular[sp];
  const double *C = sp == 0 ? s->mo_up : s->mo_dn;
  int nup = whichE == -1 ? n : 1;
  int e0 = whichE == -1 ? start : whichE;
  double *X = malloc(sizeof(double) * 5 * B);
  double *Dv = malloc(sizeof(double) * nup * n);
  double *Dd[4];
  for (int k = 0; k < 4; k++) Dd[k] = malloc(sizeof(double) * nup * n);
  for (int e = 0; e < nup; e++) {
    orc_basis(s, w->R + 3 * (e0 + e), X, X + B, X + 2 * B, X + 3 * B, X + 4 * B);
    /* Array2D::gemm with rhsIsTransposed (Array2D.h:425-504): D = X C^T. */
    for (int o = 0; o < n; o++) {
      double acc[5] = {0, 0, 0, 0, 0};
      for (int b = 0; b < B; b++) {
        double cb = C[o * B + b];
        for (int k = 0; k < 5; k++) acc[k] += X[k * B + b] * cb;
      }
      Dv[e * n + o] = acc[0];
      for (int k = 0; k < 3; k++) Dd[k][e * n + o] = acc[1 + k];
      Dd[3][e * n + o] = acc[4];
    }
  }
  int row = whichE == -1 ? -1 : whichE - start;
  int iEl = row < 0 ? 0 : row;
  for (int e = 0; e < nup; e++)
    for (int o = 0; o < n; o++) {
      for (int k = 0; k < 3; k++) w->Dx[sp][k][(iEl + e) * n + o] = Dd[k][e * n + o];
      w->Dxx[sp][(iEl + e) * n + o] = Dd[3][e * n + o];
    }
  int calcOK = 1;
  if (row < 0) {
    calcOK = orc_det_inverse(n, Dv, w->inv[sp], &w->det[sp]);
  } else {
    double ratio = orc_inverse_update_row(n, w->inv[sp], row, Dv);
    if (ratio != 0.0) w->det[sp] *= ratio;
    else calcOK = 0;
  }
  free(X); free(Dv);
  for (int k = 0; k < 4; k++) free(Dd[k]);
  if (!calcOK) { w->singular[sp] = 1; return 1; }
  double lap = 0.0;
  double *g = w->gradPR[sp];
  for (int e = 0; e < n; e++) g[3 * e] = g[3 * e + 1] = g[3 * e + 2] = 0.0;
  for (int o = 0; o < n; o++)
    for (int e = 0; e < n; e++) {
      double iv = w->inv[sp][o * n + e];
      lap += w->Dxx[sp][e * n + o] * iv;
      for (int k = 0; k < 3; k++) g[3 * e + k] += w->Dx[sp][k][e * n + o] * iv;
    }
  w->lapPR[sp] = lap;
  w->singular[sp] = 0;
  return 0;
}

/* ------------------------------------------------------------- potential -- */
/* QMCPotential_Energy::calc_P_nn/en/ee (QMCPotential_Energy.cpp:242-333). */
static double v_nn(const qmcb_system *s) {
  double v = 0.0;
  for (int i = 0; i < s->n_nuc; i++)
    for (int j = i + 1; j < s->n_nuc; j++) {
      double dx = s->nuc_xyz[3 * i] - s->nuc_xyz[3 * j],
             dy = s->nuc_xyz[3 * i + 1] - s->nuc_xyz[3 * j + 1],
             dz = s->nuc_xyz[3 * i + 2] - s->nuc_xyz[3 * j + 2];
      v += s->nuc_z[i] * s->nuc_z[j] / sqrt(dx * dx + dy * dy + dz * dz);
    }
  return v;
}

/* ------------------------------------------------------------- evaluate -- */
int orc_evaluate(const qmcb_system *s, orc_walker *w, const double *R, int whichE,
                 int drift_mod, double dt, double umrigar_a, qmcb_results *out) {
  int N = w->N, nn = w->nn, na = w->n[0];
  if (whichE == -1) memcpy(w->R, R, sizeof(double) * N * 3);
  else memcpy(w->R + 3 * whichE, R + 3 * whichE, sizeof(double) * 3);
  update_distances(s, w, whichE);

  /* QMCJastrow::evaluate (QMCJastrow.cpp:134-187): e-n then e-e. */
  jastrow_en(s, w, whichE);
  jastrow_ee(s, w, whichE);

  int sing = 0;
  int sA = slater(s, w, 0, whichE);
  int sB = slater(s, w, 1, whichE);
  sing = sA || sB;
  double det_ratio = 1.0;

  double Pen = 0.0, Pee = 0.0;
  for (int A = 0; A < nn; A++)
    for (int e = 0; e < N; e++) Pen -= s->nuc_z[A] / w->riA[e * nn + A];
  for (int i = 0; i < N; i++)
    for (int j = 0; j < i; j++) Pee += (-1.0) * (-1.0) / w->rij[i * N + j];
  double PE = v_nn(s) + Pee + Pen;

  double *grad = malloc(sizeof(double) * N * 3);
  double *mgrad = malloc(sizeof(double) * N * 3);
  double lapR = 0.0;
  double logpsi = 0.0;
  double mod_ratio = 0.0;
  int sign = 1;
  if (!sing) {
    /* calculate_Psi_quantities + update_SCF (QMCSCFJastrow.cpp:196-327), one
     * determinant with CI coefficient 1 so termPR = 1. */
    double dA = w->n[0] > 0 ? w->det[0] : 1.0, dB = w->n[1] > 0 ? w->det[1] : 1.0;
    logpsi = log(fabs(dA)) + log(fabs(dB)) + w->U;
    sign = ((dA < 0) != (dB < 0)) ? -1 : 1;
    lapR = (w->n[0] > 0 ? w->lapPR[0] : 0.0) + (w->n[1] > 0 ? w->lapPR[1] : 0.0);
    lapR += w->U_xx;
    for (int e = 0; e < N; e++)
      for (int k = 0; k < 3; k++) {
        double dx = e < na ? w->gradPR[0][3 * e + k] : w->gradPR[1][3 * (e - na) + k];
        grad[3 * e + k] = dx + w->U_x[3 * e + k];
        lapR += w->U_x[3 * e + k] * dx * 2.0;
        lapR += w->U_x[3 * e + k] * w->U_x[3 * e + k];
      }
    /* calculate_Modified_Grad_PsiRatio (QMCSCFJastrow.cpp:329-472). */
    memcpy(mgrad, grad, sizeof(double) * N * 3);
    if (drift_mod == QMCB_DRIFT_UMRIGAR93_EQUAL) {
      double m = 0.0;
      for (int i = 0; i < 3 * N; i++) m += grad[i] * grad[i];
      double f = (-1.0 + sqrt(1.0 + 2 * umrigar_a * m * dt)) / (umrigar_a * m * dt);
      for (int i = 0; i < 3 * N; i++) mgrad[i] *= f;
    } else if (drift_mod == QMCB_DRIFT_UMRIGAR93_UNEQUAL) {
      for (int i = 0; i < N; i++) {
        /* QMCMolecule::findClosestNucleusIndex: smallest r_iA, first wins */
        int cn = 0;
        for (int A = 1; A < nn; A++)
          if (w->riA[i * nn + A] < w->riA[i * nn + cn]) cn = A;
        double nv[3], vv[3], dsq = 0.0;
        for (int k = 0; k < 3; k++) {
          double t = w->R[3 * i + k] - s->nuc_xyz[3 * cn + k];
          nv[k] = t; dsq += t * t;
        }
        for (int k = 0; k < 3; k++) nv[k] *= 1.0 / sqrt(dsq);
        double v2 = 0.0;
        for (int k = 0; k < 3; k++) { vv[k] = grad[3 * i + k]; v2 += vv[k] * vv[k]; }
        if (fabs(v2) < 1e-10) { vv[0] = vv[1] = vv[2] = 0.0; }
        else for (int k = 0; k < 3; k++) vv[k] *= 1.0 / sqrt(v2);
        double Z = s->nuc_z[cn];
        double t = Z * Z * dsq;
        double dotp = nv[0] * vv[0] + nv[1] * vv[1] + nv[2] * vv[2];
        double a = 0.5 * (1.0 + dotp) + t / (10.0 * (4.0 + t));
        double f = fabs(v2) < 1e-10 ? 1.0
                   : (-1.0 + sqrt(1.0 + 2.0 * a * v2 * dt)) / (a * v2 * dt);
        for (int k = 0; k < 3; k++) mgrad[3 * i + k] *= f;
      }
    }
    double lm = 0.0, lu = 0.0;
    for (int i = 0; i < 3 * N; i++) { lm += mgrad[i] * mgrad[i]; lu += grad[i] * grad[i]; }
    lm = sqrt(lm); lu = sqrt(lu);
    mod_ratio = lm / lu; /* modificationRatio (QMCSCFJastrow.cpp:450-458) */
    if (isnan(lm) || isnan(lu) || isnan(lm / lu) || lu == 0) sing = 1;
  }
  if (out) {
    if (out->log_psi) out->log_psi[0] = logpsi;
    if (out->sign_psi) out->sign_psi[0] = (int8_t)sign;
    if (out->grad_lnpsi) memcpy(out->grad_lnpsi, grad, sizeof(double) * N * 3);
    if (out->mod_drift) memcpy(out->mod_drift, mgrad, sizeof(double) * N * 3);
    if (out->lap_ratio) out->lap_ratio[0] = lapR;
    if (out->e_local) out->e_local[0] = -0.5 * lapR + PE;
    if (out->kinetic) out->kinetic[0] = -0.5 * lapR;
    if (out->potential) out->potential[0] = PE;
    if (out->v_ne) out->v_ne[0] = Pen;
    if (out->v_ee) out->v_ee[0] = Pee;
    if (out->det_ratio) out->det_ratio[0] = det_ratio;
    if (out->jastrow_u) out->jastrow_u[0] = w->U;
    if (out->singular) out->singular[0] = (int8_t)sing;
    if (out->mod_ratio) out->mod_ratio[0] = mod_ratio;
    if (out->r2) {
      /* x2, y2, z2 of the evaluated configuration (QMCSCFJastrow.cpp:164-169) */
      out->r2[0] = out->r2[1] = out->r2[2] = 0.0;
      for (int e = 0; e < N; e++)
        for (int k = 0; k < 3; k++) out->r2[k] += w->R[3 * e + k] * w->R[3 * e + k];
    }
  }
  free(grad); free(mgrad);
  return sing;
}

/* ------------------------------------------------------- CPU VMC timing -- */
static unsigned long long sm64(unsigned long long *x) {
  unsigned long long z = (*x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double unif(unsigned long long *x) { return (sm64(x) >> 11) * (1.0 / 9007199254740992.0); }
static double gauss(unsigned long long *x) {
  double u1 = unif(x), u2 = unif(x);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

double orc_vmc_time(const qmcb_system *s, int nw, const double *R0, int n_moves,
                    double tau, unsigned long long seed, double *mean_el) {
  int N = s->n_up + s->n_dn;
  orc_walker **W = malloc(sizeof(orc_walker *) * nw);
  double *R = malloc(sizeof(double) * nw * N * 3);
  double *F = malloc(sizeof(double) * nw * N * 3), *Fn = malloc(sizeof(double) * N * 3);
  double *lp = malloc(sizeof(double) * nw), *el = malloc(sizeof(double) * nw);
  memcpy(R, R0, sizeof(double) * nw * N * 3);
  qmcb_results o;
  memset(&o, 0, sizeof(o));
  for (int w = 0; w < nw; w++) {
    W[w] = orc_walker_new(s);
    o.log_psi = lp + w; o.e_local = el + w; o.mod_drift = F + (size_t)w * N * 3;
    orc_evaluate(s, W[w], R + (size_t)w * N * 3, -1, QMCB_DRIFT_NONE, tau, 0.5, &o);
  }
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  double esum = 0.0;
  long long ns = 0;
  double sig = sqrt(tau);
  for (int m = 0; m < n_moves; m++) {
    int e = m % N;
    for (int w = 0; w < nw; w++) {
      double *Rw = R + (size_t)w * N * 3, *Fw = F + (size_t)w * N * 3;
      double old[3], fwd = 0.0;
      for (int i = 0; i < N; i++)
        if (i != e)
          for (int k = 0; k < 3; k++) fwd += tau * Fw[3 * i + k] * tau * Fw[3 * i + k];
      for (int k = 0; k < 3; k++) {
        old[k] = Rw[3 * e + k];
        double g = sig * gauss(&seed);
        Rw[3 * e + k] += tau * Fw[3 * e + k] + g;
        fwd += g * g;
      }
      double lpn, eln;
      o.log_psi = &lpn; o.e_local = &eln; o.mod_drift = Fn;
      int sing = orc_evaluate(s, W[w], Rw, e, QMCB_DRIFT_NONE, tau, 0.5, &o);
      double rev = 0.0;
      for (int i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) {
          double oldc = i == e ? old[k] : Rw[3 * i + k];
          double t = oldc - Rw[3 * i + k] - tau * Fn[3 * i + k];
          rev += t * t;
        }
      double p = sing ? 0.0 : exp(2.0 * (lpn - lp[w]) - (rev - fwd) / (2.0 * tau));
      if (p > 1.0) p = 1.0;
      if (p > unif(&seed)) {
        lp[w] = lpn; el[w] = eln;
        memcpy(Fw, Fn, sizeof(double) * N * 3);
      } else {
        for (int k = 0; k < 3; k++) Rw[3 * e + k] = old[k];
        o.log_psi = lp + w; o.e_local = el + w; o.mod_drift = Fw;
        orc_evaluate(s, W[w], Rw, e, QMCB_DRIFT_NONE, tau, 0.5, &o);
      }
      esum += el[w];
      ns++;
    }
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (mean_el) *mean_el = esum / (ns ? ns : 1);
  for (int w = 0; w < nw; w++) orc_walker_free(W[w]);
  free(W); free(R); free(F); free(Fn); free(lp); free(el);
  return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
