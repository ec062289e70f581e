/*
 * oracle.h -- CPU restatement of QMcBeaver's FP64 walker evaluation.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is compared
 * against; it is never linked into, loaded by or called from the product
 * library (qmcbeaver_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Parity status of each function is listed in oracle.c's header.
 */
#ifndef QMCB_ORACLE_H
#define QMCB_ORACLE_H

#include "../include/qmcbeaver_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_walker orc_walker;

/* Per-walker record mirroring QMCWalkerData in one_e_per_iter mode
 * (QMCWalkerData.cpp:56-161): Slater inverse/derivative storage plus the
 * Jastrow pair caches.  Fresh walkers have zero caches. */
orc_walker *orc_walker_new(const qmcb_system *s);
void orc_walker_free(orc_walker *w);
/* Deep copy (QMCWalker::createChildWalkers full copy). */
void orc_walker_copy(orc_walker *dst, const orc_walker *src);

/* QMCSCFJastrow::evaluate for one walker (QMCSCFJastrow.cpp:91-181):
 *  whichE == -1: all-electron evaluate (LU inverse), R is the full config;
 *  whichE >= 0 : single-electron evaluate: R is the full config with row
 *                whichE moved; Sherman-Morrison row update of the resident
 *                inverse (QMCSlater.cpp:713-724).
 * Results are written to `out` (num = 1 arrays; NULL members skipped). */
int orc_evaluate(const qmcb_system *s, orc_walker *w, const double *R,
                 int whichE, int drift_mod, double dt, double umrigar_a,
                 qmcb_results *out);

/* Building blocks, exposed for closed-form tests. */
/* Basis values of one point: QMCBasisFunction::evaluateBasisFunctions
 * (QMCBasisFunction.cpp:287-403); out arrays [n_basis]. */
void orc_basis(const qmcb_system *s, const double *r, double *val, double *gx,
               double *gy, double *gz, double *lap);
/* Pade u, u', u'' (PadeCorrelationFunction.cpp:64-83, Polynomial.cpp:127-149). */
void orc_pade(const qmcb_pade *f, double r, double *u, double *du, double *d2u);
/* NR LU inverse + determinant (Array2D.h:1204-1421, 1521-1548).  a[n][n] is
 * destroyed, inv[n][n] receives a^-1.  Returns 0 when singular. */
int orc_det_inverse(int n, double *a, double *inv, double *det);
/* Row-replacement Sherman-Morrison (Array2D.h:1566-1610). inv[n][n] is
 * updated in place; returns the determinant ratio. */
double orc_inverse_update_row(int n, double *inv, int row, const double *newrow);

/* Single-thread VMC timing loop used as the CPU baseline: n_moves
 * single-electron importance-sampled proposals per walker, each followed by
 * the full evaluate, Metropolis, and re-evaluate on rejection
 * (QMCWalker.cpp:179-282, 343-398, 756-780, 1578-1685).  Returns seconds. */
double orc_vmc_time(const qmcb_system *s, int n_walkers, const double *R0,
                    int n_moves, double tau, unsigned long long seed,
                    double *mean_e_local);

#ifdef __cplusplus
}
#endif
#endif
