/*
 * Driver of the reference's own Hairer integrators, compiled from the reference sources by oracle/Makefile into
 * oracle/_ref/libhairer_ref.so (target `ref`): RODAS (hairerRodas.f) and RADAU5 (hairerRadau5.f + hairerToolbox.f)
 * exactly as OSlo's wrap_rodas / wrap_radau5 call them (OSlo.f90:349-412, 485-527: IJAC = 0 full numerical Jacobian,
 * ITOL = 1 vector tolerances, IMAS = 0, IOUT = 0, WORK(1) = UROUND = 1.1e-19, H = 0, IWORK = 0 defaults), on the
 * chemistry right-hand side of the oracle (mose_chem_oracle.c, rhs_native) or on Robertson's problem.
 *
 * TEST INFRASTRUCTURE ONLY: it checks the oracle's restatement of the two integrators against the reference's code.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mose_gpu.h"

typedef void (*f_fcn)(int *n, double *x, double *y, double *f, double *rpar, int *ipar);
typedef void (*f_any)(void);

extern void radau5_(int *n, f_fcn fcn, double *x, double *y, double *xend, double *h, double *rtol, double *atol,
                    int *itol, f_any jac, int *ijac, int *mljac, int *mujac, f_any mas, int *imas, int *mlmas,
                    int *mumas, f_any solout, int *iout, double *work, int *lwork, int *iwork, int *liwork,
                    double *rpar, int *ipar, int *idid, int *nn, int *nn2, int *nn3, int *nn4, double *xsol,
                    double *hsol, double *c2m1, double *c1m1, int *mle, int *mue, int *mbjac, int *mbb, int *mdiag,
                    int *mdiff, int *mbdiag);
extern void rodas_(int *n, f_fcn fcn, int *ifcn, double *x, double *y, double *xend, double *h, double *rtol,
                   double *atol, int *itol, f_any jac, int *ijac, int *mljac, int *mujac, f_any dfx, int *idfx,
                   f_any mas, int *imas, int *mlmas, int *mumas, f_any solout, int *iout, double *work, int *lwork,
                   int *iwork, int *liwork, double *rpar, int *ipar, int *idid);

/* from mose_chem_oracle.c (linked into the same library) */
void oc_rhs(void *h, const double *Z, double *F);

static _Thread_local void *g_mech;
static _Thread_local int g_nfcn;

static void chem_fcn(int *n, double *x, double *y, double *f, double *rpar, int *ipar) {
  (void)n; (void)x; (void)rpar; (void)ipar;
  oc_rhs(g_mech, y, f);
  g_nfcn++;
}

static void rober_fcn(int *n, double *x, double *y, double *f, double *rpar, int *ipar) {
  (void)n; (void)x; (void)rpar; (void)ipar;
  f[0] = -0.04 * y[0] + 1.0e4 * y[1] * y[2];
  f[2] = 3.0e7 * y[1] * y[1];
  f[1] = -f[0] - f[2];
  g_nfcn++;
}

static void dummy(void) {}

/* one integration over [0, xend]; stats = {nstep, naccpt, nrejct, nfcn} (IWORK(16..18) / IWORK(14), RODAS and
 * RADAU5 alike) */
static int integrate(int method, int n, f_fcn fcn, double *y, double xend, const double *rtol, const double *atol,
                     int *stats) {
  double x = 0.0, h = 0.0, rt[MOSE_MAX_SPECIES + 1], at[MOSE_MAX_SPECIES + 1], rpar[1] = {0.0};
  int ipar[1] = {0}, itol = 1, ijac = 0, mljac = n, mujac = 0, imas = 0, mlmas = n, mumas = 0, iout = 0, idid = 0;
  for (int i = 0; i < n; ++i) { rt[i] = rtol[i]; at[i] = atol[i]; }
  int lwork = 4 * n * n + 12 * n + 40, liwork = 3 * n + 40;
  double *work = calloc((size_t)lwork, sizeof(double));
  int *iwork = calloc((size_t)liwork, sizeof(int));
  work[0] = 1.1e-19;
  if (method == MOSE_ODE_RADAU5) {
    lwork = 4 * n * n + 12 * n + 20;
    liwork = 3 * n + 20;
    int nn, nn2, nn3, nn4, mle, mue, mbjac, mbb, mdiag, mdiff, mbdiag;
    double xsol, hsol, c2m1, c1m1;
    radau5_(&n, fcn, &x, y, &xend, &h, rt, at, &itol, dummy, &ijac, &mljac, &mujac, dummy, &imas, &mlmas, &mumas,
            dummy, &iout, work, &lwork, iwork, &liwork, rpar, ipar, &idid, &nn, &nn2, &nn3, &nn4, &xsol, &hsol,
            &c2m1, &c1m1, &mle, &mue, &mbjac, &mbb, &mdiag, &mdiff, &mbdiag);
  } else {
    int ifcn = 0, idfx = 0;
    rodas_(&n, fcn, &ifcn, &x, y, &xend, &h, rt, at, &itol, dummy, &ijac, &mljac, &mujac, dummy, &idfx, dummy, &imas,
           &mlmas, &mumas, dummy, &iout, work, &lwork, iwork, &liwork, rpar, ipar, &idid);
  }
  if (stats) { stats[0] = iwork[15]; stats[1] = iwork[16]; stats[2] = iwork[17]; stats[3] = iwork[13]; }
  free(work);
  free(iwork);
  return idid;
}

int ref_robertson(int method, double *y, double xend, double rtol, double atol, int *stats) {
  double rt[3] = {rtol, rtol, rtol}, at[3] = {atol, atol, atol};
  return integrate(method, 3, rober_fcn, y, xend, rt, at, stats);
}

/* Chemistry_Newstate_blk (src/lib/physics/Lib_Chemistry.f90:33-76) on one cell with the reference integrator;
 * mech = an oc_create handle (its Ri are R_univ / wm); q = (species densities, p), updated; returns idid */
int ref_advance_cell(void *mech, int method, int ns, const double *Ri, double *q, double dt, const double *rtol,
                     const double *atol, double *hr, int *nstep) {
  double Z[MOSE_MAX_SPECIES + 1], rho = 0.0, R = 0.0;
  for (int s = 0; s < ns; ++s) { Z[s] = q[s]; rho = rho + Z[s]; }
  for (int s = 0; s < ns; ++s) R = R + Ri[s] * Z[s] / rho;
  Z[ns] = q[ns] / (rho * R);
  double Tb = Z[ns];
  g_mech = mech;
  int stats[4];
  int idid = integrate(method, ns + 1, chem_fcn, Z, dt, rtol, atol, stats);
  double rs = 0.0, R2 = 0.0;
  for (int s = 0; s < ns; ++s) {
    if (Z[s] <= 1e-20) Z[s] = 1e-20;
    q[s] = Z[s];
    rs = rs + Z[s];
  }
  for (int s = 0; s < ns; ++s) R2 = R2 + Ri[s] * Z[s] / rs;
  q[ns] = Z[ns] * rho * R2;
  if (hr) *hr = fabs(Z[ns] - Tb) / Tb;
  if (nstep) *nstep = stats[0];
  return idid;
}
