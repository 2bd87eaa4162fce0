/*
 * Flat entry points of the CPU oracle for ctypes (tests/ and bench.py's cpu_baseline leg only).
 * TEST INFRASTRUCTURE: see mose_oracle.h.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mose_oracle.h"

void o_fill_ghosts(const o_case *C, double *P);
void o_residual(const o_case *C, double *P, double *R, double dt);
void o_irs(const o_case *C, double *R, double dt);
double o_compute_dt(const o_case *C, const double *P);
void o_newstate(const o_case *C, double *P, const double *PO, const double *R, double dt, double rkc);
int o_step(const o_case *C, double *P, double *PO, double *R, int nsteps, double *dt_out);
double o_rlimiter(int lim, double x, double y);
void o_conv_face(const o_case *C, const double *pm1, const double *p0, const double *pp1, const double *pp2,
                 const double dlc[4], const double nrm[3], double area, double *res0);
void o_riemann(const o_case *C, int rs, double bet, const double *dl, double ul, double vl, double wl, double pl,
               double al, double dltot, const double *dr, double ur, double vr, double wr, double pr, double ar,
               double drtot, double nx, double ny, double nz, double F[5]);
double o_sound_speed(const o_case *C, const double *rhoi, double p);

/* the chemistry oracle (mose_chem_oracle.c) */
int oc_advance_m(void *h, int method, int64_t ncell, double *state, double dt, const double *rtol, const double *atol,
                 double *hr, int32_t *idid, int32_t *nstep, int nthreads);

typedef struct {
  void *mech;                          /* oc_create handle */
  int method;                          /* MOSE_ODE_RODAS / MOSE_ODE_RADAU5 */
  double rtol[ONSMAX + 1], atol[ONSMAX + 1];
  int64_t failed;                      /* cells that ended with idid < 0, summed over the calls */
} oa_chem;

typedef struct {
  o_case c;
  o_block *blk;
  mose_bc_record *rec;
  mose_farfield *ff;
  double *tabs;
  oa_chem chem;
} oa_t;

static void tables_init(o_tables *t, const mose_tables *tb, double *store) {
  memset(t, 0, sizeof *t);
  t->ns = tb->nsc;
  t->tmax = MOSE_NTAB - 1;
  size_t n = (size_t)tb->nsc * MOSE_NTAB;
  memcpy(store, tb->cp, n * sizeof(double));
  memcpy(store + n, tb->h, n * sizeof(double));
  if (tb->mu) memcpy(store + 2 * n, tb->mu, n * sizeof(double));
  if (tb->k) memcpy(store + 3 * n, tb->k, n * sizeof(double));
  t->cp = store; t->h = store + n; t->mi = store + 2 * n; t->k = store + 3 * n;
  /* Ri_tab = Runiv/wm_tab and the Wilke pair constants: Load_ThermoTransport.f90:51-67 */
  for (int s = 0; s < tb->nsc; ++s) { t->wm[s] = tb->wm[s]; t->Ri[s] = 8314.46261815324 / tb->wm[s]; }
  for (int sj = 0; sj < tb->nsc; ++sj)
    for (int si = 0; si < tb->nsc; ++si) {
      double ratio = tb->wm[si] / tb->wm[sj];
      t->mm025[si][sj] = pow(ratio, -0.25);
      t->is8[si][sj] = 1.0 / sqrt(8.0 * (1.0 + ratio));
    }
}

void *oa_create(const mose_scheme *sc, const mose_tables *tb, int nb, const int *dims, const double *dl,
                const double *m, const double *vol, const double *yn, const double *fm1, const double *fm2,
                const double *fm3, int nrec, const mose_bc_record *rec, int nff, const mose_farfield *ff,
                int nthreads) {
  if (tb->nsc < 1 || tb->nsc > ONSMAX) return NULL;
  oa_t *a = (oa_t *)calloc(1, sizeof *a);
  o_case *C = &a->c;
  o_scheme *S = &C->sc;
  S->nsc = sc->nsc; S->nrans = sc->nrans; S->nprim = sc->nsc + 4 + sc->nrans;
  S->nu = sc->nsc; S->nv = sc->nsc + 1; S->nw = sc->nsc + 2; S->np = sc->nsc + 3; S->nt = sc->nsc + 4;
  S->lim = sc->limiter; S->rs = sc->riemann; S->mco = sc->mco; S->visc = sc->viscous;
  S->rans = sc->nrans > 0 ? 2 : 0;
  S->cfl = sc->cfl; S->vnn = sc->vnn; S->prl = sc->prandtl; S->sc = sc->sc; S->sct = sc->sct; S->prt = sc->prt;
  S->n_rk = sc->n_rk; S->ndir = sc->ndir; S->point_implicit = sc->point_implicit; S->guard = sc->thermo_guard;
  S->prim = sc->integration == MOSE_INT_PRIM;
  S->irs = sc->irs != 0; S->irs_njacobi = sc->irs_njacobi; S->irs_beta = sc->irs_beta;
  a->tabs = (double *)calloc((size_t)4 * tb->nsc * MOSE_NTAB, sizeof(double));
  tables_init(&C->tab, tb, a->tabs);
  a->blk = (o_block *)calloc((size_t)nb, sizeof(o_block));
  int64_t NC = 0, NF1 = 0, NF2 = 0, NF3 = 0;
  for (int b = 0; b < nb; ++b) {
    o_block *B = &a->blk[b];
    B->n1 = dims[3 * b]; B->n2 = dims[3 * b + 1]; B->n3 = dims[3 * b + 2];
    B->pi = B->n1 + 2 * OGC; B->pj = B->n2 + 2 * OGC;
    B->boff = NC; B->f1off = NF1; B->f2off = NF2; B->f3off = NF3;
    NC += (int64_t)B->pi * B->pj * (B->n3 + 2 * OGC);
    NF1 += (int64_t)(B->n1 + 1) * B->n2 * B->n3;
    NF2 += (int64_t)B->n1 * (B->n2 + 1) * B->n3;
    NF3 += (int64_t)B->n1 * B->n2 * (B->n3 + 1);
  }
  C->nb = nb; C->blk = a->blk; C->NC = NC; C->NF1 = NF1; C->NF2 = NF2; C->NF3 = NF3;
  C->dl = dl; C->m = m; C->vol = vol; C->yn = yn; C->fm1 = fm1; C->fm2 = fm2; C->fm3 = fm3;
  a->rec = (mose_bc_record *)malloc(sizeof(mose_bc_record) * (size_t)(nrec > 0 ? nrec : 1));
  if (nrec > 0) memcpy(a->rec, rec, sizeof(mose_bc_record) * (size_t)nrec);
  C->nrec = nrec; C->rec = a->rec;
  a->ff = (mose_farfield *)malloc(sizeof(mose_farfield) * (size_t)(nff > 0 ? nff : 1));
  if (nff > 0) memcpy(a->ff, ff, sizeof(mose_farfield) * (size_t)nff);
  C->nff = nff; C->ff = a->ff;
  C->nthreads = nthreads;
  C->n_strang = 1;
  C->strang = 1.0;
  return a;
}

/* Chemistry_Newstate_blk (src/lib/physics/Lib_Chemistry.f90:33-76) over every block: the interior cells' species
 * densities and pressure gathered into [ns + 1][n], advanced over dt by the chemistry oracle, scattered back. */
static void oa_react(const o_case *C, double *P, double dt, void *ctx) {
  oa_chem *ch = (oa_chem *)ctx;
  const o_scheme *S = &C->sc;
  int64_t n = 0;
  for (int b = 0; b < C->nb; ++b) n += (int64_t)C->blk[b].n1 * C->blk[b].n2 * C->blk[b].n3;
  const int nz = S->nsc + 1;
  double *st = (double *)malloc(sizeof(double) * (size_t)nz * (size_t)n);
  int64_t *rows = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  int64_t q = 0;
  for (int b = 0; b < C->nb; ++b) {
    const o_block *B = &C->blk[b];
    for (int k = 1; k <= B->n3; ++k)
      for (int j = 1; j <= B->n2; ++j)
        for (int i = 1; i <= B->n1; ++i)
          rows[q++] = B->boff + (i + OGC - 1) + (int64_t)(j + OGC - 1) * B->pi + (int64_t)(k + OGC - 1) * B->pi * B->pj;
  }
  for (int64_t c = 0; c < n; ++c) {
    for (int s = 0; s < S->nsc; ++s) st[(int64_t)s * n + c] = P[(int64_t)s * C->NC + rows[c]];
    st[(int64_t)S->nsc * n + c] = P[(int64_t)S->np * C->NC + rows[c]];
  }
  ch->failed += oc_advance_m(ch->mech, ch->method, n, st, dt, ch->rtol, ch->atol, NULL, NULL, NULL, C->nthreads);
  for (int64_t c = 0; c < n; ++c) {
    for (int s = 0; s < S->nsc; ++s) P[(int64_t)s * C->NC + rows[c]] = st[(int64_t)s * n + c];
    P[(int64_t)S->np * C->NC + rows[c]] = st[(int64_t)S->nsc * n + c];
  }
  free(st);
  free(rows);
}

/* Chemistry on (mech != NULL): Strang-split steps, half transport, chemistry over dt, half transport
 * (set_StrangSplitting, Lib_Strang.f90:10-22); mech = NULL turns it off. */
void oa_set_chemistry(void *h, void *mech, int method, const double *rtol, const double *atol) {
  oa_t *a = (oa_t *)h;
  o_case *C = &a->c;
  if (!mech) {
    C->n_strang = 1; C->strang = 1.0; C->react = NULL; C->react_ctx = NULL;
    return;
  }
  a->chem.mech = mech;
  a->chem.method = method;
  for (int q = 0; q <= C->sc.nsc; ++q) { a->chem.rtol[q] = rtol[q]; a->chem.atol[q] = atol[q]; }
  a->chem.failed = 0;
  C->n_strang = 2; C->strang = 0.5; C->react = oa_react; C->react_ctx = &a->chem;
}

/* the chemistry oracle's failed cells (idid < 0) since oa_set_chemistry */
int64_t oa_chem_failed(void *h) { return ((oa_t *)h)->chem.failed; }

/* Chemistry_Newstate_blk alone on the interior of P (the chemistry set by oa_set_chemistry) */
void oa_react_only(void *h, double *P, double dt) {
  oa_t *a = (oa_t *)h;
  oa_react(&a->c, P, dt, &a->chem);
}

void oa_destroy(void *h) {
  oa_t *a = (oa_t *)h;
  if (!a) return;
  free(a->blk); free(a->rec); free(a->ff); free(a->tabs); free(a);
}

int64_t oa_ncell(void *h) { return ((oa_t *)h)->c.NC; }

void oa_fill_ghosts(void *h, double *P) { o_fill_ghosts(&((oa_t *)h)->c, P); }
/* dt: the time step the Spalart-Allmaras point-implicit factor reads (sa_body.inc:113-116); 0 otherwise */
void oa_residual(void *h, double *P, double *R, double dt) {
  oa_t *a = (oa_t *)h;
  o_fill_ghosts(&a->c, P);
  o_residual(&a->c, P, R, dt);
}
double oa_dt(void *h, const double *P) { return o_compute_dt(&((oa_t *)h)->c, P); }
/* the implicit residual smoothing of a residual array alone (IRS_apply) */
void oa_irs(void *h, double *R, double dt) { o_irs(&((oa_t *)h)->c, R, dt); }
void oa_newstate(void *h, double *P, const double *PO, const double *R, double dt, double rkc) {
  o_newstate(&((oa_t *)h)->c, P, PO, R, dt, rkc);
}
int oa_step(void *h, double *P, int nsteps, double *dts) {
  oa_t *a = (oa_t *)h;
  size_t n = (size_t)a->c.NC * a->c.sc.nprim;
  double *PO = (double *)malloc(n * sizeof(double)), *R = (double *)calloc(n, sizeof(double));
  int rc = o_step(&a->c, P, PO, R, nsteps, dts);
  free(PO); free(R);
  return rc;
}

/* point functions for unit tests */
double oa_limiter(int lim, double x, double y) { return o_rlimiter(lim, x, y); }
void oa_conv_face(void *h, const double *pm1, const double *p0, const double *pp1, const double *pp2,
                  const double *dlc, const double *nrm, double area, double *res0) {
  o_conv_face(&((oa_t *)h)->c, pm1, p0, pp1, pp2, dlc, nrm, area, res0);
}
void oa_riemann(void *h, int rs, const double *qL, const double *qR, const double *n, double *F) {
  /* qL/qR = species densities, u, v, w, p */
  oa_t *a = (oa_t *)h;
  int ns = a->c.sc.nsc;
  double rl = 0.0, rr = 0.0;
  for (int s = 0; s < ns; ++s) { rl += qL[s]; rr += qR[s]; }
  double al = o_sound_speed(&a->c, qL, qL[ns + 3]), ar = o_sound_speed(&a->c, qR, qR[ns + 3]);
  o_riemann(&a->c, rs, 1.0, qL, qL[ns], qL[ns + 1], qL[ns + 2], qL[ns + 3], al, rl, qR, qR[ns], qR[ns + 1], qR[ns + 2],
            qR[ns + 3], ar, rr, n[0], n[1], n[2], F);
}
double oa_sound_speed(void *h, const double *rhoi, double p) { return o_sound_speed(&((oa_t *)h)->c, rhoi, p); }
