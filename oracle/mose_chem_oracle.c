/*
 * CPU oracle of the stiff-chemistry path: FLINT's general law-of-mass-action source, the constant-volume ODE
 * right-hand side, and the host RODAS integrator that MOSE's general mechanisms run through.
 *
 * TEST INFRASTRUCTURE ONLY (see mose_oracle.h): tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, as the checker or the CPU baseline; the product never calls it.
 *
 *   oc_wdot        general()                 lib/FLINT/src/lib/Lib_Chemistry_wdot.f90:163-288
 *                  f_k_troe / f_k_lindemann  lib/FLINT/src/lib/Lib_Chemistry_falloff.f90:12-80
 *                  f_kf / f_kb               lib/FLINT/src/lib/Lib_Chemistry_data.f90:71-97
 *   oc_rhs         rhs_native                lib/FLINT/src/lib/Lib_Chemistry_rhs.f90:29-81
 *   fd_jacobian    ROSCOR, IJAC = 0, full    lib/FLINT/lib/OSlo/src/lib/hairerRodas.f:617-664
 *   oc_dec/oc_sol  DEC / SOL                 lib/FLINT/src/lib/Lib_Radau5_dev.f90:950-1035 (hairerToolbox.f)
 *   rodas          ROSCOR + ROCOE(METH=1), SLVROD/DECOMR IJOB = 1, autonomous, the wrap_rodas settings
 *                  (OSlo.f90:485-527: IFCN = 0, IJAC = 0, UROUND = 1.1e-19, H = 0 -> 1e-6)
 *                  lib/FLINT/lib/OSlo/src/lib/hairerRodas.f:540-880; coefficients as rodas_dev,
 *                  Lib_Radau5_dev.f90:710-737
 *   oc_advance     Chemistry_Newstate_blk    src/lib/physics/Lib_Chemistry.f90:33-76
 *
 * Deliberate ulp-level deviations, shared with the device so that both take the same step sequence:
 * x**0.25 of the step controller as sqrt(sqrt(x)); log10 / 10** of the Troe factor through dm_log / dm_exp; the
 * falloff concentration powers as integer powers.  Parity with the reference's own libm is unpinned at the ulp level.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mose_gpu.h"

#define CNS MOSE_MAX_SPECIES
#define CNZ (MOSE_MAX_SPECIES + 1)

typedef struct {
  int ns, na, nt, nl, tmin, tmax, rows;
  double wm[CNS], Ri[CNS];
  double *cp, *h, *ni1, *ni2, *eps, *kf, *kb, *troe, *lind;
} oc_mech;

static double *dup(const double *p, size_t n) {
  double *q = (double *)malloc((n ? n : 1) * sizeof(double));
  if (n && p) memcpy(q, p, n * sizeof(double));
  return q;
}

void *oc_create(const mose_mechanism *m) {
  if (!m || m->ns < 1 || m->ns > CNS) return NULL;
  oc_mech *c = (oc_mech *)calloc(1, sizeof *c);
  c->ns = m->ns; c->na = m->n_arrh; c->nt = m->n_troe; c->nl = m->n_lind;
  c->tmin = m->tmin; c->tmax = m->tmax; c->rows = m->tmax - m->tmin + 1;
  for (int s = 0; s < c->ns; ++s) { c->wm[s] = m->wm[s]; c->Ri[s] = 8314.46261815324 / m->wm[s]; }
  size_t R = (size_t)c->rows, nr = (size_t)(c->na + c->nt + c->nl), w = (size_t)c->ns + 1;
  c->cp = dup(m->cp, c->ns * R); c->h = dup(m->h, c->ns * R);
  c->ni1 = dup(m->ni1, nr * w); c->ni2 = dup(m->ni2, nr * w); c->eps = dup(m->eps, nr * w);
  c->kf = dup(m->kf, c->na * R); c->kb = dup(m->kb, c->na * R);
  c->troe = dup(m->troe, (size_t)c->nt * 4 * R); c->lind = dup(m->lind, (size_t)c->nl * 3 * R);
  return c;
}

void oc_destroy(void *h) {
  oc_mech *c = (oc_mech *)h;
  if (!c) return;
  free(c->cp); free(c->h); free(c->ni1); free(c->ni2); free(c->eps); free(c->kf); free(c->kb); free(c->troe);
  free(c->lind); free(c);
}

/* a + (b - a) * Tdiff on rows T1, T1 + 1 of a table (f_kf, comp_ch_tabT); rows clamped to the table */
static double tlin(const oc_mech *c, const double *tab, int T1, double Td) {
  int r1 = T1 - c->tmin, r2 = r1 + 1;
  if (r1 < 0) r1 = 0;
  if (r1 > c->rows - 1) r1 = c->rows - 1;
  if (r2 < 0) r2 = 0;
  if (r2 > c->rows - 1) r2 = c->rows - 1;
  double a = tab[r1], b = tab[r2];
  return a + (b - a) * Td;
}

/* Fortran x**n for the integral stoichiometric coefficients (nint): repeated squaring */
static double ipow(double x, int n) {
  double r = 1.0;
  int first = 1;
  while (n > 0) {
    if (n & 1) { r = first ? x : r * x; first = 0; }
    n >>= 1;
    if (n) x = x * x;
  }
  return first ? 1.0 : r;
}


/* Deterministic natural log and exp for the Troe broadening factor: the same statements on the CPU and the GPU give
 * the same bits (the libm log10 / pow of the two platforms differ in the last ulp, and the adaptive integrator turns
 * an ulp into a different step sequence).  log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 * s = (m-1)/(m+1), odd series to s^23; exp: y = k ln2 + r, |r| <= ln2/2, Taylor series to r^16, times 2^k.
 * Both are within a few ulp of the true values for the arguments met here (Fcent in (0, 1], Pr > 0). */
static inline double dm_log(double x) {
  int e;
  double m = frexp(x, &e);            /* x = m 2^e, m in [0.5, 1) */
  if (m < 0.70710678118654752440) { m = m * 2.0; e = e - 1; }
  double s = (m - 1.0) / (m + 1.0), s2 = s * s, t = 0.0;
  for (int k = 23; k >= 3; k -= 2) t = (t + 1.0 / (double)k) * s2;
  double lm = 2.0 * s * (1.0 + t);
  return (double)e * 0.69314718055994528623 + ((double)e * 2.3190468138462996154e-17 + lm);
}
static inline double dm_exp(double y) {
  double kf = floor(y * 1.44269504088896340736 + 0.5);
  double r = (y - kf * 0.69314718055994528623) - kf * 2.3190468138462996154e-17;
  double t = 1.0;
  for (int k = 16; k >= 1; --k) t = 1.0 + t * r / (double)k;
  return ldexp(t, (int)kf);
}

#define DM_LN10 2.30258509299404568402

/* f_F: the Troe broadening factor (Lib_Chemistry_falloff.f90:63-80); log10 and 10** through dm_log / dm_exp */
static double troe_F(double Pr, double Fcent) {
  double lF = dm_log(Fcent) / DM_LN10, lP = dm_log(Pr) / DM_LN10;
  double cc = -0.4 - 0.67 * lF, nn = 0.75 - 1.27 * lF;
  double f1 = (lP + cc) / (nn - 0.14 * (lP + cc));
  return dm_exp(lF / (1.0 + f1 * f1) * DM_LN10);
}

/* c**nu of the falloff classes: the integer power for integral nu (the run-time real power of the reference equals
 * it up to an ulp), pow otherwise */
static double rpow(double c, double nu) {
  double r = nearbyint(nu);
  return (r == nu && r >= 0.0 && r <= 8.0) ? ipow(c, (int)r) : pow(c, nu);
}

/* general(): omegadot(s) = sum_r Wm_s (ni2 - ni1) (rate_fwd - rate_rev) */
static void oc_wdot(const oc_mech *c, const double *roi, double temp, double *om) {
  int ns = c->ns, w = ns + 1;
  double coi[CNS];
  for (int s = 0; s < ns; ++s) coi[s] = roi[s] / c->wm[s];
  int Ti = (int)temp;
  double Td = temp - Ti;
  for (int s = 0; s < ns; ++s) om[s] = 0.0;
  for (int cls = 0; cls < 3; ++cls) {
    int n = cls == 0 ? c->na : (cls == 1 ? c->nt : c->nl);
    int base = cls == 0 ? 0 : (cls == 1 ? c->na : c->na + c->nt);
    for (int ir = 0; ir < n; ++ir) {
      const double *n1 = c->ni1 + (size_t)(base + ir) * w, *n2 = c->ni2 + (size_t)(base + ir) * w;
      const double *ep = c->eps + (size_t)(base + ir) * w;
      double coM;
      if (n1[ns] > 0.0) {
        coM = 0.0;
        for (int s = 0; s < ns; ++s) coM = coM + coi[s] * ep[s];
      } else {
        coM = 1.0;
      }
      double pf = 1.0, pr = 1.0;
      double rf, rr;
      if (cls == 0) {
        for (int s = 0; s < ns; ++s) {
          if (n1[s] != 0) pf = pf * ipow(coi[s], (int)lround(n1[s]));
          if (n2[s] != 0) pr = pr * ipow(coi[s], (int)lround(n2[s]));
        }
        rf = tlin(c, c->kf + (size_t)ir * c->rows, Ti, Td) * pf * coM;
        rr = tlin(c, c->kb + (size_t)ir * c->rows, Ti, Td) * pr * coM;
      } else {
        for (int s = 0; s < ns; ++s) {
          if (n1[s] != 0) pf = pf * rpow(coi[s], n1[s]);
          if (n2[s] != 0) pr = pr * rpow(coi[s], n2[s]);
        }
        const double *tb = cls == 1 ? c->troe + (size_t)ir * 4 * c->rows : c->lind + (size_t)ir * 3 * c->rows;
        double kinf = tlin(c, tb, Ti, Td), k0 = tlin(c, tb + c->rows, Ti, Td), kc = tlin(c, tb + 2 * c->rows, Ti, Td);
        double Pr = k0 * coM / kinf;
        double k1 = kinf * (Pr / (1.0 + Pr));
        if (cls == 1) k1 = k1 * troe_F(Pr, tlin(c, tb + 3 * c->rows, Ti, Td));
        double k2 = k1 / kc;
        rf = k1 * pf;
        rr = k2 * pr;
      }
      double net = rf - rr;
      for (int s = 0; s < ns; ++s) {
        double dn = n2[s] - n1[s];
        om[s] = om[s] + c->wm[s] * dn * net;
      }
    }
  }
}

/* rhs_native: Z = (rho_1..rho_ns, T); F_T = -sum (h_s - R_s T) w_s / sum rho_s (cp_s - R_s) */
static void oc_rhs_(const oc_mech *c, const double *Z, double *F) {
  int ns = c->ns, nz = ns + 1;
  double T = Z[ns];
  if (T < (double)c->tmin || T >= (double)c->tmax || T != T) {
    for (int q = 0; q < nz; ++q) F[q] = -1.0;
    return;
  }
  int Ti = (int)T;
  double Tf = T - (double)Ti;
  double roi[CNS], dr[CNS];
  for (int s = 0; s < ns; ++s) roi[s] = Z[s] > 0.0 ? Z[s] : 0.0;
  oc_wdot(c, roi, T, dr);
  double ei = 0.0, rcv = 0.0;
  for (int s = 0; s < ns; ++s) {
    const double *ht = c->h + (size_t)s * c->rows, *ct = c->cp + (size_t)s * c->rows;
    int r = Ti - c->tmin;
    double hv = ht[r] + Tf * (ht[r + 1] - ht[r]);
    double cv = ct[r] + Tf * (ct[r + 1] - ct[r]);
    ei = ei + (hv - c->Ri[s] * T) * dr[s];
    rcv = rcv + roi[s] * (cv - c->Ri[s]);
  }
  for (int s = 0; s < ns; ++s) F[s] = dr[s];
  F[ns] = -ei / rcv;
}

void oc_rhs(void *h, const double *Z, double *F) { oc_rhs_((const oc_mech *)h, Z, F); }
void oc_wdot_pub(void *h, const double *roi, double T, double *om) { oc_wdot((const oc_mech *)h, roi, T, om); }

/* ---------------------------------------------------------------------------------------------------------- */
/* DEC / SOL (Moler, ACM 423): column-major a[i + n*j]                                                          */
/* ---------------------------------------------------------------------------------------------------------- */
int oc_dec(int n, double *a, int *ip) {
#define A(i, j) a[(i) + (size_t)n * (j)]
  int ier = 0, k;
  ip[n - 1] = 1;
  if (n == 1) { k = n - 1; goto L70; }
  for (k = 0; k < n - 1; ++k) {
    int m = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(A(i, k)) > fabs(A(m, k))) m = i;
    ip[k] = m + 1;
    double t = A(m, k);
    if (m != k) {
      ip[n - 1] = -ip[n - 1];
      A(m, k) = A(k, k);
      A(k, k) = t;
    }
    if (t == 0.0) goto L80;
    t = 1.0 / t;
    for (int i = k + 1; i < n; ++i) A(i, k) = -A(i, k) * t;
    for (int j = k + 1; j < n; ++j) {
      t = A(m, j);
      A(m, j) = A(k, j);
      A(k, j) = t;
      if (t != 0.0)
        for (int i = k + 1; i < n; ++i) A(i, j) = A(i, j) + A(i, k) * t;
    }
  }
L70:
  k = n - 1;
  if (A(n - 1, n - 1) == 0.0) goto L80;
  return 0;
L80:
  ier = k + 1;
  ip[n - 1] = 0;
  return ier;
#undef A
}

void oc_sol(int n, const double *a, double *b, const int *ip) {
#define A(i, j) a[(i) + (size_t)n * (j)]
  if (n > 1) {
    for (int k = 0; k < n - 1; ++k) {
      int m = ip[k] - 1;
      double t = b[m];
      b[m] = b[k];
      b[k] = t;
      for (int i = k + 1; i < n; ++i) b[i] = b[i] + A(i, k) * t;
    }
    for (int kb = 1; kb <= n - 1; ++kb) {
      int km1 = n - kb;   /* 1-based km1 = n - kb, k = km1 + 1 */
      int k = km1;        /* 0-based k */
      b[k] = b[k] / A(k, k);
      double t = -b[k];
      for (int i = 0; i < km1; ++i) b[i] = b[i] + A(i, k) * t;
    }
  }
  b[0] = b[0] / A(0, 0);
#undef A
}

/* ---------------------------------------------------------------------------------------------------------- */
/* RODAS (METH = 1), autonomous, numerical full Jacobian, B = identity                                          */
/* ---------------------------------------------------------------------------------------------------------- */
typedef void (*oc_fcn)(const void *ctx, const double *y, double *f);

static const double RC2 = 0.386, RC3 = 0.21, RC4 = 0.63, RGAMMA = 0.25;
static const double A21 = 0.1544000000000000e+01, A31 = 0.9466785280815826e+00, A32 = 0.2557011698983284e+00;
static const double A41 = 0.3314825187068521e+01, A42 = 0.2896124015972201e+01, A43 = 0.9986419139977817e+00;
static const double A51 = 0.1221224509226641e+01, A52 = 0.6019134481288629e+01, A53 = 0.1253708332932087e+02;
static const double A54 = -0.6878860361058950e+00;
static const double C21 = -0.5668800000000000e+01, C31 = -0.2430093356833875e+01, C32 = -0.2063599157091915e+00;
static const double C41 = -0.1073529058151375e+00, C42 = -0.9594562251023355e+01, C43 = -0.2047028614809616e+02;
static const double C51 = 0.7496443313967647e+01, C52 = -0.1024680431464352e+02, C53 = -0.3399990352819905e+02;
static const double C54 = 0.1170890893206160e+02;
static const double C61 = 0.8083246795921522e+01, C62 = -0.7981132988064893e+01, C63 = -0.3152159432874371e+02;
static const double C64 = 0.1631930543123136e+02, C65 = -0.6058818238834054e+01;

typedef struct { int nstep, naccpt, nrejct, nfcn, njac; } oc_stats;

static int rodas(int n, oc_fcn fcn, const void *ctx, double x, double *y, double xend, const double *rtol,
                 const double *atol, oc_stats *st) {
  const double uround = 1.1e-19, safe = 0.9, fac1 = 5.0, fac2 = 1.0 / 6.0;
  const int nmax = 100000;
  double dy1[CNZ], dy[CNZ], ynew[CNZ], ak1[CNZ], ak2[CNZ], ak3[CNZ], ak4[CNZ], ak5[CNZ], ak6[CNZ], tmp[CNZ];
  double fjac[CNZ * CNZ], e[CNZ * CNZ];
  int ip[CNZ];
  memset(st, 0, sizeof *st);
  double hmax = xend - x;
  double posneg = copysign(1.0, xend - x);
  double hmaxn = fmin(fabs(hmax), fabs(xend - x));
  double h = 0.0;
  if (fabs(h) <= 10.0 * uround) h = 1.0e-6;
  h = fmin(fabs(h), hmaxn);
  h = copysign(h, posneg);
  int reject = 0, last = 0, nsing = 0;
  double hopt = h, hacc = 0.0, erracc = 1e-2;
L1:
  if (st->nstep > nmax) return -2;
  if (0.1 * fabs(h) <= fabs(x) * uround) return -3;
  if (last) { h = hopt; return 1; }
  hopt = h;
  if ((x + h * 1.0001 - xend) * posneg >= 0.0) { h = xend - x; last = 1; }
  fcn(ctx, y, dy1);
  st->nfcn++;
  st->njac++;
  for (int i = 0; i < n; ++i) {
    double ysafe = y[i];
    double delt = sqrt(uround * fmax(1.0e-5, fabs(ysafe)));
    y[i] = ysafe + delt;
    fcn(ctx, y, ak1);
    for (int j = 0; j < n; ++j) fjac[j + n * i] = (ak1[j] - dy1[j]) / delt;
    y[i] = ysafe;
  }
L2:;
  double fac = 1.0 / (h * RGAMMA);
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) e[i + n * j] = -fjac[i + n * j];
    e[j + n * j] = e[j + n * j] + fac;
  }
  if (oc_dec(n, e, ip) != 0) goto L80;
  double hc21 = C21 / h, hc31 = C31 / h, hc32 = C32 / h, hc41 = C41 / h, hc42 = C42 / h, hc43 = C43 / h;
  double hc51 = C51 / h, hc52 = C52 / h, hc53 = C53 / h, hc54 = C54 / h;
  double hc61 = C61 / h, hc62 = C62 / h, hc63 = C63 / h, hc64 = C64 / h, hc65 = C65 / h;
  (void)RC2; (void)RC3; (void)RC4;
  for (int i = 0; i < n; ++i) ak1[i] = dy1[i];
  oc_sol(n, e, ak1, ip);
  for (int i = 0; i < n; ++i) ynew[i] = y[i] + A21 * ak1[i];
  fcn(ctx, ynew, dy);
  for (int i = 0; i < n; ++i) ynew[i] = hc21 * ak1[i];
  for (int i = 0; i < n; ++i) ak2[i] = dy[i] + ynew[i];
  oc_sol(n, e, ak2, ip);
  for (int i = 0; i < n; ++i) ynew[i] = y[i] + A31 * ak1[i] + A32 * ak2[i];
  fcn(ctx, ynew, dy);
  for (int i = 0; i < n; ++i) ynew[i] = hc31 * ak1[i] + hc32 * ak2[i];
  for (int i = 0; i < n; ++i) ak3[i] = dy[i] + ynew[i];
  oc_sol(n, e, ak3, ip);
  for (int i = 0; i < n; ++i) ynew[i] = y[i] + A41 * ak1[i] + A42 * ak2[i] + A43 * ak3[i];
  fcn(ctx, ynew, dy);
  for (int i = 0; i < n; ++i) ynew[i] = hc41 * ak1[i] + hc42 * ak2[i] + hc43 * ak3[i];
  for (int i = 0; i < n; ++i) ak4[i] = dy[i] + ynew[i];
  oc_sol(n, e, ak4, ip);
  for (int i = 0; i < n; ++i) ynew[i] = y[i] + A51 * ak1[i] + A52 * ak2[i] + A53 * ak3[i] + A54 * ak4[i];
  fcn(ctx, ynew, dy);
  for (int i = 0; i < n; ++i) ak6[i] = hc52 * ak2[i] + hc54 * ak4[i] + hc51 * ak1[i] + hc53 * ak3[i];
  for (int i = 0; i < n; ++i) ak5[i] = dy[i] + ak6[i];
  oc_sol(n, e, ak5, ip);
  for (int i = 0; i < n; ++i) ynew[i] = ynew[i] + ak5[i];
  fcn(ctx, ynew, dy);
  for (int i = 0; i < n; ++i) tmp[i] = hc61 * ak1[i] + hc62 * ak2[i] + hc65 * ak5[i] + hc64 * ak4[i] + hc63 * ak3[i];
  for (int i = 0; i < n; ++i) ak6[i] = dy[i] + tmp[i];
  oc_sol(n, e, ak6, ip);
  for (int i = 0; i < n; ++i) ynew[i] = ynew[i] + ak6[i];
  st->nfcn += 5;
  st->nstep++;
  double err = 0.0;
  for (int i = 0; i < n; ++i) {
    double sk = atol[i] + rtol[i] * fmax(fabs(y[i]), fabs(ynew[i]));
    double q = ak6[i] / sk;
    err = err + q * q;
  }
  err = sqrt(err / n);
  fac = fmax(fac2, fmin(fac1, sqrt(sqrt(err)) / safe));   /* err**0.25 */
  double hnew = h / fac;
  if (err <= 1.0) {
    st->naccpt++;
    if (st->naccpt > 1) {
      double facgus = (hacc / h) * sqrt(sqrt(err * err / erracc)) / safe;
      facgus = fmax(fac2, fmin(fac1, facgus));
      fac = fmax(fac, facgus);
      hnew = h / fac;
    }
    hacc = h;
    erracc = fmax(1.0e-2, err);
    for (int i = 0; i < n; ++i) y[i] = ynew[i];
    x = x + h;
    if (fabs(hnew) > hmaxn) hnew = posneg * hmaxn;
    if (reject) hnew = posneg * fmin(fabs(hnew), fabs(h));
    reject = 0;
    h = hnew;
    goto L1;
  } else {
    reject = 1;
    last = 0;
    h = hnew;
    if (st->naccpt >= 1) st->nrejct++;
    goto L2;
  }
L80:
  nsing++;
  if (nsing >= 5) return -4;
  h = h * 0.5;
  reject = 1;
  last = 0;
  goto L2;
}

static void chem_fcn(const void *ctx, const double *y, double *f) { oc_rhs_((const oc_mech *)ctx, y, f); }

/* Robertson's problem (a check of the integrator; the OSlo test problem robertson.f90) */
static void rober_fcn(const void *ctx, const double *y, double *f) {
  (void)ctx;
  f[0] = -0.04 * y[0] + 1.0e4 * y[1] * y[2];
  f[2] = 3.0e7 * y[1] * y[1];
  f[1] = -f[0] - f[2];
}

int oc_rodas_robertson(double *y, double xend, double rtol, double atol, int *stats) {
  double rt[3] = {rtol, rtol, rtol}, at[3] = {atol, atol, atol};
  oc_stats st;
  int idid = rodas(3, rober_fcn, NULL, 0.0, y, xend, rt, at, &st);
  if (stats) { stats[0] = st.nstep; stats[1] = st.naccpt; stats[2] = st.nrejct; stats[3] = st.nfcn; }
  return idid;
}

/* one cell of Chemistry_Newstate_blk on (species densities, p); returns idid */
static int advance_cell(const oc_mech *c, double *q, double dt, const double *rtol, const double *atol, double *hr,
                        oc_stats *st) {
  int ns = c->ns;
  double Z[CNZ], rho = 0.0, R = 0.0;
  for (int s = 0; s < ns; ++s) { Z[s] = q[s]; rho = rho + Z[s]; }
  for (int s = 0; s < ns; ++s) R = R + c->Ri[s] * Z[s] / rho;
  Z[ns] = q[ns] / (rho * R);
  double Tb = Z[ns];
  int idid = rodas(ns + 1, chem_fcn, c, 0.0, Z, dt, rtol, atol, st);
  double rs = 0.0, R2 = 0.0;
  for (int s = 0; s < ns; ++s) {
    if (Z[s] <= 1e-20) Z[s] = 1e-20;
    q[s] = Z[s];
    rs = rs + Z[s];
  }
  for (int s = 0; s < ns; ++s) R2 = R2 + c->Ri[s] * Z[s] / rs;
  q[ns] = Z[ns] * rho * R2;
  if (hr) *hr = fabs(Z[ns] - Tb) / Tb;
  return idid;
}

/* state: [ns + 1][ncell] var-last (species..., p), updated in place; idid/nstep per cell (may be NULL) */
int oc_advance(void *h, int64_t ncell, double *state, double dt, const double *rtol, const double *atol, double *hr,
               int32_t *idid, int32_t *nstep, int nthreads) {
  const oc_mech *c = (const oc_mech *)h;
  int ns = c->ns;
  int nbad = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : nbad)
  for (int64_t i = 0; i < ncell; ++i) {
    double q[CNZ], hri = 0.0;
    for (int s = 0; s <= ns; ++s) q[s] = state[(int64_t)s * ncell + i];
    oc_stats st;
    int id = advance_cell(c, q, dt, rtol, atol, &hri, &st);
    for (int s = 0; s <= ns; ++s) state[(int64_t)s * ncell + i] = q[s];
    if (hr) hr[i] = hri;
    if (idid) idid[i] = id;
    if (nstep) nstep[i] = st.nstep;
    if (id < 0) nbad++;
  }
  return nbad;
}
