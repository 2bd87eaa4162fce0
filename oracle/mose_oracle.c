/*
 * CPU oracle: the time-step residual hot path of open-hydra/MOSE-GPU restated in C.
 *
 * TEST INFRASTRUCTURE ONLY (see mose_oracle.h).  Built with -O2 -ffp-contract=off so that every
 * product and sum is rounded exactly where the reference's Fortran statement rounds it; each
 * expression keeps the reference's left-to-right association.  Fortran `sign(a,b)` is copysign,
 * `x**2` is x*x, MIN/MAX of several arguments are evaluated left to right.
 */
#include "mose_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GC OGC
#define ONPRIM 24

/* ------------------------------------------------------------------------------------------------ */
/* indexing                                                                                          */
/* ------------------------------------------------------------------------------------------------ */
static inline int64_t ROW(const o_block *B, int i, int j, int k) {
  return B->boff + (int64_t)(i + GC - 1) + (int64_t)(j + GC - 1) * B->pi + (int64_t)(k + GC - 1) * B->pi * B->pj;
}
/* face rows of fm1(4,0:n1,n2,n3), fm2(4,n1,0:n2,n3), fm3(4,n1,n2,0:n3) */
static inline int64_t F1(const o_block *B, int i, int j, int k) {
  return B->f1off + i + (int64_t)(B->n1 + 1) * ((j - 1) + (int64_t)B->n2 * (k - 1));
}
static inline int64_t F2(const o_block *B, int i, int j, int k) {
  return B->f2off + (i - 1) + (int64_t)B->n1 * (j + (int64_t)(B->n2 + 1) * (k - 1));
}
static inline int64_t F3(const o_block *B, int i, int j, int k) {
  return B->f3off + (i - 1) + (int64_t)B->n1 * ((j - 1) + (int64_t)B->n2 * k);
}
#define PV(P, s, r) (P)[(int64_t)(s) * NC + (r)]

static inline int tw(const o_tables *t) { return t->tmax + 1; }
/* rows clamped to the table: only a state off the table (a failing state) reaches the clamp */
static inline double TAB(const double *tab, const o_tables *t, int s, int T) {
  T = T < 0 ? 0 : (T > t->tmax ? t->tmax : T);
  return tab[(int64_t)s * tw(t) + T];
}

static inline double min2(double a, double b) { return (b < a) ? b : a; }
static inline double max2(double a, double b) { return (b > a) ? b : a; }
static inline double min3(double a, double b, double c) { return min2(min2(a, b), c); }
static inline double max3(double a, double b, double c) { return max2(max2(a, b), c); }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* Fortran int() of a double; a NaN (only in a failing state) maps to 0 so that no table index runs wild */
static inline int fint(double x) { return (x == x) ? (int)x : 0; }

/* ------------------------------------------------------------------------------------------------ */
/* FLINT thermodynamics                                                                              */
/* ------------------------------------------------------------------------------------------------ */
/* f_Rtot: lib/FLINT/src/lib/Lib_ThermoTransport.f90:238-252 */
static double f_Rtot(const o_tables *t, const double *rhoi) {
  double rho = 0.0, r = 0.0;
  for (int s = 0; s < t->ns; ++s) rho = rho + rhoi[s];
  for (int s = 0; s < t->ns; ++s) r = r + t->Ri[s] * rhoi[s] / rho;
  return r;
}
/* co_rotot_Rtot_dev: Lib_ThermoTransport_dev.f90:105-113 */
static void co_rotot_Rtot(const o_tables *t, const double *rhoi, double *rho, double *R) {
  double r = 0.0;
  for (int s = 0; s < t->ns; ++s) r = r + rhoi[s];
  *rho = r;
  *R = f_Rtot(t, rhoi);
}
/* f_molecularWeight: Lib_ThermoTransport.f90:220-235 */
static double f_molw(const o_tables *t, const double *rhoi) {
  double rho = 0.0, r = 0.0;
  for (int s = 0; s < t->ns; ++s) rho = rho + rhoi[s];
  for (int s = 0; s < t->ns; ++s) r = r + (rhoi[s] / rho / t->wm[s]);
  return 1.0 / r;
}
/* f_tabT_expr: Lib_ThermoTransport.f90:623-635 */
static inline double tabT(const double *tab, const o_tables *t, int s, int T1, int T2, double Td) {
  if (T1 < 0 || T1 > t->tmax - 1) { T1 = T1 < 0 ? 0 : t->tmax - 1; T2 = T1 + 1; }
  double v1 = TAB(tab, t, s, T1), v2 = TAB(tab, t, s, T2);
  return v1 + (v2 - v1) * Td;
}
/* f_cp_dev: Lib_ThermoTransport_dev.f90:280-293 */
static double f_cp_dev(const o_tables *t, const double *rhoi, int T1, double Td, double rho) {
  double res = 0.0;
  for (int s = 0; s < t->ns; ++s) res = res + rhoi[s] / rho * tabT(t->cp, t, s, T1, T1 + 1, Td);
  return res;
}
/* f_ss_dev: Lib_ThermoTransport_dev.f90:70-82 (no clamp of the table index) */
static double f_ss_dev(const o_tables *t, const double *rhoi, double p, double rho, double Rtot) {
  double T = p / (Rtot * rho);
  int Ti = fint(T);
  double Td = T - Ti;
  double cp = f_cp_dev(t, rhoi, Ti, Td, rho);
  double gam = cp / (cp - Rtot);
  return sqrt(gam * Rtot * T);
}
/* H0_dev: Lib_ThermoTransport_dev.f90:54-68 */
static double H0_dev(const o_tables *t, double p, const double *rhoi, double u) {
  double rho = 0.0;
  for (int s = 0; s < t->ns; ++s) rho = rho + rhoi[s];
  double Rg = f_Rtot(t, rhoi);
  double T = p / (Rg * rho);
  int Ti = fint(T);
  double Td = T - Ti;
  double e = 0.0;
  for (int s = 0; s < t->ns; ++s) {
    double h_ = tabT(t->h, t, s, Ti, Ti + 1, Td);
    e = e + h_ * rhoi[s] / rho;
  }
  return e + 0.5 * u * u;
}
/* co_k_mi_lam_Wilke_dev + co_fiij_dev: Lib_ThermoTransport_dev.f90:120-175 */
static void wilke(const o_tables *t, const double *rhoi, double rho, int T1, double Td, double *mil, double *kl) {
  int ns = t->ns;
  double Xi[ONSMAX], mi[ONSMAX], ki[ONSMAX], sqn[ONSMAX], isd[ONSMAX], lam_den[ONSMAX];
  double Wmtot = f_molw(t, rhoi);
  for (int s = 0; s < ns; ++s) {
    Xi[s] = rhoi[s] * Wmtot / (rho * t->wm[s]);
    mi[s] = tabT(t->mi, t, s, T1, T1 + 1, Td);
    ki[s] = tabT(t->k, t, s, T1, T1 + 1, Td);
  }
  for (int i = 0; i < ns; ++i) {
    sqn[i] = sqrt(mi[i]);
    isd[i] = 1.0 / sqrt(mi[i] + 1e-20);
  }
  for (int i = 0; i < ns; ++i) {
    lam_den[i] = 0.0;
    for (int j = 0; j < ns; ++j) {
      double ratio = sqn[i] * isd[j] * t->mm025[i][j];
      double op = 1.0 + ratio;
      double fi = op * op * t->is8[i][j];
      lam_den[i] = lam_den[i] + Xi[j] * fi;
    }
  }
  double m = 0.0, k = 0.0;
  for (int s = 0; s < ns; ++s) {
    double inv = 1.0 / lam_den[s];
    m = m + Xi[s] * mi[s] * inv;
    k = k + Xi[s] * ki[s] * inv;
  }
  *mil = m;
  *kl = k;
}

double o_sound_speed(const o_case *C, const double *rhoi, double p) {
  double rho, R;
  co_rotot_Rtot(&C->tab, rhoi, &rho, &R);
  return f_ss_dev(&C->tab, rhoi, p, rho, R);
}

/* ------------------------------------------------------------------------------------------------ */
/* limiters: src/lib/numerics/space/Lib_Limiters.f90:65-97 (rlimiter; rlimiter_d in                 */
/* Mod_Convective.f90:161-191 is the same statements)                                                */
/* ------------------------------------------------------------------------------------------------ */
double o_rlimiter(int lim, double x, double y) {
  double ru, phi;
  switch (lim) {
    case 1: {
      double r = 0.0;
      if (x * y < 0.0) return r;
      r = x;
      if (fabs(x) > fabs(y)) r = y;
      return r;
    }
    case 2:
      ru = y / (x + 1e-30);
      phi = (ru + fabs(ru)) / (1.0 + fabs(ru));
      return phi * x;
    case 3:
      ru = y / (x + 1e-30);
      phi = (ru * ru + ru) / (1.0 + ru * ru);
      return phi * x;
    case 4:
      ru = y / (x + 1e-30);
      phi = max2(0.0, min3(2.0 * ru, 0.5 * (1.0 + ru), 2.0));
      return phi * x;
    case 5:
      ru = y / (x + 1e-30);
      phi = max3(0.0, min2(2.0 * ru, 1.0), min2(ru, 2.0));
      return phi * x;
    default:
      return 0.0;
  }
}

/* ------------------------------------------------------------------------------------------------ */
/* Roe averages with the table index clamped to [lo, hi]                                             */
/* conv_physics.inc:70-98 (clamp [0,NTMAX-1]), riemann_open_body.inc:679-726 (clamp [Tmin,Tmax-1]),   */
/* Lib_Riemann_HLL.f90:210-310 (no clamp: pass lo = INT_MIN)                                           */
/* ------------------------------------------------------------------------------------------------ */
typedef struct { double h0l, h0r, rho_roe, u_roe, v_roe, w_roe, a_roe, h0_roe; } roe_t;

static void roe_averages(const o_tables *t, int clamp, int lo, int hi, const double *dl, const double *dr,
                         double dltot, double drtot, double ul, double vl, double wl, double ur, double vr, double wr,
                         double pl, double pr, double Rl, double Rr, roe_t *o) {
  double inv_dltot = 1.0 / dltot, inv_drtot = 1.0 / drtot;
  double Tl = pl * inv_dltot / Rl, Tr = pr * inv_drtot / Rr;
  int Til = fint(Tl), Tir = fint(Tr);
  if (clamp) { Til = imax(lo, imin(Til, hi)); Tir = imax(lo, imin(Tir, hi)); }
  double dTl = Tl - Til, dTr = Tr - Tir;
  double srL = sqrt(dltot), srR = sqrt(drtot), inv_sr = 1.0 / (srL + srR);
  o->rho_roe = srL * srR;
  o->u_roe = (srR * ur + srL * ul) * inv_sr;
  o->v_roe = (srR * vr + srL * vl) * inv_sr;
  o->w_roe = (srR * wr + srL * wl) * inv_sr;
  double R_roe = (srR * Rr + srL * Rl) * inv_sr;
  double T_roe = (srR * Tr + srL * Tl) * inv_sr;
  double h0l = 0.0, h0r = 0.0, cv_roe = 0.0, sum_ei = 0.0;
  for (int s = 0; s < t->ns; ++s) {
    double invW = 8314.46261815324 / t->wm[s];
    double hl = TAB(t->h, t, s, Til) + (TAB(t->h, t, s, Til + 1) - TAB(t->h, t, s, Til)) * dTl;
    double hr = TAB(t->h, t, s, Tir) + (TAB(t->h, t, s, Tir + 1) - TAB(t->h, t, s, Tir)) * dTr;
    double el = hl - invW * Tl, er = hr - invW * Tr;
    h0l = h0l + hl * dl[s] * inv_dltot;
    h0r = h0r + hr * dr[s] * inv_drtot;
    double d_roe = (srR * dr[s] * inv_drtot + srL * dl[s] * inv_dltot) * inv_sr;
    double cpl = TAB(t->cp, t, s, Til) + (TAB(t->cp, t, s, Til + 1) - TAB(t->cp, t, s, Til)) * dTl;
    double cprA = TAB(t->cp, t, s, Tir), cprB = (TAB(t->cp, t, s, Tir + 1) - TAB(t->cp, t, s, Tir)) * dTr;
    cv_roe = cv_roe + d_roe * (0.5 * (cpl + cprA + cprB) - invW);
    sum_ei = sum_ei + d_roe * (srR * er + srL * el) * inv_sr;
  }
  h0l = h0l + 0.5 * (ul * ul + vl * vl + wl * wl);
  h0r = h0r + 0.5 * (ur * ur + vr * vr + wr * wr);
  o->h0l = h0l;
  o->h0r = h0r;
  o->h0_roe = (srR * h0r + srL * h0l) * inv_sr;
  double gam_roe = 1.0 + R_roe / cv_roe;
  double vel2 = o->u_roe * o->u_roe + o->v_roe * o->v_roe + o->w_roe * o->w_roe;
  o->a_roe = sqrt((gam_roe - 1.0) * (o->h0_roe - 0.5 * vel2 + cv_roe * T_roe - sum_ei));
}

/* HLLC (Batten): conv_physics.inc:87-115 == Lib_Riemann_HLL.f90:21-134 (same statements) */
/* HLLE (rs 2): conv_physics.inc:116-124 / riemann_HLLE, Lib_Riemann_HLL.f90:137-198 (the signal speeds of HLLC) */
static void hlle(double dltot, double drtot, double ul, double vl, double wl, double pl, double al, double ur,
                 double vr, double wr, double pr, double ar, double nx, double ny, double nz, const roe_t *ro,
                 double F[5]) {
  double unl = ul * nx + vl * ny + wl * nz;
  double unr = ur * nx + vr * ny + wr * nz;
  double un_roe = ro->u_roe * nx + ro->v_roe * ny + ro->w_roe * nz;
  double h0l = ro->h0l, h0r = ro->h0r;
  double S1 = min3(0.0, unl - al, un_roe - ro->a_roe);
  double S4 = max3(0.0, unr + ar, un_roe + ro->a_roe);
  double inv_denom = 1.0 / (S4 - S1);
  double FrL = dltot * unl, FrR = drtot * unr;
  F[0] = (S4 * FrL - S1 * FrR + S1 * S4 * (drtot - dltot)) * inv_denom;
  F[1] = (S4 * (FrL * ul + pl * nx) - S1 * (FrR * ur + pr * nx) + S1 * S4 * (drtot * ur - dltot * ul)) * inv_denom;
  F[2] = (S4 * (FrL * vl + pl * ny) - S1 * (FrR * vr + pr * ny) + S1 * S4 * (drtot * vr - dltot * vl)) * inv_denom;
  F[3] = (S4 * (FrL * wl + pl * nz) - S1 * (FrR * wr + pr * nz) + S1 * S4 * (drtot * wr - dltot * wl)) * inv_denom;
  F[4] = (S4 * (FrL * h0l) - S1 * (FrR * h0r) + S1 * S4 * (drtot * (h0r - pr / drtot) - dltot * (h0l - pl / dltot))) *
         inv_denom;
}

static void hllc(double dltot, double drtot, double ul, double vl, double wl, double pl, double al, double ur,
                 double vr, double wr, double pr, double ar, double nx, double ny, double nz, const roe_t *ro,
                 double F[5]) {
  double unl = ul * nx + vl * ny + wl * nz;
  double unr = ur * nx + vr * ny + wr * nz;
  double un_roe = ro->u_roe * nx + ro->v_roe * ny + ro->w_roe * nz;
  double h0l = ro->h0l, h0r = ro->h0r;
  double e0l = h0l - pl / dltot, e0r = h0r - pr / drtot;
  double S1 = min3(0.0, unl - al, un_roe - ro->a_roe);
  double S4 = max3(0.0, unr + ar, un_roe + ro->a_roe);
  double inv_denom = 1.0 / (dltot * (S1 - unl) - drtot * (S4 - unr));
  double Sstar = (pr - pl + dltot * unl * (S1 - unl) - drtot * unr * (S4 - unr)) * inv_denom;
  double pstar = dltot * (unl - S1) * (unl - Sstar) + pl;
  double Fr, Fu, Fv, Fw, FE;
  if (S1 >= 0.0) {
    Fr = dltot * unl; Fu = Fr * ul + pl * nx; Fv = Fr * vl + pl * ny; Fw = Fr * wl + pl * nz; FE = Fr * h0l;
  } else if (Sstar >= 0.0) {
    Fr = dltot * unl; Fu = Fr * ul + pl * nx; Fv = Fr * vl + pl * ny; Fw = Fr * wl + pl * nz; FE = Fr * h0l;
    inv_denom = 1.0 / (S1 - Sstar);
    double U1S = dltot * (S1 - unl) * inv_denom;
    double U2S = ((S1 - unl) * dltot * ul + (pstar - pl) * nx) * inv_denom;
    double U3S = ((S1 - unl) * dltot * vl + (pstar - pl) * ny) * inv_denom;
    double U4S = ((S1 - unl) * dltot * wl + (pstar - pl) * nz) * inv_denom;
    double U5S = ((S1 - unl) * dltot * e0l - pl * unl + pstar * Sstar) * inv_denom;
    Fr = Fr + S1 * (U1S - dltot); Fu = Fu + S1 * (U2S - dltot * ul); Fv = Fv + S1 * (U3S - dltot * vl);
    Fw = Fw + S1 * (U4S - dltot * wl); FE = FE + S1 * (U5S - dltot * e0l);
  } else if (S4 > 0.0) {
    Fr = drtot * unr; Fu = Fr * ur + pr * nx; Fv = Fr * vr + pr * ny; Fw = Fr * wr + pr * nz; FE = Fr * h0r;
    inv_denom = 1.0 / (S4 - Sstar);
    double U1S = drtot * (S4 - unr) * inv_denom;
    double U2S = ((S4 - unr) * drtot * ur + (pstar - pr) * nx) * inv_denom;
    double U3S = ((S4 - unr) * drtot * vr + (pstar - pr) * ny) * inv_denom;
    double U4S = ((S4 - unr) * drtot * wr + (pstar - pr) * nz) * inv_denom;
    double U5S = ((S4 - unr) * drtot * e0r - pr * unr + pstar * Sstar) * inv_denom;
    Fr = Fr + S4 * (U1S - drtot); Fu = Fu + S4 * (U2S - drtot * ur); Fv = Fv + S4 * (U3S - drtot * vr);
    Fw = Fw + S4 * (U4S - drtot * wr); FE = FE + S4 * (U5S - drtot * e0r);
  } else {
    Fr = drtot * unr; Fu = Fr * ur + pr * nx; Fv = Fr * vr + pr * ny; Fw = Fr * wr + pr * nz; FE = Fr * h0r;
  }
  F[0] = Fr; F[1] = Fu; F[2] = Fv; F[3] = Fw; F[4] = FE;
}

/* static enthalpy of one state with the clamped index of riemann_open_body.inc:103-111 */
static double h_open(const o_tables *t, const double *d, double p) {
  double rho = 0.0, rg = 0.0, hs = 0.0;
  for (int s = 0; s < t->ns; ++s) rho = rho + d[s];
  for (int s = 0; s < t->ns; ++s) rg = rg + t->Ri[s] * d[s] / rho;
  int lo = 0, hi = t->tmax;
  for (int s = 0; s < t->ns; ++s) {
    double T = p / (rg * rho);
    double Tc = max2((double)lo, min2(T, (double)(hi - 1)));
    int Ti = fint(Tc);
    double Td = Tc - Ti;
    int T1 = imax(lo, imin(Ti, hi)), T2 = imax(lo, imin(Ti + 1, hi));
    double hs_ = TAB(t->h, t, s, T1) + (TAB(t->h, t, s, T2) - TAB(t->h, t, s, T1)) * Td;
    hs = hs + hs_ * d[s] / rho;
  }
  return hs;
}

/* AUSM+ (rs 4): riemann_open_body.inc:100-150 */
static void ausmp(const o_tables *t, const double *dl, double ul, double vl, double wl, double pl, double al,
                  double dltot, const double *dr, double ur, double vr, double wr, double pr, double ar, double drtot,
                  double nx, double ny, double nz, double F[5]) {
  const double ab_ = 0.125, aa_ = 0.1875;
  double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  double vel = sqrt(ul * ul + vl * vl + wl * wl);
  double H_L = h_open(t, dl, pl) + 0.5 * vel * vel;
  vel = sqrt(ur * ur + vr * vr + wr * wr);
  double H_R = h_open(t, dr, pr) + 0.5 * vel * vel;
  double a_half = sqrt(al * ar);
  double ML = unl / a_half, MR = unr / a_half;
  double MplusL, pplusL, MminusR, pminusR;
  if (fabs(ML) < 1.0) {
    double a = ML + 1.0, b = ML * ML - 1.0;
    MplusL = 0.25 * (a * a) + ab_ * (b * b);
    pplusL = 0.25 * (a * a) * (2.0 - ML) + aa_ * ML * (b * b);
  } else {
    MplusL = 0.5 * (ML + fabs(ML));
    pplusL = 0.5 * (1.0 + copysign(1.0, ML));
  }
  if (fabs(MR) < 1.0) {
    double a = MR - 1.0, b = MR * MR - 1.0;
    MminusR = -(0.25 * (a * a)) - ab_ * (b * b);
    pminusR = 0.25 * (a * a) * (2.0 + MR) - aa_ * MR * (b * b);
  } else {
    MminusR = 0.5 * (MR - fabs(MR));
    pminusR = 0.5 * (1.0 - copysign(1.0, MR));
  }
  double Mhalf = MplusL + MminusR;
  double phalf = pplusL * pl + pminusR * pr;
  double mdotp = dltot * a_half * 0.5 * (Mhalf + fabs(Mhalf));
  double mdotm = drtot * a_half * 0.5 * (Mhalf - fabs(Mhalf));
  F[0] = mdotp + mdotm;
  F[1] = mdotp * ul + mdotm * ur + nx * phalf;
  F[2] = mdotp * vl + mdotm * vr + ny * phalf;
  F[3] = mdotp * wl + mdotm * wr + nz * phalf;
  F[4] = mdotp * H_L + mdotm * H_R;
}

/* LMRoe (rs 12): riemann_open_body.inc:965-1090; with riemann-options-Mco = 1 it is the classical Roe flux */
static void lmroe(const o_tables *t, double mco, const double *dl, double ul, double vl, double wl, double pl,
                  double dltot, const double *dr, double ur, double vr, double wr, double pr, double drtot, double nx,
                  double ny, double nz, const roe_t *ro, double F[5]) {
  double nmag = sqrt(nx * nx + ny * ny + nz * nz);
  if (nmag <= 0.0) { F[0] = F[1] = F[2] = F[3] = F[4] = 0.0; return; }
  double n1 = nx / nmag, n2 = ny / nmag, n3 = nz / nmag;
  double tx, ty, tz;
  if (fabs(n1) < 0.9) { tx = 1.0; ty = 0.0; tz = 0.0; } else { tx = 0.0; ty = 1.0; tz = 0.0; }
  double t1x = tx - (tx * n1 + ty * n2 + tz * n3) * n1;
  double t1y = ty - (tx * n1 + ty * n2 + tz * n3) * n2;
  double t1z = tz - (tx * n1 + ty * n2 + tz * n3) * n3;
  double tmag = sqrt(t1x * t1x + t1y * t1y + t1z * t1z);
  if (tmag <= 0.0) { t1x = -n2; t1y = n1; t1z = 0.0; tmag = sqrt(t1x * t1x + t1y * t1y + t1z * t1z); }
  t1x = t1x / tmag; t1y = t1y / tmag; t1z = t1z / tmag;
  double t2x = n2 * t1z - n3 * t1y, t2y = n3 * t1x - n1 * t1z, t2z = n1 * t1y - n2 * t1x;
  double unL = ul * n1 + vl * n2 + wl * n3, unR = ur * n1 + vr * n2 + wr * n3;
  double ut1L = ul * t1x + vl * t1y + wl * t1z, ut1R = ur * t1x + vr * t1y + wr * t1z;
  double ut2L = ul * t2x + vl * t2y + wl * t2z, ut2R = ur * t2x + vr * t2y + wr * t2z;
  double hL = h_open(t, dl, pl), hR = h_open(t, dr, pr);
  double H0L = hL + 0.5 * (ul * ul + vl * vl + wl * wl);
  double H0R = hR + 0.5 * (ur * ur + vr * vr + wr * wr);
  double FL1 = dltot * unL, FL2 = dltot * unL * ul + pl * n1, FL3 = dltot * unL * vl + pl * n2;
  double FL4 = dltot * unL * wl + pl * n3, FL5 = dltot * unL * H0L;
  double FR1 = drtot * unR, FR2 = drtot * unR * ur + pr * n1, FR3 = drtot * unR * vr + pr * n2;
  double FR4 = drtot * unR * wr + pr * n3, FR5 = drtot * unR * H0R;
  double u_roe = ro->u_roe, v_roe = ro->v_roe, w_roe = ro->w_roe, a_roe = ro->a_roe, rho_roe = ro->rho_roe;
  double h0_roe = ro->h0_roe;
  double unRoe = u_roe * n1 + v_roe * n2 + w_roe * n3;
  double u2 = u_roe * u_roe + v_roe * v_roe + w_roe * w_roe;
  double utmag = sqrt(max2(0.0, u2 - unRoe * unRoe));
  double Ma = (a_roe > 0.0) ? (fabs(unRoe) + utmag) / a_roe : 1.0;
  double phi = max2(mco, min2(1.0, Ma));
  double drho = drtot - dltot, dp = pr - pl, dun = unR - unL, dut1 = ut1R - ut1L, dut2 = ut2R - ut2L;
  double dunmod = phi * dun;
  double ab1 = fabs(unRoe - a_roe), ab2 = fabs(unRoe), ab3 = fabs(unRoe), ab4 = fabs(unRoe);
  double ab5 = fabs(unRoe + a_roe);
  double a2 = a_roe * a_roe, inva = 1.0 / a_roe, inva2 = 1.0 / a2;
  double al4 = drho - dp * inva2, al2 = rho_roe * dut1, al3 = rho_roe * dut2;
  double al1 = 0.5 * (dp * inva2 - rho_roe * dunmod * inva);
  double al5 = 0.5 * (dp * inva2 + rho_roe * dunmod * inva);
  double r1[5] = {1.0, u_roe - a_roe * n1, v_roe - a_roe * n2, w_roe - a_roe * n3, h0_roe - a_roe * unRoe};
  double r5[5] = {1.0, u_roe + a_roe * n1, v_roe + a_roe * n2, w_roe + a_roe * n3, h0_roe + a_roe * unRoe};
  double r4[5] = {1.0, u_roe, v_roe, w_roe, 0.5 * u2};
  double r2[5] = {0.0, t1x, t1y, t1z, u_roe * t1x + v_roe * t1y + w_roe * t1z};
  double r3[5] = {0.0, t2x, t2y, t2z, u_roe * t2x + v_roe * t2y + w_roe * t2z};
  double FL[5] = {FL1, FL2, FL3, FL4, FL5}, FR[5] = {FR1, FR2, FR3, FR4, FR5};
  for (int q = 0; q < 5; ++q) {
    double d = ab1 * al1 * r1[q] + ab2 * al2 * r2[q] + ab3 * al3 * r3[q] + ab4 * al4 * r4[q] + ab5 * al5 * r5[q];
    F[q] = 0.5 * (FL[q] + FR[q]) - 0.5 * d;
  }
}

/* SLAU, the part common to the interior kernels (conv_physics.inc:126-167) and the record kernels' riemann_SLAU
 * (Lib_Riemann.f90:140-198): from the sound speeds al/ar and the total enthalpies h0l/h0r, which each computes its
 * own way (below). */
static void slau(double dltot, double drtot, double ul, double vl, double wl, double pl, double al, double ur,
                 double vr, double wr, double pr, double ar, double nx, double ny, double nz, double h0l, double h0r,
                 double F[5]) {
  const double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  const double af = 0.5 * (al + ar);
  const double ml = unl / af, mr = unr / af;
  const double mt = min2(1.0, sqrt(0.5 * ((ul * ul + vl * vl + wl * wl) + (ur * ur + vr * vr + wr * wr))) / af);
  const double chi = (1.0 - mt) * (1.0 - mt);
  const double vnm = (dltot * fabs(unl) + drtot * fabs(unr)) / (dltot + drtot);
  const double frho = -max2(min2(ml, 0.0), -1.0) * min2(max2(mr, 0.0), 1.0);
  const double vnl = (1.0 - frho) * vnm + frho * fabs(unl);
  const double vnr = (1.0 - frho) * vnm + frho * fabs(unr);
  const double Fr = 0.5 * (dltot * (unl + vnl) + drtot * (unr - vnr) - (chi / af) * (pr - pl));
  double bl, br;
  if (fabs(ml) < 1.0) bl = 0.25 * (2.0 - ml) * ((ml + 1.0) * (ml + 1.0));
  else if (ml >= 0.0) bl = 1.0;
  else bl = 0.0;
  if (fabs(mr) < 1.0) br = 0.25 * (2.0 + mr) * ((mr - 1.0) * (mr - 1.0));
  else if (mr >= 0.0) br = 0.0;
  else br = 1.0;
  const double ph = 0.5 * (pl + pr) + 0.5 * (bl - br) * (pl - pr) + (1.0 - chi) * (bl + br - 1.0) * 0.5 * (pl + pr);
  F[0] = Fr;
  if (Fr >= 0.0) {
    F[1] = Fr * ul + nx * ph; F[2] = Fr * vl + ny * ph; F[3] = Fr * wl + nz * ph; F[4] = Fr * h0l;
  } else {
    F[1] = Fr * ur + nx * ph; F[2] = Fr * vr + ny * ph; F[3] = Fr * wr + nz * ph; F[4] = Fr * h0r;
  }
}

/* riemann_SLAU's total enthalpy of one state (Lib_Riemann.f90:152-166): R = sum Ri rho_i / rho, T = p / (rho R),
 * the table rows of int(T), h0 = sum h_i rho_i / rho + |u|^2 / 2 */
static double slau_h0(const o_tables *t, const double *d, double dtot, double u, double v, double w, double p) {
  double R = 0.0, h0 = 0.0;
  for (int s = 0; s < t->ns; ++s) R = R + t->Ri[s] * d[s] / dtot;
  const double Tk = p / (dtot * R);
  const int Ti = fint(Tk);
  const double dT = Tk - Ti;
  for (int s = 0; s < t->ns; ++s) h0 = h0 + tabT(t->h, t, s, Ti, Ti + 1, dT) * d[s] / dtot;
  return h0 + 0.5 * (u * u + v * v + w * w);
}

/* riemann_open (rs 4 and 12) and the Riemann dispatch of the record kernels (Lib_Riemann.f90:114-136).
 * `clamp_hllc` selects the table-index clamp of the Roe averages used by HLLC: 1 in the interior kernel
 * (conv_physics.inc:73-74), 0 in the record kernels (Lib_Riemann_HLL.f90:244-247). */
static void riemann_any(const o_case *C, int rs, int clamp_hllc, const double *dl, double ul, double vl, double wl,
                        double pl, double al, double dltot, const double *dr, double ur, double vr, double wr, double pr,
                        double ar, double drtot, double nx, double ny, double nz, double Rl, double Rr, double F[5]) {
  const o_tables *t = &C->tab;
  roe_t ro;
  if (rs == 1) {
    roe_averages(t, clamp_hllc, 0, ONTMAX - 1, dl, dr, dltot, drtot, ul, vl, wl, ur, vr, wr, pl, pr, Rl, Rr, &ro);
    hllc(dltot, drtot, ul, vl, wl, pl, al, ur, vr, wr, pr, ar, nx, ny, nz, &ro, F);
  } else if (rs == 2) {
    roe_averages(t, clamp_hllc, 0, ONTMAX - 1, dl, dr, dltot, drtot, ul, vl, wl, ur, vr, wr, pl, pr, Rl, Rr, &ro);
    hlle(dltot, drtot, ul, vl, wl, pl, al, ur, vr, wr, pr, ar, nx, ny, nz, &ro, F);
  } else if (rs == 3) {
    /* the interior kernels take h0 of the inline Roe block (conv_physics.inc:57-78), the records riemann_SLAU's own */
    double h0l, h0r;
    if (clamp_hllc) {
      roe_averages(t, 1, 0, ONTMAX - 1, dl, dr, dltot, drtot, ul, vl, wl, ur, vr, wr, pl, pr, Rl, Rr, &ro);
      h0l = ro.h0l; h0r = ro.h0r;
    } else {
      h0l = slau_h0(t, dl, dltot, ul, vl, wl, pl);
      h0r = slau_h0(t, dr, drtot, ur, vr, wr, pr);
    }
    slau(dltot, drtot, ul, vl, wl, pl, al, ur, vr, wr, pr, ar, nx, ny, nz, h0l, h0r, F);
  } else if (rs == 4) {
    ausmp(t, dl, ul, vl, wl, pl, al, dltot, dr, ur, vr, wr, pr, ar, drtot, nx, ny, nz, F);
  } else if (rs == 12) {
    /* the Roe averages of riemann_open_body.inc:679-726 take Rl/Rr = f_Rtot of the states and clamp to [Tmin,Tmax-1] */
    double rho = 0.0, R1 = 0.0, R2 = 0.0;
    for (int s = 0; s < t->ns; ++s) rho = rho + dl[s];
    for (int s = 0; s < t->ns; ++s) R1 = R1 + t->Ri[s] * dl[s] / rho;
    rho = 0.0;
    for (int s = 0; s < t->ns; ++s) rho = rho + dr[s];
    for (int s = 0; s < t->ns; ++s) R2 = R2 + t->Ri[s] * dr[s] / rho;
    roe_averages(t, 1, 0, t->tmax - 1, dl, dr, dltot, drtot, ul, vl, wl, ur, vr, wr, pl, pr, R1, R2, &ro);
    lmroe(t, C->sc.mco, dl, ul, vl, wl, pl, dltot, dr, ur, vr, wr, pr, drtot, nx, ny, nz, &ro, F);
  } else {
    F[0] = F[1] = F[2] = F[3] = F[4] = 0.0;
  }
}

void o_riemann(const o_case *C, int rs, double bet, const double *dl, double ul, double vl, double wl, double pl,
               double al, double dltot, const double *dr, double ur, double vr, double wr, double pr, double ar,
               double drtot, double nx, double ny, double nz, double F[5]) {
  (void)bet;
  double Rl = f_Rtot(&C->tab, dl), Rr = f_Rtot(&C->tab, dr);
  riemann_any(C, rs, 1, dl, ul, vl, wl, pl, al, dltot, dr, ur, vr, wr, pr, ar, drtot, nx, ny, nz, Rl, Rr, F);
}

/* ------------------------------------------------------------------------------------------------ */
/* interior convective face: conv_physics.inc:4-179 (one face of conv_cuf_i/j/k, Mod_Convective.f90)  */
/* ------------------------------------------------------------------------------------------------ */
void o_conv_face(const o_case *C, const double *pm1, const double *p0, const double *pp1, const double *pp2,
                 const double dlc[4], const double nrm[3], double area, double *res0) {
  const o_scheme *S = &C->sc;
  const o_tables *t = &C->tab;
  int nprim = S->nprim, nsc = S->nsc, np = S->np;
  double P1[ONPRIM], P4[ONPRIM];
  double l0 = 0.5 * (dlc[0] + dlc[1]), l1 = 0.5 * (dlc[1] + dlc[2]), l2 = 0.5 * (dlc[2] + dlc[3]);
  double lm = 0.5 * dlc[1], lp = 0.5 * dlc[2];
  double beta_order = 1.0;
  double idl0 = 1.0 / l0, idl1 = 1.0 / l1, idl2 = 1.0 / l2;
  for (int s = 0; s < nprim; ++s) {
    double slope0 = (p0[s] - pm1[s]) * idl0;
    double slope1 = (pp1[s] - p0[s]) * idl1;
    double slope2 = (pp2[s] - pp1[s]) * idl2;
    double slopel = o_rlimiter(S->lim, slope1, slope0);
    double sloper = o_rlimiter(S->lim, slope2, slope1);
    P1[s] = p0[s] + beta_order * slopel * lm;
    P4[s] = pp1[s] - beta_order * sloper * lp;
  }
  for (int s = 0; s < nsc; ++s) { P1[s] = max2(P1[s], 1e-20); P4[s] = max2(P4[s], 1e-20); }
  double rotot1 = 0.0, Rtot1 = 0.0, rotot4 = 0.0, Rtot4 = 0.0;
  for (int s = 0; s < nsc; ++s) rotot1 = rotot1 + P1[s];
  for (int s = 0; s < nsc; ++s) Rtot1 = Rtot1 + t->Ri[s] * P1[s] / rotot1;
  for (int s = 0; s < nsc; ++s) rotot4 = rotot4 + P4[s];
  for (int s = 0; s < nsc; ++s) Rtot4 = Rtot4 + t->Ri[s] * P4[s] / rotot4;
  double a1, a4;
  {
    double T = P1[np] / (Rtot1 * rotot1);
    int Ti = fint(T); if (Ti < 0) Ti = 0; if (Ti > ONTMAX - 1) Ti = ONTMAX - 1;
    double Td = T - Ti, cpd = 0.0;
    for (int s = 0; s < nsc; ++s) {
      double cpi = TAB(t->cp, t, s, Ti) + (TAB(t->cp, t, s, Ti + 1) - TAB(t->cp, t, s, Ti)) * Td;
      cpd = cpd + P1[s] / rotot1 * cpi;
    }
    double gamd = cpd / (cpd - Rtot1);
    a1 = sqrt(gamd * Rtot1 * T);
  }
  {
    double T = P4[np] / (Rtot4 * rotot4);
    int Ti = fint(T); if (Ti < 0) Ti = 0; if (Ti > ONTMAX - 1) Ti = ONTMAX - 1;
    double Td = T - Ti, cpd = 0.0;
    for (int s = 0; s < nsc; ++s) {
      double cpi = TAB(t->cp, t, s, Ti) + (TAB(t->cp, t, s, Ti + 1) - TAB(t->cp, t, s, Ti)) * Td;
      cpd = cpd + P4[s] / rotot4 * cpi;
    }
    double gamd = cpd / (cpd - Rtot4);
    a4 = sqrt(gamd * Rtot4 * T);
  }
  double F[5];
  riemann_any(C, S->rs, 1, P1, P1[S->nu], P1[S->nv], P1[S->nw], P1[np], a1, rotot1, P4, P4[S->nu], P4[S->nv],
              P4[S->nw], P4[np], a4, rotot4, nrm[0], nrm[1], nrm[2], Rtot1, Rtot4, F);
  double su = copysign(0.5, F[0]), sp = 0.5 + su, sm = su - 0.5;
  double ir1 = 1.0 / rotot1, ir4 = 1.0 / rotot4;
  for (int s = 0; s < nsc; ++s) res0[s] = F[0] * area * (sp * (P1[s] * ir1) - sm * (P4[s] * ir4));
  res0[S->nu] = F[1] * area; res0[S->nv] = F[2] * area; res0[S->nw] = F[3] * area; res0[np] = F[4] * area;
  for (int s = np + 1; s < nprim; ++s) res0[s] = F[0] * area * (sp * P1[s] * ir1 - sm * P4[s] * ir4);
}

/* ------------------------------------------------------------------------------------------------ */
/* viscous                                                                                           */
/* ------------------------------------------------------------------------------------------------ */
/* SA constants: src/lib/physics/turbulence/Mod_Spalart.f90:14-34 */
#define SA_CV1 7.1
#define SA_SIGMA (2.0 / 3.0)
static const double NTT = -2.0 / 3.0;   /* Mod_Fluid.f90:4 negativetwothird */

/* Spalart_Eddy_Viscosity: Mod_Spalart.f90:382-398 (fv1(chi) * rho nu~, chi = rho nu~ / mu_l) */
static double sa_mut(double rnut, double mil) {
  double chi = rnut / mil;
  double c3 = chi * chi * chi;
  return c3 / (c3 + SA_CV1 * SA_CV1 * SA_CV1) * rnut;
}

/* The face thermo of Diffusive_Face_Thermo_Fill (Mod_Diffusive.f90:291-313) from the averaged state */
typedef struct { double rho, T, mil, kl, cp; int Ti; double Td; } fthermo_t;
static void face_thermo(const o_case *C, const double *prim, fthermo_t *o) {
  const o_tables *t = &C->tab;
  double rho, Rg;
  co_rotot_Rtot(t, prim, &rho, &Rg);
  double T = prim[C->sc.np] / (rho * Rg);
  int Ti = fint(T);
  double Td = T - Ti;
  double cp = f_cp_dev(t, prim, Ti, Td, rho);
  double mil, kl;
  wilke(t, prim, rho, Ti, Td, &mil, &kl);
  if (C->sc.prl > 0.0) kl = mil * cp / C->sc.prl;
  o->rho = rho; o->T = T; o->mil = mil; o->kl = kl; o->cp = cp; o->Ti = Ti; o->Td = Td;
}

/* interior viscous face: diff_cuf_kernels.inc:11-243 + diff_physics.inc:1-161.
 * cells c[0..9] = c1..c10 of the kernel, rc/Tc the thermo caches, ax = (ia,ib,ic) */
static void diff_face(const o_case *C, const double *P, const double *rc, const double *Tc, const int64_t c[10],
                      const int ax[3], const double *M1, const double *M2, int64_t mstride, const double nrm[3],
                      double area, double wd1, double wd2, double *res0) {
  const o_scheme *S = &C->sc;
  const o_tables *t = &C->tab;
  const int64_t NC = C->NC;
  int nprim = S->nprim, nsc = S->nsc, np = S->np, nu = S->nu, nw = S->nw;
  int ia = ax[0], ib = ax[1], ic = ax[2];
  double G[ONPRIM][3], Primf[ONPRIM], Flx[ONPRIM], Dm[ONPRIM];
  double p1[ONPRIM], p2[ONPRIM];
  for (int s = 0; s < nprim; ++s) { p1[s] = PV(P, s, c[0]); p2[s] = PV(P, s, c[1]); }
  fthermo_t ft;
  {
    double prim[ONPRIM];
    for (int s = 0; s < nsc; ++s) prim[s] = 0.5 * (p1[s] + p2[s]);
    prim[np] = 0.5 * (p1[np] + p2[np]);
    face_thermo(C, prim, &ft);
  }
  double irho1 = 1.0 / rc[c[0]], irho2 = 1.0 / rc[c[1]];
  for (int s = 0; s < nsc; ++s) G[s][ia] = p2[s] * irho2 - p1[s] * irho1;
  G[np][ia] = Tc[c[1]] - Tc[c[0]];
  for (int s = nu; s <= nw; ++s) G[s][ia] = p2[s] - p1[s];
  double waldis = 0.0;
  if (nprim > np + 1) {
    for (int s = np + 1; s < nprim; ++s) G[s][ia] = p2[s] * irho2 - p1[s] * irho1;
    waldis = 0.5 * (wd1 + wd2);
  }
  for (int pass = 0; pass < 2; ++pass) {
    int a = pass ? ic : ib;
    int64_t q3 = c[2 + 4 * pass], q4 = c[3 + 4 * pass], q5 = c[4 + 4 * pass], q6 = c[5 + 4 * pass];
    double i1 = 1.0 / rc[q3], i2 = 1.0 / rc[q4], i3 = 1.0 / rc[q5], i4 = 1.0 / rc[q6];
    for (int s = 0; s < nsc; ++s)
      G[s][a] = (PV(P, s, q4) * i2 - PV(P, s, q3) * i1 + PV(P, s, q6) * i4 - PV(P, s, q5) * i3) * 0.25;
    G[np][a] = (Tc[q4] - Tc[q3] + Tc[q6] - Tc[q5]) * 0.25;
    for (int s = nu; s <= nw; ++s) G[s][a] = (PV(P, s, q4) - PV(P, s, q3) + PV(P, s, q6) - PV(P, s, q5)) * 0.25;
    for (int s = np + 1; s < nprim; ++s)
      G[s][a] = (PV(P, s, q4) * i2 - PV(P, s, q3) * i1 + PV(P, s, q6) * i4 - PV(P, s, q5) * i3) * 0.25;
  }
  double Mav[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) Mav[a][b] = 0.5 * (M1[(a + 3 * b) * mstride] + M2[(a + 3 * b) * mstride]);
  for (int iv = 0; iv < nprim; ++iv) {
    double g1 = G[iv][0], g2 = G[iv][1], g3 = G[iv][2];
    G[iv][0] = g1 * Mav[0][0] + g2 * Mav[1][0] + g3 * Mav[2][0];
    G[iv][1] = g1 * Mav[0][1] + g2 * Mav[1][1] + g3 * Mav[2][1];
    G[iv][2] = g1 * Mav[0][2] + g2 * Mav[1][2] + g3 * Mav[2][2];
  }
  for (int iv = 0; iv < nprim; ++iv) Primf[iv] = 0.5 * (p1[iv] + p2[iv]);
  double drho = ft.rho, dT = ft.T, dcp = ft.cp, dmil = ft.mil, dkl = ft.kl;
  int Ti = fint(dT);
  double Tdf = dT - Ti;
  double Gv[3][3];
  for (int d = 0; d < 3; ++d) { Gv[0][d] = G[nu][d]; Gv[1][d] = G[nu + 1][d]; Gv[2][d] = G[nu + 2][d]; }
  double mie = 0.0;
  if (S->rans == 2) {
    double chi = Primf[S->nt] / dmil;
    double c3 = chi * chi * chi;
    mie = c3 / (c3 + SA_CV1 * SA_CV1 * SA_CV1) * Primf[S->nt];
  }
  for (int s = 0; s < nsc; ++s) Dm[s] = (dmil / S->sc + mie / S->sct) / drho;
  double kappaf = dkl + mie * dcp / S->prt;
  double dmi = dmil + mie, dlam = NTT * dmi;
  double Diver = Gv[0][0] + Gv[1][1] + Gv[2][2];
  double lamDiv = dlam * Diver, rhokd = 0.0;
  double Txx = lamDiv + 2.0 * dmi * Gv[0][0] + rhokd;
  double Tyy = lamDiv + 2.0 * dmi * Gv[1][1] + rhokd;
  double Tzz = lamDiv + 2.0 * dmi * Gv[2][2] + rhokd;
  double Txy = dmi * (Gv[0][1] + Gv[1][0]), Txz = dmi * (Gv[0][2] + Gv[2][0]), Tyz = dmi * (Gv[1][2] + Gv[2][1]);
  double St1 = Txx * nrm[0] + Txy * nrm[1] + Txz * nrm[2];
  double St2 = Txy * nrm[0] + Tyy * nrm[1] + Tyz * nrm[2];
  double St3 = Txz * nrm[0] + Tyz * nrm[1] + Tzz * nrm[2];
  double DiffH = 0.0;
  for (int iv = 0; iv < nprim; ++iv) Flx[iv] = 0.0;
  for (int s = 0; s < nsc; ++s) {
    double DmGY = Dm[s] * (G[s][0] * nrm[0] + G[s][1] * nrm[1] + G[s][2] * nrm[2]);
    Flx[s] = drho * area * DmGY;
    double hh1 = TAB(t->h, t, s, Ti), hh2 = TAB(t->h, t, s, Ti + 1);
    double hh = hh1 + Tdf * (hh2 - hh1);
    DiffH = DiffH + DmGY * hh;
  }
  double sflx = 0.0;
  for (int s = 0; s < nsc; ++s) sflx = sflx + Flx[s];
  for (int s = 0; s < nsc; ++s) Flx[s] = Flx[s] - sflx * Primf[s] / drho;
  Flx[nu] = area * St1; Flx[nu + 1] = area * St2; Flx[nu + 2] = area * St3;
  Flx[np] = area * ((St1 * Primf[nu] + St2 * Primf[nu + 1] + St3 * Primf[nu + 2]) +
                    kappaf * (G[np][0] * nrm[0] + G[np][1] * nrm[1] + G[np][2] * nrm[2]) + DiffH * drho);
  if (S->rans == 2) {
    int nt = S->nt;
    double gndot = G[nt][0] * nrm[0] + G[nt][1] * nrm[1] + G[nt][2] * nrm[2];
    Flx[nt] = (dmil + Primf[nt]) * gndot * area / SA_SIGMA;
  }
  (void)waldis;
  for (int iv = 0; iv < nprim; ++iv) res0[iv] = Flx[iv];
}

/* Compute_Diffusive_Flux_St (Mod_Diffusive.f90:781-867) with Stress_Vector_Std (Mod_Fluid.f90:20-44) and
 * Spalart_RANS_Diffusive_Flux (Mod_Spalart.f90:401-418) on the face state of Diffusive_Face_State (:748-776) */
static void diff_flux_st(const o_case *C, const double *Prim, double G[][3], double area, const double *nrm,
                         const fthermo_t *ft, double *Flux) {
  const o_scheme *S = &C->sc;
  const o_tables *t = &C->tab;
  int nsc = S->nsc, np = S->np, nu = S->nu, nprim = S->nprim;
  double rho = ft->rho, cp = ft->cp, mil = ft->mil, kl = ft->kl;
  double Gv[3][3];
  for (int d = 0; d < 3; ++d) { Gv[0][d] = G[nu][d]; Gv[1][d] = G[nu + 1][d]; Gv[2][d] = G[nu + 2][d]; }
  double mie = 0.0;
  if (S->rans == 2) mie = sa_mut(Prim[S->nt], mil);
  double Dm[ONPRIM];
  for (int s = 0; s < nsc; ++s) Dm[s] = (mil / S->sc + mie / S->sct) / rho;
  double kappa = kl + mie * cp / S->prt;
  double mi = mil + mie, lam = NTT * mi;
  double Div = Gv[0][0] + Gv[1][1] + Gv[2][2];
  double lamDiv = lam * Div;
  double Txx = lamDiv + 2.0 * mi * Gv[0][0], Tyy = lamDiv + 2.0 * mi * Gv[1][1], Tzz = lamDiv + 2.0 * mi * Gv[2][2];
  double Txy = mi * (Gv[0][1] + Gv[1][0]), Txz = mi * (Gv[0][2] + Gv[2][0]), Tyz = mi * (Gv[1][2] + Gv[2][1]);
  double St[3];
  St[0] = Txx * nrm[0] + Txy * nrm[1] + Txz * nrm[2];
  St[1] = Txy * nrm[0] + Tyy * nrm[1] + Tyz * nrm[2];
  St[2] = Txz * nrm[0] + Tyz * nrm[1] + Tzz * nrm[2];
  double DiffH = 0.0;
  for (int s = 0; s < nprim; ++s) Flux[s] = 0.0;
  for (int s = 0; s < nsc; ++s) {
    double DmG = Dm[s] * (G[s][0] * nrm[0] + G[s][1] * nrm[1] + G[s][2] * nrm[2]);
    Flux[s] = rho * area * DmG;
    double hh1 = TAB(t->h, t, s, ft->Ti), hh2 = TAB(t->h, t, s, ft->Ti + 1);
    double hh = hh1 + ft->Td * (hh2 - hh1);
    DiffH = DiffH + DmG * hh;
  }
  double sf = 0.0;
  for (int s = 0; s < nsc; ++s) sf = sf + Flux[s];
  for (int s = 0; s < nsc; ++s) Flux[s] = Flux[s] - sf * Prim[s] / rho;
  for (int d = 0; d < 3; ++d) Flux[nu + d] = area * St[d];
  Flux[np] = area * ((St[0] * Prim[nu] + St[1] * Prim[nu + 1] + St[2] * Prim[nu + 2]) +
                     kappa * (G[np][0] * nrm[0] + G[np][1] * nrm[1] + G[np][2] * nrm[2]) + DiffH * rho);
  if (S->rans == 2) {
    int nt = S->nt;
    Flux[nt] = area * (mil + Prim[nt]) * (G[nt][0] * nrm[0] + G[nt][1] * nrm[1] + G[nt][2] * nrm[2]) / SA_SIGMA;
  }
}

/* ------------------------------------------------------------------------------------------------ */
/* thermo cache: Mod_Thermo_Cache.f90:291-391 (rho, R, T on every padded cell, mu_l on the interior) */
/* ------------------------------------------------------------------------------------------------ */
static int thermo_cache(const o_case *C, const double *P, double *rc, double *Rc, double *Tc, double *mc) {
  const int64_t NC = C->NC;
  const o_scheme *S = &C->sc;
  int nbad = 0;
  for (int b = 0; b < C->nb; ++b) {
    const o_block *B = &C->blk[b];
    for (int k = 1 - GC; k <= B->n3 + GC; ++k)
      for (int j = 1 - GC; j <= B->n2 + GC; ++j)
        for (int i = 1 - GC; i <= B->n1 + GC; ++i) {
          int64_t r = ROW(B, i, j, k);
          double rl[ONSMAX], rho, Rg;
          for (int s = 0; s < S->nsc; ++s) rl[s] = PV(P, s, r);
          co_rotot_Rtot(&C->tab, rl, &rho, &Rg);
          double T = PV(P, S->np, r) / (rho * Rg);
          int inner = (i >= 1 && i <= B->n1 && j >= 1 && j <= B->n2 && k >= 1 && k <= B->n3);
          if (inner && S->visc && S->guard) {
            if (!(T > 1.0 && T < 9990.0)) {
              nbad++;
              if (T != T) T = 300.0;
              T = min2(max2(T, 1.0), 9990.0);
            }
          }
          rc[r] = rho; Rc[r] = Rg; Tc[r] = T;
          if (inner && S->visc) {
            int Ti = fint(T);
            double mil, kl;
            wilke(&C->tab, rl, rho, Ti, T - Ti, &mil, &kl);
            mc[r] = mil;
          }
        }
  }
  return nbad;
}

/* ------------------------------------------------------------------------------------------------ */
/* ghosts: Lib_Extrapolation.f90 Fill_Ghost_Cell_local (:1370-1524)                                  */
/* ------------------------------------------------------------------------------------------------ */
static const int guide[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

static void record_geometry(const o_case *C, const mose_bc_record *rc, double nrm[3], double *area, double dlc4[4]);

void o_fill_ghosts(const o_case *C, double *P) {
  const int64_t NC = C->NC;
  int nprim = C->sc.nprim;
  /* pass 1 per block: zero-gradient ghosts (Ghost_ZG_Extrapolate :590-607) for 0/4/11, the mirror ghost of the
   * symmetry faces (Ghost_Symmetry :558-587: the first layer only, velocity reflected in the face) for 3 */
  for (int e = 0; e < C->nrec; ++e) {
    const mose_bc_record *rc = &C->rec[e];
    const o_block *B = &C->blk[rc->b];
    int dir = (rc->f + 1) / 2 - 1, modfm = rc->f % 2, gstep = 2 * modfm - 1;
    if (rc->typ == 0 || rc->typ == 4 || rc->typ == 11) {
      for (int g = 1; g <= GC; ++g) {
        int ig = rc->i, jg = rc->j, kg = rc->k;
        if (dir == 0) ig -= gstep * g; else if (dir == 1) jg -= gstep * g; else kg -= gstep * g;
        for (int s = 0; s < nprim; ++s) PV(P, s, ROW(B, ig, jg, kg)) = PV(P, s, ROW(B, rc->i, rc->j, rc->k));
      }
    } else if (rc->typ == 3) {
      const o_scheme *S = &C->sc;
      double nrm[3], area, dlc4[4];
      record_geometry(C, rc, nrm, &area, dlc4);
      int ig = rc->i, jg = rc->j, kg = rc->k;
      if (dir == 0) ig -= gstep; else if (dir == 1) jg -= gstep; else kg -= gstep;
      int64_t rm = ROW(B, rc->i, rc->j, rc->k), rg = ROW(B, ig, jg, kg);
      double vel[3] = {PV(P, S->nu, rm), PV(P, S->nv, rm), PV(P, S->nw, rm)};
      double udotn = vel[0] * nrm[0] + vel[1] * nrm[1] + vel[2] * nrm[2];
      for (int q = 0; q < 3; ++q) vel[q] = vel[q] - 2.0 * udotn * nrm[q];
      for (int s = 0; s < S->nsc; ++s) PV(P, s, rg) = PV(P, s, rm);
      PV(P, S->nu, rg) = vel[0]; PV(P, S->nv, rg) = vel[1]; PV(P, S->nw, rg) = vel[2];
      for (int s = S->np; s < nprim; ++s) PV(P, s, rg) = PV(P, s, rm);
    }
  }
  /* wall ghosts (first pass, `case default` of the host fill) and the cubic pass of the 400 faces
   * (Fill_Ghost_Cubic_Extrapolation_blk :859-879) */
  for (int pass = 0; pass < 2; ++pass)
    for (int e = 0; e < C->nrec; ++e) {
      const mose_bc_record *rc = &C->rec[e];
      const o_block *B = &C->blk[rc->b];
      int want = pass == 0 ? (rc->typ == 5 || rc->typ == 6) : (rc->typ == 11);
      if (!want) continue;
      int dir = (rc->f + 1) / 2 - 1, modfm = rc->f % 2, gstep = 2 * modfm - 1, step = -gstep;
      for (int g = 1; g <= GC; ++g) {
        int o[4] = {g * step, (g - 1) * step, (g - 2) * step, (g - 3) * step};
        int64_t r[4];
        for (int q = 0; q < 4; ++q) {
          int ii = rc->i, jj = rc->j, kk = rc->k;
          if (dir == 0) ii += o[q]; else if (dir == 1) jj += o[q]; else kk += o[q];
          r[q] = ROW(B, ii, jj, kk);
        }
        for (int s = 0; s < nprim; ++s)
          PV(P, s, r[0]) = 3.0 * PV(P, s, r[1]) - 3.0 * PV(P, s, r[2]) + PV(P, s, r[3]);
      }
    }
  /* connection copy: Ghost_Connection_pair :1023-1135 (donor rows are interior: no order dependence) */
  for (int e = 0; e < C->nrec; ++e) {
    const mose_bc_record *rc = &C->rec[e];
    if (rc->typ != 1) continue;
    const o_block *Bm = &C->blk[rc->b], *Bs = &C->blk[rc->bs];
    for (int g = 1; g <= GC; ++g) {
      int Ig = rc->i - guide[rc->f - 1][0] * g, Jg = rc->j - guide[rc->f - 1][1] * g, Kg = rc->k - guide[rc->f - 1][2] * g;
      int Is = rc->is + guide[rc->fs - 1][0] * (g - 1), Js = rc->js + guide[rc->fs - 1][1] * (g - 1);
      int Ks = rc->ks + guide[rc->fs - 1][2] * (g - 1);
      for (int s = 0; s < nprim; ++s) PV(P, s, ROW(Bm, Ig, Jg, Kg)) = PV(P, s, ROW(Bs, Is, Js, Ks));
    }
  }
}

/* the four transverse donor neighbours of a connection record (Ghost_Connection_pairs_B :1136-1176) */
static void donor_transverse(const mose_bc_record *rc, int q[4][3]) {
  int Is = rc->is, Js = rc->js, Ks = rc->ks, d11 = rc->d11, d12 = rc->d12, d21 = rc->d21, d22 = rc->d22;
  if (rc->fs <= 2) {
    int t[4][3] = {{Is, Js - d11, Ks - d12}, {Is, Js + d11, Ks + d12}, {Is, Js - d21, Ks - d22}, {Is, Js + d21, Ks + d22}};
    memcpy(q, t, sizeof t);
  } else if (rc->fs <= 4) {
    int t[4][3] = {{Is - d11, Js, Ks - d12}, {Is + d11, Js, Ks + d12}, {Is - d21, Js, Ks - d22}, {Is + d21, Js, Ks + d22}};
    memcpy(q, t, sizeof t);
  } else {
    int t[4][3] = {{Is - d11, Js - d12, Ks}, {Is + d11, Js + d12, Ks}, {Is - d21, Js - d22, Ks}, {Is + d21, Js + d22, Ks}};
    memcpy(q, t, sizeof t);
  }
}

/* ------------------------------------------------------------------------------------------------ */
/* boundary records: bc_record_body.inc + bc_record_apply.inc                                         */
/* ------------------------------------------------------------------------------------------------ */
static void record_geometry(const o_case *C, const mose_bc_record *rc, double nrm[3], double *area, double dlc4[4]) {
  const o_block *B = &C->blk[rc->b];
  int Im = rc->i, Jm = rc->j, Km = rc->k, Fm = rc->f;
  int dir = (Fm + 1) / 2, modfm = Fm % 2, gstep = 2 * modfm - 1;
  const double *fm;
  int64_t f, nf;
  if (dir == 1) { f = F1(B, Im - modfm, Jm, Km); fm = C->fm1; nf = C->NF1; }
  else if (dir == 2) { f = F2(B, Im, Jm - modfm, Km); fm = C->fm2; nf = C->NF2; }
  else { f = F3(B, Im, Jm, Km - modfm); fm = C->fm3; nf = C->NF3; }
  nrm[0] = fm[f]; nrm[1] = fm[nf + f]; nrm[2] = fm[2 * nf + f]; *area = fm[3 * nf + f];
  int o1 = modfm * (-gstep * GC) + (1 - modfm) * gstep, o2 = modfm * (-gstep), o3 = (1 - modfm) * (-gstep);
  int o4 = (1 - modfm) * (-gstep * GC) + modfm * gstep;
  int o[4] = {o1, o2, o3, o4};
  for (int q = 0; q < 4; ++q) {
    int ii = Im, jj = Jm, kk = Km;
    if (dir == 1) ii += o[q]; else if (dir == 2) jj += o[q]; else kk += o[q];
    dlc4[q] = C->dl[(int64_t)(dir - 1) * C->NC + ROW(B, ii, jj, kk)];
  }
}

/* BC_Extrapolation: Lib_BC.f90:1119-1163 */
static void bc_extrapolation(const o_case *C, const double *P, double *R, const mose_bc_record *rc, const double nrm[3],
                             double area) {
  const o_scheme *S = &C->sc;
  const int64_t NC = C->NC;
  int64_t r = ROW(&C->blk[rc->b], rc->i, rc->j, rc->k);
  double Prim[ONPRIM], Flux[ONPRIM];
  for (int s = 0; s < S->nprim; ++s) Prim[s] = PV(P, s, r);
  int nu = S->nu, nv = S->nv, nw = S->nw, np = S->np;
  double Un = Prim[nu] * nrm[0] + Prim[nv] * nrm[1] + Prim[nw] * nrm[2];
  double Fmass = 0.0;
  for (int s = 0; s < S->nsc; ++s) { Flux[s] = Prim[s] * Un * area; Fmass = Fmass + Flux[s]; }
  Flux[nu] = Fmass * Prim[nu] + Prim[np] * area * nrm[0];
  Flux[nv] = Fmass * Prim[nv] + Prim[np] * area * nrm[1];
  Flux[nw] = Fmass * Prim[nw] + Prim[np] * area * nrm[2];
  double vmag = sqrt(Prim[nu] * Prim[nu] + Prim[nv] * Prim[nv] + Prim[nw] * Prim[nw]);
  Flux[np] = Fmass * H0_dev(&C->tab, Prim[np], Prim, vmag);
  for (int s = np + 1; s < S->nprim; ++s) Flux[s] = Prim[s] * Un * area;
  double sgn = (double)(1 - 2 * (rc->f % 2));
  for (int s = 0; s < S->nprim; ++s) PV(R, s, r) = PV(R, s, r) + sgn * Flux[s];
}

/* BC_Connection_Eul_lean: Lib_BC.f90:1770-1864 */
static void bc_connection_eul(const o_case *C, const double *P, double *R, const mose_bc_record *rc, const double nrm[3],
                              double area, const double dlc4[4]) {
  const o_scheme *S = &C->sc;
  const int64_t NC = C->NC;
  const o_block *B = &C->blk[rc->b];
  int Im = rc->i, Jm = rc->j, Km = rc->k, Fm = rc->f;
  int dir = (Fm + 1) / 2, modfm = Fm % 2, modfm1 = 1 - modfm, modfm2 = 1 - 2 * modfm, gstep = 2 * modfm - 1;
  int o[4] = {modfm * (-gstep * GC) + modfm1 * gstep, modfm * (-gstep), modfm1 * (-gstep),
              modfm1 * (-gstep * GC) + modfm * gstep};
  int64_t r[4];
  for (int q = 0; q < 4; ++q) {
    int ii = Im, jj = Jm, kk = Km;
    if (dir == 1) ii += o[q]; else if (dir == 2) jj += o[q]; else kk += o[q];
    r[q] = ROW(B, ii, jj, kk);
  }
  double beta = 1.0;
  double Dl1 = dlc4[0] * 0.5, Dl2 = dlc4[1] * 0.5, Dl3 = dlc4[2] * 0.5, Dl4 = dlc4[3] * 0.5;
  double D12 = Dl1 + Dl2, D23 = Dl2 + Dl3, D34 = Dl3 + Dl4;
  double PL[ONPRIM], PR[ONPRIM];
  for (int s = 0; s < S->nprim; ++s) {
    double q1 = PV(P, s, r[0]), q2 = PV(P, s, r[1]), q3 = PV(P, s, r[2]), q4 = PV(P, s, r[3]);
    double d21 = (q2 - q1) / D12, d32 = (q3 - q2) / D23, d43 = (q4 - q3) / D34;
    PL[s] = q2 + beta * o_rlimiter(S->lim, d32, d21) * Dl2;
    PR[s] = q3 - beta * o_rlimiter(S->lim, d43, d32) * Dl3;
  }
  double rhoL, RL, rhoR, RR;
  co_rotot_Rtot(&C->tab, PL, &rhoL, &RL);
  double aL = f_ss_dev(&C->tab, PL, PL[S->np], rhoL, RL);
  co_rotot_Rtot(&C->tab, PR, &rhoR, &RR);
  double aR = f_ss_dev(&C->tab, PR, PR[S->np], rhoR, RR);
  double F[5];
  riemann_any(C, S->rs, 0, PL, PL[S->nu], PL[S->nv], PL[S->nw], PL[S->np], aL, rhoL, PR, PR[S->nu], PR[S->nv],
              PR[S->nw], PR[S->np], aR, rhoR, nrm[0], nrm[1], nrm[2], f_Rtot(&C->tab, PL), f_Rtot(&C->tab, PR), F);
  double su = copysign(0.5, F[0]), selL = 0.5 + su, selR = su - 0.5, Fmass = F[0] * area;
  int64_t rr = ROW(B, Im, Jm, Km);
  double m2 = (double)modfm2;
  for (int s = 0; s < S->nsc; ++s)
    PV(R, s, rr) = PV(R, s, rr) + m2 * Fmass * (selL * PL[s] / rhoL - selR * PR[s] / rhoR);
  PV(R, S->nu, rr) = PV(R, S->nu, rr) + m2 * F[1] * area;
  PV(R, S->nv, rr) = PV(R, S->nv, rr) + m2 * F[2] * area;
  PV(R, S->nw, rr) = PV(R, S->nw, rr) + m2 * F[3] * area;
  PV(R, S->np, rr) = PV(R, S->np, rr) + m2 * F[4] * area;
  for (int s = S->np + 1; s < S->nprim; ++s)
    PV(R, s, rr) = PV(R, s, rr) + m2 * Fmass * (selL * PL[s] / rhoL - selR * PR[s] / rhoR);
}

/* BC_Connection_Visc_lean: Lib_BC.f90:2113-2201 */
static void bc_connection_visc(const o_case *C, const double *P, double *R, const mose_bc_record *rc, const double nrm[3],
                               double area) {
  const o_scheme *S = &C->sc;
  const int64_t NC = C->NC;
  const o_block *B = &C->blk[rc->b], *Bs = &C->blk[rc->bs];
  int Im = rc->i, Jm = rc->j, Km = rc->k, Fm = rc->f;
  int dir = (Fm + 1) / 2, modfm2 = 1 - 2 * (Fm % 2), gstep = 2 * (Fm % 2) - 1;
  int ia = dir - 1, ib, ic;
  int cb[2][3] = {{Im, Jm, Km}, {Im, Jm, Km}}, cc[2][3] = {{Im, Jm, Km}, {Im, Jm, Km}};   /* [0] = +, [1] = - */
  if (dir == 1) { ib = 1; ic = 2; cb[0][1] = Jm + 1; cb[1][1] = Jm - 1; cc[0][2] = Km + 1; cc[1][2] = Km - 1; }
  else if (dir == 2) { ib = 0; ic = 2; cb[0][0] = Im + 1; cb[1][0] = Im - 1; cc[0][2] = Km + 1; cc[1][2] = Km - 1; }
  else { ib = 0; ic = 1; cb[0][0] = Im + 1; cb[1][0] = Im - 1; cc[0][1] = Jm + 1; cc[1][1] = Jm - 1; }
  int Ig = Im, Jg = Jm, Kg = Km;
  if (dir == 1) Ig = Im - gstep; else if (dir == 2) Jg = Jm - gstep; else Kg = Km - gstep;
  int64_t rrec = ROW(B, Im, Jm, Km), rgho = ROW(B, Ig, Jg, Kg);
  double Mm[3][3];
  for (int s = 0; s < 3; ++s)
    for (int a = 0; a < 3; ++a)
      Mm[a][s] = 0.5 * (C->m[(int64_t)(a + 3 * s) * NC + rrec] + C->m[(int64_t)(a + 3 * s) * NC + rgho]);
  int tq[4][3];
  donor_transverse(rc, tq);
  int64_t cr[10];
  cr[0] = rrec;
  cr[1] = rgho;                       /* Pg(:,1) = the first ghost layer (Ghost_Connection_pair stores the copy) */
  cr[2] = ROW(B, cb[0][0], cb[0][1], cb[0][2]);
  cr[3] = ROW(B, cb[1][0], cb[1][1], cb[1][2]);
  cr[4] = ROW(B, cc[0][0], cc[0][1], cc[0][2]);
  cr[5] = ROW(B, cc[1][0], cc[1][1], cc[1][2]);
  for (int q = 0; q < 4; ++q) cr[6 + q] = ROW(Bs, tq[q][0], tq[q][1], tq[q][2]);
  double rh[10], rg[10];
  for (int q = 0; q < 10; ++q) {
    double y[ONSMAX];
    for (int s = 0; s < S->nsc; ++s) y[s] = PV(P, s, cr[q]);
    co_rotot_Rtot(&C->tab, y, &rh[q], &rg[q]);
  }
  double G2l[ONPRIM][3], G2g[ONPRIM][3], Pavg[ONPRIM];
  for (int s = 0; s < S->nprim; ++s) {
    double pv[10], vv[10];
    for (int q = 0; q < 10; ++q) pv[q] = PV(P, s, cr[q]);
    if (s >= S->nu && s <= S->nw) for (int q = 0; q < 10; ++q) vv[q] = pv[q];
    else if (s == S->np) for (int q = 0; q < 10; ++q) vv[q] = pv[q] / (rh[q] * rg[q]);
    else for (int q = 0; q < 10; ++q) vv[q] = pv[q] / rh[q];
    double gl[3], gg[3];
    gl[ia] = (vv[1] - vv[0]) * (double)modfm2;
    gl[ib] = (vv[2] - vv[3]) * 0.5;
    gl[ic] = (vv[4] - vv[5]) * 0.5;
    gg[ia] = (vv[0] - vv[1]) * (double)(-modfm2);
    gg[ib] = (vv[7] - vv[6]) * 0.5;
    gg[ic] = (vv[9] - vv[8]) * 0.5;
    for (int kk = 0; kk < 3; ++kk) {
      G2l[s][kk] = gl[0] * Mm[0][kk] + gl[1] * Mm[1][kk] + gl[2] * Mm[2][kk];
      G2g[s][kk] = gg[0] * Mm[0][kk] + gg[1] * Mm[1][kk] + gg[2] * Mm[2][kk];
    }
    Pavg[s] = 0.5 * (pv[0] + pv[1]);
  }
  fthermo_t ft;
  face_thermo(C, Pavg, &ft);
  double Fl[ONPRIM], Fg[ONPRIM];
  diff_flux_st(C, Pavg, G2l, area, nrm, &ft, Fl);
  diff_flux_st(C, Pavg, G2g, area, nrm, &ft, Fg);
  for (int s = 0; s < S->nprim; ++s) PV(R, s, rrec) = PV(R, s, rrec) - 0.5 * (double)modfm2 * (Fl[s] + Fg[s]);
}

/* the two cells of a record along its face direction: the boundary cell and the first interior one */
static void record_cells(const o_case *C, const mose_bc_record *rc, int64_t *rb, int64_t *ri) {
  const o_block *B = &C->blk[rc->b];
  int dir = (rc->f + 1) / 2, gstep = 2 * (rc->f % 2) - 1;
  int ii = rc->i, jj = rc->j, kk = rc->k;
  if (dir == 1) ii += gstep; else if (dir == 2) jj += gstep; else kk += gstep;
  *rb = ROW(B, rc->i, rc->j, rc->k);
  *ri = ROW(B, ii, jj, kk);
}

/* Bound_Dl = dl_f(dir, record cell) / 2, Int_Dl = dl_f(dir, interior neighbour) / 2 (bc_record_body.inc:25-26) */
static void record_half_lengths(const o_case *C, const mose_bc_record *rc, double *bdl, double *idl) {
  int64_t rb, ri;
  record_cells(C, rc, &rb, &ri);
  int d = (rc->f + 1) / 2 - 1;
  *bdl = 0.5 * C->dl[(int64_t)d * C->NC + rb];
  *idl = 0.5 * C->dl[(int64_t)d * C->NC + ri];
}

/* BC_Symmetry_Eul: Lib_BC.f90:640-733 (the wall pressure flux of types 3, 5, 6 and the closed inlet of type 4;
 * momentum only).  Bound_Dl / Int_Dl = half the lengths of the boundary and interior cells (bc_record_body.inc:25-26). */
static void bc_symmetry_eul(const o_case *C, const double *P, double *R, const mose_bc_record *rc, const double nrm[3],
                            double area, double Bound_Dl, double Int_Dl) {
  const o_scheme *S = &C->sc;
  const int64_t NC = C->NC;
  int nu = S->nu, nv = S->nv, nw = S->nw, np = S->np;
  int modfm = rc->f % 2, modfm1 = 1 - modfm, modfm2 = 1 - 2 * modfm;
  int64_t rb, ri;
  record_cells(C, rc, &rb, &ri);
  double Bd[ONPRIM], In[ONPRIM], Gh[ONPRIM], P1[ONPRIM], P3[ONPRIM], Fa[ONPRIM], Sl[ONPRIM];
  for (int s = 0; s < S->nprim; ++s) { Bd[s] = PV(P, s, rb); In[s] = PV(P, s, ri); }
  double Un = Bd[nu] * nrm[0] + Bd[nv] * nrm[1] + Bd[nw] * nrm[2];
  for (int s = 0; s < S->nprim; ++s) Gh[s] = Bd[s];
  Gh[nu] = Bd[nu] - 2.0 * Un * nrm[0];
  Gh[nv] = Bd[nv] - 2.0 * Un * nrm[1];
  Gh[nw] = Bd[nw] - 2.0 * Un * nrm[2];
  for (int s = 0; s < S->nprim; ++s) {
    P1[s] = (double)modfm * Gh[s] + (double)modfm1 * In[s];
    P3[s] = (double)modfm * In[s] + (double)modfm1 * Gh[s];
  }
  double D12 = 2.0 * Bound_Dl * (double)modfm + (Bound_Dl + Int_Dl) * (double)modfm1;
  double D23 = (Bound_Dl + Int_Dl) * (double)modfm + 2.0 * Bound_Dl * (double)modfm1;
  for (int s = 0; s <= np; ++s) {
    double s32 = (P3[s] - Bd[s]) / D23, s21 = (Bd[s] - P1[s]) / D12;
    Sl[s] = o_rlimiter(S->lim, s32, s21);
  }
  double Dist = Bound_Dl * (double)modfm2;
  for (int s = 0; s <= np; ++s) Fa[s] = Bd[s] + Sl[s] * Dist;
  double rotot, Rtot;
  co_rotot_Rtot(&C->tab, Fa, &rotot, &Rtot);
  double Sound = f_ss_dev(&C->tab, Fa, Fa[np], rotot, Rtot);
  Un = Fa[nu] * nrm[0] + Fa[nv] * nrm[1] + Fa[nw] * nrm[2];
  double Vt[3] = {Fa[nu] - Un * nrm[0], Fa[nv] - Un * nrm[1], Fa[nw] - Un * nrm[2]};
  double Un1 = Un * (double)modfm2;
  double U1[3], U4[3];
  for (int q = 0; q < 3; ++q) { U1[q] = Un1 * nrm[q] + Vt[q]; U4[q] = -Un1 * nrm[q] + Vt[q]; }
  double F[5], Rg = f_Rtot(&C->tab, Fa);
  riemann_any(C, S->rs, 0, Fa, U1[0], U1[1], U1[2], Fa[np], Sound, rotot, Fa, U4[0], U4[1], U4[2], Fa[np], Sound, rotot,
              nrm[0], nrm[1], nrm[2], Rg, Rg, F);
  PV(R, nu, rb) = PV(R, nu, rb) + (double)modfm2 * F[1] * area;
  PV(R, nv, rb) = PV(R, nv, rb) + (double)modfm2 * F[2] * area;
  PV(R, nw, rb) = PV(R, nw, rb) + (double)modfm2 * F[3] * area;
}

/* Stress_Vector_Std: Mod_Fluid.f90:20-44 */
static void stress_std(const double G[3][3], const double n[3], double mi_l, double mi_t, double St[3]) {
  double mi = mi_l + mi_t, lam = NTT * mi;
  double Div = G[0][0] + G[1][1] + G[2][2], lamDiv = lam * Div;
  double Txx = lamDiv + 2.0 * mi * G[0][0], Tyy = lamDiv + 2.0 * mi * G[1][1], Tzz = lamDiv + 2.0 * mi * G[2][2];
  double Txy = mi * (G[0][1] + G[1][0]), Txz = mi * (G[0][2] + G[2][0]), Tyz = mi * (G[1][2] + G[2][1]);
  St[0] = Txx * n[0] + Txy * n[1] + Txz * n[2];
  St[1] = Txy * n[0] + Tyy * n[1] + Tyz * n[2];
  St[2] = Txz * n[0] + Tyz * n[1] + Tzz * n[2];
}

/* BC_Wall_Heat_dev (type 5, adiabatic no-slip: Lib_BC.f90:879-982) and BC_Wall_Temperature_dev (type 6, isothermal:
 * Lib_BC.f90:997-1115), with the Spalart-Allmaras wall flux of nu~ (nu~_wall = 0). */
static void bc_wall_visc(const o_case *C, const double *P, double *R, const mose_bc_record *rc, const double nrm[3],
                         double area) {
  const o_scheme *S = &C->sc;
  const o_tables *t = &C->tab;
  const int64_t NC = C->NC;
  int nu = S->nu, np = S->np, nsc = S->nsc;
  int dir = (rc->f + 1) / 2, modfm = rc->f % 2, modfm2 = 1 - 2 * modfm, modfm3 = 2 * (2 * modfm - 1);
  int64_t rb = ROW(&C->blk[rc->b], rc->i, rc->j, rc->k);
  double Pb[ONPRIM], M[3][3];
  for (int s = 0; s < S->nprim; ++s) Pb[s] = PV(P, s, rb);
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) M[a][c] = C->m[(int64_t)(a + 3 * c) * NC + rb];
  double rho, Rgas;
  co_rotot_Rtot(t, Pb, &rho, &Rgas);
  double T = Pb[np] / (rho * Rgas), mil, kl, Tw = rc->tw;
  if (rc->typ == 5) {
    int Ti = fint(T);
    wilke(t, Pb, rho, Ti, T - Ti, &mil, &kl);
  } else {
    double rho_w = Pb[np] / (Rgas * Tw), Pw[ONSMAX];
    for (int s = 0; s < nsc; ++s) Pw[s] = Pb[s] * rho_w / rho;
    int Ti = fint(Tw);
    double Td = Tw - Ti;
    wilke(t, Pw, rho_w, Ti, Td, &mil, &kl);
    if (S->prl > 0.0) kl = mil * f_cp_dev(t, Pw, Ti, Td, rho_w) / S->prl;
  }
  double Gr[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}}, G2[3][3];
  for (int s = 0; s < 3; ++s) Gr[s][dir - 1] = Pb[nu + s] * (double)modfm3;
  for (int kk = 0; kk < 3; ++kk)
    for (int ii = 0; ii < 3; ++ii) G2[ii][kk] = Gr[ii][0] * M[0][kk] + Gr[ii][1] * M[1][kk] + Gr[ii][2] * M[2][kk];
  double St[3];
  stress_std(G2, nrm, mil, 0.0, St);
  for (int q = 0; q < 3; ++q) PV(R, nu + q, rb) = PV(R, nu + q, rb) - (double)modfm2 * area * St[q];
  if (rc->typ == 6) {
    double Gd = (T - Tw) * (double)modfm3, Gn[3];
    for (int kk = 0; kk < 3; ++kk) Gn[kk] = Gd * M[dir - 1][kk];
    double Fnp = area * kl * (Gn[0] * nrm[0] + Gn[1] * nrm[1] + Gn[2] * nrm[2]);
    PV(R, np, rb) = PV(R, np, rb) - (double)modfm2 * Fnp;
  }
  if (S->rans == 2) {
    double gd = Pb[S->nt] / rho * (double)modfm3, g2[3];
    for (int kk = 0; kk < 3; ++kk) g2[kk] = gd * M[dir - 1][kk];
    double Fr = area * mil * (g2[0] * nrm[0] + g2[1] * nrm[1] + g2[2] * nrm[2]) / SA_SIGMA;
    PV(R, S->nt, rb) = PV(R, S->nt, rb) - (double)modfm2 * Fr;
  }
}

/* BC_Symmetry_Visc: Lib_BC.f90:741-861.  Velocity gradient from the cell and its mirror image along the normal, and
 * along each transverse direction from the two neighbours and their mirror images (the Fm cases :783-802) */
static void bc_sym_visc(const o_case *C, const double *P, double *R, const mose_bc_record *rc, const double nrm[3],
                        double area) {
  const o_scheme *S = &C->sc;
  const o_tables *t = &C->tab;
  const int64_t NC = C->NC;
  const o_block *B = &C->blk[rc->b];
  int nu = S->nu, np = S->np;
  int Fm = rc->f, modfm2 = 1 - 2 * (Fm % 2), Im = rc->i, Jm = rc->j, Km = rc->k, a, b, c;
  int64_t rb = ROW(B, Im, Jm, Km), n3, n4, n7, n8;
  if (Fm <= 2) {
    n3 = ROW(B, Im, Jm - 1, Km); n4 = ROW(B, Im, Jm + 1, Km); n7 = ROW(B, Im, Jm, Km - 1); n8 = ROW(B, Im, Jm, Km + 1);
    a = 0; b = 1; c = 2;
  } else if (Fm <= 4) {
    n3 = ROW(B, Im - 1, Jm, Km); n4 = ROW(B, Im + 1, Jm, Km); n7 = ROW(B, Im, Jm, Km - 1); n8 = ROW(B, Im, Jm, Km + 1);
    a = 1; b = 0; c = 2;
  } else {
    n3 = ROW(B, Im - 1, Jm, Km); n4 = ROW(B, Im + 1, Jm, Km); n7 = ROW(B, Im, Jm - 1, Km); n8 = ROW(B, Im, Jm + 1, Km);
    a = 2; b = 0; c = 1;
  }
  double Pb[ONPRIM], M[3][3];
  for (int s = 0; s < S->nprim; ++s) Pb[s] = PV(P, s, rb);
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 3; ++y) M[x][y] = C->m[(int64_t)(x + 3 * y) * NC + rb];
  double rho, Rgas;
  co_rotot_Rtot(t, Pb, &rho, &Rgas);
  double T = Pb[np] / (rho * Rgas), mil, kl;
  int Ti = fint(T);
  wilke(t, Pb, rho, Ti, T - Ti, &mil, &kl);
  double V1[3], V2[3], V3[3], V4[3], V5[3], V6[3], V7[3], V8[3], V9[3], V10[3];
  for (int q = 0; q < 3; ++q) {
    V1[q] = Pb[nu + q]; V3[q] = PV(P, nu + q, n3); V4[q] = PV(P, nu + q, n4); V7[q] = PV(P, nu + q, n7);
    V8[q] = PV(P, nu + q, n8);
  }
  double d1 = V1[0] * nrm[0] + V1[1] * nrm[1] + V1[2] * nrm[2];
  double d3 = V3[0] * nrm[0] + V3[1] * nrm[1] + V3[2] * nrm[2];
  double d4 = V4[0] * nrm[0] + V4[1] * nrm[1] + V4[2] * nrm[2];
  double d7 = V7[0] * nrm[0] + V7[1] * nrm[1] + V7[2] * nrm[2];
  double d8 = V8[0] * nrm[0] + V8[1] * nrm[1] + V8[2] * nrm[2];
  for (int q = 0; q < 3; ++q) {
    V2[q] = V1[q] - 2.0 * d1 * nrm[q];
    V5[q] = V3[q] - 2.0 * d3 * nrm[q];
    V6[q] = V4[q] - 2.0 * d4 * nrm[q];
    V9[q] = V7[q] - 2.0 * d7 * nrm[q];
    V10[q] = V8[q] - 2.0 * d8 * nrm[q];
  }
  double Gr[3][3], G2[3][3];
  for (int q = 0; q < 3; ++q) {
    Gr[q][a] = (V2[q] - V1[q]) * (double)modfm2;
    Gr[q][b] = (V4[q] - V3[q] + V6[q] - V5[q]) * 0.25;
    Gr[q][c] = (V8[q] - V7[q] + V10[q] - V9[q]) * 0.25;
  }
  for (int kk = 0; kk < 3; ++kk)
    for (int ii = 0; ii < 3; ++ii) G2[ii][kk] = Gr[ii][0] * M[0][kk] + Gr[ii][1] * M[1][kk] + Gr[ii][2] * M[2][kk];
  double mie = 0.0;
  if (S->rans == 2) {   /* Spalart_Eddy_Viscosity (Mod_Spalart.f90:382-398): fv1(chi) rho nu~ */
    double chi = Pb[S->nt] / mil, c3 = chi * chi * chi;
    mie = c3 / (c3 + SA_CV1 * SA_CV1 * SA_CV1) * Pb[S->nt];
  }
  double St[3], Vel[3];
  stress_std(G2, nrm, mil, mie, St);
  for (int q = 0; q < 3; ++q) Vel[q] = 0.5 * (V1[q] + V2[q]);
  for (int q = 0; q < 3; ++q) PV(R, nu + q, rb) = PV(R, nu + q, rb) - (double)modfm2 * area * St[q];
  PV(R, np, rb) = PV(R, np, rb) - (double)modfm2 * area * (Vel[0] * St[0] + Vel[1] * St[1] + Vel[2] * St[2]);
}

/* the face flux of the far-field closures from a face state (the common tails of BC_Outflow, Lib_BC.f90:1218-1242,
 * and BC_Subsonic_Inflow, :1489-1511; `sub` selects the operand order of the latter) */
static void farfield_tail(const o_case *C, const double *Fa, double Face_rho, double Face_Rgas, double Face_Un,
                          const double nrm[3], double Area, int sub, double *Flux) {
  const o_scheme *S = &C->sc;
  const o_tables *t = &C->tab;
  int nu = S->nu, nv = S->nv, nw = S->nw, np = S->np;
  double He = 0.0, Face_T = Fa[np] / (Face_Rgas * Face_rho);
  int Ti = (int)Face_T;
  double Td = Face_T - Ti;
  for (int s = 0; s < S->nprim; ++s) Flux[s] = 0.0;
  for (int s = 0; s < S->nsc; ++s) {
    He = He + Fa[s] / Face_rho * tabT(t->h, t, s, Ti, Ti + 1, Td);
    Flux[s] = sub ? Fa[s] * Face_Un * Area : Fa[s] * Area * Face_Un;
  }
  double Fmass = sub ? Face_rho * Face_Un * Area : Face_rho * Area * Face_Un;
  if (sub) {
    Flux[nu] = Fmass * Fa[nu] + Fa[np] * Area * nrm[0];
    Flux[nv] = Fmass * Fa[nv] + Fa[np] * Area * nrm[1];
    Flux[nw] = Fmass * Fa[nw] + Fa[np] * Area * nrm[2];
  } else {
    Flux[nu] = Fa[nu] * Fmass + Fa[np] * Area * nrm[0];
    Flux[nv] = Fa[nv] * Fmass + Fa[np] * Area * nrm[1];
    Flux[nw] = Fa[nw] * Fmass + Fa[np] * Area * nrm[2];
  }
  double vt2 = Fa[nu] * Fa[nu] + Fa[nv] * Fa[nv] + Fa[nw] * Fa[nw];
  Flux[np] = Fmass * (0.5 * vt2 + He);
  for (int s = np + 1; s < S->nprim; ++s) Flux[s] = Fa[s] / Face_rho * Fmass;
}

/* BC_Outflow: Lib_BC.f90:1168-1243 */
static void bc_outflow(const o_case *C, const double *Pr, double rho, double Rgas, double Sound, double Un, double Gamma,
                       double Riem, const double nrm[3], double Area, double p_Amb, double BC_Sign, double *Flux) {
  const o_scheme *S = &C->sc;
  int nu = S->nu, nv = S->nv, nw = S->nw, np = S->np;
  double Vt[3] = {Pr[nu] - Un * nrm[0], Pr[nv] - Un * nrm[1], Pr[nw] - Un * nrm[2]};
  double Mach_N = fabs(Un / Sound), p_exit = p_Amb, Fa[ONPRIM], Face_Un, Face_rho, Face_Rgas;
  if (Mach_N < 1.0 && p_Amb > 0.0) {
    double delta = 0.5 * (Gamma - 1.0);
    double denom = 1.0 / pow(1.0 + delta, Gamma / (Gamma - 1.0));
    double pstar = Pr[np] * denom * pow(1.0 + delta * (Mach_N * Mach_N), Gamma / (Gamma - 1.0));
    if (pstar >= p_Amb) p_exit = pstar;
    double Isen = (Gamma - 1.0) / (2.0 * Gamma);
    double Face_Sound = Sound * pow(p_exit / Pr[np], Isen);
    Face_Un = (Riem - Face_Sound) / (BC_Sign * 0.5 * (Gamma - 1.0));
    Fa[np] = p_exit;
    Face_rho = Gamma * Fa[np] / (Face_Sound * Face_Sound);
    for (int s = 0; s < S->nsc; ++s) Fa[s] = Pr[s] / rho * Face_rho;
    Fa[nu] = Vt[0] + Face_Un * nrm[0];
    Fa[nv] = Vt[1] + Face_Un * nrm[1];
    Fa[nw] = Vt[2] + Face_Un * nrm[2];
    for (int s = np + 1; s < S->nprim; ++s) Fa[s] = Pr[s] / rho * Face_rho;
    Face_Rgas = Rgas;
  } else {
    for (int s = 0; s < S->nprim; ++s) Fa[s] = Pr[s];
    Face_Un = Un; Face_rho = rho; Face_Rgas = Rgas;
  }
  farfield_tail(C, Fa, Face_rho, Face_Rgas, Face_Un, nrm, Area, 0, Flux);
}

/* BC_Supersonic_Inflow: Lib_BC.f90:1247-1330 */
static void bc_supersonic_inflow(const o_case *C, const double *P2, const double *In, int modfm, const mose_farfield *ff,
                                 const double nrm[3], double Area, double *Flux) {
  const o_scheme *S = &C->sc;
  const o_tables *t = &C->tab;
  int nu = S->nu, nv = S->nv, nw = S->nw, np = S->np, nsc = S->nsc, modfm1 = 1 - modfm, modfm2 = 1 - 2 * modfm;
  double flow_Rgas = f_Rtot(t, P2);
  double vmag2 = P2[nu] * P2[nu] + P2[nv] * P2[nv] + P2[nw] * P2[nw];
  double rsum = 0.0;
  for (int s = 0; s < nsc; ++s) rsum = rsum + P2[s];
  double flow_Sound = f_ss_dev(t, P2, P2[np], rsum, flow_Rgas);
  double flow_Mach = sqrt(vmag2) / flow_Sound;
  double rel = ff->rel_fac;
  double Mloc = rel * ff->mach + (1.0 - rel) * flow_Mach;
  double p0loc = rel * ff->p0 + (1.0 - rel) * P2[np];
  double Sup[ONPRIM], Sup_T = ff->t0, Sup_Rgas = 0.0;
  Sup[np] = p0loc;
  for (int s = 0; s < nsc; ++s) Sup_Rgas = Sup_Rgas + ff->ci[s] * t->Ri[s];
  double Sup_rho = Sup[np] / (Sup_Rgas * Sup_T);
  for (int s = 0; s < nsc; ++s) Sup[s] = Sup_rho * ff->ci[s];
  double Sup_Sound = f_ss_dev(t, Sup, Sup[np], Sup_rho, Sup_Rgas);
  Sup[nu] = Mloc * Sup_Sound * cos(ff->alpha) * cos(ff->beta);
  Sup[nv] = Mloc * Sup_Sound * sin(ff->alpha) * cos(ff->beta);
  Sup[nw] = Mloc * Sup_Sound * sin(ff->beta);
  for (int s = 0; s < S->nrans; ++s) Sup[S->nt + s] = Sup_rho * ff->rans[s];
  double L[ONPRIM], Rr[ONPRIM];
  for (int s = 0; s < S->nprim; ++s) {
    double p1 = (double)modfm * Sup[s] + (double)modfm1 * In[s];
    double p3 = (double)modfm * In[s] + (double)modfm1 * Sup[s];
    double Slope = o_rlimiter(S->lim, p3 - P2[s], P2[s] - p1);
    double Fa = P2[s] + 0.5 * Slope * (double)modfm2;
    L[s] = (double)modfm * Sup[s] + (double)modfm1 * Fa;
    Rr[s] = (double)modfm * Fa + (double)modfm1 * Sup[s];
  }
  double rho_L, Rg_L, rho_R, Rg_R;
  co_rotot_Rtot(t, L, &rho_L, &Rg_L);
  double aL = f_ss_dev(t, L, L[np], rho_L, Rg_L);
  co_rotot_Rtot(t, Rr, &rho_R, &Rg_R);
  double aR = f_ss_dev(t, Rr, Rr[np], rho_R, Rg_R);
  double F[5];
  riemann_any(C, S->rs, 0, L, L[nu], L[nv], L[nw], L[np], aL, rho_L, Rr, Rr[nu], Rr[nv], Rr[nw], Rr[np], aR, rho_R,
              nrm[0], nrm[1], nrm[2], f_Rtot(t, L), f_Rtot(t, Rr), F);
  double su = copysign(1.0, F[0]), SelL = (1.0 + su) / 2.0, SelR = (su - 1.0) / 2.0, Fmass = F[0] * Area;
  for (int s = 0; s < S->nprim; ++s) Flux[s] = 0.0;
  for (int s = 0; s < nsc; ++s) Flux[s] = Fmass * (SelL * L[s] / rho_L - SelR * Rr[s] / rho_R);
  Flux[nu] = F[1] * Area; Flux[nv] = F[2] * Area; Flux[nw] = F[3] * Area; Flux[np] = F[4] * Area;
  for (int s = np + 1; s < S->nprim; ++s) Flux[s] = Fmass * (SelL * L[s] / rho_L - SelR * Rr[s] / rho_R);
}

/* BC_Subsonic_Inflow: Lib_BC.f90:1335-1512.  Returns 1 when the face must close (the caller applies the symmetry). */
static int bc_subsonic_inflow(const o_case *C, const double *Pr, double Un, double Sound, double Gamma, double Riem,
                              const mose_farfield *ff, double BC_Sign, const double tv[3], const double nrm[3],
                              double Area, double *Flux) {
  const o_scheme *S = &C->sc;
  const o_tables *t = &C->tab;
  int nu = S->nu, nv = S->nv, nw = S->nw, np = S->np, nsc = S->nsc;
  double bv[3] = {nrm[1] * tv[2] - nrm[2] * tv[1], nrm[2] * tv[0] - nrm[0] * tv[2], nrm[0] * tv[1] - nrm[1] * tv[0]};
  double bm = sqrt(bv[0] * bv[0] + bv[1] * bv[1] + bv[2] * bv[2]);
  bv[0] = bv[0] / bm; bv[1] = bv[1] / bm; bv[2] = bv[2] / bm;
  double alpha, beta;
  if (ff->alpha == 0.0 && ff->beta == 0.0) {
    alpha = atan(nrm[1] / (nrm[0] + 1e-20));
    beta = atan(nrm[2] / (nrm[0] + 1e-20));
  } else {
    alpha = ff->alpha; beta = ff->beta;
  }
  double ta = tan(alpha), tb = tan(beta);
  double XA = nrm[1] - nrm[0] * ta, XB = tv[1] - tv[0] * ta, XC = bv[1] - bv[0] * ta;
  double XD = nrm[2] - nrm[0] * tb, XE = tv[2] - tv[0] * tb, XF = bv[2] - bv[0] * tb;
  double check = (XC * XE - XF * XB);
  double E1 = (XF * XA - XC * XD) / check, E2 = (XD * XB - XE * XA) / check;
  double Fa[ONPRIM], Face_rho, Face_Rgas, Face_Un;
  for (int s = 0; s < S->nprim; ++s) Fa[s] = 0.0;
  if (ff->mach >= 0.0) {
    double Discr = -1.0, p0t = ff->p0, Sound0 = 0.0, Coef_D = 0.0, YA = 0.0, YB = 0.0;
    int it = 0;
    Face_Rgas = 0.0;
    while (Discr < 0.0) {
      it = it + 1;
      Face_Rgas = 0.0;
      for (int s = 0; s < nsc; ++s) Face_Rgas = Face_Rgas + ff->ci[s] * t->Ri[s];
      Sound0 = sqrt(Gamma * Face_Rgas * ff->t0);
      Coef_D = (1.0 + E1 * E1 + E2 * E2) * (Gamma - 1.0) / 2.0;
      double Coef_C = pow(p0t / Pr[np], (Gamma - 1.0) / Gamma);
      double sr = Sound / Sound0;
      Coef_C = Coef_C * (sr * sr);
      double gh = (Gamma - 1.0) / 2.0;
      YA = gh * gh + Coef_C * Coef_D;
      YB = BC_Sign * Riem * ((Gamma - 1.0) / 2.0);
      double YC = Riem * Riem - Coef_C * (Sound0 * Sound0);
      Discr = YB * YB - YA * YC;
      if (Discr < 0.0 && fabs(Discr) < 1e-2) Discr = 0.0;
      if (Discr < 0.0 && fabs(Discr) > 1e-2) {
        if (ff->rel_fac == 1.0) { Discr = 0.0; break; }
        p0t = ff->rel_fac * p0t + (1.0 - ff->rel_fac) * Pr[np];
      }
      if (it > 1000) { Discr = 0.0; break; }
    }
    double Un_Sign = copysign(1.0, Un);
    double Un3 = (YB + Un_Sign * sqrt(Discr)) / YA;
    double Sound3 = sqrt(Sound0 * Sound0 - Coef_D * (Un3 * Un3));
    double p3 = p0t * pow(Sound3 / Sound0, 2.0 * Gamma / (Gamma - 1.0));
    double Ut3 = Un3 * E1, Ub3 = Un3 * E2;
    Fa[np] = p3;
    Face_rho = Gamma * Fa[np] / (Sound3 * Sound3);
    for (int s = 0; s < nsc; ++s) Fa[s] = ff->ci[s] * Face_rho;
    Fa[nu] = Un3 * nrm[0] + Ut3 * tv[0] + Ub3 * bv[0];
    Fa[nv] = Un3 * nrm[1] + Ut3 * tv[1] + Ub3 * bv[1];
    Fa[nw] = Un3 * nrm[2] + Ut3 * tv[2] + Ub3 * bv[2];
    Face_Un = Un3;
  } else {
    double mdot = ff->p0, p3 = Pr[np], Rgas3 = 0.0, cp3 = 0.0;
    int Ti = (int)ff->t0;
    double Td = ff->t0 - Ti;
    for (int s = 0; s < nsc; ++s) {
      cp3 = cp3 + ff->ci[s] * tabT(t->cp, t, s, Ti, Ti + 1, Td);
      Rgas3 = Rgas3 + ff->ci[s] * t->Ri[s];
    }
    if (ff->mach != -15.0) {
      double T3;
      if (ff->mach == -10.0) {
        double q = mdot * Rgas3 / p3, q2 = p3 / (mdot * Rgas3);
        T3 = (-cp3 + sqrt(cp3 * cp3 + 2.0 * cp3 * ff->t0 * (q * q))) * (q2 * q2);
      } else {
        T3 = ff->t0;
      }
      double rho3 = p3 / (Rgas3 * T3), Un3 = mdot / rho3, Ut3 = Un3 * E1, Ub3 = Un3 * E2;
      Face_Rgas = Rgas3; Face_rho = rho3;
      for (int s = 0; s < nsc; ++s) Fa[s] = Face_rho * ff->ci[s];
      Fa[nu] = Un3 * nrm[0] + Ut3 * tv[0] + Ub3 * bv[0];
      Fa[nv] = Un3 * nrm[1] + Ut3 * tv[1] + Ub3 * bv[1];
      Fa[nw] = Un3 * nrm[2] + Ut3 * tv[2] + Ub3 * bv[2];
      Fa[np] = p3;
      Face_Un = Un3;
    } else {
      Face_Rgas = Rgas3;
      Face_rho = p3 / (Face_Rgas * ff->t0);
      for (int s = 0; s < nsc; ++s) Fa[s] = Face_rho * ff->ci[s];
      Fa[nu] = mdot / Face_rho; Fa[nv] = 0.0; Fa[nw] = Fa[nu] * 0.25; Fa[np] = p3;
      Face_Un = Fa[nu] * nrm[0] + Fa[nv] * nrm[1] + Fa[nw] * nrm[2];
    }
  }
  for (int s = 0; s < S->nrans; ++s) Fa[S->nt + s] = ff->rans[s] * Face_rho;
  farfield_tail(C, Fa, Face_rho, Face_Rgas, Face_Un, nrm, Area, 1, Flux);
  return 0;
}

/* BC_Inflow_Outflow: Lib_BC.f90:1519-1597.  Returns 1 when the face closes to a reflective wall. */
static int bc_inflow_outflow(const o_case *C, const double *P, double *R, const mose_bc_record *rc, const double nrm[3],
                             double Area) {
  const o_scheme *S = &C->sc;
  const o_tables *t = &C->tab;
  const int64_t NC = C->NC;
  int nu = S->nu, nv = S->nv, nw = S->nw, np = S->np;
  int modfm = rc->f % 2, modfm2 = 1 - 2 * modfm;
  const mose_farfield *ff = &C->ff[rc->ff];
  int64_t rb, ri;
  record_cells(C, rc, &rb, &ri);
  double Bd[ONPRIM], In[ONPRIM], Flux[ONPRIM];
  for (int s = 0; s < S->nprim; ++s) { Bd[s] = PV(P, s, rb); In[s] = PV(P, s, ri); }
  double BC_Sign = (double)modfm2;
  double Un = Bd[nu] * nrm[0] + Bd[nv] * nrm[1] + Bd[nw] * nrm[2];
  double rho, Rg;
  co_rotot_Rtot(t, Bd, &rho, &Rg);
  double Sound = f_ss_dev(t, Bd, Bd[np], rho, Rg);
  double T = Bd[np] / (rho * Rg);
  int Ti = (int)T;
  double cp = f_cp_dev(t, Bd, Ti, T - Ti, rho);
  double Gamma = cp / (cp - Rg);
  double Riem = Sound + BC_Sign * 0.5 * (Gamma - 1.0) * Un;
  double InOut = BC_Sign * Un;
  if (ff->mach < 0.0) InOut = -1.0;
  if (InOut >= 0.0) {
    if (ff->pamb == -3.14 || ff->pamb == -2.71) return 1;
    bc_outflow(C, Bd, rho, Rg, Sound, Un, Gamma, Riem, nrm, Area, ff->pamb, BC_Sign, Flux);
  } else {
    if (ff->mach > 1.0 || ff->pamb == -2.71) {
      bc_supersonic_inflow(C, Bd, In, modfm, ff, nrm, Area, Flux);
    } else if (ff->mach < 1.0) {
      if (bc_subsonic_inflow(C, Bd, Un, Sound, Gamma, Riem, ff, BC_Sign, rc->tvec, nrm, Area, Flux)) return 1;
    } else {
      return 0;   /* Mach exactly 1: no closure applies and the reference adds nothing defined; rejected at set-up */
    }
  }
  for (int s = 0; s < S->nprim; ++s) PV(R, s, rb) = PV(R, s, rb) + (double)modfm2 * Flux[s];
  return 0;
}

/* ------------------------------------------------------------------------------------------------ */
/* Spalart-Allmaras: the wall ghost of Extrapolate_Wall_blk (Lib_Extrapolation.f90:128-190, eddy_dev 2) and the   */
/* source of SA_Source_Blk + sa_body.inc (Mod_Spalart.f90:105-151; constants :14-34)                             */
/* ------------------------------------------------------------------------------------------------ */
#define SA_CB1 0.1355
#define SA_CB2 0.622
#define SA_CW2 0.3
#define SA_CW3 2.0
#define SA_CK 0.41
#define SA_C2 0.7
#define SA_C3 0.9

static void extrapolate_wall(const o_case *C, double *P) {
  const o_scheme *S = &C->sc;
  const int64_t NC = C->NC;
  for (int e = 0; e < C->nrec; ++e) {
    const mose_bc_record *rc = &C->rec[e];
    if (rc->typ != 5 && rc->typ != 6) continue;
    const o_block *B = &C->blk[rc->b];
    int Ig = rc->i - guide[rc->f - 1][0], Jg = rc->j - guide[rc->f - 1][1], Kg = rc->k - guide[rc->f - 1][2];
    int64_t ci = ROW(B, rc->i, rc->j, rc->k), cg = ROW(B, Ig, Jg, Kg);
    for (int s = S->nu; s <= S->nw; ++s) PV(P, s, cg) = -PV(P, s, ci);
    for (int s = 0; s < S->nsc; ++s) PV(P, s, cg) = PV(P, s, ci);
    PV(P, S->np, cg) = PV(P, S->np, ci);
    PV(P, S->nt, cg) = -PV(P, S->nt, ci);
  }
}

static void sa_source(const o_case *C, const double *P, double *R, const double *rc, const double *mc, double dt) {
  const o_scheme *S = &C->sc;
  const int64_t NC = C->NC;
  const double cw1 = SA_CB1 / (SA_CK * SA_CK) + (1.0 + SA_CB2) / SA_SIGMA, cK_sq = SA_CK * SA_CK;
  const double cv1_3 = SA_CV1 * SA_CV1 * SA_CV1, cw3_6 = 64.0;
  int nt = S->nt, nu = S->nu, nsc = S->nsc;
  for (int b = 0; b < C->nb; ++b) {
    const o_block *B = &C->blk[b];
    int64_t st[3] = {1, B->pi, (int64_t)B->pi * B->pj};
#pragma omp parallel for collapse(2) schedule(static) num_threads(C->nthreads > 0 ? C->nthreads : 1)
    for (int k = 1; k <= B->n3; ++k)
      for (int j = 1; j <= B->n2; ++j)
        for (int i = 1; i <= B->n1; ++i) {
          int64_t c0 = ROW(B, i, j, k);
          double rho = rc[c0], mil = mc[c0];
          double nit = PV(P, nt, c0) / rho, chi = PV(P, nt, c0) / mil, dist = C->yn[c0];
          double M[3][3], Gt[3][3], Gv[3][3];
          for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) M[a][c] = C->m[(int64_t)(a + 3 * c) * NC + c0];
          for (int d = 0; d < 3; ++d)
            for (int a = 0; a < 3; ++a) Gt[a][d] = (PV(P, nu + a, c0 + st[d]) - PV(P, nu + a, c0 - st[d])) * 0.5;
          for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) Gv[a][c] = Gt[a][0] * M[0][c] + Gt[a][1] * M[1][c] + Gt[a][2] * M[2][c];
          double v1 = Gv[2][1] - Gv[1][2], v2 = Gv[0][2] - Gv[2][0], v3 = Gv[1][0] - Gv[0][1];
          double Om = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
          double chi3 = chi * chi * chi;
          double fv1 = chi3 / (chi3 + cv1_3);
          double fv2 = 1.0 - chi / (1.0 + chi * fv1);
          double Sbar = nit / (cK_sq * (dist * dist)) * fv2, Stilde;
          if (Sbar >= (-SA_C2 * Om)) Stilde = Om + Sbar;
          else Stilde = Om + Om * (SA_C2 * SA_C2 * Om + SA_C3 * Sbar) / ((SA_C3 - 2.0 * SA_C2) * Om - Sbar);
          if (Stilde == 0.0) Stilde = 1e-40;
          double Prod = SA_CB1 * Stilde * nit;
          double sum[6];
          for (int q = 0; q < 6; ++q) {
            int64_t r = c0 + (q % 2 ? st[q / 2] : -st[q / 2]);
            double ssum = 0.0;
            for (int s = 0; s < nsc; ++s) ssum = ssum + PV(P, s, r);
            sum[q] = ssum;
          }
          double Gn[3];
          for (int d = 0; d < 3; ++d)
            Gn[d] = (PV(P, nt, c0 + st[d]) / sum[2 * d + 1] - PV(P, nt, c0 - st[d]) / sum[2 * d]) * 0.5;
          double Gr[3];
          for (int c = 0; c < 3; ++c) Gr[c] = Gn[0] * M[0][c] + Gn[1] * M[1][c] + Gn[2] * M[2][c];
          double Gr2 = Gr[0] * Gr[0] + Gr[1] * Gr[1] + Gr[2] * Gr[2];
          double Diff = SA_CB2 / SA_SIGMA * Gr2;
          double r_ = nit / (Stilde * cK_sq * (dist * dist));
          r_ = min2(r_, 10.0);
          double r2 = r_ * r_, r6 = r2 * r2 * r2;
          double g = r_ + SA_CW2 * (r6 - r_);
          double g2 = g * g, g6 = g2 * g2 * g2;
          double fw = g * pow((1.0 + cw3_6) / (g6 + cw3_6), 1.0 / 6.0);
          double nd = nit / dist;
          double Dest = cw1 * fw * (nd * nd);
          double Src = rho * (Prod + Diff - Dest) * C->vol[c0];
          if (S->point_implicit) {
            double fnu = 1.0 / (1.0 + dt * 2.0 * cw1 * fw * nit / (dist * dist));
            Src = Src * fnu;
          }
          PV(R, nt, c0) = PV(R, nt, c0) - Src;
        }
  }
}

/* ------------------------------------------------------------------------------------------------ */
/* residual: Fluxes (conv + diff pass 1, pass 2 in the order of pass2_body.inc:1-61) then BC_Fluxes  */
/* in face order 1..6 (Mod_BC.f90:1042-1062 bc_launch_ranges)                                         */
/* ------------------------------------------------------------------------------------------------ */
void o_residual(const o_case *C, double *P, double *R, double dt) {
  const o_scheme *S = &C->sc;
  const int64_t NC = C->NC;
  int nprim = S->nprim;
  double *rc = (double *)malloc(sizeof(double) * NC * 4);
  double *Rc = rc + NC, *Tc = rc + 2 * NC, *mc = rc + 3 * NC;
  thermo_cache(C, P, rc, Rc, Tc, mc);
  for (int b = 0; b < C->nb; ++b) {
    const o_block *B = &C->blk[b];
    int n1 = B->n1, n2 = B->n2, n3 = B->n3;
    int64_t nfd[3] = {(int64_t)(n1 - 1) * n2 * n3, (int64_t)n1 * (n2 - 1) * n3, (int64_t)n1 * n2 * (n3 - 1)};
    double *fh[3], *dfh[3];
    for (int d = 0; d < 3; ++d) {
      int64_t nf = nfd[d] > 0 ? nfd[d] : 1;
      fh[d] = (double *)calloc((size_t)nf * nprim, sizeof(double));
      dfh[d] = (double *)calloc((size_t)nf * nprim, sizeof(double));
    }
    for (int d = 0; d < 3; ++d) {
      int hi1 = n1 - (d == 0), hi2 = n2 - (d == 1), hi3 = n3 - (d == 2);
      int64_t st = (d == 0) ? 1 : (d == 1) ? B->pi : (int64_t)B->pi * B->pj;
      const double *fm = d == 0 ? C->fm1 : d == 1 ? C->fm2 : C->fm3;
      int64_t NF = d == 0 ? C->NF1 : d == 1 ? C->NF2 : C->NF3;
#pragma omp parallel for collapse(2) schedule(static) num_threads(C->nthreads > 0 ? C->nthreads : 1)
      for (int k = 1; k <= hi3; ++k)
        for (int j = 1; j <= hi2; ++j)
          for (int i = 1; i <= hi1; ++i) {
            int64_t c0 = ROW(B, i, j, k);
            double pm1[ONPRIM], p0[ONPRIM], pp1[ONPRIM], pp2[ONPRIM], res0[ONPRIM], dlc[4], nrm[3];
            for (int s = 0; s < nprim; ++s) {
              pm1[s] = PV(P, s, c0 - st); p0[s] = PV(P, s, c0); pp1[s] = PV(P, s, c0 + st); pp2[s] = PV(P, s, c0 + 2 * st);
            }
            for (int q = 0; q < 4; ++q) dlc[q] = C->dl[(int64_t)d * NC + c0 + (q - 1) * st];
            int64_t fg = d == 0 ? F1(B, i, j, k) : d == 1 ? F2(B, i, j, k) : F3(B, i, j, k);
            nrm[0] = fm[fg]; nrm[1] = fm[NF + fg]; nrm[2] = fm[2 * NF + fg];
            double area = fm[3 * NF + fg];
            o_conv_face(C, pm1, p0, pp1, pp2, dlc, nrm, area, res0);
            int64_t f = (d == 0) ? (i - 1) + (int64_t)(j - 1) * (n1 - 1) + (int64_t)(k - 1) * (n1 - 1) * n2
                      : (d == 1) ? (i - 1) + (int64_t)(j - 1) * n1 + (int64_t)(k - 1) * n1 * (n2 - 1)
                                 : (i - 1) + (int64_t)(j - 1) * n1 + (int64_t)(k - 1) * n1 * n2;
            for (int s = 0; s < nprim; ++s) fh[d][(int64_t)s * nfd[d] + f] = res0[s];
            if (S->visc) {
              /* the 10-cell stencil of diff_cuf_i/j/k: c1,c2 across the face, then +-1 on axis ib at both cells,
               * then +-1 on axis ic at both cells (diff_cuf_kernels.inc:48-57,148-157,228-237) */
              int ax[3];
              int64_t sb, sc2;
              if (d == 0) { ax[0] = 0; ax[1] = 1; ax[2] = 2; sb = B->pi; sc2 = (int64_t)B->pi * B->pj; }
              else if (d == 1) { ax[0] = 1; ax[1] = 0; ax[2] = 2; sb = 1; sc2 = (int64_t)B->pi * B->pj; }
              else { ax[0] = 2; ax[1] = 0; ax[2] = 1; sb = 1; sc2 = B->pi; }
              int64_t c[10] = {c0, c0 + st, c0 - sb, c0 + sb, c0 + st - sb, c0 + st + sb,
                               c0 - sc2, c0 + sc2, c0 + st - sc2, c0 + st + sc2};
              diff_face(C, P, rc, Tc, c, ax, C->m + c0, C->m + c0 + st, NC, nrm, area, C->yn[c0], C->yn[c0 + st], res0);
              for (int s = 0; s < nprim; ++s) dfh[d][(int64_t)s * nfd[d] + f] = res0[s];
            }
          }
    }
    /* pass 2 (pass2_body.inc:1-61) */
#pragma omp parallel for collapse(2) schedule(static) num_threads(C->nthreads > 0 ? C->nthreads : 1)
    for (int k = 1; k <= n3; ++k)
      for (int j = 1; j <= n2; ++j)
        for (int i = 1; i <= n1; ++i) {
          int64_t c0 = ROW(B, i, j, k);
          double tmp[ONPRIM];
          for (int s = 0; s < nprim; ++s) tmp[s] = 0.0;
          int64_t fi_p = (i - 1) + (int64_t)(j - 1) * (n1 - 1) + (int64_t)(k - 1) * (n1 - 1) * n2;
          int64_t fj_p = (i - 1) + (int64_t)(j - 1) * n1 + (int64_t)(k - 1) * n1 * (n2 - 1);
          int64_t fk_p = (i - 1) + (int64_t)(j - 1) * n1 + (int64_t)(k - 1) * n1 * n2;
          for (int pass = 0; pass < (S->visc ? 2 : 1); ++pass) {
            double **buf = pass ? dfh : fh;
            double sg = pass ? -1.0 : 1.0;
            /* convective: tmp + right - left;  diffusive: tmp - right + left */
            if (i < n1) for (int s = 0; s < nprim; ++s) tmp[s] = pass ? tmp[s] - buf[0][s * nfd[0] + fi_p] : tmp[s] + buf[0][s * nfd[0] + fi_p];
            if (i > 1) for (int s = 0; s < nprim; ++s) tmp[s] = pass ? tmp[s] + buf[0][s * nfd[0] + fi_p - 1] : tmp[s] - buf[0][s * nfd[0] + fi_p - 1];
            if (j < n2) for (int s = 0; s < nprim; ++s) tmp[s] = pass ? tmp[s] - buf[1][s * nfd[1] + fj_p] : tmp[s] + buf[1][s * nfd[1] + fj_p];
            if (j > 1) for (int s = 0; s < nprim; ++s) tmp[s] = pass ? tmp[s] + buf[1][s * nfd[1] + fj_p - n1] : tmp[s] - buf[1][s * nfd[1] + fj_p - n1];
            if (k < n3) for (int s = 0; s < nprim; ++s) tmp[s] = pass ? tmp[s] - buf[2][s * nfd[2] + fk_p] : tmp[s] + buf[2][s * nfd[2] + fk_p];
            if (k > 1) for (int s = 0; s < nprim; ++s) tmp[s] = pass ? tmp[s] + buf[2][s * nfd[2] + fk_p - (int64_t)n1 * n2] : tmp[s] - buf[2][s * nfd[2] + fk_p - (int64_t)n1 * n2];
            (void)sg;
          }
          for (int s = 0; s < nprim; ++s) PV(R, s, c0) = tmp[s];
        }
    for (int d = 0; d < 3; ++d) { free(fh[d]); free(dfh[d]); }
  }
  /* BC_Fluxes: per block, faces 1..6 (records are grouped that way), R += per record */
  for (int e = 0; e < C->nrec; ++e) {
    const mose_bc_record *rc2 = &C->rec[e];
    double nrm[3], area, dlc4[4];
    record_geometry(C, rc2, nrm, &area, dlc4);
    switch (rc2->typ) {
      case 1:
        bc_connection_eul(C, P, R, rc2, nrm, area, dlc4);
        if (S->visc) bc_connection_visc(C, P, R, rc2, nrm, area);
        break;
      case 11:
        bc_extrapolation(C, P, R, rc2, nrm, area);
        break;
      case 3: {
        double bdl, idl;
        record_half_lengths(C, rc2, &bdl, &idl);
        bc_symmetry_eul(C, P, R, rc2, nrm, area, bdl, idl);
        if (S->visc) bc_sym_visc(C, P, R, rc2, nrm, area);
        break;
      }
      case 5:
      case 6: {
        double bdl, idl;
        record_half_lengths(C, rc2, &bdl, &idl);
        bc_symmetry_eul(C, P, R, rc2, nrm, area, bdl, idl);
        if (S->visc) bc_wall_visc(C, P, R, rc2, nrm, area);
        break;
      }
      case 4:
        if (bc_inflow_outflow(C, P, R, rc2, nrm, area)) {
          double bdl, idl;
          record_half_lengths(C, rc2, &bdl, &idl);
          bc_symmetry_eul(C, P, R, rc2, nrm, area, bdl, idl);
        }
        break;
      default:
        break;
    }
  }
  /* RANS (Mod_Explicit.f90:215-220): the wall ghost of nu~ on the state, then the source */
  if (S->rans == 2) {
    extrapolate_wall(C, P);
    sa_source(C, P, R, rc, mc, dt);
  }
  free(rc);
}

/* ------------------------------------------------------------------------------------------------ */
/* dt: Compute_dt_blk (dt_blk.inc:157-218, the non-additive rule) + the global min                    */
/* ------------------------------------------------------------------------------------------------ */
double o_compute_dt(const o_case *C, const double *P) {
  const o_scheme *S = &C->sc;
  const int64_t NC = C->NC;
  double *rc = (double *)malloc(sizeof(double) * NC * 4);
  double *Rc = rc + NC, *Tc = rc + 2 * NC, *mc = rc + 3 * NC;
  thermo_cache(C, P, rc, Rc, Tc, mc);
  double dtg = 1e5;   /* Explicit_Step sets dtglobal = 1d5 before Compute_dt (Mod_Explicit.f90:124) */
  for (int b = 0; b < C->nb; ++b) {
    const o_block *B = &C->blk[b];
    for (int k = 1; k <= B->n3; ++k)
      for (int j = 1; j <= B->n2; ++j)
        for (int i = 1; i <= B->n1; ++i) {
          int64_t c0 = ROW(B, i, j, k);
          double pl[ONPRIM];
          for (int s = 0; s < S->nprim; ++s) pl[s] = PV(P, s, c0);
          double Sound = f_ss_dev(&C->tab, pl, pl[S->np], rc[c0], Rc[c0]);
          double vel[3] = {pl[S->nu], pl[S->nv], pl[S->nw]};
          double dtc = 1e8;
          for (int d = 0; d < S->ndir; ++d) {
            double mx = C->m[(int64_t)(d + 0) * NC + c0], my = C->m[(int64_t)(d + 3) * NC + c0];
            double mz = C->m[(int64_t)(d + 6) * NC + c0];
            double nn = sqrt(mx * mx + my * my + mz * mz);   /* norm2: parity unpinned at the ulp level */
            double vx = mx / nn, vy = my / nn, vz = mz / nn;
            double vel_ = fabs(vel[0] * vx + vel[1] * vy + vel[2] * vz);
            double lambda = vel_ + Sound;
            double dtd = C->dl[(int64_t)d * NC + c0] / lambda * S->cfl;
            dtc = min2(dtd, dtc);
          }
          if (S->visc) {
            double mie = 0.0;
            if (S->rans == 2) mie = sa_mut(pl[S->nt], mc[c0]);
            double mi = mie + mc[c0];
            for (int d = 0; d < S->ndir; ++d) {
              double dl = C->dl[(int64_t)d * NC + c0];
              double dtd = (rc[c0] * (dl * dl) * S->vnn) / mi;
              dtc = min2(dtd, dtc);
            }
          }
          dtg = min2(dtc, dtg);
        }
  }
  free(rc);
  return dtg;
}

/* ------------------------------------------------------------------------------------------------ */
/* conservative update: Newstate_Conservative_blk + newstate_cons_body.inc:5-89                       */
/* ------------------------------------------------------------------------------------------------ */
static void o_newstate_cons(const o_case *C, double *P, const double *PO, const double *R, double dt, double rkc) {
  const o_scheme *S = &C->sc;
  const o_tables *t = &C->tab;
  const int64_t NC = C->NC;
  const double strang = C->strang;   /* strangcoeff: 1, or 1/2 in a Strang-split step (Lib_Strang.f90:10-22) */
  int nsc = S->nsc, np = S->np, nu = S->nu, nw = S->nw, nprim = S->nprim;
  int tmin_ = 0, tmax_ = t->tmax;
  for (int b = 0; b < C->nb; ++b) {
    const o_block *B = &C->blk[b];
#pragma omp parallel for collapse(2) schedule(static) num_threads(C->nthreads > 0 ? C->nthreads : 1)
    for (int k = 1; k <= B->n3; ++k)
      for (int j = 1; j <= B->n2; ++j)
        for (int i = 1; i <= B->n1; ++i) {
          int64_t c0 = ROW(B, i, j, k);
          double p_l[ONPRIM], po_l[ONPRIM], res_l[ONPRIM], consold[ONPRIM], cons[ONPRIM];
          for (int s = 0; s < nprim; ++s) { p_l[s] = PV(P, s, c0); po_l[s] = PV(PO, s, c0); res_l[s] = PV(R, s, c0); }
          double vol_l = C->vol[c0];
          /* T seed: the thermo cache of the stage (T = p/(rho R) of the current state) */
          double tc_l;
          {
            double rho_, Rg_;
            co_rotot_Rtot(t, p_l, &rho_, &Rg_);
            tc_l = p_l[np] / (rho_ * Rg_);
            if (S->visc && S->guard && !(tc_l > 1.0 && tc_l < 9990.0)) {
              if (tc_l != tc_l) tc_l = 300.0;
              tc_l = min2(max2(tc_l, 1.0), 9990.0);
            }
          }
          for (int pass = 0; pass < 2; ++pass) {
            const double *q = pass ? p_l : po_l;
            double *out = pass ? cons : consold;
            double rho = 0.0;
            for (int s = 0; s < nsc; ++s) rho = rho + q[s];
            double Rg = f_Rtot(t, q);
            double vmag = 0.0;
            for (int s = nu; s <= nw; ++s) vmag = vmag + q[s] * q[s];
            double Tloc = q[np] / (Rg * rho);
            int Ti = fint(Tloc); if (Ti < tmin_) Ti = tmin_; if (Ti > tmax_ - 1) Ti = tmax_ - 1;
            double Td = Tloc - Ti, hmix = 0.0;
            for (int s = 0; s < nsc; ++s) {
              double h_s = TAB(t->h, t, s, Ti) + Td * (TAB(t->h, t, s, Ti + 1) - TAB(t->h, t, s, Ti));
              hmix = hmix + h_s * q[s] / rho;
            }
            double E0 = hmix - q[np] / rho + 0.5 * vmag;
            for (int s = 0; s < nsc; ++s) out[s] = q[s];
            for (int s = nu; s <= nw; ++s) out[s] = rho * q[s];
            out[np] = rho * E0;
            for (int s = np + 1; s < nprim; ++s) out[s] = q[s];
          }
          /* with IRS the residual arrives scaled and smoothed (newstate_cons_body.inc:47) */
          if (!S->irs) for (int s = 0; s < nprim; ++s) res_l[s] = -res_l[s] / vol_l * dt * strang;
          for (int s = 0; s < nprim; ++s) cons[s] = consold[s] + (cons[s] - consold[s] + res_l[s]) * rkc;
          double rho = 0.0;
          for (int s = 0; s < nsc; ++s) rho = rho + cons[s];
          double Rg = f_Rtot(t, cons);
          for (int s = 0; s < nsc; ++s) p_l[s] = cons[s];
          for (int s = nu; s <= nw; ++s) p_l[s] = cons[s] / rho;
          double vmag = 0.0;
          for (int s = nu; s <= nw; ++s) vmag = vmag + (cons[s] / rho) * (cons[s] / rho);
          double energy = cons[np] / rho - 0.5 * vmag;
          double diff_en = 1e3, T_it = tc_l;
          int it = 1;
          while ((fabs(diff_en / energy) >= 1e-6) && (it <= 20)) {
            int Ti = fint(T_it); if (Ti < tmin_) Ti = tmin_; if (Ti > tmax_ - 1) Ti = tmax_ - 1;
            double Td = T_it - Ti, cv = -Rg, enth = 0.0;
            for (int s = 0; s < nsc; ++s) {
              double cp_s = TAB(t->cp, t, s, Ti) + Td * (TAB(t->cp, t, s, Ti + 1) - TAB(t->cp, t, s, Ti));
              double h_s = TAB(t->h, t, s, Ti) + Td * (TAB(t->h, t, s, Ti + 1) - TAB(t->h, t, s, Ti));
              cv = cv + cons[s] / rho * cp_s;
              enth = enth + cons[s] / rho * h_s;
            }
            double p_it = rho * Rg * T_it;
            diff_en = fabs((enth - p_it / rho) - energy);
            T_it = T_it + (energy - (enth - p_it / rho)) / cv;
            it = it + 1;
          }
          p_l[np] = rho * Rg * T_it;
          for (int s = np + 1; s < nprim; ++s) p_l[s] = cons[s];
          for (int s = 0; s < nsc; ++s) if (p_l[s] < 1e-20) p_l[s] = 1e-20;
          if (S->rans) for (int s = S->nt; s < S->nt + S->nrans; ++s) if (p_l[s] < 1e-10) p_l[s] = 1e-10;
          for (int s = 0; s < nprim; ++s) PV(P, s, c0) = p_l[s];
        }
  }
}

/* ------------------------------------------------------------------------------------------------ */
/* primitive update: Newstate_Primitive_blk (Lib_Newstate.f90:340-531), integration-variables = prim   */
/* ------------------------------------------------------------------------------------------------ */
/* One cell: Compute_Prim_Residual (the conservative residual turned into primitive increments through the inverse
 * of d(cons)/d(prim), Weiss-Smith form with Theta = (drho/dp)_T = 1/(R T)), the scaling dt/vol*strang, RK_stage on
 * the primitives, and the floors of Check_And_Fix_State.  The thermodynamics of the current state: co_rotot_Rtot,
 * f_cp_dev, H0_dev and f_tabT_expr (table rows clamped as everywhere in the oracle). */
static void o_prim_cell(const o_tables *t, const o_scheme *S, double *p_l, const double *po_l, double *res_l, double dt_,
                 double vol_, double strang, double rkc) {
  const int nsc = S->nsc, nu = S->nu, nv = S->nv, nw = S->nw, np = S->np, nprim = S->nprim;
  double rho, Rgas;
  co_rotot_Rtot(t, p_l, &rho, &Rgas);
  const double temperature = p_l[np] / (rho * Rgas);
  const int T_i = fint(temperature);
  const double Tdiff = temperature - T_i;
  const double inv_rho = 1.0 / rho;
  const double vel[3] = {p_l[nu], p_l[nv], p_l[nw]};
  const double modvel = sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);   /* norm2 */
  const double q2 = modvel * modvel;
  const double cp = f_cp_dev(t, p_l, T_i, Tdiff, rho);
  const double rho_T = -rho / temperature;
  const double rho_p = 1.0 / (Rgas * temperature);
  const double enthalpy = H0_dev(t, p_l[np], p_l, modvel);
  const double Theta = rho_p;   /* prim (prec would take 1/(beta2 a^2) + 1/(cp T)) */
  const double Denom = 1.0 / (rho * cp * Theta + rho_T);
  if (nsc == 1) {
    /* invGamma: (drho, drhoV, drhoE) -> (dp, du, dv, dw, dT) */
    double invG[5][5];
    invG[0][0] = (rho_T * (enthalpy - q2) + rho * cp) * Denom;
    invG[0][1] = rho_T * vel[0] * Denom;
    invG[0][2] = rho_T * vel[1] * Denom;
    invG[0][3] = rho_T * vel[2] * Denom;
    invG[0][4] = -rho_T * Denom;
    for (int r = 1; r <= 3; ++r) {
      for (int c = 0; c < 5; ++c) invG[r][c] = 0.0;
      invG[r][0] = -vel[r - 1] * inv_rho;
      invG[r][r] = inv_rho;
    }
    invG[4][0] = -(Theta * (enthalpy - q2) - 1.0) * Denom;
    invG[4][1] = -vel[0] * Theta * Denom;
    invG[4][2] = -vel[1] * Theta * Denom;
    invG[4][3] = -vel[2] * Theta * Denom;
    invG[4][4] = Theta * Denom;
    const double cres[5] = {-res_l[0], -res_l[nu], -res_l[nv], -res_l[nw], -res_l[np]};
    double pres[5];
    for (int r = 0; r < 5; ++r) {
      pres[r] = 0.0;
      for (int c = 0; c < 5; ++c) pres[r] = pres[r] + invG[r][c] * cres[c];
    }
    res_l[nu] = pres[1]; res_l[nv] = pres[2]; res_l[nw] = pres[3];
    res_l[np] = pres[0];
    res_l[0] = rho_p * pres[0] + rho_T * pres[4];   /* density from dp and dT */
  } else {
    /* several species: the mass-fraction sensitivities enter the p and T rows */
    double Yi[ONSMAX], drdyi[ONSMAX], dhdyi[ONSMAX], dyi[ONSMAX], Rs[ONPRIM + 1], dPa[ONPRIM + 1];
    for (int s = 0; s < nsc - 1; ++s) {
      Yi[s] = p_l[s] * inv_rho;
      if (Yi[s] <= 1e-9) Yi[s] = 0.0;
      drdyi[s] = -rho * (t->Ri[s] - t->Ri[nsc - 1]) / Rgas;
      dhdyi[s] = tabT(t->h, t, s, T_i, T_i + 1, Tdiff) - tabT(t->h, t, nsc - 1, T_i, T_i + 1, Tdiff);
    }
    double sres = 0.0;
    for (int s = 0; s < nsc; ++s) sres = sres + res_l[s];
    Rs[1] = -sres;
    Rs[2] = -res_l[nu]; Rs[3] = -res_l[nv]; Rs[4] = -res_l[nw]; Rs[5] = -res_l[np];
    for (int s = 0; s < nsc - 1; ++s) Rs[6 + s] = -res_l[s];
    double sYd = 0.0, sdY = 0.0, sR1 = 0.0, sR2 = 0.0;
    for (int s = 0; s < nsc - 1; ++s) {
      sYd = sYd + Yi[s] * drdyi[s];
      sdY = sdY + dhdyi[s] * Yi[s];
      sR1 = sR1 + Rs[6 + s] * (rho_T * dhdyi[s] - cp * drdyi[s]);
      sR2 = sR2 + Rs[6 + s] * (Theta * dhdyi[s] + drdyi[s] * inv_rho);
    }
    const double sv = vel[0] * Rs[2] + vel[1] * Rs[3] + vel[2] * Rs[4];
    dPa[1] = (Rs[1] * (cp * (rho + sYd) - rho_T * (sdY + q2 - enthalpy)) + sR1 - rho_T * Rs[5] + rho_T * sv) * Denom;
    dPa[2] = -Rs[1] * vel[0] * inv_rho + Rs[2] * inv_rho;
    dPa[3] = -Rs[1] * vel[1] * inv_rho + Rs[3] * inv_rho;
    dPa[4] = -Rs[1] * vel[2] * inv_rho + Rs[4] * inv_rho;
    dPa[5] = (Rs[1] * (Theta * (sdY - enthalpy + q2) + sYd * inv_rho + 1.0) - sR2 + Theta * (Rs[5] - sv)) * Denom;
    for (int s = 0; s < nsc - 1; ++s) dPa[6 + s] = (Rs[6 + s] - Rs[1] * Yi[s]) * inv_rho;
    double sdrdy = 0.0, dyN = 0.0;
    for (int s = 0; s < nsc - 1; ++s) {
      dyi[s] = dPa[6 + s];
      sdrdy = sdrdy + drdyi[s] * dyi[s];
      dyN = dyN + dyi[s];
    }
    const double drho = rho_p * dPa[1] + rho_T * dPa[5] + sdrdy;
    dyN = -dyN;
    const double drhoN = rho * dyN + p_l[nsc - 1] * inv_rho * drho;
    for (int s = 0; s < nsc - 1; ++s) res_l[s] = rho * dyi[s] + Yi[s] * drho;
    res_l[nsc - 1] = drhoN;
    res_l[nu] = dPa[2]; res_l[nv] = dPa[3]; res_l[nw] = dPa[4];
    res_l[np] = dPa[1];
  }
  for (int s = np + 1; s < nprim; ++s) res_l[s] = -res_l[s];   /* turbulence variables, not transformed */
  for (int s = 0; s < nprim; ++s) res_l[s] = res_l[s] * dt_ / vol_ * strang;
  for (int s = 0; s < nprim; ++s) p_l[s] = po_l[s] + (p_l[s] - po_l[s] + res_l[s]) * rkc;
  for (int s = 0; s < nsc; ++s) if (p_l[s] < 1e-20) p_l[s] = 1e-20;
  if (S->rans) for (int s = S->nt; s < S->nt + S->nrans; ++s) if (p_l[s] < 1e-10) p_l[s] = 1e-10;
}

static void o_newstate_prim(const o_case *C, double *P, const double *PO, const double *R, double dt, double rkc) {
  const o_scheme *S = &C->sc;
  const int64_t NC = C->NC;
  const int nprim = S->nprim;
  for (int b = 0; b < C->nb; ++b) {
    const o_block *B = &C->blk[b];
#pragma omp parallel for collapse(2) schedule(static) num_threads(C->nthreads > 0 ? C->nthreads : 1)
    for (int k = 1; k <= B->n3; ++k)
      for (int j = 1; j <= B->n2; ++j)
        for (int i = 1; i <= B->n1; ++i) {
          const int64_t c0 = ROW(B, i, j, k);
          double p_l[ONPRIM], po_l[ONPRIM], res_l[ONPRIM];
          for (int s = 0; s < nprim; ++s) { p_l[s] = PV(P, s, c0); po_l[s] = PV(PO, s, c0); res_l[s] = PV(R, s, c0); }
          o_prim_cell(&C->tab, S, p_l, po_l, res_l, dt, C->vol[c0], C->strang, rkc);
          for (int s = 0; s < nprim; ++s) PV(P, s, c0) = p_l[s];
        }
  }
}

/* the state update of the case's integration variables (Assign_Integration_Variables, Lib_Newstate.f90:43-60) */
void o_newstate(const o_case *C, double *P, const double *PO, const double *R, double dt, double rkc) {
  if (C->sc.prim) o_newstate_prim(C, P, PO, R, dt, rkc);
  else o_newstate_cons(C, P, PO, R, dt, rkc);
}

/* ------------------------------------------------------------------------------------------------ */
/* implicit residual smoothing: IRS_apply (Lib_IRS.f90:53-96)                                         */
/* ------------------------------------------------------------------------------------------------ */
/* The conservative update's scaling R = -R/vol*dt*strang of every variable (irs_scale_blk :99-121; the primitive
 * update smooths the raw residual), then per computational direction: the first ghost layer of the residual
 * (irs_ghost_residual :125-172: zero gradient on every block face, then each connection record's ghost from its donor
 * cell) and njacobi Jacobi sweeps R* <- (R + beta (R*(c-off) + R*(c+off))) / (1 + 2 beta) on the interior, started
 * from R* = R over the whole block (irs_smooth_blk :194-235); R <- R*.  Only the flow variables 0..np are smoothed. */
void o_irs(const o_case *C, double *R, double dt) {
  const o_scheme *S = &C->sc;
  const int64_t NC = C->NC;
  const int nprim = S->nprim, nf = S->np + 1;
  if (!S->prim) {
    for (int b = 0; b < C->nb; ++b) {
      const o_block *B = &C->blk[b];
      for (int k = 1; k <= B->n3; ++k)
        for (int j = 1; j <= B->n2; ++j)
          for (int i = 1; i <= B->n1; ++i) {
            const int64_t c0 = ROW(B, i, j, k);
            const double vol_ = C->vol[c0];
            for (int s = 0; s < nprim; ++s) PV(R, s, c0) = -PV(R, s, c0) / vol_ * dt * C->strang;
          }
    }
  }
  double *rs1 = (double *)malloc(sizeof(double) * (size_t)NC * nf);
  double *rs2 = (double *)malloc(sizeof(double) * (size_t)NC * nf);
  for (int d = 0; d < S->ndir; ++d) {
    for (int b = 0; b < C->nb; ++b) {
      const o_block *B = &C->blk[b];
      const int n[3] = {B->n1, B->n2, B->n3};
      for (int f = 1; f <= 6; ++f) {
        const int a = (f - 1) / 2, lo = (f % 2) == 1;
        const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;   /* the two directions along the face */
        for (int q = 1; q <= n[v]; ++q)
          for (int p = 1; p <= n[u]; ++p) {
            int cm[3], cg[3];
            cm[a] = lo ? 1 : n[a]; cm[u] = p; cm[v] = q;
            for (int e = 0; e < 3; ++e) cg[e] = cm[e] - guide[f - 1][e];
            const int64_t rm = ROW(B, cm[0], cm[1], cm[2]), rg = ROW(B, cg[0], cg[1], cg[2]);
            for (int s = 0; s < nf; ++s) PV(R, s, rg) = PV(R, s, rm);
          }
      }
    }
    for (int e = 0; e < C->nrec; ++e) {
      const mose_bc_record *rc = &C->rec[e];
      if (rc->typ != 1) continue;
      const o_block *Bm = &C->blk[rc->b], *Bs = &C->blk[rc->bs];
      const int64_t rg = ROW(Bm, rc->i - guide[rc->f - 1][0], rc->j - guide[rc->f - 1][1], rc->k - guide[rc->f - 1][2]);
      const int64_t rd = ROW(Bs, rc->is, rc->js, rc->ks);
      for (int s = 0; s < nf; ++s) PV(R, s, rg) = PV(R, s, rd);
    }
    for (int b = 0; b < C->nb; ++b) {
      const o_block *B = &C->blk[b];
      const int64_t off = d == 0 ? 1 : (d == 1 ? (int64_t)B->pi : (int64_t)B->pi * B->pj);
      const int64_t nflat = (int64_t)B->pi * B->pj * (B->n3 + 2 * OGC);
      for (int s = 0; s < nf; ++s)
        for (int64_t c = 0; c < nflat; ++c) rs1[(int64_t)s * NC + B->boff + c] = PV(R, s, B->boff + c);
      for (int jac = 0; jac < S->irs_njacobi; ++jac) {
        for (int k = 1; k <= B->n3; ++k)
          for (int j = 1; j <= B->n2; ++j)
            for (int i = 1; i <= B->n1; ++i) {
              const int64_t c0 = ROW(B, i, j, k);
              for (int s = 0; s < nf; ++s)
                rs2[(int64_t)s * NC + c0] = (PV(R, s, c0) + S->irs_beta * (rs1[(int64_t)s * NC + c0 - off] +
                                             rs1[(int64_t)s * NC + c0 + off])) / (1.0 + 2.0 * S->irs_beta);
            }
        for (int k = 1; k <= B->n3; ++k)
          for (int j = 1; j <= B->n2; ++j)
            for (int i = 1; i <= B->n1; ++i) {
              const int64_t c0 = ROW(B, i, j, k);
              for (int s = 0; s < nf; ++s) rs1[(int64_t)s * NC + c0] = rs2[(int64_t)s * NC + c0];
            }
      }
      for (int k = 1; k <= B->n3; ++k)
        for (int j = 1; j <= B->n2; ++j)
          for (int i = 1; i <= B->n1; ++i) {
            const int64_t c0 = ROW(B, i, j, k);
            for (int s = 0; s < nf; ++s) PV(R, s, c0) = rs1[(int64_t)s * NC + c0];
          }
    }
  }
  free(rs1);
  free(rs2);
}

/* RKcoeff(3,3): src/lib/numerics/time/explicit/Lib_RK.f90:6-9 */
static double rk_coeff(int n_rk, int irk) {
  static const double c[3][3] = {{1.0, 0.0, 0.0}, {1.0, 0.5, 0.0}, {1.0, 0.25, 2.0 / 3.0}};
  return c[n_rk - 1][irk - 1];
}

/* Explicit_Step (Mod_Explicit.f90:52-397), time accurate: dt from the state at the start of the step, then the
 * strang loop (:155-294): per substage PO <- P (Copy_State) and n_RK stages of ghosts -> residual -> conservative
 * update with res = -R/vol dt strangcoeff; with chemistry (n_strang = 2, strangcoeff = 1/2) the first substage is
 * followed by Chemistry_Newstate_blk over the whole dt on the resident state (:282-294). */
int o_step(const o_case *C, double *P, double *PO, double *R, int nsteps, double *dt_out) {
  const int64_t NC = C->NC;
  int nprim = C->sc.nprim;
  const int n_strang = C->n_strang > 0 ? C->n_strang : 1;
  for (int n = 0; n < nsteps; ++n) {
    double dt = o_compute_dt(C, P);
    if (dt_out) dt_out[n] = dt;
    for (int i_strang = 1; i_strang <= n_strang; ++i_strang) {
      memcpy(PO, P, sizeof(double) * NC * nprim);
      for (int irk = 1; irk <= C->sc.n_rk; ++irk) {
        o_fill_ghosts(C, P);
        o_residual(C, P, R, dt);
        if (C->sc.irs) o_irs(C, R, dt);
        o_newstate(C, P, PO, R, dt, rk_coeff(C->sc.n_rk, irk));
      }
      if (i_strang + n_strang == 3 && C->react) C->react(C, P, dt, C->react_ctx);
    }
  }
  return 0;
}
