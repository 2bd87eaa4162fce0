// Device physics of the residual path: table thermodynamics, limiters, MUSCL faces, HLLC / AUSM+ / Roe
// (LMRoe) fluxes, the viscous face flux, the boundary records and the conservative update of one cell.
//
// Each function keeps the floating-point association of the reference statement it cites, and the library is
// compiled with -ffp-contract=off, so a face or a cell gives the bits the reference's double-precision code
// gives for the same operands.  NS (species) and NR (RANS variables) are compile-time so every per-variable
// array lives in registers.
#pragma once

#include <hip/hip_runtime.h>

#include "mose_internal.h"

namespace mose {

constexpr int GC = MOSE_GC;
constexpr int NTMAX = MOSE_NTMAX;
constexpr int NTAB = MOSE_NTAB;
constexpr double RUNIV = MOSE_RUNIV;

template <int NS, int NR>
struct Vars {
  static constexpr int ns = NS, nu = NS, nv = NS + 1, nw = NS + 2, np = NS + 3, nt = NS + 4, nprim = NS + 4 + NR;
};

// ---------------------------------------------------------------------------------------------------------
// small helpers with the reference's semantics
// ---------------------------------------------------------------------------------------------------------
__device__ __forceinline__ double dmin2(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax2(double a, double b) { return (b > a) ? b : a; }
__device__ __forceinline__ int ftrunc(double x) { return (x == x) ? static_cast<int>(x) : 0; }   // Fortran int()
// Table rows are clamped to the table: a state off the table (T < 0 or T > 10000 K, only met in a failing state
// that the state check reports) reads the edge rows instead of memory outside the table.  For every T on the
// table the row is the reference's.
__device__ __forceinline__ int trow(int T) { return T < 0 ? 0 : (T > NTMAX ? NTMAX : T); }
__device__ __forceinline__ double tab(const double* t, int s, int T) { return t[s * NTAB + trow(T)]; }
// f_tabT_expr (Lib_ThermoTransport.f90:623-635): v1 + (v2 - v1) * Td
__device__ __forceinline__ double tabT(const double* t, int s, int T, double Td) {
  int T1 = T < 0 ? 0 : (T > NTMAX - 1 ? NTMAX - 1 : T);
  double v1 = t[s * NTAB + T1], v2 = t[s * NTAB + T1 + 1];
  return v1 + (v2 - v1) * Td;
}

// ---------------------------------------------------------------------------------------------------------
// FLINT thermodynamics (Lib_ThermoTransport.f90 / Lib_ThermoTransport_dev.f90)
// ---------------------------------------------------------------------------------------------------------
template <int NS>
__device__ __forceinline__ double rho_sum(const double* r) {
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < NS; ++q) s = s + r[q];
  return s;
}
// f_Rtot (:238-252)
template <int NS>
__device__ __forceinline__ double f_Rtot(const mose_dtab& T, const double* r) {
  double rho = rho_sum<NS>(r), R = 0.0;
#pragma unroll
  for (int q = 0; q < NS; ++q) R = R + T.Ri[q] * r[q] / rho;
  return R;
}
// f_cp_dev (_dev.f90:280-293)
template <int NS>
__device__ __forceinline__ double f_cp(const mose_dtab& T, const double* r, int Ti, double Td, double rho) {
  double c = 0.0;
#pragma unroll
  for (int q = 0; q < NS; ++q) c = c + r[q] / rho * tabT(T.cp, q, Ti, Td);
  return c;
}
// f_ss_dev (_dev.f90:70-82), no clamp of the index
template <int NS>
__device__ __forceinline__ double f_ss(const mose_dtab& T, const double* r, double p, double rho, double Rt) {
  double Tk = p / (Rt * rho);
  int Ti = ftrunc(Tk);
  double cp = f_cp<NS>(T, r, Ti, Tk - Ti, rho);
  double gam = cp / (cp - Rt);
  return sqrt(gam * Rt * Tk);
}
// H0_dev (_dev.f90:54-68)
template <int NS>
__device__ __forceinline__ double H0(const mose_dtab& T, double p, const double* r, double u) {
  double rho = rho_sum<NS>(r), Rg = f_Rtot<NS>(T, r);
  double Tk = p / (Rg * rho);
  int Ti = ftrunc(Tk);
  double Td = Tk - Ti, e = 0.0;
#pragma unroll
  for (int q = 0; q < NS; ++q) e = e + tabT(T.h, q, Ti, Td) * r[q] / rho;
  return e + 0.5 * u * u;
}
// co_k_mi_lam_Wilke_dev + co_fiij_dev (_dev.f90:120-175); f_molecularWeight (:220-235)
template <int NS>
__device__ __forceinline__ void wilke(const mose_dtab& T, const double* r, double rho, int Ti, double Td, double& mil,
                                      double& kl) {
  double rs = rho_sum<NS>(r), mw = 0.0;
#pragma unroll
  for (int q = 0; q < NS; ++q) mw = mw + (r[q] / rs / T.wm[q]);
  double Wmtot = 1.0 / mw;
  double X[NS], mi[NS], ki[NS], sqn[NS], isd[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    X[q] = r[q] * Wmtot / (rho * T.wm[q]);
    mi[q] = tabT(T.mi, q, Ti, Td);
    ki[q] = tabT(T.k, q, Ti, Td);
  }
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    sqn[q] = sqrt(mi[q]);
    isd[q] = 1.0 / sqrt(mi[q] + 1e-20);
  }
  double m = 0.0, k = 0.0;
#pragma unroll
  for (int a = 0; a < NS; ++a) {
    double den = 0.0;
#pragma unroll
    for (int b = 0; b < NS; ++b) {
      double op = 1.0 + sqn[a] * isd[b] * T.mm025[a][b];
      den = den + X[b] * (op * op * T.is8[a][b]);
    }
    double inv = 1.0 / den;
    m = m + X[a] * mi[a] * inv;
    k = k + X[a] * ki[a] * inv;
  }
  mil = m;
  kl = k;
}

// ---------------------------------------------------------------------------------------------------------
// limiters: Lib_Limiters.f90:65-97
// ---------------------------------------------------------------------------------------------------------
__device__ __forceinline__ double limiter(int lim, double x, double y) {
  double ru, phi;
  switch (lim) {
    case 1: {
      if (x * y < 0.0) return 0.0;
      return (fabs(x) > fabs(y)) ? y : x;
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
      phi = dmax2(0.0, dmin2(dmin2(2.0 * ru, 0.5 * (1.0 + ru)), 2.0));
      return phi * x;
    case 5:
      ru = y / (x + 1e-30);
      phi = dmax2(dmax2(0.0, dmin2(2.0 * ru, 1.0)), dmin2(ru, 2.0));
      return phi * x;
    default:
      return 0.0;
  }
}

// ---------------------------------------------------------------------------------------------------------
// one side of a face: species densities + u v w p, and its mixture rho / R / sound speed
// ---------------------------------------------------------------------------------------------------------
struct Roe {
  double h0l, h0r, rho, u, v, w, a, h0;
};

// Roe averages: conv_physics.inc:70-98 (clamp [0, NTMAX-1]), riemann_open_body.inc:679-726 (clamp
// [Tmin, Tmax-1] = [0, 9999] for the 0..10000 tables), Lib_Riemann_HLL.f90:210-310 (no clamp)
template <int NS>
__device__ __forceinline__ Roe roe_avg(const mose_dtab& T, bool clamp, const double* dl, const double* dr, double dlt,
                                       double drt, double ul, double vl, double wl, double ur, double vr, double wr,
                                       double pl, double pr, double Rl, double Rr) {
  Roe o;
  double idl = 1.0 / dlt, idr = 1.0 / drt;
  double Tl = pl * idl / Rl, Tr = pr * idr / Rr;
  int Til = ftrunc(Tl), Tir = ftrunc(Tr);
  if (clamp) {
    Til = Til < 0 ? 0 : (Til > NTMAX - 1 ? NTMAX - 1 : Til);
    Tir = Tir < 0 ? 0 : (Tir > NTMAX - 1 ? NTMAX - 1 : Tir);
  }
  double dTl = Tl - Til, dTr = Tr - Tir;
  double sL = sqrt(dlt), sR = sqrt(drt), isr = 1.0 / (sL + sR);
  o.rho = sL * sR;
  o.u = (sR * ur + sL * ul) * isr;
  o.v = (sR * vr + sL * vl) * isr;
  o.w = (sR * wr + sL * wl) * isr;
  double R_roe = (sR * Rr + sL * Rl) * isr, T_roe = (sR * Tr + sL * Tl) * isr;
  double h0l = 0.0, h0r = 0.0, cv = 0.0, sei = 0.0;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    double invW = RUNIV / T.wm[q];
    double hl = tab(T.h, q, Til) + (tab(T.h, q, Til + 1) - tab(T.h, q, Til)) * dTl;
    double hr = tab(T.h, q, Tir) + (tab(T.h, q, Tir + 1) - tab(T.h, q, Tir)) * dTr;
    double el = hl - invW * Tl, er = hr - invW * Tr;
    h0l = h0l + hl * dl[q] * idl;
    h0r = h0r + hr * dr[q] * idr;
    double dro = (sR * dr[q] * idr + sL * dl[q] * idl) * isr;
    double cpl = tab(T.cp, q, Til) + (tab(T.cp, q, Til + 1) - tab(T.cp, q, Til)) * dTl;
    cv = cv + dro * (0.5 * (cpl + tab(T.cp, q, Tir) + (tab(T.cp, q, Tir + 1) - tab(T.cp, q, Tir)) * dTr) - invW);
    sei = sei + dro * (sR * er + sL * el) * isr;
  }
  h0l = h0l + 0.5 * (ul * ul + vl * vl + wl * wl);
  h0r = h0r + 0.5 * (ur * ur + vr * vr + wr * wr);
  o.h0l = h0l;
  o.h0r = h0r;
  o.h0 = (sR * h0r + sL * h0l) * isr;
  double gam = 1.0 + R_roe / cv;
  double vel2 = o.u * o.u + o.v * o.v + o.w * o.w;
  o.a = sqrt((gam - 1.0) * (o.h0 - 0.5 * vel2 + cv * T_roe - sei));
  return o;
}

struct F5 {
  double r, u, v, w, e;
};

// HLLC (Batten): conv_physics.inc:87-115 / Lib_Riemann_HLL.f90:21-134
// HLLE (rs 2): conv_physics.inc:116-124 in the interior kernels, riemann_HLLE (Lib_Riemann_HLL.f90:137-198) in the
// record kernels: the two-wave HLL flux between the Davis/Einfeldt signal speeds of HLLC
__device__ __forceinline__ F5 hlle(double dlt, double drt, double ul, double vl, double wl, double pl, double al,
                                   double ur, double vr, double wr, double pr, double ar, double nx, double ny, double nz,
                                   const Roe& ro) {
  double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  double unr_ = ro.u * nx + ro.v * ny + ro.w * nz;
  double S1 = dmin2(dmin2(0.0, unl - al), unr_ - ro.a);
  double S4 = dmax2(dmax2(0.0, unr + ar), unr_ + ro.a);
  double idn = 1.0 / (S4 - S1);
  double frl = dlt * unl, frr = drt * unr;
  F5 F;
  F.r = (S4 * frl - S1 * frr + S1 * S4 * (drt - dlt)) * idn;
  F.u = (S4 * (frl * ul + pl * nx) - S1 * (frr * ur + pr * nx) + S1 * S4 * (drt * ur - dlt * ul)) * idn;
  F.v = (S4 * (frl * vl + pl * ny) - S1 * (frr * vr + pr * ny) + S1 * S4 * (drt * vr - dlt * vl)) * idn;
  F.w = (S4 * (frl * wl + pl * nz) - S1 * (frr * wr + pr * nz) + S1 * S4 * (drt * wr - dlt * wl)) * idn;
  F.e = (S4 * (frl * ro.h0l) - S1 * (frr * ro.h0r) + S1 * S4 * (drt * (ro.h0r - pr / drt) - dlt * (ro.h0l - pl / dlt))) *
        idn;
  return F;
}

__device__ __forceinline__ F5 hllc(double dlt, double drt, double ul, double vl, double wl, double pl, double al,
                                   double ur, double vr, double wr, double pr, double ar, double nx, double ny, double nz,
                                   const Roe& ro) {
  double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  double unr_ = ro.u * nx + ro.v * ny + ro.w * nz;
  double e0l = ro.h0l - pl / dlt, e0r = ro.h0r - pr / drt;
  double S1 = dmin2(dmin2(0.0, unl - al), unr_ - ro.a);
  double S4 = dmax2(dmax2(0.0, unr + ar), unr_ + ro.a);
  double idn = 1.0 / (dlt * (S1 - unl) - drt * (S4 - unr));
  double Ss = (pr - pl + dlt * unl * (S1 - unl) - drt * unr * (S4 - unr)) * idn;
  double ps = dlt * (unl - S1) * (unl - Ss) + pl;
  F5 F;
  if (S1 >= 0.0 || Ss >= 0.0) {
    F.r = dlt * unl; F.u = F.r * ul + pl * nx; F.v = F.r * vl + pl * ny; F.w = F.r * wl + pl * nz; F.e = F.r * ro.h0l;
    if (!(S1 >= 0.0)) {
      double id = 1.0 / (S1 - Ss);
      double U1 = dlt * (S1 - unl) * id;
      double U2 = ((S1 - unl) * dlt * ul + (ps - pl) * nx) * id;
      double U3 = ((S1 - unl) * dlt * vl + (ps - pl) * ny) * id;
      double U4 = ((S1 - unl) * dlt * wl + (ps - pl) * nz) * id;
      double U5 = ((S1 - unl) * dlt * e0l - pl * unl + ps * Ss) * id;
      F.r = F.r + S1 * (U1 - dlt); F.u = F.u + S1 * (U2 - dlt * ul); F.v = F.v + S1 * (U3 - dlt * vl);
      F.w = F.w + S1 * (U4 - dlt * wl); F.e = F.e + S1 * (U5 - dlt * e0l);
    }
  } else {
    F.r = drt * unr; F.u = F.r * ur + pr * nx; F.v = F.r * vr + pr * ny; F.w = F.r * wr + pr * nz; F.e = F.r * ro.h0r;
    if (S4 > 0.0) {
      double id = 1.0 / (S4 - Ss);
      double U1 = drt * (S4 - unr) * id;
      double U2 = ((S4 - unr) * drt * ur + (ps - pr) * nx) * id;
      double U3 = ((S4 - unr) * drt * vr + (ps - pr) * ny) * id;
      double U4 = ((S4 - unr) * drt * wr + (ps - pr) * nz) * id;
      double U5 = ((S4 - unr) * drt * e0r - pr * unr + ps * Ss) * id;
      F.r = F.r + S4 * (U1 - drt); F.u = F.u + S4 * (U2 - drt * ur); F.v = F.v + S4 * (U3 - drt * vr);
      F.w = F.w + S4 * (U4 - drt * wr); F.e = F.e + S4 * (U5 - drt * e0r);
    }
  }
  return F;
}

// static enthalpy with the clamped index of riemann_open_body.inc:103-111
template <int NS>
__device__ __forceinline__ double h_open(const mose_dtab& T, const double* d, double p) {
  double rho = rho_sum<NS>(d), rg = 0.0, hs = 0.0;
#pragma unroll
  for (int q = 0; q < NS; ++q) rg = rg + T.Ri[q] * d[q] / rho;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    double Tk = p / (rg * rho);
    double Tc = dmax2(0.0, dmin2(Tk, (double)(NTMAX - 1)));
    int Ti = ftrunc(Tc);
    double Td = Tc - Ti;
    int T1 = Ti < 0 ? 0 : (Ti > NTMAX ? NTMAX : Ti), T2 = Ti + 1 < 0 ? 0 : (Ti + 1 > NTMAX ? NTMAX : Ti + 1);
    double hs_ = tab(T.h, q, T1) + (tab(T.h, q, T2) - tab(T.h, q, T1)) * Td;
    hs = hs + hs_ * d[q] / rho;
  }
  return hs;
}

// AUSM+ (rs 4): riemann_open_body.inc:100-150
template <int NS>
__device__ __forceinline__ F5 ausmp(const mose_dtab& T, const double* dl, double ul, double vl, double wl, double pl,
                                    double al, double dlt, const double* dr, double ur, double vr, double wr, double pr,
                                    double ar, double drt, double nx, double ny, double nz) {
  const double ab = 0.125, aa = 0.1875;
  double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  double vel = sqrt(ul * ul + vl * vl + wl * wl);
  double HL = h_open<NS>(T, dl, pl) + 0.5 * vel * vel;
  vel = sqrt(ur * ur + vr * vr + wr * wr);
  double HR = h_open<NS>(T, dr, pr) + 0.5 * vel * vel;
  double ah = sqrt(al * ar), ML = unl / ah, MR = unr / ah;
  double MpL, ppL, MmR, pmR;
  if (fabs(ML) < 1.0) {
    double a = ML + 1.0, b = ML * ML - 1.0;
    MpL = 0.25 * (a * a) + ab * (b * b);
    ppL = 0.25 * (a * a) * (2.0 - ML) + aa * ML * (b * b);
  } else {
    MpL = 0.5 * (ML + fabs(ML));
    ppL = 0.5 * (1.0 + copysign(1.0, ML));
  }
  if (fabs(MR) < 1.0) {
    double a = MR - 1.0, b = MR * MR - 1.0;
    MmR = -(0.25 * (a * a)) - ab * (b * b);
    pmR = 0.25 * (a * a) * (2.0 + MR) - aa * MR * (b * b);
  } else {
    MmR = 0.5 * (MR - fabs(MR));
    pmR = 0.5 * (1.0 - copysign(1.0, MR));
  }
  double Mh = MpL + MmR, ph = ppL * pl + pmR * pr;
  double mp = dlt * ah * 0.5 * (Mh + fabs(Mh)), mm = drt * ah * 0.5 * (Mh - fabs(Mh));
  F5 F;
  F.r = mp + mm;
  F.u = mp * ul + mm * ur + nx * ph;
  F.v = mp * vl + mm * vr + ny * ph;
  F.w = mp * wl + mm * wr + nz * ph;
  F.e = mp * HL + mm * HR;
  return F;
}

// Roe flux with the Rieper low-Mach factor (LMRoe, rs 12): riemann_open_body.inc:965-1090.  With
// riemann-options-Mco = 1 the factor is 1 and this is the classical Roe flux.
template <int NS>
__device__ __forceinline__ F5 lmroe(const mose_dtab& T, double mco, const double* dl, double ul, double vl, double wl,
                                    double pl, double dlt, const double* dr, double ur, double vr, double wr, double pr,
                                    double drt, double nx, double ny, double nz, const Roe& ro) {
  F5 F;
  double nm = sqrt(nx * nx + ny * ny + nz * nz);
  if (nm <= 0.0) { F.r = F.u = F.v = F.w = F.e = 0.0; return F; }
  double n1 = nx / nm, n2 = ny / nm, n3 = nz / nm;
  double tx = 1.0, ty = 0.0, tz = 0.0;
  if (!(fabs(n1) < 0.9)) { tx = 0.0; ty = 1.0; }
  double t1x = tx - (tx * n1 + ty * n2 + tz * n3) * n1;
  double t1y = ty - (tx * n1 + ty * n2 + tz * n3) * n2;
  double t1z = tz - (tx * n1 + ty * n2 + tz * n3) * n3;
  double tm = sqrt(t1x * t1x + t1y * t1y + t1z * t1z);
  if (tm <= 0.0) { t1x = -n2; t1y = n1; t1z = 0.0; tm = sqrt(t1x * t1x + t1y * t1y + t1z * t1z); }
  t1x = t1x / tm; t1y = t1y / tm; t1z = t1z / tm;
  double t2x = n2 * t1z - n3 * t1y, t2y = n3 * t1x - n1 * t1z, t2z = n1 * t1y - n2 * t1x;
  double unL = ul * n1 + vl * n2 + wl * n3, unR = ur * n1 + vr * n2 + wr * n3;
  double u1L = ul * t1x + vl * t1y + wl * t1z, u1R = ur * t1x + vr * t1y + wr * t1z;
  double u2L = ul * t2x + vl * t2y + wl * t2z, u2R = ur * t2x + vr * t2y + wr * t2z;
  double H0L = h_open<NS>(T, dl, pl) + 0.5 * (ul * ul + vl * vl + wl * wl);
  double H0R = h_open<NS>(T, dr, pr) + 0.5 * (ur * ur + vr * vr + wr * wr);
  double FL[5] = {dlt * unL, dlt * unL * ul + pl * n1, dlt * unL * vl + pl * n2, dlt * unL * wl + pl * n3, dlt * unL * H0L};
  double FR[5] = {drt * unR, drt * unR * ur + pr * n1, drt * unR * vr + pr * n2, drt * unR * wr + pr * n3, drt * unR * H0R};
  double un = ro.u * n1 + ro.v * n2 + ro.w * n3;
  double u2 = ro.u * ro.u + ro.v * ro.v + ro.w * ro.w;
  double ut = sqrt(dmax2(0.0, u2 - un * un));
  double Ma = (ro.a > 0.0) ? (fabs(un) + ut) / ro.a : 1.0;
  double phi = dmax2(mco, dmin2(1.0, Ma));
  double drho = drt - dlt, dp = pr - pl, dun = unR - unL, dut1 = u1R - u1L, dut2 = u2R - u2L;
  double dunmod = phi * dun;
  double ab1 = fabs(un - ro.a), ab2 = fabs(un), ab5 = fabs(un + ro.a);
  double a2 = ro.a * ro.a, ia = 1.0 / ro.a, ia2 = 1.0 / a2;
  double al4 = drho - dp * ia2, al2 = ro.rho * dut1, al3 = ro.rho * dut2;
  double al1 = 0.5 * (dp * ia2 - ro.rho * dunmod * ia);
  double al5 = 0.5 * (dp * ia2 + ro.rho * dunmod * ia);
  double r1[5] = {1.0, ro.u - ro.a * n1, ro.v - ro.a * n2, ro.w - ro.a * n3, ro.h0 - ro.a * un};
  double r5[5] = {1.0, ro.u + ro.a * n1, ro.v + ro.a * n2, ro.w + ro.a * n3, ro.h0 + ro.a * un};
  double r4[5] = {1.0, ro.u, ro.v, ro.w, 0.5 * u2};
  double r2[5] = {0.0, t1x, t1y, t1z, ro.u * t1x + ro.v * t1y + ro.w * t1z};
  double r3[5] = {0.0, t2x, t2y, t2z, ro.u * t2x + ro.v * t2y + ro.w * t2z};
  double out[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    double d = ab1 * al1 * r1[q] + ab2 * al2 * r2[q] + ab2 * al3 * r3[q] + ab2 * al4 * r4[q] + ab5 * al5 * r5[q];
    out[q] = 0.5 * (FL[q] + FR[q]) - 0.5 * d;
  }
  F.r = out[0]; F.u = out[1]; F.v = out[2]; F.w = out[3]; F.e = out[4];
  return F;
}

// SLAU, the part common to the interior kernels (conv_physics.inc:126-167) and riemann_SLAU (Lib_Riemann.f90:140-198):
// from the sound speeds and the total enthalpies h0l/h0r, which each computes its own way (riemann below)
__device__ __forceinline__ F5 slau(double dlt, double drt, double ul, double vl, double wl, double pl, double al,
                                   double ur, double vr, double wr, double pr, double ar, double nx, double ny, double nz,
                                   double h0l, double h0r) {
  const double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  const double af = 0.5 * (al + ar);
  const double ml = unl / af, mr = unr / af;
  const double mt = dmin2(1.0, sqrt(0.5 * ((ul * ul + vl * vl + wl * wl) + (ur * ur + vr * vr + wr * wr))) / af);
  const double chi = (1.0 - mt) * (1.0 - mt);
  const double vnm = (dlt * fabs(unl) + drt * fabs(unr)) / (dlt + drt);
  const double frho = -dmax2(dmin2(ml, 0.0), -1.0) * dmin2(dmax2(mr, 0.0), 1.0);
  const double vnl = (1.0 - frho) * vnm + frho * fabs(unl);
  const double vnr = (1.0 - frho) * vnm + frho * fabs(unr);
  F5 F;
  F.r = 0.5 * (dlt * (unl + vnl) + drt * (unr - vnr) - (chi / af) * (pr - pl));
  double bl, br;
  if (fabs(ml) < 1.0) bl = 0.25 * (2.0 - ml) * ((ml + 1.0) * (ml + 1.0));
  else if (ml >= 0.0) bl = 1.0;
  else bl = 0.0;
  if (fabs(mr) < 1.0) br = 0.25 * (2.0 + mr) * ((mr - 1.0) * (mr - 1.0));
  else if (mr >= 0.0) br = 0.0;
  else br = 1.0;
  const double ph = 0.5 * (pl + pr) + 0.5 * (bl - br) * (pl - pr) + (1.0 - chi) * (bl + br - 1.0) * 0.5 * (pl + pr);
  if (F.r >= 0.0) {
    F.u = F.r * ul + nx * ph; F.v = F.r * vl + ny * ph; F.w = F.r * wl + nz * ph; F.e = F.r * h0l;
  } else {
    F.u = F.r * ur + nx * ph; F.v = F.r * vr + ny * ph; F.w = F.r * wr + nz * ph; F.e = F.r * h0r;
  }
  return F;
}

// riemann_SLAU's total enthalpy of one state (Lib_Riemann.f90:152-166)
template <int NS>
__device__ __forceinline__ double slau_h0(const mose_dtab& T, const double* d, double dtot, double u, double v,
                                          double w, double p) {
  double R = 0.0, h0 = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) R = R + T.Ri[s] * d[s] / dtot;
  const double Tk = p / (dtot * R);
  const int Ti = ftrunc(Tk);
  const double dT = Tk - Ti;
#pragma unroll
  for (int s = 0; s < NS; ++s) h0 = h0 + tabT(T.h, s, Ti, dT) * d[s] / dtot;
  return h0 + 0.5 * (u * u + v * v + w * w);
}

// The Riemann dispatch (conv_physics.inc:103-183; Lib_Riemann.f90:114-136).  hllc_clamp: the interior
// kernel clamps the table index of the HLLC Roe averages, the record kernels do not.  RS != 0 fixes the
// solver at compile time (one kernel per solver keeps the register allocation of the other two out).
template <int NS, int RS = 0>
__device__ __forceinline__ F5 riemann(const mose_dtab& T, int rs_, double mco, bool hllc_clamp, const double* dl,
                                      double ul, double vl, double wl, double pl, double al, double dlt, const double* dr,
                                      double ur, double vr, double wr, double pr, double ar, double drt, double nx,
                                      double ny, double nz, double Rl, double Rr) {
  const int rs = RS != 0 ? RS : rs_;
  if (rs == 12) {
    double rl = f_Rtot<NS>(T, dl), rr = f_Rtot<NS>(T, dr);
    Roe ro = roe_avg<NS>(T, true, dl, dr, dlt, drt, ul, vl, wl, ur, vr, wr, pl, pr, rl, rr);
    return lmroe<NS>(T, mco, dl, ul, vl, wl, pl, dlt, dr, ur, vr, wr, pr, drt, nx, ny, nz, ro);
  }
  if (rs == 4) return ausmp<NS>(T, dl, ul, vl, wl, pl, al, dlt, dr, ur, vr, wr, pr, ar, drt, nx, ny, nz);
  if (rs == 3) {
    // SLAU: the interior kernels take h0 of the inline Roe block (conv_physics.inc:57-78), the records
    // riemann_SLAU's own
    double h0l, h0r;
    if (hllc_clamp) {
      Roe ro = roe_avg<NS>(T, true, dl, dr, dlt, drt, ul, vl, wl, ur, vr, wr, pl, pr, Rl, Rr);
      h0l = ro.h0l; h0r = ro.h0r;
    } else {
      h0l = slau_h0<NS>(T, dl, dlt, ul, vl, wl, pl);
      h0r = slau_h0<NS>(T, dr, drt, ur, vr, wr, pr);
    }
    return slau(dlt, drt, ul, vl, wl, pl, al, ur, vr, wr, pr, ar, nx, ny, nz, h0l, h0r);
  }
  Roe ro = roe_avg<NS>(T, hllc_clamp, dl, dr, dlt, drt, ul, vl, wl, ur, vr, wr, pl, pr, Rl, Rr);
  if (rs == 2) return hlle(dlt, drt, ul, vl, wl, pl, al, ur, vr, wr, pr, ar, nx, ny, nz, ro);
  return hllc(dlt, drt, ul, vl, wl, pl, al, ur, vr, wr, pr, ar, nx, ny, nz, ro);
}

// ---------------------------------------------------------------------------------------------------------
// interior convective face: conv_physics.inc:4-179.  q* = the 4-cell stencil (cells i-1, i, i+1, i+2 of face
// i+1/2), dlc = their lengths along the face direction, res = +Flux of the face.
// ---------------------------------------------------------------------------------------------------------
// The limited half-slope of one cell along one direction: lim(s(c,c+1), s(c-1,c)) * (dl_c / 2), with the slopes
// s(a,b) = (q_b - q_a) * (1 / (0.5 (dl_a + dl_b))) of conv_physics.inc:5-17.  The left state of face c+1/2 is
// q_c + delta_c and the right state of face c-1/2 is q_c - delta_c: both faces use the same product
// `beta * lim(slope1, slope0) * (0.5 dl(0))` of the reference, so it is computed once per cell.
__device__ __forceinline__ double half_slope(int lim, double qm, double q0, double qp, double idlm, double idlp,
                                             double hdl) {
  double s0 = (q0 - qm) * idlm, s1 = (qp - q0) * idlp;
  return 1.0 * limiter(lim, s1, s0) * hdl;
}

// The flux of an interior face from its reconstructed states L (left) and R (right): conv_physics.inc:34-179
template <int NS, int NR, int RS = 0>
__device__ __forceinline__ void conv_face_lr(const mose_dtab& T, int rs, double mco, double* L, double* R, double nx,
                                             double ny, double nz, double area, double* res);

template <int NS, int NR>
__device__ __forceinline__ void conv_face(const mose_dtab& T, int lim, int rs, double mco, const double* qm,
                                          const double* q0, const double* q1, const double* q2, const double* dlc,
                                          double nx, double ny, double nz, double area, double* res) {
  using V = Vars<NS, NR>;
  double L[V::nprim], R[V::nprim];
  double l0 = 0.5 * (dlc[0] + dlc[1]), l1 = 0.5 * (dlc[1] + dlc[2]), l2 = 0.5 * (dlc[2] + dlc[3]);
  double lm = 0.5 * dlc[1], lp = 0.5 * dlc[2];
  double i0 = 1.0 / l0, i1 = 1.0 / l1, i2 = 1.0 / l2;
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) {
    double s0 = (q0[s] - qm[s]) * i0, s1 = (q1[s] - q0[s]) * i1, s2 = (q2[s] - q1[s]) * i2;
    L[s] = q0[s] + 1.0 * limiter(lim, s1, s0) * lm;
    R[s] = q1[s] - 1.0 * limiter(lim, s2, s1) * lp;
  }
  conv_face_lr<NS, NR>(T, rs, mco, L, R, nx, ny, nz, area, res);
}

template <int NS, int NR, int RS>
__device__ __forceinline__ void conv_face_lr(const mose_dtab& T, int rs, double mco, double* L, double* R, double nx,
                                             double ny, double nz, double area, double* res) {
  using V = Vars<NS, NR>;
#pragma unroll
  for (int s = 0; s < NS; ++s) { L[s] = dmax2(L[s], 1e-20); R[s] = dmax2(R[s], 1e-20); }
  double rL = rho_sum<NS>(L), rR = rho_sum<NS>(R), RL = 0.0, RR = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) RL = RL + T.Ri[s] * L[s] / rL;
#pragma unroll
  for (int s = 0; s < NS; ++s) RR = RR + T.Ri[s] * R[s] / rR;
  double aL, aR;
  {
    double Tk = L[V::np] / (RL * rL);
    int Ti = ftrunc(Tk); Ti = Ti < 0 ? 0 : (Ti > NTMAX - 1 ? NTMAX - 1 : Ti);
    double Td = Tk - Ti, cp = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) cp = cp + L[s] / rL * (tab(T.cp, s, Ti) + (tab(T.cp, s, Ti + 1) - tab(T.cp, s, Ti)) * Td);
    double g = cp / (cp - RL);
    aL = sqrt(g * RL * Tk);
  }
  {
    double Tk = R[V::np] / (RR * rR);
    int Ti = ftrunc(Tk); Ti = Ti < 0 ? 0 : (Ti > NTMAX - 1 ? NTMAX - 1 : Ti);
    double Td = Tk - Ti, cp = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) cp = cp + R[s] / rR * (tab(T.cp, s, Ti) + (tab(T.cp, s, Ti + 1) - tab(T.cp, s, Ti)) * Td);
    double g = cp / (cp - RR);
    aR = sqrt(g * RR * Tk);
  }
  F5 F = riemann<NS, RS>(T, rs, mco, true, L, L[V::nu], L[V::nv], L[V::nw], L[V::np], aL, rL, R, R[V::nu], R[V::nv],
                     R[V::nw], R[V::np], aR, rR, nx, ny, nz, RL, RR);
  double su = copysign(0.5, F.r), sp = 0.5 + su, sm = su - 0.5;
  double iL = 1.0 / rL, iR = 1.0 / rR;
#pragma unroll
  for (int s = 0; s < NS; ++s) res[s] = F.r * area * (sp * (L[s] * iL) - sm * (R[s] * iR));
  res[V::nu] = F.u * area; res[V::nv] = F.v * area; res[V::nw] = F.w * area; res[V::np] = F.e * area;
#pragma unroll
  for (int s = V::np + 1; s < V::nprim; ++s) res[s] = F.r * area * (sp * L[s] * iL - sm * R[s] * iR);
}

// ---------------------------------------------------------------------------------------------------------
// viscous
// ---------------------------------------------------------------------------------------------------------
constexpr double SA_CV1 = 7.1;
constexpr double SA_SIGMA = 2.0 / 3.0;
constexpr double NTT = -2.0 / 3.0;   // Mod_Fluid.f90:4

struct FaceThermo {
  double rho, T, mil, kl, cp, Td;
  int Ti;
};
// Diffusive_Face_Thermo_Fill / Diffusive_Face_State (Mod_Diffusive.f90:291-313, :748-776) on the averaged state
template <int NS>
__device__ __forceinline__ FaceThermo face_thermo(const mose_dtab& T, const double* prim, double p, double prl) {
  FaceThermo o;
  double rho = rho_sum<NS>(prim), Rg = f_Rtot<NS>(T, prim);
  double Tk = p / (rho * Rg);
  int Ti = ftrunc(Tk);
  double Td = Tk - Ti;
  double cp = f_cp<NS>(T, prim, Ti, Td, rho);
  double mil, kl;
  wilke<NS>(T, prim, rho, Ti, Td, mil, kl);
  if (prl > 0.0) kl = mil * cp / prl;
  o.rho = rho; o.T = Tk; o.mil = mil; o.kl = kl; o.cp = cp; o.Ti = Ti; o.Td = Td;
  return o;
}

struct ViscConst {
  double sc, sct, prt, prl;
};

// diff_physics.inc:59-161 from the computational-space gradients g[var][axis] (already assembled) and the
// averaged metric: the metric transform, the stress, heat and species fluxes of an interior face.
template <int NS, int NR>
__device__ __forceinline__ void diff_flux_interior(const mose_dtab& T, const ViscConst& vc, double (*G)[3],
                                                   const double (*Mav)[3], const double* Pf, const FaceThermo& ft,
                                                   double nx, double ny, double nz, double area, double* Flx) {
  using V = Vars<NS, NR>;
#pragma unroll
  for (int iv = 0; iv < V::nprim; ++iv) {
    double g1 = G[iv][0], g2 = G[iv][1], g3 = G[iv][2];
    G[iv][0] = g1 * Mav[0][0] + g2 * Mav[1][0] + g3 * Mav[2][0];
    G[iv][1] = g1 * Mav[0][1] + g2 * Mav[1][1] + g3 * Mav[2][1];
    G[iv][2] = g1 * Mav[0][2] + g2 * Mav[1][2] + g3 * Mav[2][2];
  }
  double drho = ft.rho, dmil = ft.mil;
  int Ti = ftrunc(ft.T);
  double Tdf = ft.T - Ti;
  double mie = 0.0;
  if (NR > 0) {
    double chi = Pf[V::nt] / dmil, c3 = chi * chi * chi;
    mie = c3 / (c3 + SA_CV1 * SA_CV1 * SA_CV1) * Pf[V::nt];
  }
  double Dm = (dmil / vc.sc + mie / vc.sct) / drho;
  double kap = ft.kl + mie * ft.cp / vc.prt;
  double dmi = dmil + mie, lamDiv = NTT * dmi * (G[V::nu][0] + G[V::nv][1] + G[V::nw][2]);
  double Txx = lamDiv + 2.0 * dmi * G[V::nu][0] + 0.0;
  double Tyy = lamDiv + 2.0 * dmi * G[V::nv][1] + 0.0;
  double Tzz = lamDiv + 2.0 * dmi * G[V::nw][2] + 0.0;
  double Txy = dmi * (G[V::nu][1] + G[V::nv][0]), Txz = dmi * (G[V::nu][2] + G[V::nw][0]);
  double Tyz = dmi * (G[V::nv][2] + G[V::nw][1]);
  double S1 = Txx * nx + Txy * ny + Txz * nz, S2 = Txy * nx + Tyy * ny + Tyz * nz, S3 = Txz * nx + Tyz * ny + Tzz * nz;
  double DiffH = 0.0, sf = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double DmG = Dm * (G[s][0] * nx + G[s][1] * ny + G[s][2] * nz);
    Flx[s] = drho * area * DmG;
    double h1 = tab(T.h, s, Ti), h2 = tab(T.h, s, Ti + 1);
    DiffH = DiffH + DmG * (h1 + Tdf * (h2 - h1));
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) sf = sf + Flx[s];
#pragma unroll
  for (int s = 0; s < NS; ++s) Flx[s] = Flx[s] - sf * Pf[s] / drho;
  Flx[V::nu] = area * S1; Flx[V::nv] = area * S2; Flx[V::nw] = area * S3;
  Flx[V::np] = area * ((S1 * Pf[V::nu] + S2 * Pf[V::nv] + S3 * Pf[V::nw]) +
                       kap * (G[V::np][0] * nx + G[V::np][1] * ny + G[V::np][2] * nz) + DiffH * drho);
  if (NR > 0) {
    double gn = G[V::nt][0] * nx + G[V::nt][1] * ny + G[V::nt][2] * nz;
    Flx[V::nt] = (dmil + Pf[V::nt]) * gn * area / SA_SIGMA;
  }
}

// Compute_Diffusive_Flux_St (Mod_Diffusive.f90:781-867) + Stress_Vector_Std (Mod_Fluid.f90:20-44) +
// Spalart_RANS_Diffusive_Flux (Mod_Spalart.f90:401-418): the record version of the face flux
template <int NS, int NR>
__device__ __forceinline__ void diff_flux_st(const mose_dtab& T, const ViscConst& vc, const double (*G)[3],
                                             const double* Pf, const FaceThermo& ft, double nx, double ny, double nz,
                                             double area, double* Fx) {
  using V = Vars<NS, NR>;
  double rho = ft.rho, mil = ft.mil;
  double mie = 0.0;
  if (NR > 0) {
    double chi = Pf[V::nt] / mil, c3 = chi * chi * chi;
    mie = c3 / (c3 + SA_CV1 * SA_CV1 * SA_CV1) * Pf[V::nt];
  }
  double Dm = (mil / vc.sc + mie / vc.sct) / rho;
  double kap = ft.kl + mie * ft.cp / vc.prt;
  double mi = mil + mie, lamDiv = NTT * mi * (G[V::nu][0] + G[V::nv][1] + G[V::nw][2]);
  double Txx = lamDiv + 2.0 * mi * G[V::nu][0], Tyy = lamDiv + 2.0 * mi * G[V::nv][1];
  double Tzz = lamDiv + 2.0 * mi * G[V::nw][2];
  double Txy = mi * (G[V::nu][1] + G[V::nv][0]), Txz = mi * (G[V::nu][2] + G[V::nw][0]);
  double Tyz = mi * (G[V::nv][2] + G[V::nw][1]);
  double S1 = Txx * nx + Txy * ny + Txz * nz, S2 = Txy * nx + Tyy * ny + Tyz * nz, S3 = Txz * nx + Tyz * ny + Tzz * nz;
  double DiffH = 0.0, sf = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double DmG = Dm * (G[s][0] * nx + G[s][1] * ny + G[s][2] * nz);
    Fx[s] = rho * area * DmG;
    double h1 = tab(T.h, s, ft.Ti), h2 = tab(T.h, s, ft.Ti + 1);
    DiffH = DiffH + DmG * (h1 + ft.Td * (h2 - h1));
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) sf = sf + Fx[s];
#pragma unroll
  for (int s = 0; s < NS; ++s) Fx[s] = Fx[s] - sf * Pf[s] / rho;
  Fx[V::nu] = area * S1; Fx[V::nv] = area * S2; Fx[V::nw] = area * S3;
  Fx[V::np] = area * ((S1 * Pf[V::nu] + S2 * Pf[V::nv] + S3 * Pf[V::nw]) +
                      kap * (G[V::np][0] * nx + G[V::np][1] * ny + G[V::np][2] * nz) + DiffH * rho);
  if (NR > 0) Fx[V::nt] = area * (mil + Pf[V::nt]) * (G[V::nt][0] * nx + G[V::nt][1] * ny + G[V::nt][2] * nz) / SA_SIGMA;
}

// ---------------------------------------------------------------------------------------------------------
// conservative update of one cell: newstate_cons_body.inc:5-89 (rkc = RKcoeff, strang = strangcoeff)
// returns false when the new state fails State_Check_blk (p <= 0 or non-finite)
// ---------------------------------------------------------------------------------------------------------
template <int NS, int NR>
__device__ __forceinline__ void prim2cons(const mose_dtab& T, const double* q, double* out) {
  using V = Vars<NS, NR>;
  double rho = rho_sum<NS>(q), Rg = f_Rtot<NS>(T, q);
  double vm = 0.0;
#pragma unroll
  for (int s = V::nu; s <= V::nw; ++s) vm = vm + q[s] * q[s];
  double Tl = q[V::np] / (Rg * rho);
  int Ti = ftrunc(Tl); Ti = Ti < 0 ? 0 : (Ti > NTMAX - 1 ? NTMAX - 1 : Ti);
  double Td = Tl - Ti, hm = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) hm = hm + (tab(T.h, s, Ti) + Td * (tab(T.h, s, Ti + 1) - tab(T.h, s, Ti))) * q[s] / rho;
  double E0 = hm - q[V::np] / rho + 0.5 * vm;
#pragma unroll
  for (int s = 0; s < NS; ++s) out[s] = q[s];
#pragma unroll
  for (int s = V::nu; s <= V::nw; ++s) out[s] = rho * q[s];
  out[V::np] = rho * E0;
#pragma unroll
  for (int s = V::np + 1; s < V::nprim; ++s) out[s] = q[s];
}

// PRESCALED: res already holds -R/vol*dt*strang, smoothed (implicit residual smoothing, newstate_cons_body.inc:47)
template <int NS, int NR, bool PRESCALED = false>
__device__ __forceinline__ void update_cell(const mose_dtab& T, const double* p, const double* po, const double* res,
                                            double vol, double dt, double strang, double rkc, double Tseed, double* out) {
  using V = Vars<NS, NR>;
  double cold[V::nprim], c[V::nprim];
  prim2cons<NS, NR>(T, po, cold);
  prim2cons<NS, NR>(T, p, c);
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) {
    double r = PRESCALED ? res[s] : -res[s] / vol * dt * strang;
    c[s] = cold[s] + (c[s] - cold[s] + r) * rkc;
  }
  double rho = rho_sum<NS>(c), Rg = f_Rtot<NS>(T, c);
#pragma unroll
  for (int s = 0; s < NS; ++s) out[s] = c[s];
  double vm = 0.0;
#pragma unroll
  for (int s = V::nu; s <= V::nw; ++s) {
    out[s] = c[s] / rho;
    vm = vm + (c[s] / rho) * (c[s] / rho);
  }
  double en = c[V::np] / rho - 0.5 * vm;
  double de = 1e3, Tit = Tseed;
  int it = 1;
  while ((fabs(de / en) >= 1e-6) && (it <= 20)) {
    int Ti = ftrunc(Tit); Ti = Ti < 0 ? 0 : (Ti > NTMAX - 1 ? NTMAX - 1 : Ti);
    double Td = Tit - Ti, cv = -Rg, enth = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double cps = tab(T.cp, s, Ti) + Td * (tab(T.cp, s, Ti + 1) - tab(T.cp, s, Ti));
      double hs = tab(T.h, s, Ti) + Td * (tab(T.h, s, Ti + 1) - tab(T.h, s, Ti));
      cv = cv + c[s] / rho * cps;
      enth = enth + c[s] / rho * hs;
    }
    double pit = rho * Rg * Tit;
    de = fabs((enth - pit / rho) - en);
    Tit = Tit + (en - (enth - pit / rho)) / cv;
    it = it + 1;
  }
  out[V::np] = rho * Rg * Tit;
#pragma unroll
  for (int s = V::np + 1; s < V::nprim; ++s) out[s] = c[s];
#pragma unroll
  for (int s = 0; s < NS; ++s) if (out[s] < 1e-20) out[s] = 1e-20;
#pragma unroll
  for (int s = V::nt; s < V::nprim; ++s) if (out[s] < 1e-10) out[s] = 1e-10;
}

// Newstate_Primitive_blk (Lib_Newstate.f90:340-531), integration-variables = prim: Compute_Prim_Residual turns the
// conservative residual res into primitive increments through the inverse of d(cons)/d(prim) at the current state p
// (Weiss-Smith form, Theta = (drho/dp)_T = 1/(R T)); the increments are scaled by dt/vol*strang and RK_stage runs on
// the primitives; then the floors of Check_And_Fix_State.  In the statement order of the reference (the zero entries
// of the one-species inverse are multiplied as it multiplies them).
template <int NS, int NR>
__device__ __forceinline__ void update_cell_prim(const mose_dtab& T, const double* p, const double* po,
                                                 const double* res, double vol, double dt, double strang, double rkc,
                                                 double* out) {
  using V = Vars<NS, NR>;
  constexpr int NP = V::nprim;
  double r[NP];
#pragma unroll
  for (int s = 0; s < NP; ++s) r[s] = res[s];
  const double rho = rho_sum<NS>(p), Rgas = f_Rtot<NS>(T, p);
  const double temperature = p[V::np] / (rho * Rgas);
  const int Ti = ftrunc(temperature);
  const double Tdiff = temperature - Ti;
  const double inv_rho = 1.0 / rho;
  const double vel[3] = {p[V::nu], p[V::nv], p[V::nw]};
  const double modvel = sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);   // norm2
  const double q2 = modvel * modvel;
  const double cp = f_cp<NS>(T, p, Ti, Tdiff, rho);
  const double rho_T = -rho / temperature;
  const double rho_p = 1.0 / (Rgas * temperature);
  const double enthalpy = H0<NS>(T, p[V::np], p, modvel);
  const double Theta = rho_p;
  const double Denom = 1.0 / (rho * cp * Theta + rho_T);
  if constexpr (NS == 1) {
    // invGamma: (drho, drhoV, drhoE) -> (dp, du, dv, dw, dT)
    double invG[5][5];
    invG[0][0] = (rho_T * (enthalpy - q2) + rho * cp) * Denom;
    invG[0][1] = rho_T * vel[0] * Denom;
    invG[0][2] = rho_T * vel[1] * Denom;
    invG[0][3] = rho_T * vel[2] * Denom;
    invG[0][4] = -rho_T * Denom;
#pragma unroll
    for (int q = 1; q <= 3; ++q) {
#pragma unroll
      for (int c = 0; c < 5; ++c) invG[q][c] = 0.0;
      invG[q][0] = -vel[q - 1] * inv_rho;
      invG[q][q] = inv_rho;
    }
    invG[4][0] = -(Theta * (enthalpy - q2) - 1.0) * Denom;
    invG[4][1] = -vel[0] * Theta * Denom;
    invG[4][2] = -vel[1] * Theta * Denom;
    invG[4][3] = -vel[2] * Theta * Denom;
    invG[4][4] = Theta * Denom;
    const double cres[5] = {-r[0], -r[V::nu], -r[V::nv], -r[V::nw], -r[V::np]};
    double pres[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      pres[q] = 0.0;
#pragma unroll
      for (int c = 0; c < 5; ++c) pres[q] = pres[q] + invG[q][c] * cres[c];
    }
    r[V::nu] = pres[1]; r[V::nv] = pres[2]; r[V::nw] = pres[3];
    r[V::np] = pres[0];
    r[0] = rho_p * pres[0] + rho_T * pres[4];   // density from dp and dT
  } else {
    // several species: the mass-fraction sensitivities enter the p and T rows
    constexpr int M = NS - 1;
    double Yi[M], drdyi[M], dhdyi[M], dyi[M], Rs[NP + 1], dPa[NP + 1];
#pragma unroll
    for (int s = 0; s < M; ++s) {
      Yi[s] = p[s] * inv_rho;
      if (Yi[s] <= 1e-9) Yi[s] = 0.0;
      drdyi[s] = -rho * (T.Ri[s] - T.Ri[NS - 1]) / Rgas;
      dhdyi[s] = tabT(T.h, s, Ti, Tdiff) - tabT(T.h, NS - 1, Ti, Tdiff);
    }
    double sres = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) sres = sres + r[s];
    Rs[1] = -sres;
    Rs[2] = -r[V::nu]; Rs[3] = -r[V::nv]; Rs[4] = -r[V::nw]; Rs[5] = -r[V::np];
#pragma unroll
    for (int s = 0; s < M; ++s) Rs[6 + s] = -r[s];
    double sYd = 0.0, sdY = 0.0, sR1 = 0.0, sR2 = 0.0;
#pragma unroll
    for (int s = 0; s < M; ++s) {
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
#pragma unroll
    for (int s = 0; s < M; ++s) dPa[6 + s] = (Rs[6 + s] - Rs[1] * Yi[s]) * inv_rho;
    double sdrdy = 0.0, dyN = 0.0;
#pragma unroll
    for (int s = 0; s < M; ++s) {
      dyi[s] = dPa[6 + s];
      sdrdy = sdrdy + drdyi[s] * dyi[s];
      dyN = dyN + dyi[s];
    }
    const double drho = rho_p * dPa[1] + rho_T * dPa[5] + sdrdy;
    dyN = -dyN;
    const double drhoN = rho * dyN + p[NS - 1] * inv_rho * drho;
#pragma unroll
    for (int s = 0; s < M; ++s) r[s] = rho * dyi[s] + Yi[s] * drho;
    r[NS - 1] = drhoN;
    r[V::nu] = dPa[2]; r[V::nv] = dPa[3]; r[V::nw] = dPa[4];
    r[V::np] = dPa[1];
  }
#pragma unroll
  for (int s = V::np + 1; s < NP; ++s) r[s] = -r[s];   // turbulence variables: not transformed
#pragma unroll
  for (int s = 0; s < NP; ++s) r[s] = r[s] * dt / vol * strang;
#pragma unroll
  for (int s = 0; s < NP; ++s) out[s] = po[s] + (p[s] - po[s] + r[s]) * rkc;
#pragma unroll
  for (int s = 0; s < NS; ++s) if (out[s] < 1e-20) out[s] = 1e-20;
#pragma unroll
  for (int s = V::nt; s < NP; ++s) if (out[s] < 1e-10) out[s] = 1e-10;
}

}  // namespace mose
