// Device twins of the wall and far-field boundary records and of the Spalart-Allmaras cell source.
//
//   bc_symmetry      BC_Symmetry_Eul (Lib_BC.f90:640-733): the pressure flux of the symmetry faces (type 3), of the
//                    no-slip walls (types 5, 6) and of a far-field face that closes (type 4 with error = 1)
//   bc_wall_visc     BC_Wall_Heat_dev (type 5, Lib_BC.f90:879-982) / BC_Wall_Temperature_dev (type 6, :997-1115)
//   bc_sym_visc      BC_Symmetry_Visc (type 3, Lib_BC.f90:741-861)
//   bc_farfield      BC_Inflow_Outflow (Lib_BC.f90:1519-1597) with BC_Outflow (:1168-1243),
//                    BC_Supersonic_Inflow (:1247-1330) and BC_Subsonic_Inflow (:1335-1512)
//   sa_source        SA_Source_Blk + sa_body.inc (Mod_Spalart.f90:105-151), after the wall ghost of
//                    Extrapolate_Wall_blk (Lib_Extrapolation.f90:128-190)
//
// Each keeps the operand order of the statement it cites (the library is built with -ffp-contract=off).
#pragma once

#include "mose_physics.hpp"

namespace mose {

// SA constants (Mod_Spalart.f90:14-34)
constexpr double SA_CB1 = 0.1355, SA_CB2 = 0.622, SA_CW2 = 0.3, SA_CK = 0.41, SA_C2 = 0.7, SA_C3 = 0.9;

// The record's two cells along the face direction and the stencil data every wall / far-field record needs
struct RecCells {
  int64_t c, ci;        // boundary cell, interior neighbour (Im + gstep along the face direction)
  int dir, modfm;       // 1-based direction, mod(Fm, 2)
};

template <int NS, int NR>
__device__ __forceinline__ void bc_symmetry(const mose_dtab& T, const double* P, int64_t NC, const double* dl, int lim,
                                            int rs, double mco, const RecCells& rc, double nx, double ny, double nz,
                                            double area, double* e1) {
  using V = Vars<NS, NR>;
  constexpr int NP = V::nprim;
  int modfm = rc.modfm, modfm1 = 1 - modfm, modfm2 = 1 - 2 * modfm;
  const double* dlp = dl + (int64_t)(rc.dir - 1) * NC;
  double Bdl = 0.5 * dlp[rc.c], Idl = 0.5 * dlp[rc.ci];
  double Bd[NP], Gh[NP], Fa[V::np + 1];
#pragma unroll
  for (int s = 0; s < NP; ++s) Bd[s] = P[s * NC + rc.c];
  double Un = Bd[V::nu] * nx + Bd[V::nv] * ny + Bd[V::nw] * nz;
#pragma unroll
  for (int s = 0; s < NP; ++s) Gh[s] = Bd[s];
  Gh[V::nu] = Bd[V::nu] - 2.0 * Un * nx;
  Gh[V::nv] = Bd[V::nv] - 2.0 * Un * ny;
  Gh[V::nw] = Bd[V::nw] - 2.0 * Un * nz;
  double D12 = 2.0 * Bdl * (double)modfm + (Bdl + Idl) * (double)modfm1;
  double D23 = (Bdl + Idl) * (double)modfm + 2.0 * Bdl * (double)modfm1;
  double Dist = Bdl * (double)modfm2;
#pragma unroll
  for (int s = 0; s <= V::np; ++s) {
    double In = P[s * NC + rc.ci];
    double p1 = (double)modfm * Gh[s] + (double)modfm1 * In;
    double p3 = (double)modfm * In + (double)modfm1 * Gh[s];
    double s32 = (p3 - Bd[s]) / D23, s21 = (Bd[s] - p1) / D12;
    Fa[s] = Bd[s] + limiter(lim, s32, s21) * Dist;
  }
  double rot = rho_sum<NS>(Fa), Rt = f_Rtot<NS>(T, Fa);
  double snd = f_ss<NS>(T, Fa, Fa[V::np], rot, Rt);
  Un = Fa[V::nu] * nx + Fa[V::nv] * ny + Fa[V::nw] * nz;
  double Vt0 = Fa[V::nu] - Un * nx, Vt1 = Fa[V::nv] - Un * ny, Vt2 = Fa[V::nw] - Un * nz;
  double Un1 = Un * (double)modfm2;
  double U10 = Un1 * nx + Vt0, U11 = Un1 * ny + Vt1, U12 = Un1 * nz + Vt2;
  double U40 = -Un1 * nx + Vt0, U41 = -Un1 * ny + Vt1, U42 = -Un1 * nz + Vt2;
  F5 F = riemann<NS>(T, rs, mco, false, Fa, U10, U11, U12, Fa[V::np], snd, rot, Fa, U40, U41, U42, Fa[V::np], snd, rot,
                     nx, ny, nz, Rt, Rt);
  e1[V::nu] = (double)modfm2 * F.u * area;
  e1[V::nv] = (double)modfm2 * F.v * area;
  e1[V::nw] = (double)modfm2 * F.w * area;
}

// Stress_Vector_Std (Mod_Fluid.f90:20-44)
__device__ __forceinline__ void stress_std(const double G[3][3], double nx, double ny, double nz, double mil, double mit,
                                           double* St) {
  double mi = mil + mit, lam = (-2.0 / 3.0) * mi;
  double Div = G[0][0] + G[1][1] + G[2][2], lamDiv = lam * Div;
  double Txx = lamDiv + 2.0 * mi * G[0][0], Tyy = lamDiv + 2.0 * mi * G[1][1], Tzz = lamDiv + 2.0 * mi * G[2][2];
  double Txy = mi * (G[0][1] + G[1][0]), Txz = mi * (G[0][2] + G[2][0]), Tyz = mi * (G[1][2] + G[2][1]);
  St[0] = Txx * nx + Txy * ny + Txz * nz;
  St[1] = Txy * nx + Tyy * ny + Tyz * nz;
  St[2] = Txz * nx + Tyz * ny + Tzz * nz;
}

template <int NS, int NR>
__device__ __forceinline__ void bc_wall_visc(const mose_dtab& T, const double* P, int64_t NC, const double* m,
                                             double prl, int typ, double Tw, const RecCells& rc, double nx, double ny,
                                             double nz, double area, double* e2) {
  using V = Vars<NS, NR>;
  int dir = rc.dir, modfm2 = 1 - 2 * rc.modfm, modfm3 = 2 * (2 * rc.modfm - 1);
  double Pb[V::nprim], M[3][3];
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) Pb[s] = P[s * NC + rc.c];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) M[a][c] = m[(int64_t)(a + 3 * c) * NC + rc.c];
  double rho = rho_sum<NS>(Pb), Rg = f_Rtot<NS>(T, Pb);
  double Tk = Pb[V::np] / (rho * Rg), mil, kl;
  if (typ == MOSE_BC_WALL_ADIABATIC) {
    int Ti = ftrunc(Tk);
    wilke<NS>(T, Pb, rho, Ti, Tk - Ti, mil, kl);
  } else {
    double rw = Pb[V::np] / (Rg * Tw), Pw[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) Pw[s] = Pb[s] * rw / rho;
    int Ti = ftrunc(Tw);
    double Td = Tw - Ti;
    wilke<NS>(T, Pw, rw, Ti, Td, mil, kl);
    if (prl > 0.0) kl = mil * f_cp<NS>(T, Pw, Ti, Td, rw) / prl;
  }
  double Gr[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}}, G2[3][3];
#pragma unroll
  for (int s = 0; s < 3; ++s) Gr[s][dir - 1] = Pb[V::nu + s] * (double)modfm3;
#pragma unroll
  for (int kk = 0; kk < 3; ++kk)
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) G2[ii][kk] = Gr[ii][0] * M[0][kk] + Gr[ii][1] * M[1][kk] + Gr[ii][2] * M[2][kk];
  double St[3];
  stress_std(G2, nx, ny, nz, mil, 0.0, St);
#pragma unroll
  for (int q = 0; q < 3; ++q) e2[V::nu + q] = -((double)modfm2 * area * St[q]);
  if (typ == MOSE_BC_WALL_ISOTHERMAL) {
    double Gd = (Tk - Tw) * (double)modfm3;
    double g0 = Gd * M[dir - 1][0], g1 = Gd * M[dir - 1][1], g2 = Gd * M[dir - 1][2];
    double Fnp = area * kl * (g0 * nx + g1 * ny + g2 * nz);
    e2[V::np] = -((double)modfm2 * Fnp);
  }
  if (NR > 0) {
    double gd = Pb[V::nt] / rho * (double)modfm3;
    double g0 = gd * M[dir - 1][0], g1 = gd * M[dir - 1][1], g2 = gd * M[dir - 1][2];
    double Fr = area * mil * (g0 * nx + g1 * ny + g2 * nz) / SA_SIGMA;
    e2[V::nt] = -((double)modfm2 * Fr);
  }
}

// BC_Symmetry_Visc (Lib_BC.f90:741-861): the viscous stress of a symmetry face.  The velocity gradient takes the cell
// and its mirror image across the face along the normal, and along each transverse direction the two neighbours and
// their mirror images (stencil of the Fm cases :783-802).  sb, sc: the row strides of the transverse directions b, c.
template <int NS, int NR>
__device__ __forceinline__ void bc_sym_visc(const mose_dtab& T, const double* P, int64_t NC, const double* m,
                                            const RecCells& rc, int64_t sb, int64_t sc, double nx, double ny,
                                            double nz, double area, double* e2) {
  using V = Vars<NS, NR>;
  const int modfm2 = 1 - 2 * rc.modfm;
  const int a = rc.dir - 1, b = rc.dir == 1 ? 1 : 0, c = rc.dir == 3 ? 1 : 2;
  double Pb[V::nprim], M[3][3];
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) Pb[s] = P[s * NC + rc.c];
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int y = 0; y < 3; ++y) M[x][y] = m[(int64_t)(x + 3 * y) * NC + rc.c];
  double rho = rho_sum<NS>(Pb), Rg = f_Rtot<NS>(T, Pb);
  double Tk = Pb[V::np] / (rho * Rg), mil, kl;
  int Ti = ftrunc(Tk);
  wilke<NS>(T, Pb, rho, Ti, Tk - Ti, mil, kl);
  // Vel3, Vel4 (direction b) and Vel7, Vel8 (direction c), then the mirror images Vel5, Vel6, Vel9, Vel10
  const int64_t nbr[4] = {rc.c - sb, rc.c + sb, rc.c - sc, rc.c + sc};
  double v1[3], v2[3], vn[4][3], vm[4][3];
#pragma unroll
  for (int q = 0; q < 3; ++q) v1[q] = Pb[V::nu + q];
#pragma unroll
  for (int n = 0; n < 4; ++n)
#pragma unroll
    for (int q = 0; q < 3; ++q) vn[n][q] = P[(V::nu + q) * NC + nbr[n]];
  const double d1 = v1[0] * nx + v1[1] * ny + v1[2] * nz;
  v2[0] = v1[0] - 2.0 * d1 * nx;
  v2[1] = v1[1] - 2.0 * d1 * ny;
  v2[2] = v1[2] - 2.0 * d1 * nz;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const double d = vn[n][0] * nx + vn[n][1] * ny + vn[n][2] * nz;
    vm[n][0] = vn[n][0] - 2.0 * d * nx;
    vm[n][1] = vn[n][1] - 2.0 * d * ny;
    vm[n][2] = vn[n][2] - 2.0 * d * nz;
  }
  double Gr[3][3], G2[3][3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    Gr[q][a] = (v2[q] - v1[q]) * (double)modfm2;
    Gr[q][b] = (vn[1][q] - vn[0][q] + vm[1][q] - vm[0][q]) * 0.25;
    Gr[q][c] = (vn[3][q] - vn[2][q] + vm[3][q] - vm[2][q]) * 0.25;
  }
#pragma unroll
  for (int kk = 0; kk < 3; ++kk)
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) G2[ii][kk] = Gr[ii][0] * M[0][kk] + Gr[ii][1] * M[1][kk] + Gr[ii][2] * M[2][kk];
  double mie = 0.0;
  if (NR > 0) {   // Spalart_Eddy_Viscosity (Mod_Spalart.f90:382-398)
    double chi = Pb[V::nt] / mil, c3 = chi * chi * chi;
    mie = c3 / (c3 + SA_CV1 * SA_CV1 * SA_CV1) * Pb[V::nt];
  }
  double St[3];
  stress_std(G2, nx, ny, nz, mil, mie, St);
  const double u0 = 0.5 * (v1[0] + v2[0]), u1 = 0.5 * (v1[1] + v2[1]), u2 = 0.5 * (v1[2] + v2[2]);
#pragma unroll
  for (int q = 0; q < 3; ++q) e2[V::nu + q] = -((double)modfm2 * area * St[q]);
  e2[V::np] = -((double)modfm2 * area * (u0 * St[0] + u1 * St[1] + u2 * St[2]));
}

// the flux of a far-field face from its face state (the tails of BC_Outflow :1218-1242 and BC_Subsonic_Inflow
// :1489-1511; sub selects the operand order of the latter)
template <int NS, int NR>
__device__ __forceinline__ void ff_tail(const mose_dtab& T, const double* Fa, double Frho, double FR, double FUn,
                                        double nx, double ny, double nz, double A, bool sub, double* Fl) {
  using V = Vars<NS, NR>;
  double He = 0.0, FT = Fa[V::np] / (FR * Frho);
  int Ti = ftrunc(FT);
  double Td = FT - Ti;
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) Fl[s] = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    He = He + Fa[s] / Frho * tabT(T.h, s, Ti, Td);
    Fl[s] = sub ? Fa[s] * FUn * A : Fa[s] * A * FUn;
  }
  double Fm = sub ? Frho * FUn * A : Frho * A * FUn;
  if (sub) {
    Fl[V::nu] = Fm * Fa[V::nu] + Fa[V::np] * A * nx;
    Fl[V::nv] = Fm * Fa[V::nv] + Fa[V::np] * A * ny;
    Fl[V::nw] = Fm * Fa[V::nw] + Fa[V::np] * A * nz;
  } else {
    Fl[V::nu] = Fa[V::nu] * Fm + Fa[V::np] * A * nx;
    Fl[V::nv] = Fa[V::nv] * Fm + Fa[V::np] * A * ny;
    Fl[V::nw] = Fa[V::nw] * Fm + Fa[V::np] * A * nz;
  }
  double vt2 = Fa[V::nu] * Fa[V::nu] + Fa[V::nv] * Fa[V::nv] + Fa[V::nw] * Fa[V::nw];
  Fl[V::np] = Fm * (0.5 * vt2 + He);
#pragma unroll
  for (int s = V::np + 1; s < V::nprim; ++s) Fl[s] = Fa[s] / Frho * Fm;
}

// BC_Inflow_Outflow.  Returns false when the face closes to a reflective wall (error = 1).
template <int NS, int NR>
__device__ bool bc_farfield(const mose_dtab& T, const double* P, int64_t NC, int lim, int rs, double mco,
                            const mose_farfield& ff, const double* tv, const RecCells& rc, double nx, double ny,
                            double nz, double A, double* e1) {
  using V = Vars<NS, NR>;
  constexpr int NP = V::nprim;
  int modfm = rc.modfm, modfm1 = 1 - modfm, modfm2 = 1 - 2 * modfm;
  double Bd[NP], Fl[NP], Fa[NP];
#pragma unroll
  for (int s = 0; s < NP; ++s) Bd[s] = P[s * NC + rc.c];
  double Sg = (double)modfm2;
  double Un = Bd[V::nu] * nx + Bd[V::nv] * ny + Bd[V::nw] * nz;
  double rho = rho_sum<NS>(Bd), Rg = f_Rtot<NS>(T, Bd);
  double snd = f_ss<NS>(T, Bd, Bd[V::np], rho, Rg);
  double Tk = Bd[V::np] / (rho * Rg);
  int Ti = static_cast<int>(Tk);
  double cp = f_cp<NS>(T, Bd, Ti, Tk - Ti, rho);
  double G = cp / (cp - Rg);
  double Riem = snd + Sg * 0.5 * (G - 1.0) * Un;
  double InOut = Sg * Un;
  if (ff.mach < 0.0) InOut = -1.0;
  if (InOut >= 0.0) {
    if (ff.pamb == -3.14 || ff.pamb == -2.71) return false;
    // BC_Outflow
    double Vt0 = Bd[V::nu] - Un * nx, Vt1 = Bd[V::nv] - Un * ny, Vt2 = Bd[V::nw] - Un * nz;
    double MN = fabs(Un / snd), pe = ff.pamb, FUn, Frho, FR;
    if (MN < 1.0 && ff.pamb > 0.0) {
      double dlt = 0.5 * (G - 1.0);
      double den = 1.0 / pow(1.0 + dlt, G / (G - 1.0));
      double pst = Bd[V::np] * den * pow(1.0 + dlt * (MN * MN), G / (G - 1.0));
      if (pst >= ff.pamb) pe = pst;
      double Is = (G - 1.0) / (2.0 * G);
      double Fs = snd * pow(pe / Bd[V::np], Is);
      FUn = (Riem - Fs) / (Sg * 0.5 * (G - 1.0));
      Fa[V::np] = pe;
      Frho = G * Fa[V::np] / (Fs * Fs);
#pragma unroll
      for (int s = 0; s < NS; ++s) Fa[s] = Bd[s] / rho * Frho;
      Fa[V::nu] = Vt0 + FUn * nx;
      Fa[V::nv] = Vt1 + FUn * ny;
      Fa[V::nw] = Vt2 + FUn * nz;
#pragma unroll
      for (int s = V::np + 1; s < NP; ++s) Fa[s] = Bd[s] / rho * Frho;
      FR = Rg;
    } else {
#pragma unroll
      for (int s = 0; s < NP; ++s) Fa[s] = Bd[s];
      FUn = Un; Frho = rho; FR = Rg;
    }
    ff_tail<NS, NR>(T, Fa, Frho, FR, FUn, nx, ny, nz, A, false, Fl);
  } else if (ff.mach > 1.0 || ff.pamb == -2.71) {
    // BC_Supersonic_Inflow
    double fR = f_Rtot<NS>(T, Bd);
    double v2 = Bd[V::nu] * Bd[V::nu] + Bd[V::nv] * Bd[V::nv] + Bd[V::nw] * Bd[V::nw];
    double fsnd = f_ss<NS>(T, Bd, Bd[V::np], rho_sum<NS>(Bd), fR);
    double fM = sqrt(v2) / fsnd;
    double rel = ff.rel_fac;
    double Ml = rel * ff.mach + (1.0 - rel) * fM;
    double p0l = rel * ff.p0 + (1.0 - rel) * Bd[V::np];
    double Sup[NP], SR = 0.0;
    Sup[V::np] = p0l;
#pragma unroll
    for (int s = 0; s < NS; ++s) SR = SR + ff.ci[s] * T.Ri[s];
    double Srho = Sup[V::np] / (SR * ff.t0);
#pragma unroll
    for (int s = 0; s < NS; ++s) Sup[s] = Srho * ff.ci[s];
    double Ssnd = f_ss<NS>(T, Sup, Sup[V::np], Srho, SR);
    Sup[V::nu] = Ml * Ssnd * cos(ff.alpha) * cos(ff.beta);
    Sup[V::nv] = Ml * Ssnd * sin(ff.alpha) * cos(ff.beta);
    Sup[V::nw] = Ml * Ssnd * sin(ff.beta);
#pragma unroll
    for (int s = 0; s < NR; ++s) Sup[V::nt + s] = Srho * ff.rans[s];
    double L[NP], R[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      double In = P[s * NC + rc.ci];
      double p1 = (double)modfm * Sup[s] + (double)modfm1 * In;
      double p3 = (double)modfm * In + (double)modfm1 * Sup[s];
      double Sl = limiter(lim, p3 - Bd[s], Bd[s] - p1);
      double Fc = Bd[s] + 0.5 * Sl * (double)modfm2;
      L[s] = (double)modfm * Sup[s] + (double)modfm1 * Fc;
      R[s] = (double)modfm * Fc + (double)modfm1 * Sup[s];
    }
    double rL = rho_sum<NS>(L), RL = f_Rtot<NS>(T, L), rR = rho_sum<NS>(R), RR = f_Rtot<NS>(T, R);
    double aL = f_ss<NS>(T, L, L[V::np], rL, RL), aR = f_ss<NS>(T, R, R[V::np], rR, RR);
    F5 F = riemann<NS>(T, rs, mco, false, L, L[V::nu], L[V::nv], L[V::nw], L[V::np], aL, rL, R, R[V::nu], R[V::nv],
                       R[V::nw], R[V::np], aR, rR, nx, ny, nz, RL, RR);
    double su = copysign(1.0, F.r), SL = (1.0 + su) / 2.0, SR_ = (su - 1.0) / 2.0, Fm = F.r * A;
#pragma unroll
    for (int s = 0; s < NP; ++s) Fl[s] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) Fl[s] = Fm * (SL * L[s] / rL - SR_ * R[s] / rR);
    Fl[V::nu] = F.u * A; Fl[V::nv] = F.v * A; Fl[V::nw] = F.w * A; Fl[V::np] = F.e * A;
#pragma unroll
    for (int s = V::np + 1; s < NP; ++s) Fl[s] = Fm * (SL * L[s] / rL - SR_ * R[s] / rR);
  } else if (ff.mach < 1.0) {
    // BC_Subsonic_Inflow
    double b0 = ny * tv[2] - nz * tv[1], b1 = nz * tv[0] - nx * tv[2], b2 = nx * tv[1] - ny * tv[0];
    double bm = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
    b0 = b0 / bm; b1 = b1 / bm; b2 = b2 / bm;
    double al, be;
    if (ff.alpha == 0.0 && ff.beta == 0.0) {
      al = atan(ny / (nx + 1e-20));
      be = atan(nz / (nx + 1e-20));
    } else {
      al = ff.alpha; be = ff.beta;
    }
    double ta = tan(al), tb = tan(be);
    double XA = ny - nx * ta, XB = tv[1] - tv[0] * ta, XC = b1 - b0 * ta;
    double XD = nz - nx * tb, XE = tv[2] - tv[0] * tb, XF = b2 - b0 * tb;
    double chk = (XC * XE - XF * XB);
    double E1 = (XF * XA - XC * XD) / chk, E2 = (XD * XB - XE * XA) / chk;
    double Frho, FR, FUn;
#pragma unroll
    for (int s = 0; s < NP; ++s) Fa[s] = 0.0;
    if (ff.mach >= 0.0) {
      double Discr = -1.0, p0t = ff.p0, S0 = 0.0, CD = 0.0, YA = 0.0, YB = 0.0;
      int it = 0;
      FR = 0.0;
      while (Discr < 0.0) {
        it = it + 1;
        FR = 0.0;
#pragma unroll
        for (int s = 0; s < NS; ++s) FR = FR + ff.ci[s] * T.Ri[s];
        S0 = sqrt(G * FR * ff.t0);
        CD = (1.0 + E1 * E1 + E2 * E2) * (G - 1.0) / 2.0;
        double CC = pow(p0t / Bd[V::np], (G - 1.0) / G);
        double sr = snd / S0;
        CC = CC * (sr * sr);
        double gh = (G - 1.0) / 2.0;
        YA = gh * gh + CC * CD;
        YB = Sg * Riem * ((G - 1.0) / 2.0);
        double YC = Riem * Riem - CC * (S0 * S0);
        Discr = YB * YB - YA * YC;
        if (Discr < 0.0 && fabs(Discr) < 1e-2) Discr = 0.0;
        if (Discr < 0.0 && fabs(Discr) > 1e-2) {
          if (ff.rel_fac == 1.0) { Discr = 0.0; break; }
          p0t = ff.rel_fac * p0t + (1.0 - ff.rel_fac) * Bd[V::np];
        }
        if (it > 1000) { Discr = 0.0; break; }
      }
      double UnS = copysign(1.0, Un);
      double Un3 = (YB + UnS * sqrt(Discr)) / YA;
      double S3 = sqrt(S0 * S0 - CD * (Un3 * Un3));
      double p3 = p0t * pow(S3 / S0, 2.0 * G / (G - 1.0));
      double Ut3 = Un3 * E1, Ub3 = Un3 * E2;
      Fa[V::np] = p3;
      Frho = G * Fa[V::np] / (S3 * S3);
#pragma unroll
      for (int s = 0; s < NS; ++s) Fa[s] = ff.ci[s] * Frho;
      Fa[V::nu] = Un3 * nx + Ut3 * tv[0] + Ub3 * b0;
      Fa[V::nv] = Un3 * ny + Ut3 * tv[1] + Ub3 * b1;
      Fa[V::nw] = Un3 * nz + Ut3 * tv[2] + Ub3 * b2;
      FUn = Un3;
    } else {
      double md = ff.p0, p3 = Bd[V::np], R3 = 0.0, cp3 = 0.0;
      int Tj = static_cast<int>(ff.t0);
      double Tdj = ff.t0 - Tj;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        cp3 = cp3 + ff.ci[s] * tabT(T.cp, s, Tj, Tdj);
        R3 = R3 + ff.ci[s] * T.Ri[s];
      }
      if (ff.mach != -15.0) {
        double T3;
        if (ff.mach == -10.0) {
          double q = md * R3 / p3, q2 = p3 / (md * R3);
          T3 = (-cp3 + sqrt(cp3 * cp3 + 2.0 * cp3 * ff.t0 * (q * q))) * (q2 * q2);
        } else {
          T3 = ff.t0;
        }
        double r3 = p3 / (R3 * T3), Un3 = md / r3, Ut3 = Un3 * E1, Ub3 = Un3 * E2;
        FR = R3; Frho = r3;
#pragma unroll
        for (int s = 0; s < NS; ++s) Fa[s] = Frho * ff.ci[s];
        Fa[V::nu] = Un3 * nx + Ut3 * tv[0] + Ub3 * b0;
        Fa[V::nv] = Un3 * ny + Ut3 * tv[1] + Ub3 * b1;
        Fa[V::nw] = Un3 * nz + Ut3 * tv[2] + Ub3 * b2;
        Fa[V::np] = p3;
        FUn = Un3;
      } else {
        FR = R3;
        Frho = p3 / (FR * ff.t0);
#pragma unroll
        for (int s = 0; s < NS; ++s) Fa[s] = Frho * ff.ci[s];
        Fa[V::nu] = md / Frho; Fa[V::nv] = 0.0; Fa[V::nw] = Fa[V::nu] * 0.25; Fa[V::np] = p3;
        FUn = Fa[V::nu] * nx + Fa[V::nv] * ny + Fa[V::nw] * nz;
      }
    }
#pragma unroll
    for (int s = 0; s < NR; ++s) Fa[V::nt + s] = ff.rans[s] * Frho;
    ff_tail<NS, NR>(T, Fa, Frho, FR, FUn, nx, ny, nz, A, true, Fl);
  } else {
    return true;   // Mach exactly 1 is refused when the case is created
  }
#pragma unroll
  for (int s = 0; s < NP; ++s) e1[s] = (double)modfm2 * Fl[s];
  return true;
}

// SA_Source_Blk + sa_body.inc for one interior cell: the increment -Source of R(nt).  nb[q] (q = 2d + {0: minus,
// 1: plus}) are the six neighbour rows; wall[q] marks a neighbour that is the wall ghost of Extrapolate_Wall_blk
// (species and p copied, velocity and rho nu~ negated, taken from the cell itself).
template <int NS, int NR>
__device__ __forceinline__ double sa_source(const double* P, int64_t NC, const double* m, int64_t c0,
                                            const int64_t nb[6], const bool wall[6], double rho, double mil,
                                            double dist, double vol, bool pimp, double dt) {
  using V = Vars<NS, NR>;
  const double cw1 = SA_CB1 / (SA_CK * SA_CK) + (1.0 + SA_CB2) / SA_SIGMA, cK2 = SA_CK * SA_CK;
  const double cv13 = SA_CV1 * SA_CV1 * SA_CV1, cw36 = 64.0;
  double nt0 = P[V::nt * NC + c0];
  double nit = nt0 / rho, chi = nt0 / mil;
  double M[3][3], Gt[3][3], Gv[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) M[a][c] = m[(int64_t)(a + 3 * c) * NC + c0];
  double ntv[6], sum[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    int64_t r = wall[q] ? c0 : nb[q];
    double sg = wall[q] ? -1.0 : 1.0;
    ntv[q] = wall[q] ? -P[V::nt * NC + r] : P[V::nt * NC + r];
    double ss = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) ss = ss + P[s * NC + r];
    sum[q] = ss;
    (void)sg;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double vp = P[(V::nu + a) * NC + (wall[2 * d + 1] ? c0 : nb[2 * d + 1])];
      double vm = P[(V::nu + a) * NC + (wall[2 * d] ? c0 : nb[2 * d])];
      if (wall[2 * d + 1]) vp = -vp;
      if (wall[2 * d]) vm = -vm;
      Gt[a][d] = (vp - vm) * 0.5;
    }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int a = 0; a < 3; ++a) Gv[a][c] = Gt[a][0] * M[0][c] + Gt[a][1] * M[1][c] + Gt[a][2] * M[2][c];
  double v1 = Gv[2][1] - Gv[1][2], v2 = Gv[0][2] - Gv[2][0], v3 = Gv[1][0] - Gv[0][1];
  double Om = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
  double chi3 = chi * chi * chi;
  double fv1 = chi3 / (chi3 + cv13);
  double fv2 = 1.0 - chi / (1.0 + chi * fv1);
  double Sbar = nit / (cK2 * (dist * dist)) * fv2, St;
  if (Sbar >= (-SA_C2 * Om)) St = Om + Sbar;
  else St = Om + Om * (SA_C2 * SA_C2 * Om + SA_C3 * Sbar) / ((SA_C3 - 2.0 * SA_C2) * Om - Sbar);
  if (St == 0.0) St = 1e-40;
  double Prod = SA_CB1 * St * nit;
  double Gn[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) Gn[d] = (ntv[2 * d + 1] / sum[2 * d + 1] - ntv[2 * d] / sum[2 * d]) * 0.5;
  double Gr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) Gr[c] = Gn[0] * M[0][c] + Gn[1] * M[1][c] + Gn[2] * M[2][c];
  double Gr2 = Gr[0] * Gr[0] + Gr[1] * Gr[1] + Gr[2] * Gr[2];
  double Diff = SA_CB2 / SA_SIGMA * Gr2;
  double r_ = nit / (St * cK2 * (dist * dist));
  r_ = dmin2(r_, 10.0);
  double r2 = r_ * r_, r6 = r2 * r2 * r2;
  double g = r_ + SA_CW2 * (r6 - r_);
  double g2 = g * g, g6 = g2 * g2 * g2;
  double fw = g * pow((1.0 + cw36) / (g6 + cw36), 1.0 / 6.0);
  double nd = nit / dist;
  double Dest = cw1 * fw * (nd * nd);
  double Src = rho * (Prod + Diff - Dest) * vol;
  if (pimp) {
    double fnu = 1.0 / (1.0 + dt * 2.0 * cw1 * fw * nit / (dist * dist));
    Src = Src * fnu;
  }
  return Src;
}

}  // namespace mose
