// The performance formulation of the fused stage (MOSE_PATH_FAST): the same discretisation as the parity kernels, in
// fewer operations.  The translation unit that instantiates it (mose_stage_fast.hip) is compiled with fused
// multiply-add contraction and reciprocal multiplication; the parity kernels keep the reference's roundings.  The
// results differ from the parity path by rounding only: the tests hold every parity case of the fused path to a
// relative L-infinity <= 1e-12 on this path too.
//
// What changes:
//   half_slope_fast  the van Leer and van Albada limiters with one division in place of two (below)
//   conv_face_fast  LMRoe (riemann_open_body.inc:965-1090) and its Roe averages (:679-726) evaluated once per face:
//                   one reciprocal per side density, the face's species fractions, enthalpies and mixture gas
//                   constant shared by the averages and the physical fluxes (the reference recomputes them in
//                   h_open), and the two shear waves summed as rho |u_n| (du - du_n n): al2 r2 + al3 r3 over the
//                   orthonormal tangents t1, t2 is that vector, so the tangent frame (two square roots, six
//                   divisions) is not built.  With Mco >= 1 the low-Mach factor is Mco itself (max(Mco, min(1, Ma)))
//                   and Ma is not formed.  The other Riemann solvers run their parity formulas, contracted.
//   face_thermo_fast  one species: Wilke's mixture rule returns the species' own mu and k (phi_11 = 1 up to the
//                   1e-20 of co_fiij), so the table values are used directly.
#pragma once

#include "mose_physics.hpp"

namespace mose {

// The limiters of Lib_Limiters.f90:65-97 with one division: the ratio forms phi(ru) x, ru = y / (x + 1e-30), of van
// Leer and van Albada multiplied through by the denominator's |x'| or x'^2 (x' = x + 1e-30):
//   van Leer    (ru + |ru|) / (1 + |ru|) x = (y sgn(x') + |y|) / (|x'| + |y|) x
//   van Albada  (ru^2 + ru) / (1 + ru^2) x = (y^2 + y x') / (x'^2 + y^2) x
// minmod has no division; MC and superbee keep their one division.
__device__ __forceinline__ double limiter_fast(int lim, double x, double y) {
  if (lim == 2) {
    const double xp = x + 1e-30;
    return ((xp < 0.0 ? -y : y) + fabs(y)) / (fabs(xp) + fabs(y)) * x;
  }
  if (lim == 3) {
    const double xp = x + 1e-30;
    return (y * y + y * xp) / (xp * xp + y * y) * x;
  }
  return limiter(lim, x, y);
}

// half_slope (mose_physics.hpp) with limiter_fast
__device__ __forceinline__ double half_slope_fast(int lim, double qm, double q0, double qp, double idlm, double idlp,
                                                  double hdl) {
  const double s0 = (q0 - qm) * idlm, s1 = (qp - q0) * idlp;
  return limiter_fast(lim, s1, s0) * hdl;
}

template <int NS, int NR>
__device__ __forceinline__ void conv_face_lmroe_fast(const mose_dtab& T, double mco, double* L, double* R, double nx,
                                                     double ny, double nz, double area, double* res) {
  using V = Vars<NS, NR>;
#pragma unroll
  for (int s = 0; s < NS; ++s) { L[s] = dmax2(L[s], 1e-20); R[s] = dmax2(R[s], 1e-20); }
  const double rL = rho_sum<NS>(L), rR = rho_sum<NS>(R);
  const double irL = 1.0 / rL, irR = 1.0 / rR;
  double YL[NS], YR[NS], RL = 0.0, RR = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    YL[s] = L[s] * irL;
    YR[s] = R[s] * irR;
    RL = RL + T.Ri[s] * YL[s];
    RR = RR + T.Ri[s] * YR[s];
  }
  const double ul = L[V::nu], vl = L[V::nv], wl = L[V::nw], pl = L[V::np];
  const double ur = R[V::nu], vr = R[V::nv], wr = R[V::nw], pr = R[V::np];
  const double Tl = pl * irL / RL, Tr = pr * irR / RR;
  int Til = ftrunc(Tl), Tir = ftrunc(Tr);
  Til = Til < 0 ? 0 : (Til > NTMAX - 1 ? NTMAX - 1 : Til);
  Tir = Tir < 0 ? 0 : (Tir > NTMAX - 1 ? NTMAX - 1 : Tir);
  const double dTl = Tl - Til, dTr = Tr - Tir;
  // Roe averages with the weights sqrt(rho_side) / (sqrt(rho_L) + sqrt(rho_R))
  const double sL = sqrt(rL), sR = sqrt(rR), isr = 1.0 / (sL + sR);
  const double wL = sL * isr, wR = sR * isr;
  const double rho = sL * sR;
  const double u = wR * ur + wL * ul, v = wR * vr + wL * vl, w = wR * wr + wL * wl;
  const double Rroe = wR * RR + wL * RL, Troe = wR * Tr + wL * Tl;
  double h0l = 0.0, h0r = 0.0, cv = 0.0, sei = 0.0;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const double* ht = T.h + q * NTAB;
    const double* ct = T.cp + q * NTAB;
    const double hl = ht[Til] + (ht[Til + 1] - ht[Til]) * dTl, hr = ht[Tir] + (ht[Tir + 1] - ht[Tir]) * dTr;
    const double cpl = ct[Til] + (ct[Til + 1] - ct[Til]) * dTl, cpr = ct[Tir] + (ct[Tir + 1] - ct[Tir]) * dTr;
    h0l = h0l + hl * YL[q];
    h0r = h0r + hr * YR[q];
    const double dro = wR * YR[q] + wL * YL[q];
    cv = cv + dro * (0.5 * (cpl + cpr) - T.Ri[q]);
    sei = sei + dro * (wR * (hr - T.Ri[q] * Tr) + wL * (hl - T.Ri[q] * Tl));
  }
  h0l = h0l + 0.5 * (ul * ul + vl * vl + wl * wl);
  h0r = h0r + 0.5 * (ur * ur + vr * vr + wr * wr);
  const double h0 = wR * h0r + wL * h0l;
  const double vel2 = u * u + v * v + w * w;
  const double a = sqrt(Rroe / cv * (h0 - 0.5 * vel2 + cv * Troe - sei));
  // LMRoe in the face's normal frame
  const double nm2 = nx * nx + ny * ny + nz * nz;
  if (!(nm2 > 0.0)) {
#pragma unroll
    for (int s = 0; s < V::nprim; ++s) res[s] = 0.0;
    return;
  }
  const double inm = 1.0 / sqrt(nm2);
  const double n1 = nx * inm, n2 = ny * inm, n3 = nz * inm;
  const double unL = ul * n1 + vl * n2 + wl * n3, unR = ur * n1 + vr * n2 + wr * n3;
  const double mL = rL * unL, mR = rR * unR;
  const double un = u * n1 + v * n2 + w * n3;
  double phi = mco;
  if (!(mco >= 1.0)) {
    const double ut = sqrt(dmax2(0.0, vel2 - un * un));
    const double Ma = (a > 0.0) ? (fabs(un) + ut) / a : 1.0;
    phi = dmax2(mco, dmin2(1.0, Ma));
  }
  const double dp = pr - pl, dun = unR - unL;
  const double dutx = (ur - ul) - dun * n1, duty = (vr - vl) - dun * n2, dutz = (wr - wl) - dun * n3;
  const double ia = 1.0 / a, ia2 = ia * ia;
  const double pa = dp * ia2, ua = rho * phi * dun * ia;
  const double c1 = fabs(un - a) * (0.5 * (pa - ua)), c5 = fabs(un + a) * (0.5 * (pa + ua));
  const double c4 = fabs(un) * ((rR - rL) - pa), cs = fabs(un) * rho;
  const double c15 = c1 + c5, cd = (c5 - c1) * a;
  const double d0 = c15 + c4;
  const double du = d0 * u + cd * n1 + cs * dutx;
  const double dv = d0 * v + cd * n2 + cs * duty;
  const double dw = d0 * w + cd * n3 + cs * dutz;
  const double de = c15 * h0 + cd * un + c4 * (0.5 * vel2) + cs * (u * dutx + v * duty + w * dutz);
  F5 F;
  F.r = 0.5 * (mL + mR) - 0.5 * d0;
  F.u = 0.5 * ((mL * ul + pl * n1) + (mR * ur + pr * n1)) - 0.5 * du;
  F.v = 0.5 * ((mL * vl + pl * n2) + (mR * vr + pr * n2)) - 0.5 * dv;
  F.w = 0.5 * ((mL * wl + pl * n3) + (mR * wr + pr * n3)) - 0.5 * dw;
  F.e = 0.5 * (mL * h0l + mR * h0r) - 0.5 * de;
  // flux assembly (conv_physics.inc:168-179)
  const double su = copysign(0.5, F.r), sp = 0.5 + su, sm = su - 0.5;
  const double fa = F.r * area;
#pragma unroll
  for (int s = 0; s < NS; ++s) res[s] = fa * (sp * YL[s] - sm * YR[s]);
  res[V::nu] = F.u * area; res[V::nv] = F.v * area; res[V::nw] = F.w * area; res[V::np] = F.e * area;
#pragma unroll
  for (int s = V::np + 1; s < V::nprim; ++s) res[s] = fa * (sp * L[s] * irL - sm * R[s] * irR);
}

// The face flux of the performance path: LMRoe in its short form, the other solvers as the parity path states them.
template <int NS, int NR, int RS>
__device__ __forceinline__ void conv_face_fast(const mose_dtab& T, int rs, double mco, double* L, double* R, double nx,
                                               double ny, double nz, double area, double* res) {
  if (RS == MOSE_RS_LMROE || (RS == 0 && rs == MOSE_RS_LMROE))
    conv_face_lmroe_fast<NS, NR>(T, mco, L, R, nx, ny, nz, area, res);
  else
    conv_face_lr<NS, NR, RS>(T, rs, mco, L, R, nx, ny, nz, area, res);
}

// Diffusive face thermo of the performance path: one species takes mu and k from its tables (Wilke's rule reduces
// to them), several species run the parity path's mixture rule.
template <int NS>
__device__ __forceinline__ FaceThermo face_thermo_fast(const mose_dtab& T, const double* prim, double p, double prl) {
  if constexpr (NS == 1) {
    FaceThermo o;
    const double rho = prim[0];
    const double Tk = p / (rho * T.Ri[0]);
    const int Ti = ftrunc(Tk);
    const double Td = Tk - Ti;
    const double cp = tabT(T.cp, 0, Ti, Td);
    const double mil = tabT(T.mi, 0, Ti, Td);
    o.rho = rho; o.T = Tk; o.mil = mil; o.cp = cp; o.Ti = Ti; o.Td = Td;
    o.kl = prl > 0.0 ? mil * cp / prl : tabT(T.k, 0, Ti, Td);
    return o;
  } else {
    return face_thermo<NS>(T, prim, p, prl);
  }
}

}  // namespace mose
