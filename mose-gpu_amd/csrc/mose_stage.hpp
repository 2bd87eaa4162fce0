// The fused stage kernel: one workgroup per tile of TI x TJ x TK cells of any block of the case: 512 threads for a
// tile of 256 or 512 cells, 256 threads for a tile of 128 (two threads per cell of the smaller tiles).  The face and
// half-slope loops spread over all the threads; the per-cell phases run on the first TI TJ TK.
//
//   phase A  every interior face of the tile (TI+1)TJTK + TI(TJ+1)TK + TITJ(TK+1): the convective flux into LDS;
//            the per-cell (rho, 1/rho, T) and the primitives of the tile and its one-cell halo into LDS (viscous
//            cases), so that the viscous faces read their stencils from LDS
//   phase B  each cell sums its six convective faces in the fixed order of pass2_body.inc
//   phase C  every interior face of the tile: the viscous flux into the same LDS buffer
//   phase D  each cell subtracts its six viscous faces (pass2_body.inc), adds the boundary-record increments in face
//            order 1..6, and applies the RK stage update (newstate_cons_body.inc) into the output buffer;
//            on the last stage it also takes the CFL/VNN limit of its new state for the next step's dt.
//
// Each face is computed once per tile (the faces on a tile edge twice: once by each tile), no face buffer goes to
// HBM, and every sum keeps the reference's order, so the result is bitwise that of the per-face kernels.
// The tiles of a case (of every block, of any size) are listed in device memory (block, i0, j0, k0) and served by
// one launch per stage; the workgroup -> tile map deals a contiguous run of tiles to each XCD (k_stage below).
#pragma once

#include "mose_kernels.hpp"

namespace mose {

struct Tile {
  int b, i0, j0, k0;
};

struct StageArgs {
  const Tile* tiles;
  int ntiles;
  int TI, TJ, TK;
  int nslot;          // faces per tile (LDS slots)
  int nhs;            // half-slope slots: (TI+2)TJTK + TI(TJ+2)TK + TITJ(TK+2)
  int ncache;         // (TI+2)(TJ+2)(TK+2) cells of the viscous cell cache
  int r0;             // doubles of LDS region 0 (half slopes, then the viscous cell cache)
  int last;           // 1 on the last RK stage: take the dt limit of the new state
  double rkc;
  unsigned long long* dtnext;
  double* rpart;      // last stage: per-tile sums of the squared step changes [ntiles][nprim - NS + 1]
};

// The index geometry of a tile.  STI/STJ/STK > 0 fix the shape at compile time (the decode of slot indices then
// divides by constants: multiply and shift instead of the integer division by a run-time value), 0 takes it from the
// launch arguments.  r0 repeats the host's sizing of LDS region 0 (stage_r0).
struct TileGeo {
  int TI, TJ, TK, HI, HJ, nhs, NI, NJ, nslot, CI, CJ, ncache, r0;
};
template <int NP, bool VISC, int STI, int STJ, int STK>
__device__ __forceinline__ TileGeo tile_geo(const StageArgs& A) {
  TileGeo g;
  g.TI = STI ? STI : A.TI;
  g.TJ = STJ ? STJ : A.TJ;
  g.TK = STK ? STK : A.TK;
  g.HI = (g.TI + 2) * g.TJ * g.TK;
  g.HJ = g.TI * (g.TJ + 2) * g.TK;
  g.nhs = g.HI + g.HJ + g.TI * g.TJ * (g.TK + 2);
  g.NI = (g.TI + 1) * g.TJ * g.TK;
  g.NJ = g.TI * (g.TJ + 1) * g.TK;
  g.nslot = g.NI + g.NJ + g.TI * g.TJ * (g.TK + 1);
  g.CI = g.TI + 2;
  g.CJ = g.TJ + 2;
  g.ncache = g.CI * g.CJ * (g.TK + 2);
  int r0 = g.nhs * NP;
  if (VISC) r0 = r0 > (3 + NP) * g.ncache ? r0 : (3 + NP) * g.ncache;
  g.r0 = (r0 + 31) & ~31;
  return g;
}

// One interior viscous face of direction D (diff_cuf_i/j/k + diff_physics.inc) with the cell values (rho, 1/rho, T)
// and the primitives of its stencil taken from the tile's LDS cache; (i,j,k) = the face's left cell, (a,b,c) = its
// tile-local face coordinates.  The face's averaged state, thermo and metric come first; then the gradient of one
// variable at a time is assembled in computational space, taken to physical space and reduced to its flux terms, so
// that only the three velocity gradients are held at once.  Every value is the same expression as in
// diff_flux_interior (diff_physics.inc:59-161), evaluated in the same order.
template <int NS, int NR, int D>
__device__ __forceinline__ void visc_face(const CaseDev& C, const mose_blk& B, const Tile& tl, const TileGeo& G_,
                                          const double* cache, int i, int j, int k, int a, int b, int c, double* res) {
  using V = Vars<NS, NR>;
  constexpr int NP = V::nprim;
  constexpr int ib = D == 0 ? 1 : 0, ic = D == 2 ? 1 : 2;
  const int64_t NC = C.NC;
  const mose_dtab& T = *C.tab;
  const int CI = G_.CI, CJ = G_.CJ, NCa = G_.ncache;
  const int64_t sa[3] = {1, (int64_t)B.pi, (int64_t)B.pi * B.pj};
  const int ca[3] = {1, CI, CI * CJ};
  const int kst = ca[D], kb = ca[ib], kc = ca[ic];
  // cache index of the left cell: tile-local + 1 (halo), one back along D
  const int k1 = (a + 1 - (D == 0)) + CI * ((b + 1 - (D == 1)) + CJ * (c + 1 - (D == 2)));
  const int k2 = k1 + kst;
  const int64_t c0 = row(B, i, j, k), c1 = c0 + sa[D];
  // the primitives of the tile and its one-cell halo, copied into LDS after the cell scalars: P_s at [(3 + s) NCa]
  const double* Pc = cache + 3 * NCa;
  const double* Ir = cache + NCa;
  const double* Tc = cache + 2 * NCa;
  double Pf[NP];
#pragma unroll
  for (int s = 0; s < NP; ++s) Pf[s] = 0.5 * (Pc[s * NCa + k1] + Pc[s * NCa + k2]);
  const FaceThermo ft = face_thermo<NS>(T, Pf, Pf[V::np], C.vc.prl);
  const double* fm = D == 0 ? C.fm1 : (D == 1 ? C.fm2 : C.fm3);
  const int64_t NFg = D == 0 ? C.NF1 : (D == 1 ? C.NF2 : C.NF3);
  const int64_t fg = D == 0 ? B.f1off + i + (int64_t)(B.n1 + 1) * ((j - 1) + (int64_t)B.n2 * (k - 1))
                   : D == 1 ? B.f2off + (i - 1) + (int64_t)B.n1 * (j + (int64_t)(B.n2 + 1) * (k - 1))
                            : B.f3off + (i - 1) + (int64_t)B.n1 * ((j - 1) + (int64_t)B.n2 * k);
  const double nx = fm[fg], ny = fm[NFg + fg], nz = fm[2 * NFg + fg], area = fm[3 * NFg + fg];
  double Mav[3][3];
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int y = 0; y < 3; ++y) Mav[x][y] = 0.5 * (C.m[(int64_t)(x + 3 * y) * NC + c0] + C.m[(int64_t)(x + 3 * y) * NC + c1]);
  const double drho = ft.rho, dmil = ft.mil;
  const int Ti = ftrunc(ft.T);
  const double Tdf = ft.T - Ti;
  double mie = 0.0;
  if (NR > 0) {
    double chi = Pf[V::nt] / dmil, c3 = chi * chi * chi;
    mie = c3 / (c3 + SA_CV1 * SA_CV1 * SA_CV1) * Pf[V::nt];
  }
  const double Dm = (dmil / C.vc.sc + mie / C.vc.sct) / drho;
  const double kap = ft.kl + mie * ft.cp / C.vc.prt;
  const double dmi = dmil + mie;
  // the physical-space gradient of variable s: kind 0 = P_s / rho, 1 = T, 2 = P_s
  auto grad = [&](int s, int kind, double* g) {
    const double* Q = kind == 1 ? Tc : Pc + s * NCa;
    double r[3];
    if (kind == 0) r[D] = Q[k2] * Ir[k2] - Q[k1] * Ir[k1];
    else r[D] = Q[k2] - Q[k1];
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      const int kd = pass ? kc : kb;
      const int k3 = k1 - kd, k4 = k1 + kd, k5 = k2 - kd, k6 = k2 + kd;
      double t;
      if (kind == 0) t = (Q[k4] * Ir[k4] - Q[k3] * Ir[k3] + Q[k6] * Ir[k6] - Q[k5] * Ir[k5]) * 0.25;
      else t = (Q[k4] - Q[k3] + Q[k6] - Q[k5]) * 0.25;
      r[pass ? ic : ib] = t;
    }
#pragma unroll
    for (int y = 0; y < 3; ++y) g[y] = r[0] * Mav[0][y] + r[1] * Mav[1][y] + r[2] * Mav[2][y];
  };
  double DiffH = 0.0, sf = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double g[3];
    grad(s, 0, g);
    double DmG = Dm * (g[0] * nx + g[1] * ny + g[2] * nz);
    res[s] = drho * area * DmG;
    double h1 = tab(T.h, s, Ti), h2 = tab(T.h, s, Ti + 1);
    DiffH = DiffH + DmG * (h1 + Tdf * (h2 - h1));
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) sf = sf + res[s];
#pragma unroll
  for (int s = 0; s < NS; ++s) res[s] = res[s] - sf * Pf[s] / drho;
  double qT;
  {
    double g[3];
    grad(0, 1, g);
    qT = kap * (g[0] * nx + g[1] * ny + g[2] * nz);
  }
  if (NR > 0) {
    double g[3];
    grad(V::nt, 0, g);
    double gn = g[0] * nx + g[1] * ny + g[2] * nz;
    res[V::nt] = (dmil + Pf[V::nt]) * gn * area / SA_SIGMA;
  }
  double gu[3], gv[3], gw[3];
  grad(V::nu, 2, gu);
  grad(V::nv, 2, gv);
  grad(V::nw, 2, gw);
  double lamDiv = NTT * dmi * (gu[0] + gv[1] + gw[2]);
  double Txx = lamDiv + 2.0 * dmi * gu[0] + 0.0;
  double Tyy = lamDiv + 2.0 * dmi * gv[1] + 0.0;
  double Tzz = lamDiv + 2.0 * dmi * gw[2] + 0.0;
  double Txy = dmi * (gu[1] + gv[0]), Txz = dmi * (gu[2] + gw[0]);
  double Tyz = dmi * (gv[2] + gw[1]);
  double S1 = Txx * nx + Txy * ny + Txz * nz, S2 = Txy * nx + Tyy * ny + Tyz * nz, S3 = Txz * nx + Tyz * ny + Tzz * nz;
  res[V::nu] = area * S1; res[V::nv] = area * S2; res[V::nw] = area * S3;
  res[V::np] = area * ((S1 * Pf[V::nu] + S2 * Pf[V::nv] + S3 * Pf[V::nw]) + qT + DiffH * drho);
  (void)tl;
}

// One tile of one RK stage (the phases above).  Every thread of the workgroup calls it with the same tile.
template <int NS, int NR, bool VISC, int RS, int STI, int STJ, int STK>
__device__ __forceinline__ void stage_tile(const CaseDev& C, const StageArgs& A, int tidx) {
  using V = Vars<NS, NR>;
  constexpr int NP = V::nprim;
  extern __shared__ double sm[];
  int tid = threadIdx.x;
  const Tile tl = A.tiles[tidx];
  const mose_blk B = C.blk[tl.b];
  const mose_dtab& T = *C.tab;
  const int64_t NC = C.NC;
  const double* P = C.Pin;
  const TileGeo G = tile_geo<NP, VISC, STI, STJ, STK>(A);
  const int TI = G.TI, TJ = G.TJ, TK = G.TK;
  const int NI = G.NI, NJ = G.NJ;
  const int NF = G.nslot;
  double* Hs = sm;                       // region 0: [NP][nhs] half-slopes (phase A)
  double* cache = sm;                    // region 0: [3 + NP][ncache] rho, 1/rho, T, P (phases B-D)
  double* Fb = sm + G.r0;                // region 1: [NP][NF] face fluxes
  const int64_t si = 1, sj = B.pi, sk = (int64_t)B.pi * B.pj;

  // ---------------- phase A0: limited half-slopes of the tile's cells (+ one cell along each direction) --------
  const int HI = G.HI, HJ = G.HJ;
  for (int q = tid; q < G.nhs; q += blockDim.x) {
    int d, a, b, c;
    if (q < HI) { d = 0; a = q % (TI + 2); int r = q / (TI + 2); b = r % TJ; c = r / TJ; }
    else if (q < HI + HJ) { int g = q - HI; d = 1; a = g % TI; int r = g / TI; b = r % (TJ + 2); c = r / (TJ + 2); }
    else { int g = q - HI - HJ; d = 2; a = g % TI; int r = g / TI; b = r % TJ; c = r / TJ; }
    int i = tl.i0 + a - (d == 0), j = tl.j0 + b - (d == 1), k = tl.k0 + c - (d == 2);
    int n_d = d == 0 ? B.n1 : (d == 1 ? B.n2 : B.n3);
    int l_d = d == 0 ? i : (d == 1 ? j : k);
    if (n_d < 2 || l_d < 1 || l_d > n_d || i > B.n1 || j > B.n2 || k > B.n3) continue;
    int64_t st = d == 0 ? si : (d == 1 ? sj : sk);
    int64_t c0 = row(B, i, j, k);
    const double* dlp = C.dl + (int64_t)d * NC;
    double dm = dlp[c0 - st], d0 = dlp[c0], dp = dlp[c0 + st];
    double idm = 1.0 / (0.5 * (dm + d0)), idp = 1.0 / (0.5 * (d0 + dp)), hd = 0.5 * d0;
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      Hs[s * G.nhs + q] = half_slope(C.lim, P[s * NC + c0 - st], P[s * NC + c0], P[s * NC + c0 + st], idm, idp, hd);
    }
  }
  __syncthreads();

  // ---------------- phase A1: convective faces from the half-slopes ----------------
  for (int f = tid; f < NF; f += blockDim.x) {
    int d, a, b, c;
    if (f < NI) { d = 0; a = f % (TI + 1); int r = f / (TI + 1); b = r % TJ; c = r / TJ; }
    else if (f < NI + NJ) { int g = f - NI; d = 1; a = g % TI; int r = g / TI; b = r % (TJ + 1); c = r / (TJ + 1); }
    else { int g = f - NI - NJ; d = 2; a = g % TI; int r = g / TI; b = r % TJ; c = r / TJ; }
    // the face's left cell (the reference's face index): the cell before local (a,b,c) along d
    int i = tl.i0 + a - (d == 0), j = tl.j0 + b - (d == 1), k = tl.k0 + c - (d == 2);
    int n_d = d == 0 ? B.n1 : (d == 1 ? B.n2 : B.n3);
    int l_d = d == 0 ? i : (d == 1 ? j : k);
    bool ok = (l_d >= 1 && l_d <= n_d - 1) && i <= B.n1 && j <= B.n2 && k <= B.n3;
    if (!ok) continue;
    int64_t st = d == 0 ? si : (d == 1 ? sj : sk);
    int64_t c0 = row(B, i, j, k);
    // half-slope slots of the left cell (local a-1 along d, +1 halo -> a) and of the right cell
    int hl = d == 0 ? a + (TI + 2) * (b + TJ * c) : (d == 1 ? HI + a + TI * (b + (TJ + 2) * c) : HI + HJ + a + TI * (b + TJ * c));
    int hr = hl + (d == 0 ? 1 : (d == 1 ? TI : TI * TJ));
    double L[NP], R[NP], res[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      L[s] = P[s * NC + c0] + Hs[s * G.nhs + hl];
      R[s] = P[s * NC + c0 + st] - Hs[s * G.nhs + hr];
    }
    const double* fm = d == 0 ? C.fm1 : (d == 1 ? C.fm2 : C.fm3);
    int64_t NFg = d == 0 ? C.NF1 : (d == 1 ? C.NF2 : C.NF3);
    int64_t fg = d == 0 ? B.f1off + i + (int64_t)(B.n1 + 1) * ((j - 1) + (int64_t)B.n2 * (k - 1))
               : d == 1 ? B.f2off + (i - 1) + (int64_t)B.n1 * (j + (int64_t)(B.n2 + 1) * (k - 1))
                        : B.f3off + (i - 1) + (int64_t)B.n1 * ((j - 1) + (int64_t)B.n2 * k);
    conv_face_lr<NS, NR, RS>(T, C.rs, C.mco, L, R, fm[fg], fm[NFg + fg], fm[2 * NFg + fg], fm[3 * NFg + fg], res);
#pragma unroll
    for (int s = 0; s < NP; ++s) Fb[s * NF + f] = res[s];
  }
  __syncthreads();
  if (VISC) {
    // the half-slopes are dead: region 0 now holds (rho, 1/rho, T) of the tile and its one-cell halo
    const int CI = TI + 2, CJ = TJ + 2;
    for (int q = tid; q < G.ncache; q += blockDim.x) {
      int a = q % CI, r = q / CI, b = r % CJ, c = r / CJ;
      int i = tl.i0 + a - 1, j = tl.j0 + b - 1, k = tl.k0 + c - 1;
      if (i < 1 - GC || i > B.n1 + GC || j < 1 - GC || j > B.n2 + GC || k < 1 - GC || k > B.n3 + GC) continue;
      int64_t r0 = row(B, i, j, k);
      double y[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) y[s] = P[s * NC + r0];
      double rho = rho_sum<NS>(y);
      cache[q] = rho;
      cache[G.ncache + q] = 1.0 / rho;
      cache[2 * G.ncache + q] = P[V::np * NC + r0] / (rho * f_Rtot<NS>(T, y));
#pragma unroll
      for (int s = 0; s < NP; ++s) cache[(3 + s) * G.ncache + q] = P[s * NC + r0];
    }
  }

  // ---------------- phase B: convective sums ----------------
  int ca = tid % TI, cb = (tid / TI) % TJ, cc = tid / (TI * TJ);
  int ci = tl.i0 + ca, cj = tl.j0 + cb, ck = tl.k0 + cc;
  bool mine = (tid < TI * TJ * TK) && ci <= B.n1 && cj <= B.n2 && ck <= B.n3;
  int sIm = ca + (TI + 1) * (cb + TJ * cc), sJm = NI + ca + TI * (cb + (TJ + 1) * cc), sKm = NI + NJ + ca + TI * (cb + TJ * cc);
  int sIp = sIm + 1, sJp = sJm + TI, sKp = sKm + TI * TJ;
  double tmp[NP];
#pragma unroll
  for (int s = 0; s < NP; ++s) tmp[s] = 0.0;
  if (mine) {
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      double t = tmp[s];
      if (ci < B.n1) t = t + Fb[s * NF + sIp];
      if (ci > 1) t = t - Fb[s * NF + sIm];
      if (cj < B.n2) t = t + Fb[s * NF + sJp];
      if (cj > 1) t = t - Fb[s * NF + sJm];
      if (ck < B.n3) t = t + Fb[s * NF + sKp];
      if (ck > 1) t = t - Fb[s * NF + sKm];
      tmp[s] = t;
    }
  }

  // ---------------- phase C: viscous faces ----------------
  if (VISC) {
    __syncthreads();
    for (int f = tid; f < NF; f += blockDim.x) {
      int d, a, b, c;
      if (f < NI) { d = 0; a = f % (TI + 1); int r = f / (TI + 1); b = r % TJ; c = r / TJ; }
      else if (f < NI + NJ) { int g = f - NI; d = 1; a = g % TI; int r = g / TI; b = r % (TJ + 1); c = r / (TJ + 1); }
      else { int g = f - NI - NJ; d = 2; a = g % TI; int r = g / TI; b = r % TJ; c = r / TJ; }
      int i = tl.i0 + a - (d == 0), j = tl.j0 + b - (d == 1), k = tl.k0 + c - (d == 2);
      int n_d = d == 0 ? B.n1 : (d == 1 ? B.n2 : B.n3);
      int l_d = d == 0 ? i : (d == 1 ? j : k);
      bool ok = (l_d >= 1 && l_d <= n_d - 1) && i <= B.n1 && j <= B.n2 && k <= B.n3;
      if (!ok) continue;
      // local cache coordinates of the left cell c1 (tile-local + 1)
      double res[NP];
      if (d == 0) visc_face<NS, NR, 0>(C, B, tl, G, cache, i, j, k, a, b, c, res);
      else if (d == 1) visc_face<NS, NR, 1>(C, B, tl, G, cache, i, j, k, a, b, c, res);
      else visc_face<NS, NR, 2>(C, B, tl, G, cache, i, j, k, a, b, c, res);
#pragma unroll
      for (int s = 0; s < NP; ++s) Fb[s * NF + f] = res[s];
    }
    __syncthreads();
    if (mine) {
#pragma unroll
      for (int s = 0; s < NP; ++s) {
        double t = tmp[s];
        if (ci < B.n1) t = t - Fb[s * NF + sIp];
        if (ci > 1) t = t + Fb[s * NF + sIm];
        if (cj < B.n2) t = t - Fb[s * NF + sJp];
        if (cj > 1) t = t + Fb[s * NF + sJm];
        if (ck < B.n3) t = t - Fb[s * NF + sKp];
        if (ck > 1) t = t + Fb[s * NF + sKm];
        tmp[s] = t;
      }
    }
  }

  // ---------------- phase D: boundary increments, update, next dt ----------------
  double dtl = 1e300;
  double dres[NP - NS + 1];
#pragma unroll
  for (int q = 0; q < NP - NS + 1; ++q) dres[q] = 0.0;
  if (mine) {
    bool on[6] = {ci == 1, ci == B.n1, cj == 1, cj == B.n2, ck == 1, ck == B.n3};
#pragma unroll
    for (int f = 1; f <= 6; ++f) {
      if (!on[f - 1]) continue;
      int64_t pos = B.bpoff[f - 1] + plane_pos(B, f, ci, cj, ck);
#pragma unroll
      for (int s = 0; s < NP; ++s) {
        tmp[s] = tmp[s] + C.binc[(int64_t)s * C.NBP + pos];
        tmp[s] = tmp[s] + C.binc[(int64_t)(NP + s) * C.NBP + pos];
      }
    }
    int64_t c0 = row(B, ci, cj, ck);
    double p[NP], po[NP], out[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) { p[s] = P[s * NC + c0]; po[s] = C.PO[s * NC + c0]; }
    double Tseed;
    if (VISC) {
      const int CI = TI + 2, CJ = TJ + 2;
      Tseed = cache[2 * G.ncache + (ca + 1) + CI * ((cb + 1) + CJ * (cc + 1))];
      if (C.guard && !(Tseed > 1.0 && Tseed < 9990.0)) {
        flag_bad(C, tl.b, ci, cj, ck, 1);
        if (Tseed != Tseed) Tseed = 300.0;
        Tseed = dmin2(dmax2(Tseed, 1.0), 9990.0);
      }
    } else {
      double rho = rho_sum<NS>(p), Rg = f_Rtot<NS>(T, p);
      Tseed = p[V::np] / (rho * Rg);
    }
    if (NR > 0) {
      double rho = cache[(ca + 1) + (TI + 2) * ((cb + 1) + (TJ + 2) * (cc + 1))];
      cell_sa<NS, NR>(C, B, ci, cj, ck, c0, rho, cell_mil<NS>(T, p, rho, Tseed), tmp);
    }
    if (C.prim) update_cell_prim<NS, NR>(T, p, po, tmp, C.vol[c0], *C.dt, C.strang, A.rkc, out);
    else update_cell<NS, NR>(T, p, po, tmp, C.vol[c0], *C.dt, C.strang, A.rkc, Tseed, out);
    bool good = out[V::np] >= 0.0;
#pragma unroll
    for (int s = 0; s < NP; ++s) good = good && isfinite(out[s]);
    if (!good) flag_bad(C, tl.b, ci, cj, ck, 0);
#pragma unroll
    for (int s = 0; s < NP; ++s) C.Pout[s * NC + c0] = out[s];
    if (A.last) {
      dtl = cell_dt<NS, NR>(C, c0, out);
      step_change<NS, NR>(out, po, dres);
    }
  }
  if (A.last) {
    block_sum_store<NP - NS + 1>(dres, A.rpart + (int64_t)tidx * (NP - NS + 1));
    block_min_atomic(dtl, A.dtnext);
  }
}

// The workgroup -> tile map: workgroup w runs on XCD w % 8 (workgroups are dealt to the 8 XCDs round robin); XCD x
// gets the contiguous run of tiles [x n / 8, (x + 1) n / 8), so neighbouring tiles share that XCD's L2.  A run-time
// atomic tile queue here (dynamic assignment with stealing) raised the kernel's register allocation from 230 to
// 263 VGPRs and halved its occupancy, so the map is static; the tiles of every block of the case, of any size,
// are still served by this one launch.
__device__ __forceinline__ int xcd_tile(int wg, int n) {
  int x = wg & 7, q = wg >> 3;
  int per = n >> 3, rem = n & 7;
  int base = x * per + (x < rem ? x : rem);
  int cnt = per + (x < rem ? 1 : 0);
  return q < cnt ? base + q : -1;
}

// BS = the workgroup size (at least the cells of the tile).  Larger tiles recompute fewer halo half-slopes and faces
// and re-read less of their neighbours; MI355X's 160 KB of LDS per workgroup holds an 8 x 8 x 8 tile's half-slopes
// and face fluxes.  Smaller tiles leave room for two workgroups per CU.
// SH selects a tile shape fixed at compile time: 0 = from the launch arguments, 1 = 8 x 8 x 4, 2 = 8 x 8 x 8.
// Without RANS the 8 x 8 x 4 kernels take at most 128 VGPRs (127 for the viscous LMRoe one), so two 512-thread
// workgroups share a CU (4 waves per SIMD).  An amdgpu_waves_per_eu(4) attribute gives the same count but a schedule
// that hides less latency: 6.41 against 5.96 ms per stage on the 256^3 Taylor-Green case.
template <int NS, int NR, bool VISC, int RS, int BS, int SH>
__global__ void __launch_bounds__(BS) k_stage(CaseDev C, StageArgs A) {
  const int t = xcd_tile(blockIdx.x, A.ntiles);
  if (t < 0) return;
  constexpr int STI = SH ? 8 : 0, STJ = SH ? 8 : 0, STK = SH == 1 ? 4 : (SH == 2 ? 8 : 0);
  stage_tile<NS, NR, VISC, RS, STI, STJ, STK>(C, A, t);
}

// The launch of one stage over a case's tiles: the kernel of the case's solver, viscosity and tile shape (the
// shapes 8 x 8 x 4 and 8 x 8 x 8 fixed at compile time, the others read at run time; 128-cell tiles run 256 threads,
// two per cell, the others 512).
template <int NS, int NR, int BS, int SH>
void launch_stage_shape(const CaseDev& D, const StageArgs& A, int rs, bool visc, unsigned grid, size_t lds,
                        hipStream_t s) {
  if (visc) {
    if (rs == MOSE_RS_LMROE) k_stage<NS, NR, true, MOSE_RS_LMROE, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else if (rs == MOSE_RS_AUSMP) k_stage<NS, NR, true, MOSE_RS_AUSMP, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else if (rs == MOSE_RS_HLLE) k_stage<NS, NR, true, MOSE_RS_HLLE, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else if (rs == MOSE_RS_SLAU) k_stage<NS, NR, true, MOSE_RS_SLAU, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else k_stage<NS, NR, true, MOSE_RS_HLLC, BS, SH><<<grid, BS, lds, s>>>(D, A);
  } else {
    if (rs == MOSE_RS_LMROE) k_stage<NS, NR, false, MOSE_RS_LMROE, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else if (rs == MOSE_RS_AUSMP) k_stage<NS, NR, false, MOSE_RS_AUSMP, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else if (rs == MOSE_RS_HLLE) k_stage<NS, NR, false, MOSE_RS_HLLE, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else if (rs == MOSE_RS_SLAU) k_stage<NS, NR, false, MOSE_RS_SLAU, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else k_stage<NS, NR, false, MOSE_RS_HLLC, BS, SH><<<grid, BS, lds, s>>>(D, A);
  }
}

template <int NS, int NR>
void launch_stage_tiles(const CaseDev& D, const StageArgs& A, int rs, bool visc, unsigned grid, size_t lds,
                        hipStream_t s) {
  const int TI = A.TI, TJ = A.TJ, TK = A.TK;
  if constexpr (NS >= 9) {
    // nine species: the 8 x 8 x 4 and 8 x 8 x 8 tiles do not fit the LDS, every shape is read at run time
    if (TI * TJ * TK == 128) launch_stage_shape<NS, NR, 256, 0>(D, A, rs, visc, grid, lds, s);
    else launch_stage_shape<NS, NR, 512, 0>(D, A, rs, visc, grid, lds, s);
    return;
  }
  if (TI * TJ * TK == 128) launch_stage_shape<NS, NR, 256, 0>(D, A, rs, visc, grid, lds, s);
  else if (TI == 8 && TJ == 8 && TK == 4) launch_stage_shape<NS, NR, 512, 1>(D, A, rs, visc, grid, lds, s);
  else if (TI == 8 && TJ == 8 && TK == 8) launch_stage_shape<NS, NR, 512, 2>(D, A, rs, visc, grid, lds, s);
  else launch_stage_shape<NS, NR, 512, 0>(D, A, rs, visc, grid, lds, s);
}

// The performance path (mose_stage_fast.hpp), defined in its own contracted translation unit (mose_stage_fast.hip)
// for every instantiated (NS, NR); its LDS per tile in doubles: the primitives of the tile's one-cell halo box and
// of six slabs one cell further out, the viscous cell scalars (T alone for one species without RANS, else rho,
// 1/rho, T), the face slots.
template <int NS, int NR>
void launch_stage_fast(const CaseDev& D, const StageArgs& A, int rs, bool visc, unsigned grid, size_t lds,
                       hipStream_t s);

inline size_t fast_lds_doubles(int ns, int nr, int nprim, bool visc, int TI, int TJ, int TK) {
  const size_t nbox = (size_t)(TI + 2) * (TJ + 2) * (TK + 2);
  const size_t next = nbox + 2 * ((size_t)TJ * TK + (size_t)TI * TK + (size_t)TI * TJ);
  const size_t nslot = (size_t)(TI + 1) * TJ * TK + (size_t)TI * (TJ + 1) * TK + (size_t)TI * TJ * (TK + 1);
  const size_t ncc = visc ? ((ns == 1 && nr == 0) ? 1 : 3) : 0;
  // one species without RANS on 8 x 8 x 4 tiles: 128 doubles for the staged cell metrics of the viscous faces
  const size_t mst = (visc && ns == 1 && nr == 0 && TI == 8 && TJ == 8 && TK == 4) ? 128 : 0;
  return (size_t)nprim * next + ncc * nbox + mst + (size_t)nprim * nslot;
}

}  // namespace mose
