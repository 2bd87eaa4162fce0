// The fused stage of the performance path (MOSE_PATH_FAST).  Same tiles, same faces, same sums as the parity stage
// (mose_stage.hpp), reorganised around one staging pass:
//
//   phase L  every thread loads its share of the primitives of the tile, of its one-cell halo box and of six slabs
//            one cell further out along each direction (the outer cells of the half-slope stencils) into LDS, once,
//            with all of its loads in flight together; viscous cases also store the cells' temperature there
//            (one species without RANS: T alone; otherwise rho, 1/rho, T as the parity path's cache)
//   phase A  every interior face: the two limited half-slopes of its cells from the LDS stencil (no half-slope
//            buffer: each cell's half-slope is formed by the two faces that use it), the flux into LDS
//   phase B  each cell sums its six convective faces (pass2_body.inc order)
//   phase C  every interior face: the viscous flux from the same LDS primitives
//   phase D  viscous sums, boundary increments, the stage update from the LDS primitives, next dt
//
// Global memory is read by phase L (the primitives), by the faces (lengths and face metrics) and by phase D (the
// step base, the volume).  Reciprocals use the hardware estimate refined by two Newton steps (rcp_nr); one species
// without RANS drops the species-diffusion terms, which are zero in exact arithmetic (the gradient of Y = 1).  The
// results differ from the parity path by rounding only; the tests hold them to 1e-12 on every parity family.
#pragma once

#include "mose_fast.hpp"
#include "mose_stage.hpp"

namespace mose {

// The LDS geometry of a tile: the box of the tile and its one-cell halo (CI x CJ x (TK + 2) cells, indexed as the
// parity path's cell cache), then the slabs (side 0 and side 1 of direction 0, of direction 1, of direction 2), then
// the face slots of the parity path.
struct FastGeo {
  int TI, TJ, TK, CI, CJ, nbox, s0, s1, s2, next, NI, NJ, nslot;
};
template <int STI, int STJ, int STK>
__device__ __forceinline__ FastGeo fast_geo(const StageArgs& A) {
  FastGeo g;
  g.TI = STI ? STI : A.TI;
  g.TJ = STJ ? STJ : A.TJ;
  g.TK = STK ? STK : A.TK;
  g.CI = g.TI + 2;
  g.CJ = g.TJ + 2;
  g.nbox = g.CI * g.CJ * (g.TK + 2);
  g.s0 = g.TJ * g.TK;
  g.s1 = g.TI * g.TK;
  g.s2 = g.TI * g.TJ;
  g.next = g.nbox + 2 * (g.s0 + g.s1 + g.s2);
  g.NI = (g.TI + 1) * g.TJ * g.TK;
  g.NJ = g.TI * (g.TJ + 1) * g.TK;
  g.nslot = g.NI + g.NJ + g.TI * g.TJ * (g.TK + 1);
  return g;
}

// cell-scalar arrays of the viscous cache: T alone for one species without RANS, else rho, 1/rho, T
template <int NS, int NR, bool VISC>
struct FastCache {
  static constexpr int n = VISC ? ((NS == 1 && NR == 0) ? 1 : 3) : 0;
  static constexpr int iT = n == 1 ? 0 : 2;
};

// 1/x to about one ulp: the hardware estimate and two Newton-Raphson steps r' = r + r (1 - x r), each of which
// squares the relative error.  Every use has a positive normal argument (a density, a length, a sum of squares...).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(e, r, r);
  e = fma(-x, r, 1.0);
  return fma(e, r, r);
}

// The limiters of Lib_Limiters.f90:65-97 in their one-division forms (limiter_fast) with the reciprocal above
__device__ __forceinline__ double limiter_nr(int lim, double x, double y) {
  if (lim == 2) {
    const double xp = x + 1e-30;
    return ((xp < 0.0 ? -y : y) + fabs(y)) * rcp_nr(fabs(xp) + fabs(y)) * x;
  }
  if (lim == 3) {
    const double xp = x + 1e-30;
    return (y * y + y * xp) * rcp_nr(xp * xp + y * y) * x;
  }
  return limiter(lim, x, y);
}

// One interior viscous face of direction D (diff_physics.inc:59-161) from the LDS stencil: Pl = the primitives
// (variable stride nx), Cc = the cell scalars (stride nb); k1/k2 = the box indices of the face's left/right cells,
// kb/kc = the box strides of the two other directions; c0/c1 = the cells' rows (metric), or, with MG, mg = the
// metric averaged over the face's two cells, 0.5 (m_left + m_right), component q = x + 3 y at mg[q].
template <int NS, int NR, int D, bool MG = false>
__device__ __forceinline__ void visc_face_lds(const CaseDev& C, const double* Pl, int nx_, const double* Cc, int nb,
                                              int k1, int k2, int kb, int kc, int64_t c0, int64_t c1, double nx,
                                              double ny, double nz, double area, double* res,
                                              const double* mg = nullptr) {
  using V = Vars<NS, NR>;
  using FC = FastCache<NS, NR, true>;
  constexpr int NP = V::nprim;
  constexpr int ib = D == 0 ? 1 : 0, ic = D == 2 ? 1 : 2;
  const int64_t NC = C.NC;
  const mose_dtab& T = *C.tab;
  const double* Tc = Cc + FC::iT * nb;
  const double* Ir = Cc + nb;     // 1/rho (three-array cache only)
  double Pf[NP];
#pragma unroll
  for (int s = 0; s < NP; ++s) Pf[s] = 0.5 * (Pl[s * nx_ + k1] + Pl[s * nx_ + k2]);
  const FaceThermo ft = face_thermo_fast<NS>(T, Pf, Pf[V::np], C.vc.prl);
  double Mav[3][3];
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int y = 0; y < 3; ++y) {
      if constexpr (MG) Mav[x][y] = mg[x + 3 * y];
      else Mav[x][y] = 0.5 * (C.m[(int64_t)(x + 3 * y) * NC + c0] + C.m[(int64_t)(x + 3 * y) * NC + c1]);
    }
  const double drho = ft.rho, dmil = ft.mil, ird = rcp_nr(drho);
  double mie = 0.0;
  if (NR > 0) {
    double chi = Pf[V::nt] * rcp_nr(dmil), c3 = chi * chi * chi;
    mie = c3 * rcp_nr(c3 + SA_CV1 * SA_CV1 * SA_CV1) * Pf[V::nt];
  }
  const double kap = ft.kl + mie * ft.cp / C.vc.prt;
  const double dmi = dmil + mie;
  // the physical-space gradient of variable s: kind 0 = P_s / rho, 1 = T, 2 = P_s
  auto grad = [&](int s, int kind, double* g) {
    const double* Q = kind == 1 ? Tc : Pl + s * nx_;
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
  double DiffH = 0.0;
  if constexpr (!(NS == 1 && NR == 0)) {
    const double Dm = (dmil / C.vc.sc + mie / C.vc.sct) * ird;
    const int Ti = ftrunc(ft.T);
    const double Tdf = ft.T - Ti;
    double sf = 0.0;
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
    for (int s = 0; s < NS; ++s) res[s] = res[s] - sf * Pf[s] * ird;
  } else {
    res[0] = 0.0;   // one species: Y = 1 everywhere, no diffusion flux and no enthalpy it carries
  }
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
    res[V::nt] = (dmil + Pf[V::nt]) * gn * area * (1.0 / SA_SIGMA);
  }
  double gu[3], gv[3], gw[3];
  grad(V::nu, 2, gu);
  grad(V::nv, 2, gv);
  grad(V::nw, 2, gw);
  double lamDiv = NTT * dmi * (gu[0] + gv[1] + gw[2]);
  double Txx = lamDiv + 2.0 * dmi * gu[0];
  double Tyy = lamDiv + 2.0 * dmi * gv[1];
  double Tzz = lamDiv + 2.0 * dmi * gw[2];
  double Txy = dmi * (gu[1] + gv[0]), Txz = dmi * (gu[2] + gw[0]);
  double Tyz = dmi * (gv[2] + gw[1]);
  double S1 = Txx * nx + Txy * ny + Txz * nz, S2 = Txy * nx + Tyy * ny + Tyz * nz, S3 = Txz * nx + Tyz * ny + Tzz * nz;
  res[V::nu] = area * S1; res[V::nv] = area * S2; res[V::nw] = area * S3;
  res[V::np] = area * ((S1 * Pf[V::nu] + S2 * Pf[V::nv] + S3 * Pf[V::nw]) + qT + DiffH * drho);
}

// newstate_cons_body.inc:5-89 (update_cell) with reciprocals in place of its divisions; w = dt strang / vol
template <int NS, int NR>
__device__ __forceinline__ void prim2cons_nr(const mose_dtab& T, const double* q, double* out) {
  using V = Vars<NS, NR>;
  const double rho = rho_sum<NS>(q), ir = rcp_nr(rho);
  double Rg = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) Rg = Rg + T.Ri[s] * q[s] * ir;
  double vm = 0.0;
#pragma unroll
  for (int s = V::nu; s <= V::nw; ++s) vm = vm + q[s] * q[s];
  const double Tl = q[V::np] * ir * rcp_nr(Rg);
  int Ti = ftrunc(Tl); Ti = Ti < 0 ? 0 : (Ti > NTMAX - 1 ? NTMAX - 1 : Ti);
  const double Td = Tl - Ti;
  double hm = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) hm = hm + (tab(T.h, s, Ti) + Td * (tab(T.h, s, Ti + 1) - tab(T.h, s, Ti))) * q[s] * ir;
  const double E0 = hm - q[V::np] * ir + 0.5 * vm;
#pragma unroll
  for (int s = 0; s < NS; ++s) out[s] = q[s];
#pragma unroll
  for (int s = V::nu; s <= V::nw; ++s) out[s] = rho * q[s];
  out[V::np] = rho * E0;
#pragma unroll
  for (int s = V::np + 1; s < V::nprim; ++s) out[s] = q[s];
}

template <int NS, int NR>
__device__ __forceinline__ void update_cell_nr(const mose_dtab& T, const double* p, const double* po, const double* res,
                                               double w, double rkc, double Tseed, double* out) {
  using V = Vars<NS, NR>;
  double cold[V::nprim], c[V::nprim];
  prim2cons_nr<NS, NR>(T, po, cold);
  prim2cons_nr<NS, NR>(T, p, c);
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) c[s] = cold[s] + (c[s] - cold[s] - res[s] * w) * rkc;
  const double rho = rho_sum<NS>(c), ir = rcp_nr(rho);
  double Rg = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) Rg = Rg + T.Ri[s] * c[s] * ir;
#pragma unroll
  for (int s = 0; s < NS; ++s) out[s] = c[s];
  double vm = 0.0;
#pragma unroll
  for (int s = V::nu; s <= V::nw; ++s) {
    out[s] = c[s] * ir;
    vm = vm + out[s] * out[s];
  }
  const double en = c[V::np] * ir - 0.5 * vm;
  double de = 1e3, Tit = Tseed;
  int it = 1;
  while ((fabs(de) >= 1e-6 * fabs(en)) && (it <= 20)) {
    int Ti = ftrunc(Tit); Ti = Ti < 0 ? 0 : (Ti > NTMAX - 1 ? NTMAX - 1 : Ti);
    double Td = Tit - Ti, cv = -Rg, enth = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double cps = tab(T.cp, s, Ti) + Td * (tab(T.cp, s, Ti + 1) - tab(T.cp, s, Ti));
      double hs = tab(T.h, s, Ti) + Td * (tab(T.h, s, Ti + 1) - tab(T.h, s, Ti));
      cv = cv + c[s] * ir * cps;
      enth = enth + c[s] * ir * hs;
    }
    const double eit = enth - Rg * Tit;     // enth - p / rho at Tit
    de = fabs(eit - en);
    Tit = Tit + (en - eit) * rcp_nr(cv);
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

// One tile of one RK stage on the performance path (the phases above); BS threads, every thread calls it.
template <int NS, int NR, bool VISC, int RS, int BS, int STI, int STJ, int STK>
__device__ __forceinline__ void stage_tile_fast(const CaseDev& C, const StageArgs& A, int tidx) {
  using V = Vars<NS, NR>;
  using FC = FastCache<NS, NR, VISC>;
  constexpr int NP = V::nprim;
  extern __shared__ double sm[];
  const int tid = threadIdx.x;
  const Tile tl = A.tiles[tidx];
  const mose_blk B = C.blk[tl.b];
  const mose_dtab& T = *C.tab;
  const int64_t NC = C.NC;
  const double* P = C.Pin;
  const FastGeo G = fast_geo<STI, STJ, STK>(A);
  const int TI = G.TI, TJ = G.TJ, TK = G.TK, CI = G.CI, CJ = G.CJ, NX = G.next, NB = G.nbox;
  const int NI = G.NI, NJ = G.NJ, NF = G.nslot;
  double* Pl = sm;                              // [NP][next]
  double* Cc = sm + NP * NX;                    // [FC::n][nbox]
  // one species without RANS on 8 x 8 x 4 tiles: phase C reads the tile's cell metrics from LDS (below), and 128
  // doubles between Cc and Fb make room for them
  constexpr bool MST = VISC && NS == 1 && NR == 0 && STI == 8 && STJ == 8 && STK == 4;
  double* Fb = Cc + FC::n * NB + (MST ? 128 : 0);   // [NP][nslot]
  const int64_t si = 1, sj = B.pi, sk = (int64_t)B.pi * B.pj;
  // The i-stride of the viscous faces' cell-metric loads, for one species without RANS as a run-time value (B.pi is
  // never 0): with the constant 1 the compiler merges each pair m[c0], m[c0 + 1] into one 16-byte load and issues the
  // nine merged loads one after another, each waiting for the last (3.4 % of the 256^3 Taylor-Green stage); as two
  // 8-byte loads they are issued together, as on the j- and k-faces.
  const int64_t siv = (NS == 1 && NR == 0) ? (int64_t)(B.pi > 0) : si;

  // ---------------- phase L: the primitives of the box and the slabs into LDS ----------------
  // up to LU cells per thread per pass, every load of the pass issued before the first LDS store (LU x NP doubles
  // in flight per thread: fewer cells per pass for more variables); more than eight variables take one cell per pass
  constexpr int LU = NP <= 5 ? 4 : (NP <= 8 ? 2 : 1);
  if constexpr (LU == 1) {
  for (int q = tid; q < NX; q += BS) {
    int la, lb, lc;
    if (q < NB) {
      la = q % CI - 1; const int r = q / CI; lb = r % CJ - 1; lc = r / CJ - 1;
    } else {
      int g = q - NB;
      if (g < 2 * G.s0) {
        const int side = g >= G.s0; g -= side * G.s0;
        la = side ? TI + 1 : -2; lb = g % TJ; lc = g / TJ;
      } else if ((g -= 2 * G.s0) < 2 * G.s1) {
        const int side = g >= G.s1; g -= side * G.s1;
        lb = side ? TJ + 1 : -2; la = g % TI; lc = g / TI;
      } else {
        g -= 2 * G.s1;
        const int side = g >= G.s2; g -= side * G.s2;
        lc = side ? TK + 1 : -2; la = g % TI; lb = g / TI;
      }
    }
    const int i = tl.i0 + la, j = tl.j0 + lb, k = tl.k0 + lc;
    if (i < 1 - GC || i > B.n1 + GC || j < 1 - GC || j > B.n2 + GC || k < 1 - GC || k > B.n3 + GC) continue;
    const int64_t r0 = row(B, i, j, k);
    double y[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) y[s] = P[s * NC + r0];
#pragma unroll
    for (int s = 0; s < NP; ++s) Pl[s * NX + q] = y[s];
    if constexpr (VISC) {
      if (q < NB) {
        if constexpr (FC::n == 1) {
          Cc[q] = y[V::np] * rcp_nr(y[0] * T.Ri[0]);
        } else {
          const double rho = rho_sum<NS>(y), ir = rcp_nr(rho);
          double Rg = 0.0;
#pragma unroll
          for (int s = 0; s < NS; ++s) Rg = Rg + T.Ri[s] * y[s] * ir;
          Cc[q] = rho;
          Cc[NB + q] = ir;
          Cc[2 * NB + q] = y[V::np] * ir * rcp_nr(Rg);
        }
      }
    }
  }
  } else {
  for (int q0 = tid; q0 < NX; q0 += LU * BS) {
    double y[LU][NP];
    int qs[LU];
    bool ok[LU];
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int q = q0 + u * BS;
      qs[u] = q;
      int la = 0, lb = 0, lc = 0;
      if (q < NB) {
        la = q % CI - 1; const int r = q / CI; lb = r % CJ - 1; lc = r / CJ - 1;
      } else {
        int g = q - NB;
        if (g < 2 * G.s0) {
          const int side = g >= G.s0; g -= side * G.s0;
          la = side ? TI + 1 : -2; lb = g % TJ; lc = g / TJ;
        } else if ((g -= 2 * G.s0) < 2 * G.s1) {
          const int side = g >= G.s1; g -= side * G.s1;
          lb = side ? TJ + 1 : -2; la = g % TI; lc = g / TI;
        } else {
          g -= 2 * G.s1;
          const int side = g >= G.s2; g -= side * G.s2;
          lc = side ? TK + 1 : -2; la = g % TI; lb = g / TI;
        }
      }
      const int i = tl.i0 + la, j = tl.j0 + lb, k = tl.k0 + lc;
      ok[u] = q < NX && !(i < 1 - GC || i > B.n1 + GC || j < 1 - GC || j > B.n2 + GC || k < 1 - GC || k > B.n3 + GC);
      const int64_t r0 = ok[u] ? row(B, i, j, k) : 0;
#pragma unroll
      for (int s = 0; s < NP; ++s) y[u][s] = ok[u] ? P[s * NC + r0] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      if (!ok[u]) continue;
      const int q = qs[u];
#pragma unroll
      for (int s = 0; s < NP; ++s) Pl[s * NX + q] = y[u][s];
      if constexpr (VISC) {
        if (q < NB) {
          if constexpr (FC::n == 1) {
            Cc[q] = y[u][V::np] * rcp_nr(y[u][0] * T.Ri[0]);
          } else {
            const double rho = rho_sum<NS>(y[u]), ir = rcp_nr(rho);
            double Rg = 0.0;
#pragma unroll
            for (int s = 0; s < NS; ++s) Rg = Rg + T.Ri[s] * y[u][s] * ir;
            Cc[q] = rho;
            Cc[NB + q] = ir;
            Cc[2 * NB + q] = y[u][V::np] * ir * rcp_nr(Rg);
          }
        }
      }
    }
  }
  }
  __syncthreads();

  // ---------------- phase A: convective faces ----------------
  for (int f = tid; f < NF; f += BS) {
    int d, a, b, c;
    if (f < NI) { d = 0; a = f % (TI + 1); int r = f / (TI + 1); b = r % TJ; c = r / TJ; }
    else if (f < NI + NJ) { int g = f - NI; d = 1; a = g % TI; int r = g / TI; b = r % (TJ + 1); c = r / (TJ + 1); }
    else { int g = f - NI - NJ; d = 2; a = g % TI; int r = g / TI; b = r % TJ; c = r / TJ; }
    const int i = tl.i0 + a - (d == 0), j = tl.j0 + b - (d == 1), k = tl.k0 + c - (d == 2);
    const int n_d = d == 0 ? B.n1 : (d == 1 ? B.n2 : B.n3);
    const int l_d = d == 0 ? i : (d == 1 ? j : k);
    if (!((l_d >= 1 && l_d <= n_d - 1) && i <= B.n1 && j <= B.n2 && k <= B.n3)) continue;
    // box indices of the four cells of the stencil along d: the outer two from the slabs at the tile's edges
    const int e = d == 0 ? a : (d == 1 ? b : c);
    const int Tn = d == 0 ? TI : (d == 1 ? TJ : TK);
    const int kst = d == 0 ? 1 : (d == 1 ? CI : CI * CJ);
    const int kL = (a + 1 - (d == 0)) + CI * ((b + 1 - (d == 1)) + CJ * (c + 1 - (d == 2)));
    const int kR = kL + kst;
    const int sl = d == 0 ? NB + b + TJ * c : (d == 1 ? NB + 2 * G.s0 + a + TI * c : NB + 2 * (G.s0 + G.s1) + a + TI * b);
    const int ss = d == 0 ? G.s0 : (d == 1 ? G.s1 : G.s2);
    const int kLL = e >= 1 ? kL - kst : sl;
    const int kRR = e < Tn ? kR + kst : sl + ss;
    const int64_t st = d == 0 ? si : (d == 1 ? sj : sk);
    const int64_t c0 = row(B, i, j, k);
    const double* dlp = C.dl + (int64_t)d * NC;
    const double dLL = dlp[c0 - st], dL = dlp[c0], dR = dlp[c0 + st], dRR = dlp[c0 + 2 * st];
    const double* fm = d == 0 ? C.fm1 : (d == 1 ? C.fm2 : C.fm3);
    const int64_t NFg = d == 0 ? C.NF1 : (d == 1 ? C.NF2 : C.NF3);
    const int64_t fg = d == 0 ? B.f1off + i + (int64_t)(B.n1 + 1) * ((j - 1) + (int64_t)B.n2 * (k - 1))
                     : d == 1 ? B.f2off + (i - 1) + (int64_t)B.n1 * (j + (int64_t)(B.n2 + 1) * (k - 1))
                              : B.f3off + (i - 1) + (int64_t)B.n1 * ((j - 1) + (int64_t)B.n2 * k);
    const double fnx = fm[fg], fny = fm[NFg + fg], fnz = fm[2 * NFg + fg], fa = fm[3 * NFg + fg];
    const double i0 = rcp_nr(0.5 * (dLL + dL)), i1 = rcp_nr(0.5 * (dL + dR)), i2 = rcp_nr(0.5 * (dR + dRR));
    const double hL = 0.5 * dL, hR = 0.5 * dR;
    double L[NP], R[NP], res[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const double* q = Pl + s * NX;
      const double qLL = q[kLL], qL = q[kL], qR = q[kR], qRR = q[kRR];
      L[s] = qL + limiter_nr(C.lim, (qR - qL) * i1, (qL - qLL) * i0) * hL;
      R[s] = qR - limiter_nr(C.lim, (qRR - qR) * i2, (qR - qL) * i1) * hR;
    }
    conv_face_fast<NS, NR, RS>(T, C.rs, C.mco, L, R, fnx, fny, fnz, fa, res);
#pragma unroll
    for (int s = 0; s < NP; ++s) Fb[s * NF + f] = res[s];
  }
  __syncthreads();

  // ---------------- phase B: convective sums ----------------
  const int ca = tid % TI, cb = (tid / TI) % TJ, cc = tid / (TI * TJ);
  const int ci = tl.i0 + ca, cj = tl.j0 + cb, ck = tl.k0 + cc;
  const bool mine = (tid < TI * TJ * TK) && ci <= B.n1 && cj <= B.n2 && ck <= B.n3;
  const int sIm = ca + (TI + 1) * (cb + TJ * cc), sJm = NI + ca + TI * (cb + (TJ + 1) * cc),
            sKm = NI + NJ + ca + TI * (cb + TJ * cc);
  const int sIp = sIm + 1, sJp = sJm + TI, sKp = sKm + TI * TJ;
  const int kc0 = (ca + 1) + CI * ((cb + 1) + CJ * (cc + 1));     // the cell's box index
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

  // ---------------- the tile's cell metrics into LDS (MST) ----------------
  // During phase C the slabs of Pl (read only by phase A), the species row of Fb (one species carries no diffusion
  // flux) and the 128 doubles before Fb hold the nine metric components of the tile's 256 cells: components 0..4 in
  // the slabs of the five variables (256 doubles each), 5..8 in the 1024 doubles from Cc + NB.  The faces with both
  // cells in the tile (640 of 896) then take their metric from LDS instead of 18 global loads; the 256 on the tile's
  // sides load the outer cell's.  Threads 0..255 store components 0..4 while phase B runs; threads 256..511 hold
  // 5..8 in registers until phase B's reads of Fb are over.
  double mreg[4] = {0.0, 0.0, 0.0, 0.0};
  if constexpr (MST) {
    const int l = tid & 255, mi = tl.i0 + (l & 7), mj = tl.j0 + ((l >> 3) & 7), mk = tl.k0 + (l >> 6);
    const bool inb = mi <= B.n1 && mj <= B.n2 && mk <= B.n3;
    const int64_t cm = inb ? row(B, mi, mj, mk) : 0;
    if (tid < 256) {
#pragma unroll
      for (int q = 0; q < 5; ++q) Pl[q * NX + NB + l] = inb ? C.m[(int64_t)q * NC + cm] : 0.0;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) mreg[q] = inb ? C.m[(int64_t)(5 + q) * NC + cm] : 0.0;
    }
  }

  // ---------------- phase C: viscous faces ----------------
  if constexpr (MST) {
    double* Mx = Cc + NB;   // metric components 5..8 of the tile's cells
    __syncthreads();
    if (tid >= 256) {
#pragma unroll
      for (int q = 0; q < 4; ++q) Mx[q * 256 + (tid & 255)] = mreg[q];
    }
    __syncthreads();
    for (int g = tid; g < NF; g += BS) {
      // the faces with both cells in the tile first (224 i-, 224 j-, 192 k-faces), then the 256 on the tile's sides,
      // so that every wavefront takes one of the two kinds; side 0 = the lower cell outside, 1 = the upper
      int d, a, b, c, side = -1;
      if (g < 224) { d = 0; a = 1 + g % 7; b = (g / 7) & 7; c = g / 56; }
      else if (g < 448) { const int h = g - 224; d = 1; a = h & 7; b = 1 + (h >> 3) % 7; c = h / 56; }
      else if (g < 640) { const int h = g - 448; d = 2; a = h & 7; b = (h >> 3) & 7; c = 1 + (h >> 6); }
      else {
        const int h = g - 640;
        if (h < 64) { d = 0; side = h >> 5; b = h & 7; c = (h & 31) >> 3; a = side ? 8 : 0; }
        else if (h < 128) { d = 1; side = (h - 64) >> 5; a = h & 7; c = ((h - 64) & 31) >> 3; b = side ? 8 : 0; }
        else { d = 2; side = (h - 128) >> 6; a = h & 7; b = ((h - 128) & 63) >> 3; c = side ? 4 : 0; }
      }
      const int f = d == 0 ? a + 9 * (b + 8 * c) : (d == 1 ? NI + a + 8 * (b + 9 * c) : NI + NJ + a + 8 * (b + 8 * c));
      const int i = tl.i0 + a - (d == 0), j = tl.j0 + b - (d == 1), k = tl.k0 + c - (d == 2);
      const int n_d = d == 0 ? B.n1 : (d == 1 ? B.n2 : B.n3);
      const int l_d = d == 0 ? i : (d == 1 ? j : k);
      if (!((l_d >= 1 && l_d <= n_d - 1) && i <= B.n1 && j <= B.n2 && k <= B.n3)) continue;
      const int k1 = (a + 1 - (d == 0)) + CI * ((b + 1 - (d == 1)) + CJ * (c + 1 - (d == 2)));
      const int64_t c0 = row(B, i, j, k);
      const double* fm = d == 0 ? C.fm1 : (d == 1 ? C.fm2 : C.fm3);
      const int64_t NFg = d == 0 ? C.NF1 : (d == 1 ? C.NF2 : C.NF3);
      const int64_t fg = d == 0 ? B.f1off + i + (int64_t)(B.n1 + 1) * ((j - 1) + (int64_t)B.n2 * (k - 1))
                       : d == 1 ? B.f2off + (i - 1) + (int64_t)B.n1 * (j + (int64_t)(B.n2 + 1) * (k - 1))
                                : B.f3off + (i - 1) + (int64_t)B.n1 * ((j - 1) + (int64_t)B.n2 * k);
      const double fnx = fm[fg], fny = fm[NFg + fg], fnz = fm[2 * NFg + fg], fa = fm[3 * NFg + fg];
      // the tile cells of the face (lower lo, upper up) and the face-averaged metric
      const int lo = (a - (d == 0)) + 8 * ((b - (d == 1)) + 8 * (c - (d == 2)));
      const int up = lo + (d == 0 ? 1 : (d == 1 ? 8 : 64));
      auto mt = [&](int q, int l) { return q < 5 ? Pl[q * NX + NB + l] : Mx[(q - 5) * 256 + l]; };
      double mg[9];
      if (side < 0) {
#pragma unroll
        for (int q = 0; q < 9; ++q) mg[q] = 0.5 * (mt(q, lo) + mt(q, up));
      } else {
        const int lin = side ? lo : up;
        const int64_t cout = side ? c0 + (d == 0 ? si : (d == 1 ? sj : sk)) : c0;
#pragma unroll
        for (int q = 0; q < 9; ++q) mg[q] = 0.5 * (mt(q, lin) + C.m[(int64_t)q * NC + cout]);
      }
      double res[NP];
      if (d == 0) visc_face_lds<NS, NR, 0, true>(C, Pl, NX, Cc, NB, k1, k1 + 1, CI, CI * CJ, c0, c0, fnx, fny, fnz, fa, res, mg);
      else if (d == 1) visc_face_lds<NS, NR, 1, true>(C, Pl, NX, Cc, NB, k1, k1 + CI, 1, CI * CJ, c0, c0, fnx, fny, fnz, fa, res, mg);
      else visc_face_lds<NS, NR, 2, true>(C, Pl, NX, Cc, NB, k1, k1 + CI * CJ, 1, CI, c0, c0, fnx, fny, fnz, fa, res, mg);
      // the species row of Fb holds metric components: one species has no diffusion flux (res[0] = 0)
#pragma unroll
      for (int s = 1; s < NP; ++s) Fb[s * NF + f] = res[s];
    }
    __syncthreads();
    if (mine) {
#pragma unroll
      for (int s = 1; s < NP; ++s) {
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
  } else if constexpr (VISC) {
    __syncthreads();
    for (int f = tid; f < NF; f += BS) {
      int d, a, b, c;
      if (f < NI) { d = 0; a = f % (TI + 1); int r = f / (TI + 1); b = r % TJ; c = r / TJ; }
      else if (f < NI + NJ) { int g = f - NI; d = 1; a = g % TI; int r = g / TI; b = r % (TJ + 1); c = r / (TJ + 1); }
      else { int g = f - NI - NJ; d = 2; a = g % TI; int r = g / TI; b = r % TJ; c = r / TJ; }
      const int i = tl.i0 + a - (d == 0), j = tl.j0 + b - (d == 1), k = tl.k0 + c - (d == 2);
      const int n_d = d == 0 ? B.n1 : (d == 1 ? B.n2 : B.n3);
      const int l_d = d == 0 ? i : (d == 1 ? j : k);
      if (!((l_d >= 1 && l_d <= n_d - 1) && i <= B.n1 && j <= B.n2 && k <= B.n3)) continue;
      const int k1 = (a + 1 - (d == 0)) + CI * ((b + 1 - (d == 1)) + CJ * (c + 1 - (d == 2)));
      const int64_t c0 = row(B, i, j, k);
      const double* fm = d == 0 ? C.fm1 : (d == 1 ? C.fm2 : C.fm3);
      const int64_t NFg = d == 0 ? C.NF1 : (d == 1 ? C.NF2 : C.NF3);
      const int64_t fg = d == 0 ? B.f1off + i + (int64_t)(B.n1 + 1) * ((j - 1) + (int64_t)B.n2 * (k - 1))
                       : d == 1 ? B.f2off + (i - 1) + (int64_t)B.n1 * (j + (int64_t)(B.n2 + 1) * (k - 1))
                                : B.f3off + (i - 1) + (int64_t)B.n1 * ((j - 1) + (int64_t)B.n2 * k);
      const double fnx = fm[fg], fny = fm[NFg + fg], fnz = fm[2 * NFg + fg], fa = fm[3 * NFg + fg];
      double res[NP];
      if (d == 0) visc_face_lds<NS, NR, 0>(C, Pl, NX, Cc, NB, k1, k1 + 1, CI, CI * CJ, c0, c0 + siv, fnx, fny, fnz, fa, res);
      else if (d == 1) visc_face_lds<NS, NR, 1>(C, Pl, NX, Cc, NB, k1, k1 + CI, 1, CI * CJ, c0, c0 + sj, fnx, fny, fnz, fa, res);
      else visc_face_lds<NS, NR, 2>(C, Pl, NX, Cc, NB, k1, k1 + CI * CJ, 1, CI, c0, c0 + sk, fnx, fny, fnz, fa, res);
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
    const bool on[6] = {ci == 1, ci == B.n1, cj == 1, cj == B.n2, ck == 1, ck == B.n3};
#pragma unroll
    for (int f = 1; f <= 6; ++f) {
      if (!on[f - 1]) continue;
      const int64_t pos = B.bpoff[f - 1] + plane_pos(B, f, ci, cj, ck);
#pragma unroll
      for (int s = 0; s < NP; ++s) {
        tmp[s] = tmp[s] + C.binc[(int64_t)s * C.NBP + pos];
        tmp[s] = tmp[s] + C.binc[(int64_t)(NP + s) * C.NBP + pos];
      }
    }
    const int64_t c0 = row(B, ci, cj, ck);
    double p[NP], po[NP], out[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) { p[s] = Pl[s * NX + kc0]; po[s] = C.PO[s * NC + c0]; }
    const double w = *C.dt * C.strang * rcp_nr(C.vol[c0]);
    double Tseed;
    if constexpr (VISC) {
      Tseed = Cc[FC::iT * NB + kc0];
      if (C.guard && !(Tseed > 1.0 && Tseed < 9990.0)) {
        flag_bad(C, tl.b, ci, cj, ck, 1);
        if (Tseed != Tseed) Tseed = 300.0;
        Tseed = dmin2(dmax2(Tseed, 1.0), 9990.0);
      }
    } else {
      const double rho = rho_sum<NS>(p), Rg = f_Rtot<NS>(T, p);
      Tseed = p[V::np] / (rho * Rg);
    }
    if constexpr (NR > 0) {
      const double rho = VISC ? Cc[kc0] : rho_sum<NS>(p);
      cell_sa<NS, NR>(C, B, ci, cj, ck, c0, rho, cell_mil<NS>(T, p, rho, Tseed), tmp);
    }
    // the primitive update below nine species (the nine-species kernels take the conservative one only)
    if constexpr (NS < 9) {
      if (C.prim) update_cell_prim<NS, NR>(T, p, po, tmp, C.vol[c0], *C.dt, C.strang, A.rkc, out);
      else update_cell_nr<NS, NR>(T, p, po, tmp, w, A.rkc, Tseed, out);
    } else {
      update_cell_nr<NS, NR>(T, p, po, tmp, w, A.rkc, Tseed, out);
    }
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

// The performance path's kernel: the parity kernel's workgroup -> tile map (xcd_tile), BS threads, the tile shape
// fixed at compile time for SH = 1 (8 x 8 x 4) and 2 (8 x 8 x 8), read at run time for SH = 0.  The one-species
// 8 x 8 x 4 kernel without RANS asks for 4 waves per SIMD (at most 128 VGPRs): two of its 75 KB workgroups share a CU.
template <int NS, int NR, bool VISC, int RS, int BS, int SH>
__global__ void __launch_bounds__(BS, (SH == 1 && NS == 1 && NR == 0) ? 4 : 1) k_stage_fast(CaseDev C, StageArgs A) {
  const int t = xcd_tile(blockIdx.x, A.ntiles);
  if (t < 0) return;
  constexpr int STI = SH ? 8 : 0, STJ = SH ? 8 : 0, STK = SH == 1 ? 4 : (SH == 2 ? 8 : 0);
  stage_tile_fast<NS, NR, VISC, RS, BS, STI, STJ, STK>(C, A, t);
}

}  // namespace mose
