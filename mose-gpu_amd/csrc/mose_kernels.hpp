// Kernels of the residual path on the flat var-last arena.
//
//   k_ghost      one thread per boundary record: the physical ghost layers (zero gradient then cubic for the
//                extrapolation faces, cubic for walls) and the block-to-block / periodic halo copy
//   k_bcrec      one thread per boundary record: the record's two residual increments (Euler part, viscous
//                part) into the boundary-face planes
//   k_face<D>    one thread per interior face of direction D: convective and viscous face fluxes
//   k_update     one thread per interior cell: the residual in the fixed order of the reference, the
//                boundary increments in face order, the RK stage update and the state check
//   k_dt         one thread per interior cell: the CFL / VNN limit and a wavefront + atomic min
#pragma once

#include "mose_bcdev.hpp"
#include "mose_physics.hpp"

namespace mose {

struct CaseDev {
  // state buffers
  const double* Pin;
  double* Pout;
  const double* PO;
  double* Pg;          // buffer whose ghosts are filled (= Pin)
  int64_t NC, NF1, NF2, NF3;
  // geometry
  const double *dl, *m, *vol, *yn, *fm1, *fm2, *fm3;
  // interior face buffers: Fc[d][(s*NFI[d]) + f], Fd alike
  double* Fc[3];
  double* Fd[3];
  int64_t NFI[3];
  // boundary planes: inc[(q*nprim+s)*NBP + bpoff + pos], q = 0 Euler, 1 viscous
  double* binc;
  int64_t NBP;
  const mose_blk* blk;
  const mose_bc_record* rec;
  const mose_farfield* ff;       // far-field table of the type-4 records
  const signed char* btyp;       // record type of each boundary-plane position (walls mark the SA wall ghost)
  int nrec;
  int lim, rs, visc, guard, pimp;
  double mco;
  ViscConst vc;
  const mose_dtab* tab;
  double* dt;            // dt of the current step (device scalar)
  double strang;         // strangcoeff of the update: 1, or 1/2 in a Strang-split step (Lib_Strang.f90:10-22)
  unsigned long long* dtmin;   // min reduction slot (positive doubles as bits)
  unsigned int* bad;     // [0] count of failing cells, [1] thermo guard count
  unsigned long long* first_bad;
  double cfl, vnn;
  int ndir;
  int prim;              // integration-variables = prim: the primitive update (update_cell_prim); in the padding after ndir
};

__device__ __forceinline__ int64_t row(const mose_blk& B, int i, int j, int k) {
  return B.boff + (int64_t)(i + GC - 1) + (int64_t)(j + GC - 1) * B.pi + (int64_t)(k + GC - 1) * B.pi * B.pj;
}
__device__ __forceinline__ int64_t plane_pos(const mose_blk& B, int f, int i, int j, int k) {
  if (f <= 2) return (j - 1) + (int64_t)B.n2 * (k - 1);
  if (f <= 4) return (i - 1) + (int64_t)B.n1 * (k - 1);
  return (i - 1) + (int64_t)B.n1 * (j - 1);
}

__constant__ int c_guide[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

// the unit normal and area of a record's boundary face: fm<dir> of the face at index (cell - modfm) along the face
// direction (bc_record_body.inc:9-24)
__device__ __forceinline__ void rec_face(const CaseDev& C, const mose_blk& B, const mose_bc_record& r, double& nx,
                                         double& ny, double& nz, double& area) {
  const int dir = (r.f + 1) / 2, modfm = r.f % 2;
  const double* fm;
  int64_t f, NF;
  if (dir == 1) { fm = C.fm1; NF = C.NF1; f = B.f1off + (r.i - modfm) + (int64_t)(B.n1 + 1) * ((r.j - 1) + (int64_t)B.n2 * (r.k - 1)); }
  else if (dir == 2) { fm = C.fm2; NF = C.NF2; f = B.f2off + (r.i - 1) + (int64_t)B.n1 * ((r.j - modfm) + (int64_t)(B.n2 + 1) * (r.k - 1)); }
  else { fm = C.fm3; NF = C.NF3; f = B.f3off + (r.i - 1) + (int64_t)B.n1 * ((r.j - 1) + (int64_t)B.n2 * (r.k - modfm)); }
  nx = fm[f]; ny = fm[NF + f]; nz = fm[2 * NF + f]; area = fm[3 * NF + f];
}

// ------------------------------------------------------------------------------------------------------------
// ghosts (Fill_Ghost_Cell_local, Lib_Extrapolation.f90:1370-1524)
// ------------------------------------------------------------------------------------------------------------
template <int NS, int NR>
__global__ void __launch_bounds__(256) k_ghost(CaseDev C) {
  using V = Vars<NS, NR>;
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= C.nrec) return;
  const mose_bc_record r = C.rec[e];
  const mose_blk B = C.blk[r.b];
  double* P = C.Pg;
  const int64_t NC = C.NC;
  int dir = (r.f + 1) / 2 - 1, modfm = r.f % 2, gstep = 2 * modfm - 1;
  int64_t st = dir == 0 ? 1 : (dir == 1 ? (int64_t)B.pi : (int64_t)B.pi * B.pj);
  int64_t c = row(B, r.i, r.j, r.k);
  // ghost g of the face lies at c - gstep*g*st
  if (r.typ == MOSE_BC_CONNECTION) {
    // Ghost_Connection_pair (:1023-1135): ghost g <- donor layer g-1 (donor rows are interior)
    const mose_blk S = C.blk[r.bs];
    for (int g = 1; g <= GC; ++g) {
      int64_t dst = row(B, r.i - c_guide[r.f - 1][0] * g, r.j - c_guide[r.f - 1][1] * g, r.k - c_guide[r.f - 1][2] * g);
      int64_t src = row(S, r.is + c_guide[r.fs - 1][0] * (g - 1), r.js + c_guide[r.fs - 1][1] * (g - 1),
                        r.ks + c_guide[r.fs - 1][2] * (g - 1));
#pragma unroll
      for (int s = 0; s < V::nprim; ++s) P[s * NC + dst] = P[s * NC + src];
    }
    return;
  }
  if (r.typ == MOSE_BC_SYMMETRY) {
    // Ghost_Symmetry (Lib_Extrapolation.f90:558-587): the first ghost layer only, the cell with its velocity mirrored
    // in the face.  The second layer is not written; no stencil of the residual reads it behind a symmetry face.
    double nx, ny, nz, area;
    rec_face(C, B, r, nx, ny, nz, area);
    const int64_t dst = c - (int64_t)gstep * st;
    double u = P[V::nu * NC + c], v = P[V::nv * NC + c], w = P[V::nw * NC + c];
    double udn = u * nx + v * ny + w * nz;
    u = u - 2.0 * udn * nx;
    v = v - 2.0 * udn * ny;
    w = w - 2.0 * udn * nz;
#pragma unroll
    for (int s = 0; s < NS; ++s) P[s * NC + dst] = P[s * NC + c];
    P[V::nu * NC + dst] = u;
    P[V::nv * NC + dst] = v;
    P[V::nw * NC + dst] = w;
#pragma unroll
    for (int s = V::np; s < V::nprim; ++s) P[s * NC + dst] = P[s * NC + c];
    return;
  }
  int ext = dir == 0 ? B.n1 : (dir == 1 ? B.n2 : B.n3);
  bool zg = (r.typ == MOSE_BC_EXTRAPOLATION || r.typ == MOSE_BC_FARFIELD || r.typ == 0);
  bool cub = (r.typ == MOSE_BC_EXTRAPOLATION && ext >= 3) || r.typ == MOSE_BC_WALL_ADIABATIC ||
             r.typ == MOSE_BC_WALL_ISOTHERMAL;
  if (zg) {
    // Ghost_ZG_Extrapolate (:590-607)
    for (int g = 1; g <= GC; ++g) {
      int64_t dst = c - (int64_t)gstep * g * st;
#pragma unroll
      for (int s = 0; s < V::nprim; ++s) P[s * NC + dst] = P[s * NC + c];
    }
  }
  if (cub) {
    // Ghost_Extrapolate (:616-640), layer g from layers g-1, g-2, g-3 (sequential in g).  On a one-cell direction
    // the cubic pass of the 400 faces reproduces the zero-gradient values exactly (3x - 3x + x), so it is skipped.
    int64_t stp = -(int64_t)gstep * st;
    for (int g = 1; g <= GC; ++g) {
      int64_t d0 = c + g * stp, d1 = c + (g - 1) * stp, d2 = c + (g - 2) * stp, d3 = c + (g - 3) * stp;
#pragma unroll
      for (int s = 0; s < V::nprim; ++s)
        P[s * NC + d0] = 3.0 * P[s * NC + d1] - 3.0 * P[s * NC + d2] + P[s * NC + d3];
    }
  }
}

// ------------------------------------------------------------------------------------------------------------
// boundary records (bc_record_body.inc + bc_record_apply.inc; Lib_BC.f90)
// ------------------------------------------------------------------------------------------------------------
template <int NS, int NR>
__global__ void __launch_bounds__(128) k_bcrec(CaseDev C) {
  using V = Vars<NS, NR>;
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= C.nrec) return;
  const mose_bc_record r = C.rec[e];
  if (r.typ == MOSE_BC_NONE) return;   // no flux: its plane entries stay at the zeros of the case's creation
  const mose_blk B = C.blk[r.b];
  const mose_dtab& T = *C.tab;
  const double* P = C.Pin;
  const int64_t NC = C.NC;
  int Im = r.i, Jm = r.j, Km = r.k, Fm = r.f;
  int dir = (Fm + 1) / 2, modfm = Fm % 2, gstep = 2 * modfm - 1, modfm2 = 1 - 2 * modfm;
  int64_t st = dir == 1 ? 1 : (dir == 2 ? (int64_t)B.pi : (int64_t)B.pi * B.pj);
  int64_t c = row(B, Im, Jm, Km);
  double nx, ny, nz, area;
  rec_face(C, B, r, nx, ny, nz, area);
  double e1[V::nprim], e2[V::nprim];
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) { e1[s] = 0.0; e2[s] = 0.0; }
  if (r.typ == MOSE_BC_EXTRAPOLATION) {
    // BC_Extrapolation (Lib_BC.f90:1119-1163): the physical flux of the boundary cell
    double q[V::nprim], Fl[V::nprim];
#pragma unroll
    for (int s = 0; s < V::nprim; ++s) q[s] = P[s * NC + c];
    double Un = q[V::nu] * nx + q[V::nv] * ny + q[V::nw] * nz, Fm_ = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) { Fl[s] = q[s] * Un * area; Fm_ = Fm_ + Fl[s]; }
    Fl[V::nu] = Fm_ * q[V::nu] + q[V::np] * area * nx;
    Fl[V::nv] = Fm_ * q[V::nv] + q[V::np] * area * ny;
    Fl[V::nw] = Fm_ * q[V::nw] + q[V::np] * area * nz;
    double vmag = sqrt(q[V::nu] * q[V::nu] + q[V::nv] * q[V::nv] + q[V::nw] * q[V::nw]);
    Fl[V::np] = Fm_ * H0<NS>(T, q[V::np], q, vmag);
#pragma unroll
    for (int s = V::np + 1; s < V::nprim; ++s) Fl[s] = q[s] * Un * area;
    double sg = (double)(1 - 2 * modfm);
#pragma unroll
    for (int s = 0; s < V::nprim; ++s) e1[s] = sg * Fl[s];
  } else if (r.typ == MOSE_BC_CONNECTION) {
    // BC_Connection_Eul_lean (Lib_BC.f90:1770-1864): MUSCL from the ghosts, Riemann flux of the case
    int o1 = modfm * (-gstep * GC) + (1 - modfm) * gstep, o2 = modfm * (-gstep), o3 = (1 - modfm) * (-gstep);
    int o4 = (1 - modfm) * (-gstep * GC) + modfm * gstep;
    int64_t r1 = c + o1 * st, r2 = c + o2 * st, r3 = c + o3 * st, r4 = c + o4 * st;
    const double* dlp = C.dl + (int64_t)(dir - 1) * NC;
    double Dl1 = dlp[r1] * 0.5, Dl2 = dlp[r2] * 0.5, Dl3 = dlp[r3] * 0.5, Dl4 = dlp[r4] * 0.5;
    double D12 = Dl1 + Dl2, D23 = Dl2 + Dl3, D34 = Dl3 + Dl4;
    double L[V::nprim], R[V::nprim];
#pragma unroll
    for (int s = 0; s < V::nprim; ++s) {
      double q1 = P[s * NC + r1], q2 = P[s * NC + r2], q3 = P[s * NC + r3], q4 = P[s * NC + r4];
      double d21 = (q2 - q1) / D12, d32 = (q3 - q2) / D23, d43 = (q4 - q3) / D34;
      L[s] = q2 + 1.0 * limiter(C.lim, d32, d21) * Dl2;
      R[s] = q3 - 1.0 * limiter(C.lim, d43, d32) * Dl3;
    }
    double rL = rho_sum<NS>(L), RL = f_Rtot<NS>(T, L), rR = rho_sum<NS>(R), RR = f_Rtot<NS>(T, R);
    double aL = f_ss<NS>(T, L, L[V::np], rL, RL), aR = f_ss<NS>(T, R, R[V::np], rR, RR);
    F5 F = riemann<NS>(T, C.rs, C.mco, false, L, L[V::nu], L[V::nv], L[V::nw], L[V::np], aL, rL, R, R[V::nu], R[V::nv],
                       R[V::nw], R[V::np], aR, rR, nx, ny, nz, RL, RR);
    double su = copysign(0.5, F.r), sL = 0.5 + su, sR = su - 0.5, Fms = F.r * area, m2 = (double)modfm2;
#pragma unroll
    for (int s = 0; s < NS; ++s) e1[s] = m2 * Fms * (sL * L[s] / rL - sR * R[s] / rR);
    e1[V::nu] = m2 * F.u * area; e1[V::nv] = m2 * F.v * area; e1[V::nw] = m2 * F.w * area; e1[V::np] = m2 * F.e * area;
#pragma unroll
    for (int s = V::np + 1; s < V::nprim; ++s) e1[s] = m2 * Fms * (sL * L[s] / rL - sR * R[s] / rR);
    if (C.visc) {
      // BC_Connection_Visc_lean (Lib_BC.f90:2113-2201): the face gradient on the receiver side and on the donor
      // side (transverse donor neighbours, Ghost_Connection_pairs_B :1136-1176), the two record fluxes averaged
      const mose_blk S = C.blk[r.bs];
      int ia = dir - 1, ib = dir == 1 ? 1 : 0, ic = dir == 3 ? 1 : 2;
      int64_t sb = (ib == 0) ? 1 : (int64_t)B.pi;
      int64_t sc = (ic == 1) ? (int64_t)B.pi : (int64_t)B.pi * B.pj;
      int64_t cg = c - (int64_t)gstep * st;
      int tq[4][3];
      {
        int Is = r.is, Js = r.js, Ks = r.ks;
        if (r.fs <= 2) {
          int t[4][3] = {{Is, Js - r.d11, Ks - r.d12}, {Is, Js + r.d11, Ks + r.d12}, {Is, Js - r.d21, Ks - r.d22}, {Is, Js + r.d21, Ks + r.d22}};
          for (int a = 0; a < 4; ++a) for (int b = 0; b < 3; ++b) tq[a][b] = t[a][b];
        } else if (r.fs <= 4) {
          int t[4][3] = {{Is - r.d11, Js, Ks - r.d12}, {Is + r.d11, Js, Ks + r.d12}, {Is - r.d21, Js, Ks - r.d22}, {Is + r.d21, Js, Ks + r.d22}};
          for (int a = 0; a < 4; ++a) for (int b = 0; b < 3; ++b) tq[a][b] = t[a][b];
        } else {
          int t[4][3] = {{Is - r.d11, Js - r.d12, Ks}, {Is + r.d11, Js + r.d12, Ks}, {Is - r.d21, Js - r.d22, Ks}, {Is + r.d21, Js + r.d22, Ks}};
          for (int a = 0; a < 4; ++a) for (int b = 0; b < 3; ++b) tq[a][b] = t[a][b];
        }
      }
      int64_t cr[10] = {c, cg, c + sb, c - sb, c + sc, c - sc, row(S, tq[0][0], tq[0][1], tq[0][2]),
                        row(S, tq[1][0], tq[1][1], tq[1][2]), row(S, tq[2][0], tq[2][1], tq[2][2]),
                        row(S, tq[3][0], tq[3][1], tq[3][2])};
      double Mm[3][3];
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) Mm[a][b] = 0.5 * (C.m[(int64_t)(a + 3 * b) * NC + c] + C.m[(int64_t)(a + 3 * b) * NC + cg]);
      double rh[10], rg[10];
#pragma unroll
      for (int q = 0; q < 10; ++q) {
        double y[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) y[s] = P[s * NC + cr[q]];
        rh[q] = rho_sum<NS>(y);
        rg[q] = f_Rtot<NS>(T, y);
      }
      double Gl[V::nprim][3], Gg[V::nprim][3], Pa[V::nprim];
#pragma unroll
      for (int s = 0; s < V::nprim; ++s) {
        double vv[10];
#pragma unroll
        for (int q = 0; q < 10; ++q) {
          double pv = P[s * NC + cr[q]];
          vv[q] = (s >= V::nu && s <= V::nw) ? pv : (s == V::np ? pv / (rh[q] * rg[q]) : pv / rh[q]);
        }
        double gl[3], gg[3];
        gl[ia] = (vv[1] - vv[0]) * (double)modfm2;
        gl[ib] = (vv[2] - vv[3]) * 0.5;
        gl[ic] = (vv[4] - vv[5]) * 0.5;
        gg[ia] = (vv[0] - vv[1]) * (double)(-modfm2);
        gg[ib] = (vv[7] - vv[6]) * 0.5;
        gg[ic] = (vv[9] - vv[8]) * 0.5;
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
          Gl[s][kk] = gl[0] * Mm[0][kk] + gl[1] * Mm[1][kk] + gl[2] * Mm[2][kk];
          Gg[s][kk] = gg[0] * Mm[0][kk] + gg[1] * Mm[1][kk] + gg[2] * Mm[2][kk];
        }
        Pa[s] = 0.5 * (P[s * NC + cr[0]] + P[s * NC + cr[1]]);
      }
      FaceThermo ft = face_thermo<NS>(T, Pa, Pa[V::np], C.vc.prl);
      double Fl[V::nprim], Fg[V::nprim];
      diff_flux_st<NS, NR>(T, C.vc, Gl, Pa, ft, nx, ny, nz, area, Fl);
      diff_flux_st<NS, NR>(T, C.vc, Gg, Pa, ft, nx, ny, nz, area, Fg);
#pragma unroll
      for (int s = 0; s < V::nprim; ++s) e2[s] = -(0.5 * (double)modfm2 * (Fl[s] + Fg[s]));
    }
  }
  if (r.typ == MOSE_BC_SYMMETRY) {
    RecCells rc{c, c + gstep * st, dir, modfm};
    bc_symmetry<NS, NR>(T, P, NC, C.dl, C.lim, C.rs, C.mco, rc, nx, ny, nz, area, e1);
    if (C.visc) {
      // transverse strides in the order of BC_Symmetry_Visc's stencil: (j, k), (i, k), (i, j)
      const int64_t sj = B.pi, sk = (int64_t)B.pi * B.pj;
      bc_sym_visc<NS, NR>(T, P, NC, C.m, rc, dir == 1 ? sj : 1, dir == 3 ? sj : sk, nx, ny, nz, area, e2);
    }
  }
  if (r.typ == MOSE_BC_WALL_ADIABATIC || r.typ == MOSE_BC_WALL_ISOTHERMAL || r.typ == MOSE_BC_FARFIELD) {
    RecCells rc{c, c + gstep * st, dir, modfm};
    bool closed = false;
    if (r.typ == MOSE_BC_FARFIELD)
      closed = !bc_farfield<NS, NR>(T, P, NC, C.lim, C.rs, C.mco, C.ff[r.ff], r.tvec, rc, nx, ny, nz, area, e1);
    if (r.typ != MOSE_BC_FARFIELD || closed) bc_symmetry<NS, NR>(T, P, NC, C.dl, C.lim, C.rs, C.mco, rc, nx, ny, nz, area, e1);
    if (C.visc && r.typ != MOSE_BC_FARFIELD)
      bc_wall_visc<NS, NR>(T, P, NC, C.m, C.vc.prl, r.typ, r.tw, rc, nx, ny, nz, area, e2);
  }
  int64_t pos = B.bpoff[Fm - 1] + plane_pos(B, Fm, Im, Jm, Km);
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) {
    C.binc[(int64_t)s * C.NBP + pos] = e1[s];
    C.binc[(int64_t)(V::nprim + s) * C.NBP + pos] = e2[s];
  }
}

// ------------------------------------------------------------------------------------------------------------
// interior faces of direction D (conv_cuf_i/j/k + diff_cuf_i/j/k): one thread per face of one block
// ------------------------------------------------------------------------------------------------------------
template <int NS, int NR, int D>
__global__ void __launch_bounds__(128) k_face(CaseDev C, int b) {
  using V = Vars<NS, NR>;
  const mose_blk B = C.blk[b];
  int hi1 = B.n1 - (D == 0), hi2 = B.n2 - (D == 1), hi3 = B.n3 - (D == 2);
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t nf = (int64_t)hi1 * hi2 * hi3;
  if (t >= nf) return;
  int i = (int)(t % hi1) + 1;
  int64_t rem = t / hi1;
  int j = (int)(rem % hi2) + 1, k = (int)(rem / hi2) + 1;
  const mose_dtab& T = *C.tab;
  const double* P = C.Pin;
  const int64_t NC = C.NC;
  int64_t st = D == 0 ? 1 : (D == 1 ? (int64_t)B.pi : (int64_t)B.pi * B.pj);
  int64_t c0 = row(B, i, j, k);
  double qm[V::nprim], q0[V::nprim], q1[V::nprim], q2[V::nprim], res[V::nprim], dlc[4];
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) {
    qm[s] = P[s * NC + c0 - st]; q0[s] = P[s * NC + c0]; q1[s] = P[s * NC + c0 + st]; q2[s] = P[s * NC + c0 + 2 * st];
  }
  const double* dlp = C.dl + (int64_t)D * NC;
#pragma unroll
  for (int q = 0; q < 4; ++q) dlc[q] = dlp[c0 + (q - 1) * st];
  const double* fm = D == 0 ? C.fm1 : (D == 1 ? C.fm2 : C.fm3);
  int64_t NF = D == 0 ? C.NF1 : (D == 1 ? C.NF2 : C.NF3);
  int64_t fg = D == 0 ? B.f1off + i + (int64_t)(B.n1 + 1) * ((j - 1) + (int64_t)B.n2 * (k - 1))
             : D == 1 ? B.f2off + (i - 1) + (int64_t)B.n1 * (j + (int64_t)(B.n2 + 1) * (k - 1))
                      : B.f3off + (i - 1) + (int64_t)B.n1 * ((j - 1) + (int64_t)B.n2 * k);
  double nx = fm[fg], ny = fm[NF + fg], nz = fm[2 * NF + fg], area = fm[3 * NF + fg];
  conv_face<NS, NR>(T, C.lim, C.rs, C.mco, qm, q0, q1, q2, dlc, nx, ny, nz, area, res);
  int64_t fl = B.ifoff[D] + t;   // the reference's face numbering f is i fastest over the same ranges
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) C.Fc[D][(int64_t)s * C.NFI[D] + fl] = res[s];
  if (!C.visc) return;
  // diff_cuf_kernels.inc: c1,c2 across the face, +-1 on axis ib at both cells, +-1 on axis ic at both cells
  constexpr int ia = D, ib = D == 0 ? 1 : 0, ic = D == 2 ? 1 : 2;
  int64_t sb = ib == 0 ? 1 : (int64_t)B.pi;
  int64_t sc = ic == 1 ? (int64_t)B.pi : (int64_t)B.pi * B.pj;
  int64_t c1 = c0 + st;
  double G[V::nprim][3];
  double rh1 = rho_sum<NS>(q0), rh2 = rho_sum<NS>(q1);
  double T1 = q0[V::np] / (rh1 * f_Rtot<NS>(T, q0)), T2 = q1[V::np] / (rh2 * f_Rtot<NS>(T, q1));
  double ir1 = 1.0 / rh1, ir2 = 1.0 / rh2;
#pragma unroll
  for (int s = 0; s < NS; ++s) G[s][ia] = q1[s] * ir2 - q0[s] * ir1;
  G[V::np][ia] = T2 - T1;
#pragma unroll
  for (int s = V::nu; s <= V::nw; ++s) G[s][ia] = q1[s] - q0[s];
#pragma unroll
  for (int s = V::nt; s < V::nprim; ++s) G[s][ia] = q1[s] * ir2 - q0[s] * ir1;
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    int ax = pass ? ic : ib;
    int64_t sd = pass ? sc : sb;
    int64_t a3 = c0 - sd, a4 = c0 + sd, a5 = c1 - sd, a6 = c1 + sd;
    double r3[NS], r4[NS], r5[NS], r6[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) { r3[s] = P[s * NC + a3]; r4[s] = P[s * NC + a4]; r5[s] = P[s * NC + a5]; r6[s] = P[s * NC + a6]; }
    double h3 = rho_sum<NS>(r3), h4 = rho_sum<NS>(r4), h5 = rho_sum<NS>(r5), h6 = rho_sum<NS>(r6);
    double t3 = P[V::np * NC + a3] / (h3 * f_Rtot<NS>(T, r3)), t4 = P[V::np * NC + a4] / (h4 * f_Rtot<NS>(T, r4));
    double t5 = P[V::np * NC + a5] / (h5 * f_Rtot<NS>(T, r5)), t6 = P[V::np * NC + a6] / (h6 * f_Rtot<NS>(T, r6));
    double i3 = 1.0 / h3, i4 = 1.0 / h4, i5 = 1.0 / h5, i6 = 1.0 / h6;
#pragma unroll
    for (int s = 0; s < NS; ++s) G[s][ax] = (r4[s] * i4 - r3[s] * i3 + r6[s] * i6 - r5[s] * i5) * 0.25;
    G[V::np][ax] = (t4 - t3 + t6 - t5) * 0.25;
#pragma unroll
    for (int s = V::nu; s <= V::nw; ++s)
      G[s][ax] = (P[s * NC + a4] - P[s * NC + a3] + P[s * NC + a6] - P[s * NC + a5]) * 0.25;
#pragma unroll
    for (int s = V::nt; s < V::nprim; ++s)
      G[s][ax] = (P[s * NC + a4] * i4 - P[s * NC + a3] * i3 + P[s * NC + a6] * i6 - P[s * NC + a5] * i5) * 0.25;
  }
  double Mav[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b2 = 0; b2 < 3; ++b2) Mav[a][b2] = 0.5 * (C.m[(int64_t)(a + 3 * b2) * NC + c0] + C.m[(int64_t)(a + 3 * b2) * NC + c1]);
  double Pf[V::nprim], ps[NS];
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) Pf[s] = 0.5 * (q0[s] + q1[s]);
#pragma unroll
  for (int s = 0; s < NS; ++s) ps[s] = Pf[s];
  FaceThermo ft = face_thermo<NS>(T, ps, Pf[V::np], C.vc.prl);
  diff_flux_interior<NS, NR>(T, C.vc, G, Mav, Pf, ft, nx, ny, nz, area, res);
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) C.Fd[D][(int64_t)s * C.NFI[D] + fl] = res[s];
}

// ------------------------------------------------------------------------------------------------------------
// per cell: the residual (pass2_body.inc:1-61, then the boundary increments in face order 1..6), the stage
// update (newstate_cons_body.inc) and the state check (State_Check_blk, Lib_Newstate.f90:200-231)
// ------------------------------------------------------------------------------------------------------------
template <int NS, int NR>
__device__ __forceinline__ void cell_residual(const CaseDev& C, const mose_blk& B, int i, int j, int k, double* tmp) {
  using V = Vars<NS, NR>;
  int n1 = B.n1, n2 = B.n2, n3 = B.n3;
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) tmp[s] = 0.0;
  int64_t fi = B.ifoff[0] + (i - 1) + (int64_t)(j - 1) * (n1 - 1) + (int64_t)(k - 1) * (n1 - 1) * n2;
  int64_t fj = B.ifoff[1] + (i - 1) + (int64_t)(j - 1) * n1 + (int64_t)(k - 1) * n1 * (n2 - 1);
  int64_t fk = B.ifoff[2] + (i - 1) + (int64_t)(j - 1) * n1 + (int64_t)(k - 1) * n1 * n2;
  const double *Ci = C.Fc[0], *Cj = C.Fc[1], *Ck = C.Fc[2];
  int64_t Ni = C.NFI[0], Nj = C.NFI[1], Nk = C.NFI[2];
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) {
    double t = tmp[s];
    if (i < n1) t = t + Ci[s * Ni + fi];
    if (i > 1) t = t - Ci[s * Ni + fi - 1];
    if (j < n2) t = t + Cj[s * Nj + fj];
    if (j > 1) t = t - Cj[s * Nj + fj - n1];
    if (k < n3) t = t + Ck[s * Nk + fk];
    if (k > 1) t = t - Ck[s * Nk + fk - (int64_t)n1 * n2];
    tmp[s] = t;
  }
  if (C.visc) {
    const double *Di = C.Fd[0], *Dj = C.Fd[1], *Dk = C.Fd[2];
#pragma unroll
    for (int s = 0; s < V::nprim; ++s) {
      double t = tmp[s];
      if (i < n1) t = t - Di[s * Ni + fi];
      if (i > 1) t = t + Di[s * Ni + fi - 1];
      if (j < n2) t = t - Dj[s * Nj + fj];
      if (j > 1) t = t + Dj[s * Nj + fj - n1];
      if (k < n3) t = t - Dk[s * Nk + fk];
      if (k > 1) t = t + Dk[s * Nk + fk - (int64_t)n1 * n2];
      tmp[s] = t;
    }
  }
  // boundary records, face order 1..6 (BC_Fluxes launch order, Mod_BC.f90:1042-1062)
  bool on[6] = {i == 1, i == n1, j == 1, j == n2, k == 1, k == n3};
#pragma unroll
  for (int f = 1; f <= 6; ++f) {
    if (!on[f - 1]) continue;
    int64_t pos = B.bpoff[f - 1] + plane_pos(B, f, i, j, k);
#pragma unroll
    for (int s = 0; s < V::nprim; ++s) {
      tmp[s] = tmp[s] + C.binc[(int64_t)s * C.NBP + pos];
      tmp[s] = tmp[s] + C.binc[(int64_t)(V::nprim + s) * C.NBP + pos];
    }
  }
}

__device__ __forceinline__ void flag_bad(const CaseDev& C, int b, int i, int j, int k, int which) {
  atomicAdd(&C.bad[which], 1u);
  unsigned long long key = ((unsigned long long)(unsigned)b << 48) | ((unsigned long long)(unsigned)k << 32) |
                           ((unsigned long long)(unsigned)j << 16) | (unsigned long long)(unsigned)i;
  atomicMin(C.first_bad, key);
}

// The squared changes of one cell over the step (Compute_Residual_blk, Mod_Diagnostic.f90:63-106): the species
// sum, then u, v, w, p and the RANS variables.
template <int NS, int NR>
__device__ __forceinline__ void step_change(const double* pnew, const double* pold, double* r) {
  using V = Vars<NS, NR>;
  double d = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) d = d + (pnew[s] - pold[s]);
  r[0] = d * d;
#pragma unroll
  for (int s = V::nu; s < V::nprim; ++s) {
    double e = pnew[s] - pold[s];
    r[s - NS + 1] = e * e;
  }
}

// Workgroup sums of N values in a fixed order (wavefront butterfly, then the waves in order); thread 0 stores them.
// Every thread of the workgroup must call it.
template <int N>
__device__ __forceinline__ void block_sum_store(const double* v, double* out) {
  __shared__ double ws[16][N > 0 ? N : 1];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) x = x + __shfl_xor(x, o, 64);
    if (lane == 0) ws[w][q] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double t = ws[0][q];
      for (int k = 1; k < nw; ++k) t = t + ws[k][q];
      out[q] = t;
    }
  }
  __syncthreads();
}

// sums of the per-workgroup partials in index order (one thread per component)
static __global__ void k_rnorm_final(const double* part, int64_t nparts, int nres, double* out) {
  int q = threadIdx.x;
  if (blockIdx.x != 0 || q >= nres) return;
  double t = 0.0;
  for (int64_t i = 0; i < nparts; ++i) t = t + part[i * nres + q];
  out[q] = t;
}

// field_state_hash (MOSE_Field_m.f90:254-284): sums of the low and of the high 32 bits of every interior value
template <int NS, int NR>
__global__ void __launch_bounds__(256) k_hash(const double* P, int64_t NC, mose_blk B, unsigned long long* h) {
  using V = Vars<NS, NR>;
  __shared__ unsigned long long s1[256], s2[256];
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long a = 0, b = 0;
  if (t < (int64_t)B.n1 * B.n2 * B.n3) {
    int i = (int)(t % B.n1) + 1;
    int64_t rem = t / B.n1;
    int j = (int)(rem % B.n2) + 1, k = (int)(rem / B.n2) + 1;
    int64_t c = row(B, i, j, k);
#pragma unroll
    for (int s = 0; s < V::nprim; ++s) {
      unsigned long long x = (unsigned long long)__double_as_longlong(P[s * NC + c]);
      a += x & 0xffffffffull;
      b += x >> 32;
    }
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) { s1[threadIdx.x] += s1[threadIdx.x + o]; s2[threadIdx.x] += s2[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&h[0], s1[0]);
    atomicAdd(&h[1], s2[0]);
  }
}

// The Spalart-Allmaras source of one interior cell into tmp(nt) (Mod_Explicit.f90:215-220: after the boundary
// records).  rho and mil are the cell's thermo-cache values (Mod_Thermo_Cache.f90:291-350).
template <int NS, int NR>
__device__ __forceinline__ void cell_sa(const CaseDev& C, const mose_blk& B, int i, int j, int k, int64_t c0, double rho,
                                        double mil, double* tmp) {
  if constexpr (NR > 0) {
    using V = Vars<NS, NR>;
    const int64_t st[3] = {1, (int64_t)B.pi, (int64_t)B.pi * B.pj};
    const bool on[6] = {i == 1, i == B.n1, j == 1, j == B.n2, k == 1, k == B.n3};
    int64_t nb[6];
    bool wall[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      nb[q] = c0 + ((q & 1) ? st[q >> 1] : -st[q >> 1]);
      wall[q] = false;
      if (on[q]) {
        signed char t = C.btyp[B.bpoff[q] + plane_pos(B, q + 1, i, j, k)];
        wall[q] = (t == MOSE_BC_WALL_ADIABATIC || t == MOSE_BC_WALL_ISOTHERMAL);
      }
    }
    double src = sa_source<NS, NR>(C.Pin, C.NC, C.m, c0, nb, wall, rho, mil, C.yn[c0], C.vol[c0], C.pimp != 0, *C.dt);
    tmp[V::nt] = tmp[V::nt] - src;
  } else {
    (void)C; (void)B; (void)i; (void)j; (void)k; (void)c0; (void)rho; (void)mil; (void)tmp;
  }
}

// the laminar viscosity of the thermo cache from the (guarded) cell temperature
template <int NS>
__device__ __forceinline__ double cell_mil(const mose_dtab& T, const double* p, double rho, double Tk) {
  int Ti = ftrunc(Tk);
  double mil, kl;
  wilke<NS>(T, p, rho, Ti, Tk - Ti, mil, kl);
  return mil;
}

template <int NS, int NR>
__device__ __forceinline__ void update_one(const CaseDev& C, const mose_blk& B, int b, int64_t t, double rkc,
                                           double* dres, const double* Rin);

// one thread per interior cell of block b: residual, SA source, RK stage update, state check; on the last stage
// also the per-workgroup sums of the squared step changes (rpart != nullptr).  Rin != nullptr: the residual is read
// from Rin (smoothed by IRS, SA source included; scaled too for the conservative update) instead of the faces.
template <int NS, int NR>
__global__ void __launch_bounds__(256) k_update(CaseDev C, int b, double rkc, double* rpart, int64_t part0,
                                                const double* Rin) {
  using V = Vars<NS, NR>;
  const mose_blk B = C.blk[b];
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dres[V::nprim - NS + 1];
#pragma unroll
  for (int q = 0; q < V::nprim - NS + 1; ++q) dres[q] = 0.0;
  if (t < (int64_t)B.n1 * B.n2 * B.n3) update_one<NS, NR>(C, B, b, t, rkc, rpart ? dres : nullptr, Rin);
  if (rpart) block_sum_store<V::nprim - NS + 1>(dres, rpart + (part0 + blockIdx.x) * (V::nprim - NS + 1));
}

template <int NS, int NR>
__device__ __forceinline__ void update_one(const CaseDev& C, const mose_blk& B, int b, int64_t t, double rkc,
                                           double* dres, const double* Rin) {
  using V = Vars<NS, NR>;
  int i = (int)(t % B.n1) + 1;
  int64_t rem = t / B.n1;
  int j = (int)(rem % B.n2) + 1, k = (int)(rem / B.n2) + 1;
  const mose_dtab& T = *C.tab;
  const int64_t NC = C.NC;
  int64_t c0 = row(B, i, j, k);
  double res[V::nprim], p[V::nprim], po[V::nprim], out[V::nprim];
  if (Rin) {
#pragma unroll
    for (int s = 0; s < V::nprim; ++s) res[s] = Rin[s * NC + c0];
  } else {
    cell_residual<NS, NR>(C, B, i, j, k, res);
  }
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) { p[s] = C.Pin[s * NC + c0]; po[s] = C.PO[s * NC + c0]; }
  double rho = rho_sum<NS>(p), Rg = f_Rtot<NS>(T, p);
  double Tseed = p[V::np] / (rho * Rg);
  if (C.visc && C.guard && !(Tseed > 1.0 && Tseed < 9990.0)) {
    flag_bad(C, b, i, j, k, 1);
    if (Tseed != Tseed) Tseed = 300.0;
    Tseed = dmin2(dmax2(Tseed, 1.0), 9990.0);
  }
  if (NR > 0 && !Rin) cell_sa<NS, NR>(C, B, i, j, k, c0, rho, cell_mil<NS>(T, p, rho, Tseed), res);
  if (C.prim) update_cell_prim<NS, NR>(T, p, po, res, C.vol[c0], *C.dt, C.strang, rkc, out);
  else if (Rin) update_cell<NS, NR, true>(T, p, po, res, C.vol[c0], *C.dt, C.strang, rkc, Tseed, out);
  else update_cell<NS, NR>(T, p, po, res, C.vol[c0], *C.dt, C.strang, rkc, Tseed, out);
  bool ok = out[V::np] >= 0.0;
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) ok = ok && isfinite(out[s]);
  if (!ok) flag_bad(C, b, i, j, k, 0);
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) C.Pout[s * NC + c0] = out[s];
  if (dres) step_change<NS, NR>(out, po, dres);
}

// the residual alone (for checks): R = pass 2 + boundary increments
template <int NS, int NR>
__global__ void __launch_bounds__(256) k_residual_out(CaseDev C, int b, double* R) {
  using V = Vars<NS, NR>;
  const mose_blk B = C.blk[b];
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B.n1 * B.n2 * B.n3) return;
  int i = (int)(t % B.n1) + 1;
  int64_t rem = t / B.n1;
  int j = (int)(rem % B.n2) + 1, k = (int)(rem / B.n2) + 1;
  double res[V::nprim];
  cell_residual<NS, NR>(C, B, i, j, k, res);
  int64_t c0 = row(B, i, j, k);
  if (NR > 0) {
    const mose_dtab& T = *C.tab;
    double p[V::nprim];
#pragma unroll
    for (int s = 0; s < V::nprim; ++s) p[s] = C.Pin[s * C.NC + c0];
    double rho = rho_sum<NS>(p), Tk = p[V::np] / (rho * f_Rtot<NS>(T, p));
    if (C.guard && !(Tk > 1.0 && Tk < 9990.0)) {
      if (Tk != Tk) Tk = 300.0;
      Tk = dmin2(dmax2(Tk, 1.0), 9990.0);
    }
    cell_sa<NS, NR>(C, B, i, j, k, c0, rho, cell_mil<NS>(T, p, rho, Tk), res);
  }
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) R[s * C.NC + c0] = res[s];
}

// ------------------------------------------------------------------------------------------------------------
// implicit residual smoothing: IRS_apply (Lib_IRS.f90:53-96) on the residual array R (var-last, the arena's rows)
// ------------------------------------------------------------------------------------------------------------
// irs_scale_blk (:99-121): R = -R / vol * dt * strang on the interior cells, every variable (conservative update only)
template <int NS, int NR>
__global__ void __launch_bounds__(256) k_irs_scale(CaseDev C, int b, double* R) {
  using V = Vars<NS, NR>;
  const mose_blk B = C.blk[b];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B.n1 * B.n2 * B.n3) return;
  const int i = (int)(t % B.n1) + 1;
  const int64_t rem = t / B.n1;
  const int j = (int)(rem % B.n2) + 1, k = (int)(rem / B.n2) + 1;
  const int64_t c0 = row(B, i, j, k);
  const double vol = C.vol[c0], dt = *C.dt;
#pragma unroll
  for (int s = 0; s < V::nprim; ++s) R[s * C.NC + c0] = -R[s * C.NC + c0] / vol * dt * C.strang;
}

// irs_ghost_residual (:125-152): the first ghost layer behind every face of block b takes its boundary cell's value
// (the flow variables 0..nf-1); one thread per boundary cell-face
static __global__ void __launch_bounds__(256) k_irs_ghost_faces(const mose_blk* blk, int b, double* R, int64_t NC, int nf) {
  const mose_blk B = blk[b];
  const int64_t a1 = (int64_t)B.n2 * B.n3, a2 = (int64_t)B.n1 * B.n3, a3 = (int64_t)B.n1 * B.n2;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k, f;
  if (t < 2 * a1) { f = t < a1 ? 1 : 2; t -= (f - 1) * a1; j = (int)(t % B.n2) + 1; k = (int)(t / B.n2) + 1; i = f == 1 ? 1 : B.n1; }
  else if ((t -= 2 * a1) < 2 * a2) { f = t < a2 ? 3 : 4; t -= (f - 3) * a2; i = (int)(t % B.n1) + 1; k = (int)(t / B.n1) + 1; j = f == 3 ? 1 : B.n2; }
  else if ((t -= 2 * a2) < 2 * a3) { f = t < a3 ? 5 : 6; t -= (f - 5) * a3; i = (int)(t % B.n1) + 1; j = (int)(t / B.n1) + 1; k = f == 5 ? 1 : B.n3; }
  else return;
  const int64_t cm = row(B, i, j, k);
  const int64_t cg = row(B, i - c_guide[f - 1][0], j - c_guide[f - 1][1], k - c_guide[f - 1][2]);
  for (int s = 0; s < nf; ++s) R[s * NC + cg] = R[s * NC + cm];
}

// irs_ghost_conn (:175-192): the ghost behind each connection record's receiver cell takes the donor cell's value
static __global__ void __launch_bounds__(256) k_irs_ghost_conn(const mose_blk* blk, const mose_bc_record* rec, int nrec,
                                                        double* R, int64_t NC, int nf) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nrec) return;
  const mose_bc_record r = rec[e];
  if (r.typ != MOSE_BC_CONNECTION) return;
  const mose_blk Bm = blk[r.b], Bs = blk[r.bs];
  const int64_t cg = row(Bm, r.i - c_guide[r.f - 1][0], r.j - c_guide[r.f - 1][1], r.k - c_guide[r.f - 1][2]);
  const int64_t cd = row(Bs, r.is, r.js, r.ks);
  for (int s = 0; s < nf; ++s) R[s * NC + cg] = R[s * NC + cd];
}

// irs_smooth_blk (:194-235), its three loops: R* = R over the block's whole arena, one Jacobi sweep on the interior
// (rs2 = (R + beta (rs1(c - off) + rs1(c + off))) / (1 + 2 beta)), and an interior copy
static __global__ void __launch_bounds__(256) k_irs_copy_block(mose_blk B, const double* src, double* dst, int64_t NC, int nf) {
  const int64_t n = (int64_t)B.pi * B.pj * (B.n3 + 2 * GC);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  for (int s = 0; s < nf; ++s) dst[s * NC + B.boff + t] = src[s * NC + B.boff + t];
}

static __global__ void __launch_bounds__(256) k_irs_sweep(mose_blk B, const double* R, const double* rs1, double* rs2,
                                                   int64_t NC, int nf, int64_t off, double beta) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B.n1 * B.n2 * B.n3) return;
  const int i = (int)(t % B.n1) + 1;
  const int64_t rem = t / B.n1;
  const int j = (int)(rem % B.n2) + 1, k = (int)(rem / B.n2) + 1;
  const int64_t c0 = row(B, i, j, k);
  for (int s = 0; s < nf; ++s)
    rs2[s * NC + c0] = (R[s * NC + c0] + beta * (rs1[s * NC + c0 - off] + rs1[s * NC + c0 + off])) / (1.0 + 2.0 * beta);
}

static __global__ void __launch_bounds__(256) k_irs_copy_interior(mose_blk B, const double* src, double* dst, int64_t NC,
                                                           int nf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B.n1 * B.n2 * B.n3) return;
  const int i = (int)(t % B.n1) + 1;
  const int64_t rem = t / B.n1;
  const int j = (int)(rem % B.n2) + 1, k = (int)(rem / B.n2) + 1;
  const int64_t c0 = row(B, i, j, k);
  for (int s = 0; s < nf; ++s) dst[s * NC + c0] = src[s * NC + c0];
}

// ------------------------------------------------------------------------------------------------------------
// dt: Compute_dt_blk (dt_blk.inc:157-218), min over the cells by wavefront shuffles + one atomic per block
// ------------------------------------------------------------------------------------------------------------
template <int NS, int NR>
__device__ __forceinline__ double cell_dt(const CaseDev& C, int64_t c0, const double* pl) {
  using V = Vars<NS, NR>;
  const mose_dtab& T = *C.tab;
  const int64_t NC = C.NC;
  double rho = rho_sum<NS>(pl), Rg = f_Rtot<NS>(T, pl);
  double snd = f_ss<NS>(T, pl, pl[V::np], rho, Rg);
  double dtc = 1e8;
  for (int d = 0; d < C.ndir; ++d) {
    double mx = C.m[(int64_t)d * NC + c0], my = C.m[(int64_t)(d + 3) * NC + c0], mz = C.m[(int64_t)(d + 6) * NC + c0];
    double nn = sqrt(mx * mx + my * my + mz * mz);
    double vx = mx / nn, vy = my / nn, vz = mz / nn;
    double lam = fabs(pl[V::nu] * vx + pl[V::nv] * vy + pl[V::nw] * vz) + snd;
    double dtd = C.dl[(int64_t)d * NC + c0] / lam * C.cfl;
    dtc = dmin2(dtd, dtc);
  }
  if (C.visc) {
    double Tk = pl[V::np] / (rho * Rg);
    int Ti = ftrunc(Tk);
    double mil, kl;
    wilke<NS>(T, pl, rho, Ti, Tk - Ti, mil, kl);
    double mie = 0.0;
    if (NR > 0) {
      double chi = pl[V::nt] / mil, c3 = chi * chi * chi;
      mie = c3 / (c3 + SA_CV1 * SA_CV1 * SA_CV1) * pl[V::nt];
    }
    double mi = mie + mil;
    for (int d = 0; d < C.ndir; ++d) {
      double dl = C.dl[(int64_t)d * NC + c0];
      dtc = dmin2((rho * (dl * dl) * C.vnn) / mi, dtc);
    }
  }
  return dtc;
}

__device__ __forceinline__ void block_min_atomic(double v, unsigned long long* slot) {
  // wavefront (64 lanes) min, then across the waves of the block through LDS, one atomic per block
  for (int o = 32; o > 0; o >>= 1) v = dmin2(v, __shfl_xor(v, o, 64));
  __shared__ double wmin[16];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wmin[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmin[0];
    int nw = (blockDim.x + 63) >> 6;
    for (int q = 1; q < nw; ++q) m = dmin2(m, wmin[q]);
    atomicMin(slot, (unsigned long long)__double_as_longlong(m));
  }
}

template <int NS, int NR>
__global__ void __launch_bounds__(256) k_dt(CaseDev C, int b) {
  using V = Vars<NS, NR>;
  const mose_blk B = C.blk[b];
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 1e300;
  if (t < (int64_t)B.n1 * B.n2 * B.n3) {
    int i = (int)(t % B.n1) + 1;
    int64_t rem = t / B.n1;
    int j = (int)(rem % B.n2) + 1, k = (int)(rem / B.n2) + 1;
    int64_t c0 = row(B, i, j, k);
    double pl[V::nprim];
#pragma unroll
    for (int s = 0; s < V::nprim; ++s) pl[s] = C.Pin[s * C.NC + c0];
    v = cell_dt<NS, NR>(C, c0, pl);
  }
  block_min_atomic(v, C.dtmin);
}

// dt <- min(1e5, reduction) (Explicit_Step sets dtglobal = 1d5 before Compute_dt, Mod_Explicit.f90:124)
static __global__ void k_dt_init(unsigned long long* slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *slot = (unsigned long long)__double_as_longlong(1e5);
}
static __global__ void k_dt_final(const unsigned long long* slot, double* dt, double* hist, int n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double v = __longlong_as_double((long long)*slot);
    *dt = v;
    if (hist) hist[n] = v;
  }
}

}  // namespace mose
