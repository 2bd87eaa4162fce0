/*
 * CPU oracle for the MOSE-GPU time-step residual hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, statement by statement and in the
 * reference's floating-point order, the double-precision routines of open-hydra/MOSE-GPU that
 * the MI355X path replaces.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or the CPU baseline; the product never calls it.
 *
 * Every function cites the reference file:line it follows.  Parity of each function is pinned
 * by the checks in tests/test_oracle_*.py (closed forms, invariants, Toro's exact Sod solution,
 * brute force); a function without such a check says "parity unpinned" in its comment.
 *
 * Layouts (0-based in C; comments use the reference's 1-based i,j,k with ghosts 1-GC..n+GC):
 *   arena   NC rows = the padded cells of all blocks, block b from row boff[b];
 *           row(b,i,j,k) = boff + (i+GC-1) + (j+GC-1)*pi + (k+GC-1)*pi*pj, pi = n1+2GC, pj = n2+2GC
 *           (src/lib/gpu/MOSE_Field_m.f90:42-57,103-109: P_flat(cell,var), var-last, i fastest)
 *   state   P[s*NC + row]
 *   geometry (same arena rows)  dl[d*NC+row] d=0..2, m[(a+3b)*NC+row] = m_f(a+1,b+1,.), vol[row], yn[row]
 *   faces   fm1(4,0:n1,n2,n3) per block, concatenated over blocks (offset f1off[b]), stored as
 *           q-planes fm1[q*NF1 + f1off[b] + i + (n1+1)*((j-1) + n2*(k-1))]; fm2, fm3 alike
 */
#ifndef MOSE_ORACLE_H
#define MOSE_ORACLE_H

#include <stdint.h>

#include "../include/mose_gpu.h"   /* mose_bc_record, mose_scheme, mose_tables: the same host structs as the product */

#define OGC 2            /* ghost layers, src/lib/base/Global_m.f90 gc = 2 */
#define ONTMAX 10000     /* src/lib/numerics/fluxes/Mod_Convective.f90:36 NTMAX */
#define ONSMAX 16

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int ns;            /* species */
  int tmax;          /* rows 0..tmax (FLINT sets Tmin = 0 when the file starts at 1 K, Load_ThermoTransport.f90:80-96) */
  double wm[ONSMAX], Ri[ONSMAX];
  const double *cp, *h, *mi, *k;   /* [s*(tmax+1) + T] */
  double mm025[ONSMAX][ONSMAX];    /* (Wm(i)/Wm(j))**(-0.25)   Load_ThermoTransport.f90:57-66 */
  double is8[ONSMAX][ONSMAX];      /* 1/sqrt(8(1+Wm(i)/Wm(j))) */
} o_tables;

typedef struct {
  int nsc, nrans, nprim;
  int nu, nv, nw, np, nt;       /* 0-based variable indices */
  int lim, rs;                  /* limiter id 1..6 (Lib_Limiters.f90:65-97), Riemann id (1 HLLC, 4 AUSM+, 12 LMRoe) */
  double mco;                   /* riemann-options-Mco */
  int visc;                     /* navier-stokes */
  int rans;                     /* 0 none, 2 Spalart-Allmaras */
  double cfl, vnn, prl, sc, sct, prt;
  int n_rk;                     /* 1..3 */
  int ndir;                     /* computational directions (Lib_Metrics.f90:139-196) */
  int point_implicit;           /* SA Patankar factor (sa_body.inc:113-116) */
  int guard;                    /* thermo guard on (Mod_Thermo_Cache.f90:323-335) */
  int prim;                     /* integration-variables = prim: Newstate_Primitive_blk (Lib_Newstate.f90:340-531) */
  int irs, irs_njacobi;         /* implicit residual smoothing (Lib_IRS.f90) and its Jacobi sweeps */
  double irs_beta;
} o_scheme;

typedef struct {
  int n1, n2, n3;
  int64_t boff;                 /* first arena row of the block */
  int pi, pj;                   /* n1+2GC, n2+2GC */
  int64_t f1off, f2off, f3off;  /* offsets of the block in the face arenas */
} o_block;



struct o_case_s;
/* the chemistry of a Strang-split step: react the interior cells of P in place over dt (Chemistry_Newstate_blk) */
typedef void (*o_react_fn)(const struct o_case_s *C, double *P, double dt, void *ctx);

typedef struct o_case_s {
  o_scheme sc;
  o_tables tab;
  int nb;
  const o_block *blk;
  int64_t NC, NF1, NF2, NF3;
  const double *dl, *m, *vol, *yn, *fm1, *fm2, *fm3;
  int nrec;
  const mose_bc_record *rec;    /* grouped by block, then face 1..6 */
  int nff;
  const mose_farfield *ff;      /* far-field table of the type-4 records (rec->ff) */
  int nthreads;
  /* Strang splitting (Lib_Strang.f90:10-22): n_strang = 1, strang = 1 without chemistry; n_strang = 2, strang = 1/2
   * and react != NULL with it */
  int n_strang;
  double strang;
  o_react_fn react;
  void *react_ctx;
} o_case;

#ifdef __cplusplus
}
#endif
#endif
