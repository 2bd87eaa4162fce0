// Per-cell stiff chemistry on the device: FLINT's general law-of-mass-action source, the constant-volume ODE of
// rhs_native, and the RODAS integrator with the numerical Jacobian of the host path (wrap_rodas -> hairerRodas.f),
// one team of NS + 1 lanes per cell.
//
//   net_rate    general()                 lib/FLINT/src/lib/Lib_Chemistry_wdot.f90:163-288 (+ falloff :12-80)
//   team_rhs    rhs_native                lib/FLINT/src/lib/Lib_Chemistry_rhs.f90:29-81
//   team_dec_lds / team_sol_lds  DEC / SOL  lib/FLINT/src/lib/Lib_Radau5_dev.f90:950-1035
//   k_chem      ROSCOR (METH = 1, IJAC = 0, autonomous) hairerRodas.f:540-880 inside Chemistry_Newstate_blk
//               (src/lib/physics/Lib_Chemistry.f90:33-76)
//
// A team of NS + 1 lanes integrates one cell (row r of every matrix and component r of every vector on lane r).
// Work distribution: a persistent grid; each team takes its next cell from an atomic queue when its cell is done
// (uneven stiffness -> uneven step counts), and every iteration of the team loop does one step attempt.
#pragma once

#include <hip/hip_runtime.h>

#include "mose_internal.h"
#include "mose_libm.h"

namespace mose {

struct ChemDev {
  int ns, na, nt, nl, tmin, rows;
  double wm[MOSE_MAX_SPECIES], Ri[MOSE_MAX_SPECIES];
  const double *cp, *h, *ni1, *ni2, *eps, *kf, *kb, *troe, *lind;
};

struct ChemArgs {
  double* state;            // rows of `stride` doubles: species densities in rows 0..NS-1, p in row prow
  const long long* rows;    // the column of cell c in state (a flow case's interior rows), or null: column c
  int prow;                 // the row of p: NS for a batch, NS + 3 on a flow case's arena
  const double* dtp;        // the case's dt on the device (a Strang step reacts over the step's dt), or null: dt
  double* hr;               // [ncell]
  int* idid;                // [ncell]
  int* nstep;               // [ncell]
  unsigned long long* queue;   // [0] next cell
  unsigned long long* stats;   // [0] steps, [1] rejected, [2] rhs calls, [3] failed, [4] max steps,
                               // [5] ~0 - (first failed cell), 0 if none
  long long ncell;
  long long stride;          // row stride of state (>= ncell)
  double dt;
  double rtol[MOSE_MAX_SPECIES + 1], atol[MOSE_MAX_SPECIES + 1];   // RADAU5: the transformed tolerances
  double rk[12];             // RADAU5: c1, c2, c1 - 1, c2 - 1, c1 - c2, dd1, dd2, dd3, 1/u1, alph, beta, FNEWT
};

__device__ __forceinline__ double ctab(const ChemDev& M, const double* t, int T1, double Td) {
  int r1 = T1 - M.tmin, r2 = r1 + 1;
  r1 = r1 < 0 ? 0 : (r1 > M.rows - 1 ? M.rows - 1 : r1);
  r2 = r2 < 0 ? 0 : (r2 > M.rows - 1 ? M.rows - 1 : r2);
  double a = t[r1], b = t[r2];
  return a + (b - a) * Td;
}

// Fortran x**n for a small non-negative integer n (repeated squaring, as the run-time integer power)
__device__ __forceinline__ double chem_ipow(double x, int n) {
  double r = 1.0;
  bool first = true;
  while (n > 0) {
    if (n & 1) { r = first ? x : r * x; first = false; }
    n >>= 1;
    if (n) x = x * x;
  }
  return first ? 1.0 : r;
}


/* Deterministic natural log for the Troe broadening factor's log10, the statements of the host side: x = m 2^e with
 * m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1), odd series to s^23.  Within a few ulp of the true
 * value for the arguments met here (Fcent in (0, 1], Pr > 0). */
__device__ __forceinline__ double dm_log(double x) {
  int e;
  double m = frexp(x, &e);            /* x = m 2^e, m in [0.5, 1) */
  if (m < 0.70710678118654752440) { m = m * 2.0; e = e - 1; }
  double s = (m - 1.0) / (m + 1.0), s2 = s * s, t = 0.0;
  for (int k = 23; k >= 3; k -= 2) t = (t + 1.0 / (double)k) * s2;
  double lm = 2.0 * s * (1.0 + t);
  return (double)e * 0.69314718055994528623 + ((double)e * 2.3190468138462996154e-17 + lm);
}

constexpr double DM_LN10 = 2.30258509299404568402;

// f_F (Lib_Chemistry_falloff.f90:63-80); log10 through dm_log, 10** through mose_pow (glibc's pow, as the reference)
__device__ __forceinline__ double troe_F(double Pr, double Fcent) {
  double lF = dm_log(Fcent) / DM_LN10, lP = dm_log(Pr) / DM_LN10;
  double cc = -0.4 - 0.67 * lF, nn = 0.75 - 1.27 * lF;
  double f1 = (lP + cc) / (nn - 0.14 * (lP + cc));
  return mose_pow(10.0, lF / (1.0 + f1 * f1));
}

// c**nu of the falloff classes: integer power for integral nu, pow otherwise
__device__ __forceinline__ double rpow(double c, double nu) {
  double r = rint(nu);
  return (r == nu && r >= 0.0 && r <= 8.0) ? chem_ipow(c, (int)r) : mose_pow(c, nu);
}

// ---------------------------------------------------------------------------------------------------------------
// A team of NS + 1 lanes integrates one cell: lane r of the team owns row r of the Jacobian and of the LU factors and
// component r of every vector.  Reactions are spread over the lanes (reaction q on lane q % (NS + 1)); every sum the sequential statements form (species
// production over the reactions, the energy sums, the pivot search, the error norm) is gathered into each lane
// (through the team's LDS area or by shuffles) and added in the same order, so the team reproduces the sequential
// statements bit for bit.
// ---------------------------------------------------------------------------------------------------------------
constexpr int SLOT = 16;    // species slots of the LDS tables (ns <= 16)
constexpr int NRMAX = 32;   // reactions per mechanism in these kernels

// A team is NS + 1 lanes (one per ODE component); a wavefront holds 64 / (NS + 1) teams and its remaining lanes idle.
template <int NS>
struct TeamShape {
  static constexpr int size = NS + 1;
  static constexpr int per_wave = 64 / size;
  static constexpr int per_wg = 4 * per_wave;   // 256-lane workgroups
};

__device__ __forceinline__ double tsh(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ int tshi(int v, int src) { return __shfl(v, src, 64); }

// Compact form of the mechanism's reactions, built once per workgroup in LDS: the orders nu'(q, s) and nu''(q, s)
// as 4-bit fields when every one is an integer in [0, 8] (the powers of both rate classes are then integer powers:
// (int)rint(nu) for the Arrhenius class, the integral branch of rpow for the falloff classes), the third-body flag
// (nu'(q, ns+1) > 0) and the third-body efficiencies.
struct RxnLds {
  unsigned long long o1[NRMAX], o2[NRMAX];
  unsigned int c1[NRMAX], c2[NRMAX];   // the nonzero orders of each side as bytes (species << 4 | order), ascending
  int flags[NRMAX];          // bit 0: third body; bit 1: the orders are packed in o1 / o2; bit 2: and in c1 / c2
  double eps[NRMAX * SLOT];  // eps(q, s) at [q * SLOT + s]
};

template <int NS>
__device__ __forceinline__ void rxn_build(const ChemDev& M, RxnLds& X) {
  constexpr int W = NS + 1;
  const int nr = M.na + M.nt + M.nl;
  for (int q = threadIdx.x; q < NRMAX; q += blockDim.x) {
    int fl = 0;
    unsigned long long o1 = 0ull, o2 = 0ull;
    unsigned int c1 = 0u, c2 = 0u;
    if (q < nr) {
      const double* n1 = M.ni1 + (size_t)q * W;
      const double* n2 = M.ni2 + (size_t)q * W;
      if (n1[NS] > 0.0) fl |= 1;
      bool packed = true;
      for (int s = 0; s < NS; ++s) {
        const double a = n1[s], b = n2[s];
        if (a == rint(a) && a >= 0.0 && a <= 8.0 && b == rint(b) && b >= 0.0 && b <= 8.0) {
          o1 |= (unsigned long long)(int)a << (4 * s);
          o2 |= (unsigned long long)(int)b << (4 * s);
        } else {
          packed = false;
        }
      }
      if (packed) fl |= 2;
      // at most four species with a nonzero order on each side: the byte lists (SLOT = 16 species fit 4 bits)
      int k1 = 0, k2 = 0;
      for (int s = 0; packed && s < NS; ++s) {
        const int a = (int)((o1 >> (4 * s)) & 15ull), b = (int)((o2 >> (4 * s)) & 15ull);
        if (a != 0) { if (k1 < 4) c1 |= (unsigned)(s << 4 | a) << (8 * k1); ++k1; }
        if (b != 0) { if (k2 < 4) c2 |= (unsigned)(s << 4 | b) << (8 * k2); ++k2; }
      }
      if (packed && k1 <= 4 && k2 <= 4) fl |= 4;
    }
    X.o1[q] = o1;
    X.o2[q] = o2;
    X.c1[q] = c1;
    X.c2[q] = c2;
    X.flags[q] = fl;
  }
  for (int i = threadIdx.x; i < NRMAX * SLOT; i += blockDim.x) {
    const int q = i / SLOT, s = i % SLOT;
    X.eps[i] = (q < nr && s < NS) ? M.eps[(size_t)q * W + s] : 0.0;
  }
}

// rate of progress of reaction q (the body of the loops of general(), Lib_Chemistry_wdot.f90:185-285)
template <int NS>
__device__ __forceinline__ double net_rate(const ChemDev& M, const RxnLds& X, int q, const double* coi, int Ti,
                                           double Td) {
  constexpr int W = NS + 1;
  const int cls = q < M.na ? 0 : (q < M.na + M.nt ? 1 : 2);
  const int fl = X.flags[q];
  double coM;
  if (fl & 1) {
    coM = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) coM = coM + coi[s] * X.eps[q * SLOT + s];
  } else {
    coM = 1.0;
  }
  const double* n1 = M.ni1 + (size_t)q * W;
  const double* n2 = M.ni2 + (size_t)q * W;
  double pf = 1.0, pr = 1.0, rf, rr;
  if (fl & 4) {
    // the same products as the loop below, over the nonzero orders only, in the same species order
    const unsigned c1 = X.c1[q], c2 = X.c2[q];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const unsigned v = (c1 >> (8 * e)) & 255u, w = (c2 >> (8 * e)) & 255u;
      if (v != 0u) pf = pf * chem_ipow(coi[v >> 4], (int)(v & 15u));
      if (w != 0u) pr = pr * chem_ipow(coi[w >> 4], (int)(w & 15u));
    }
  } else if (fl & 2) {
    const unsigned long long o1 = X.o1[q], o2 = X.o2[q];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int a = (int)((o1 >> (4 * s)) & 15ull), b = (int)((o2 >> (4 * s)) & 15ull);
      if (a != 0) pf = pf * chem_ipow(coi[s], a);
      if (b != 0) pr = pr * chem_ipow(coi[s], b);
    }
  } else if (cls == 0) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double a = n1[s], b = n2[s];
      if (a != 0) pf = pf * chem_ipow(coi[s], (int)rint(a));
      if (b != 0) pr = pr * chem_ipow(coi[s], (int)rint(b));
    }
  } else {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double a = n1[s], b = n2[s];
      if (a != 0) pf = pf * rpow(coi[s], a);
      if (b != 0) pr = pr * rpow(coi[s], b);
    }
  }
  if (cls == 0) {
    rf = ctab(M, M.kf + (size_t)q * M.rows, Ti, Td) * pf * coM;
    rr = ctab(M, M.kb + (size_t)q * M.rows, Ti, Td) * pr * coM;
  } else {
    const double* tb = cls == 1 ? M.troe + (size_t)(q - M.na) * 4 * M.rows : M.lind + (size_t)(q - M.na - M.nt) * 3 * M.rows;
    double kinf = ctab(M, tb, Ti, Td), k0 = ctab(M, tb + M.rows, Ti, Td), kc = ctab(M, tb + 2 * M.rows, Ti, Td);
    double Pr = k0 * coM / kinf;
    double k1 = kinf * (Pr / (1.0 + Pr));
    if (cls == 1) k1 = k1 * troe_F(Pr, ctab(M, tb + 3 * M.rows, Ti, Td));
    double k2 = k1 / kc;
    rf = k1 * pf;
    rr = k2 * pr;
  }
  return rf - rr;
}

// The exchange area of one team in LDS.  A team's lanes are in one wavefront, so a write by one lane is seen by the
// others after a wavefront-scope fence (no workgroup barrier).
struct TeamLds {
  double z[SLOT];          // the argument Z of the rhs, one component per lane
  double c[SLOT];          // the concentrations max(Z_s, 0) / W_s
  double net[NRMAX];       // rates of progress, reaction q at [q]
  double ei[SLOT];         // per-species terms of the energy sums
  double rc[SLOT];
  double q[SLOT];          // squared scaled errors
};

__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// rhs_native (Lib_Chemistry_rhs.f90:29-81) for the team: z = this lane's component of Z, returns this lane's F.
// cw[q * SLOT + s] = wm(s) * (nu''(q, s) - nu'(q, s)), the first product of the production statement of general()
// (Lib_Chemistry_wdot.f90:280-284), so that omega_s = sum over q in order of cw * net_q rounds as that statement.
template <int NS>
__device__ __forceinline__ double team_rhs(const ChemDev& M, const RxnLds& X, const double* cw, TeamLds& L, int r,
                                           double z) {
  constexpr int NZ = NS + 1;
  const int rs = r < NS ? r : 0;
  L.z[r] = z;
  L.c[r] = (z > 0.0 ? z : 0.0) / M.wm[rs];   // coi(s) on lane s
  wsync();
  const double T = L.z[NS];
  if (T < (double)M.tmin || T >= (double)(M.tmin + M.rows - 1) || T != T) return -1.0;
  const int Ti = static_cast<int>(T);
  const double Tf = T - (double)Ti;
  // the concentrations are read from the team's LDS copy where they are used: a register copy of all NS of them
  // stayed live through the reaction loop and pushed the kernel into 27 more spilled VGPRs
  const double* coi = L.c;
  const int nr = M.na + M.nt + M.nl;
#pragma unroll 1
  for (int q0 = 0; q0 < nr; q0 += NZ) {
    const int q = q0 + r;
    if (q < nr) L.net[q] = net_rate<NS>(M, X, q, coi, Ti, Tf);
  }
  wsync();
  // omega_s on lane s: production summed over the reactions in order
  double om = 0.0;
#pragma unroll
  for (int q = 0; q < NRMAX; ++q)
    if (q < nr) om = om + cw[q * SLOT + rs] * L.net[q];
  // the energy sums over the species, in order, on the lane of the T equation
  const int row = Ti - M.tmin;
  const double* ht = M.h + (size_t)rs * M.rows;
  const double* ct = M.cp + (size_t)rs * M.rows;
  const double h0 = ht[row], h1 = ht[row + 1], c0 = ct[row], c1 = ct[row + 1];
  const double hv = h0 + Tf * (h1 - h0);
  const double cv = c0 + Tf * (c1 - c0);
  L.ei[r] = (hv - M.Ri[rs] * T) * om;
  L.rc[r] = (z > 0.0 ? z : 0.0) * (cv - M.Ri[rs]);   // roi(s) (cv - R_s) on lane s
  wsync();
  if (r == NS) {
    double ei = 0.0, rcv = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      ei = ei + L.ei[s];
      rcv = rcv + L.rc[s];
    }
    return -ei / rcv;
  }
  return r < NS ? om : 0.0;
}

// DEC (Moler, ACM 423; Lib_Radau5_dev.f90:950-997) with row r of the matrix on lane r: e[j] = a(r, j).  Returns
// ier (0: regular); the pivot row of column k is ip(k) = (ipk >> 4k) & 15, the same on every lane of the team
// (ip(N), the sign of the determinant, is not kept: SOL does not read it).  The pivot column and the two exchanged
// rows pass through the team's LDS area (one 64-bit LDS access per value; a shuffle of a double costs two
// bpermutes): L.q holds a column, L.ei the pivot row m and L.rc row k (none of them is live across a DEC).
template <int N>
__device__ __forceinline__ int team_dec_lds(double (&e)[N], unsigned long long& ipk, TeamLds& L, int r) {
  static_assert(N <= 16, "4-bit pivot indices");
  ipk = 0ull;
#pragma unroll
  for (int k = 0; k < N - 1; ++k) {
    L.q[r] = e[k];
    wsync();
    int m = k;
    double am = fabs(L.q[k]);
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      double v = fabs(L.q[i]);
      if (v > am) { m = i; am = v; }
    }
    ipk |= (unsigned long long)m << (4 * k);
    double t = L.q[m];
    const double akk = L.q[k];
    if (r == m) {
#pragma unroll
      for (int j = k + 1; j < N; ++j) L.ei[j] = e[j];
    }
    if (r == k) {
#pragma unroll
      for (int j = k + 1; j < N; ++j) L.rc[j] = e[j];
    }
    wsync();
    if (m != k) e[k] = (r == k) ? t : ((r == m) ? akk : e[k]);
    if (t == 0.0) return k + 1;
    t = 1.0 / t;
    if (r > k) e[k] = -e[k] * t;
#pragma unroll
    for (int j = k + 1; j < N; ++j) {
      const double amj = L.ei[j], akj = L.rc[j];
      e[j] = (r == k) ? amj : ((r == m) ? akj : e[j]);
      if (amj != 0.0 && r > k) e[j] = e[j] + e[k] * amj;
    }
  }
  L.q[r] = e[N - 1];
  wsync();
  if (L.q[N - 1] == 0.0) return N;
  return 0;
}

// SOL (Lib_Radau5_dev.f90:1003-1035) on the team: b = this lane's component; the exchanges go through L.q
template <int N>
__device__ __forceinline__ double team_sol_lds(const double (&e)[N], unsigned long long ipk, double b, TeamLds& L,
                                               int r) {
#pragma unroll
  for (int k = 0; k < N - 1; ++k) {
    const int m = (int)((ipk >> (4 * k)) & 15ull);
    L.q[r] = b;
    wsync();
    const double bm = L.q[m], bk = L.q[k];
    const double t = bm;
    b = (r == m) ? bk : b;
    b = (r == k) ? t : b;
    if (r > k) b = b + e[k] * t;
  }
#pragma unroll
  for (int kb = 1; kb <= N - 1; ++kb) {
    const int k = N - kb;
    if (r == k) b = b / e[k];
    L.q[r] = b;
    wsync();
    const double t = -L.q[k];
    if (r < k) b = b + e[k] * t;
  }
  if (r == 0) b = b / e[0];
  return b;
}

// RODAS4 (METH = 1) coefficients (ROCOE; the values of rodas_dev, Lib_Radau5_dev.f90:710-737)
struct RodasC {
  static constexpr double GAMMA = 0.25;
  static constexpr double A21 = 0.1544000000000000e+01, A31 = 0.9466785280815826e+00, A32 = 0.2557011698983284e+00;
  static constexpr double A41 = 0.3314825187068521e+01, A42 = 0.2896124015972201e+01, A43 = 0.9986419139977817e+00;
  static constexpr double A51 = 0.1221224509226641e+01, A52 = 0.6019134481288629e+01, A53 = 0.1253708332932087e+02;
  static constexpr double A54 = -0.6878860361058950e+00;
  static constexpr double C21 = -0.5668800000000000e+01, C31 = -0.2430093356833875e+01, C32 = -0.2063599157091915e+00;
  static constexpr double C41 = -0.1073529058151375e+00, C42 = -0.9594562251023355e+01, C43 = -0.2047028614809616e+02;
  static constexpr double C51 = 0.7496443313967647e+01, C52 = -0.1024680431464352e+02, C53 = -0.3399990352819905e+02;
  static constexpr double C54 = 0.1170890893206160e+02;
  static constexpr double C61 = 0.8083246795921522e+01, C62 = -0.7981132988064893e+01, C63 = -0.3152159432874371e+02;
  static constexpr double C64 = 0.1631930543123136e+02, C65 = -0.6058818238834054e+01;
};

// One team of NS + 1 lanes per cell, persistent over the cell queue.  The team loop is a state machine in which every
// iteration makes exactly one rhs evaluation (the only call site, so all teams of a wavefront share it whatever
// phase of the step they are in):
//   EV_F0        f(y) at the start of a step
//   EV_J + i     f(y + delta e_i), column i of the numerical Jacobian
//   EV_S + k-2   f at the input of stage k = 2..6
// and then does the bookkeeping that the result enables: the next Jacobian column, DEC + SOL, a stage SOL, the error
// control with its accept / reject, the start of the next step, or the end of the cell.
enum : int { EV_NONE = -1, EV_F0 = 0, EV_J = 1 };

// The workgroup's LDS copies of the mechanism: the compact reactions and cw (see team_rhs)
template <int NS>
__device__ __forceinline__ void chem_setup(const ChemDev& M, double* cw, RxnLds& X) {
  rxn_build<NS>(M, X);
  const int nr = M.na + M.nt + M.nl;
  for (int i = threadIdx.x; i < NRMAX * SLOT; i += blockDim.x) {
    const int q = i / SLOT, s = i % SLOT;
    double v = 0.0;
    if (q < nr && s < NS) {
      double dn = M.ni2[(size_t)q * (NS + 1) + s] - M.ni1[(size_t)q * (NS + 1) + s];
      v = M.wm[s] * dn;
    }
    cw[i] = v;
  }
  __syncthreads();
}

// Chemistry_Newstate_blk, entry (src/lib/physics/Lib_Chemistry.f90:55-60): Z = (rho_i, T = p / (rho R)); returns
// this lane's component, rho and T of the cell in rho0 / Tb
template <int NS>
__device__ __forceinline__ double cell_load(const ChemDev& M, const ChemArgs& A, int cell, int base, int r, int rc,
                                            double& rho0, double& Tb) {
  const long long col = A.rows ? A.rows[cell] : (long long)cell;
  const double q = A.state[(long long)(rc < NS ? rc : A.prow) * A.stride + col];
  const double p = A.state[(long long)A.prow * A.stride + col];
  double rho = 0.0, Rm = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) rho = rho + tsh(q, base + s);
#pragma unroll
  for (int s = 0; s < NS; ++s) Rm = Rm + M.Ri[s] * tsh(q, base + s) / rho;
  const double T = p / (rho * Rm);
  rho0 = rho;
  Tb = T;
  return r < NS ? q : T;
}

// Chemistry_Newstate_blk, exit (:62-70): floors, p = T rho R, hr; and the batch statistics
template <int NS>
__device__ __forceinline__ void cell_store(const ChemDev& M, const ChemArgs& A, int cell, int base, int r, double y,
                                           double rho0, double Tb, int idid, int nst, int nrej, int nfcn) {
  const double yf = (r < NS && y <= 1e-20) ? 1e-20 : y;
  double rs = 0.0, R2 = 0.0;
  const double T = tsh(y, base + NS);
#pragma unroll
  for (int s = 0; s < NS; ++s) rs = rs + tsh(yf, base + s);
#pragma unroll
  for (int s = 0; s < NS; ++s) R2 = R2 + M.Ri[s] * tsh(yf, base + s) / rs;
  const long long col = A.rows ? A.rows[cell] : (long long)cell;
  if (r < NS) A.state[(long long)r * A.stride + col] = yf;
  if (r == NS) {
    A.state[(long long)A.prow * A.stride + col] = T * rho0 * R2;
    A.hr[cell] = fabs(T - Tb) / Tb;
    A.idid[cell] = idid;
    A.nstep[cell] = nst;
  }
  if (r == 0) {
    atomicAdd(&A.stats[0], (unsigned long long)nst);
    atomicAdd(&A.stats[1], (unsigned long long)nrej);
    atomicAdd(&A.stats[2], (unsigned long long)nfcn);
    if (idid < 0) {
      atomicAdd(&A.stats[3], 1ull);
      atomicMax(&A.stats[5], ~0ull - (unsigned long long)cell);   // the lowest failed cell gives the largest value
    }
    atomicMax(&A.stats[4], (unsigned long long)nst);
  }
}

// Four waves per SIMD: the spills this forces cost less than the latency that fewer waves leave exposed (measured
// on 262144 cells of the hydrogen-air mechanism: 1.53e4 cells/s at two waves, 1.76e4 at four).
template <int NS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_chem(ChemDev M, ChemArgs A) {
  constexpr int NZ = NS + 1;
  constexpr int EV_S = EV_J + NZ;   // EV_S + (k - 2): the rhs of stage k
  static_assert(NZ <= SLOT, "one lane per ODE component");
  using TSh = TeamShape<NS>;
  using R = RodasC;
  const double uround = 1.1e-19, safe = 0.9, fac1 = 5.0, fac2 = 1.0 / 6.0;
  const int nmax = 100000;
  const double dtend = A.dtp ? *A.dtp : A.dt;   // XEND of every cell

  __shared__ double cw[NRMAX * SLOT];
  __shared__ TeamLds tl[TSh::per_wg];
  __shared__ RxnLds X;
  chem_setup<NS>(M, cw, X);

  const int wl = threadIdx.x & 63;
  const int tw = wl / NZ;                // the team within the wavefront
  if (tw >= TSh::per_wave) return;       // the lanes past the last whole team (no barrier follows)
  const int base = tw * NZ;              // lane (in the wave) of the team's row 0
  const int r = wl - base;               // this lane's row / component
  const int rc = r;
  TeamLds& L = tl[(threadIdx.x >> 6) * TSh::per_wave + tw];

  int cell = -1, ev = EV_NONE;
  bool reject = false, last = false;
  int nsing = 0, nst = 0, nacc = 0, nrej = 0, nfcn = 0, idid = 0;
  double x = 0.0, h = 0.0, hopt = 0.0, hacc = 0.0, erracc = 1e-2, hmaxn = 0.0, posneg = 1.0;
  double rho0 = 0.0, Tb = 0.0;
  double y = 0.0, dy1 = 0.0, yn = 0.0, zin = 0.0, delt = 0.0;
  double ak1 = 0.0, ak2 = 0.0, ak3 = 0.0, ak4 = 0.0, ak5 = 0.0;
  double fjr[NZ];                         // row r of the Jacobian (kept for rejected steps)
  double e[NZ];                           // row r of the LU factors
  unsigned long long ipk = 0ull;          // pivot rows of DEC
#pragma unroll
  for (int j = 0; j < NZ; ++j) { fjr[j] = 0.0; e[j] = 0.0; }

  // label 1 of ROSCOR: the checks at the start of a step; true when the integration ends here
  auto step_start = [&]() -> bool {
    if (nst > nmax) { idid = -2; return true; }
    if (0.1 * fabs(h) <= fabs(x) * uround) { idid = -3; return true; }
    if (last) { h = hopt; idid = 1; return true; }
    hopt = h;
    if ((x + h * 1.0001 - dtend) * posneg >= 0.0) { h = dtend - x; last = true; }
    zin = y;
    ev = EV_F0;
    return false;
  };
  // the perturbed argument of Jacobian column i: y(i) + sqrt(uround max(1e-5, |y(i)|)) (IJAC = 0)
  auto column_arg = [&](int i) {
    double yi = tsh(y, base + i);
    delt = sqrt(uround * fmax(1.0e-5, fabs(yi)));
    zin = (r == i) ? yi + delt : y;
  };

  while (true) {
    if (ev == EV_NONE) {
      int c = 0;
      if (r == 0) c = (int)atomicAdd(A.queue, 1ull);
      c = tshi(c, base);
      if ((long long)c >= A.ncell) break;
      cell = c;
      y = cell_load<NS>(M, A, cell, base, r, rc, rho0, Tb);
      // ROSCOR initial preparations (x = 0, xend = dt, H = 0 -> 1e-6)
      x = 0.0;
      double hmax = dtend - x;
      posneg = copysign(1.0, dtend - x);
      hmaxn = fmin(fabs(hmax), fabs(dtend - x));
      h = 0.0;
      if (fabs(h) <= 10.0 * uround) h = 1.0e-6;
      h = fmin(fabs(h), hmaxn);
      h = copysign(h, posneg);
      reject = false; last = false; nsing = 0; nst = 0; nacc = 0; nrej = 0; nfcn = 0;
      hopt = h; hacc = 0.0; erracc = 1e-2;
      idid = 0;
      (void)step_start();   // a fresh cell always proceeds
    }

    const double f = team_rhs<NS>(M, X, cw, L, r, zin);

    bool fin = false, need_dec = false;
    if (ev == EV_F0) {
      dy1 = f;
      nfcn++;
      column_arg(0);
      ev = EV_J;
    } else if (ev < EV_S) {
      const int i = ev - EV_J;
      const double col = (f - dy1) / delt;
#pragma unroll
      for (int j = 0; j < NZ; ++j) fjr[j] = (j == i) ? col : fjr[j];
      if (i + 1 < NZ) {
        column_arg(i + 1);
        ev++;
      } else {
        nfcn += NZ;
        need_dec = true;
      }
    } else {
      const int k = ev - EV_S + 2;   // the stage whose rhs came back in f
      double b;
      if (k == 2) b = f + (R::C21 / h) * ak1;
      else if (k == 3) b = f + ((R::C31 / h) * ak1 + (R::C32 / h) * ak2);
      else if (k == 4) b = f + ((R::C41 / h) * ak1 + (R::C42 / h) * ak2 + (R::C43 / h) * ak3);
      else if (k == 5) b = f + ((R::C52 / h) * ak2 + (R::C54 / h) * ak4 + (R::C51 / h) * ak1 + (R::C53 / h) * ak3);
      else b = f + ((R::C61 / h) * ak1 + (R::C62 / h) * ak2 + (R::C65 / h) * ak5 + (R::C64 / h) * ak4 +
                    (R::C63 / h) * ak3);
      const double ak = team_sol_lds<NZ>(e, ipk, b, L, r);
      if (k == 2) { ak2 = ak; yn = y + R::A31 * ak1 + R::A32 * ak2; }
      else if (k == 3) { ak3 = ak; yn = y + R::A41 * ak1 + R::A42 * ak2 + R::A43 * ak3; }
      else if (k == 4) { ak4 = ak; yn = y + R::A51 * ak1 + R::A52 * ak2 + R::A53 * ak3 + R::A54 * ak4; }
      else if (k == 5) { ak5 = ak; yn = yn + ak5; }
      if (k < 6) {
        zin = yn;
        ev++;
      } else {
        // error estimation and step size control (ak6 = ak)
        yn = yn + ak;
        nfcn += 5;
        nst++;
        const double sk = A.atol[rc] + A.rtol[rc] * fmax(fabs(y), fabs(yn));
        const double ql = ak / sk;
        L.q[r] = ql * ql;
        wsync();
        double err = 0.0;
#pragma unroll
        for (int i = 0; i < NZ; ++i) err = err + L.q[i];
        err = sqrt(err / NZ);
        double fac = fmax(fac2, fmin(fac1, mose_pow(err, 0.25) / safe));   // err**0.25 (libm pow)
        double hnew = h / fac;
        if (err <= 1.0) {
          nacc++;
          if (nacc > 1) {
            double facgus = (hacc / h) * mose_pow(err * err / erracc, 0.25) / safe;
            facgus = fmax(fac2, fmin(fac1, facgus));
            fac = fmax(fac, facgus);
            hnew = h / fac;
          }
          hacc = h;
          erracc = fmax(1.0e-2, err);
          y = yn;
          x = x + h;
          if (fabs(hnew) > hmaxn) hnew = posneg * hmaxn;
          if (reject) hnew = posneg * fmin(fabs(hnew), fabs(h));
          reject = false;
          h = hnew;
          fin = step_start();
        } else {
          reject = true;
          last = false;
          h = hnew;
          if (nacc >= 1) nrej++;
          need_dec = true;
        }
      }
    }

    if (need_dec) {
      // label 2: decompose E = 1/(h gamma) - J; a singular E halves h (at most 4 times)
      while (true) {
        const double fac = 1.0 / (h * R::GAMMA);
#pragma unroll
        for (int j = 0; j < NZ; ++j) e[j] = (j == r) ? -fjr[j] + fac : -fjr[j];
        if (team_dec_lds<NZ>(e, ipk, L, r) == 0) break;
        nsing++;
        if (nsing >= 5) { idid = -4; fin = true; break; }
        h = h * 0.5;
        reject = true;
        last = false;
      }
      if (!fin) {
        ak1 = team_sol_lds<NZ>(e, ipk, dy1, L, r);
        yn = y + R::A21 * ak1;
        zin = yn;
        ev = EV_S;
      }
    }

    if (fin) {
      cell_store<NS>(M, A, cell, base, r, y, rho0, Tb, idid, nst, nrej, nfcn);
      ev = EV_NONE;
    }
  }
}

}  // namespace mose
