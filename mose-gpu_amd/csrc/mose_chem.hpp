// Per-cell stiff chemistry on the device: FLINT's general law-of-mass-action source, the constant-volume ODE of
// rhs_native, and the RODAS integrator with the numerical Jacobian of the host path (wrap_rodas -> hairerRodas.f),
// one lane per cell.
//
//   wdot        general()                 lib/FLINT/src/lib/Lib_Chemistry_wdot.f90:163-288 (+ falloff :12-80)
//   chem_rhs    rhs_native                lib/FLINT/src/lib/Lib_Chemistry_rhs.f90:29-81
//   lu_dec/sol  DEC / SOL                 lib/FLINT/src/lib/Lib_Radau5_dev.f90:950-1035
//   k_chem      ROSCOR (METH = 1, IJAC = 0, autonomous) hairerRodas.f:540-880 inside Chemistry_Newstate_blk
//               (src/lib/physics/Lib_Chemistry.f90:33-76)
//
// The LU factors live in registers: the pivot row is a run-time index, so every row access goes through a chain
// of selects over the compile-time rows (no dynamically indexed register array, no scratch).  The Jacobian that a
// rejected step reuses is kept in a per-lane slot of a global buffer laid out [NZ*NZ][lanes] (coalesced).
//
// Work distribution: a persistent grid; each lane takes its next cell from an atomic queue when its cell is done
// (uneven stiffness -> uneven step counts), and every iteration of the lane loop does one step attempt.
#pragma once

#include <hip/hip_runtime.h>

#include "mose_internal.h"

namespace mose {

struct ChemDev {
  int ns, na, nt, nl, tmin, rows;
  double wm[MOSE_MAX_SPECIES], Ri[MOSE_MAX_SPECIES];
  const double *cp, *h, *ni1, *ni2, *eps, *kf, *kb, *troe, *lind;
};

struct ChemArgs {
  double* state;            // [NS+1][ncell]: species densities, p
  double* hr;               // [ncell]
  int* idid;                // [ncell]
  int* nstep;               // [ncell]
  double* fjac;             // [NZ*NZ][nlanes]
  unsigned long long* queue;   // [0] next cell
  unsigned long long* stats;   // [0] steps, [1] rejected, [2] rhs calls, [3] failed, [4] max steps
  long long ncell;
  long long stride;          // row stride of state (>= ncell)
  long long nlanes;
  double dt;
  double rtol[MOSE_MAX_SPECIES + 1], atol[MOSE_MAX_SPECIES + 1];
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


/* Deterministic natural log and exp for the Troe broadening factor: the same statements on the CPU and the GPU give
 * the same bits (the libm log10 / pow of the two platforms differ in the last ulp, and the adaptive integrator turns
 * an ulp into a different step sequence).  log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 * s = (m-1)/(m+1), odd series to s^23; exp: y = k ln2 + r, |r| <= ln2/2, Taylor series to r^16, times 2^k.
 * Both are within a few ulp of the true values for the arguments met here (Fcent in (0, 1], Pr > 0). */
__device__ __forceinline__ double dm_log(double x) {
  int e;
  double m = frexp(x, &e);            /* x = m 2^e, m in [0.5, 1) */
  if (m < 0.70710678118654752440) { m = m * 2.0; e = e - 1; }
  double s = (m - 1.0) / (m + 1.0), s2 = s * s, t = 0.0;
  for (int k = 23; k >= 3; k -= 2) t = (t + 1.0 / (double)k) * s2;
  double lm = 2.0 * s * (1.0 + t);
  return (double)e * 0.69314718055994528623 + ((double)e * 2.3190468138462996154e-17 + lm);
}
__device__ __forceinline__ double dm_exp(double y) {
  double kf = floor(y * 1.44269504088896340736 + 0.5);
  double r = (y - kf * 0.69314718055994528623) - kf * 2.3190468138462996154e-17;
  double t = 1.0;
  for (int k = 16; k >= 1; --k) t = 1.0 + t * r / (double)k;
  return ldexp(t, (int)kf);
}

constexpr double DM_LN10 = 2.30258509299404568402;

// f_F (Lib_Chemistry_falloff.f90:63-80); log10 and 10** through dm_log / dm_exp, the statements of the host side
__device__ __forceinline__ double troe_F(double Pr, double Fcent) {
  double lF = dm_log(Fcent) / DM_LN10, lP = dm_log(Pr) / DM_LN10;
  double cc = -0.4 - 0.67 * lF, nn = 0.75 - 1.27 * lF;
  double f1 = (lP + cc) / (nn - 0.14 * (lP + cc));
  return dm_exp(lF / (1.0 + f1 * f1) * DM_LN10);
}

// c**nu of the falloff classes: integer power for integral nu, pow otherwise
__device__ __forceinline__ double rpow(double c, double nu) {
  double r = rint(nu);
  return (r == nu && r >= 0.0 && r <= 8.0) ? chem_ipow(c, (int)r) : pow(c, nu);
}

template <int NS>
__device__ void wdot(const ChemDev& M, const double* roi, double temp, double* om) {
  constexpr int W = NS + 1;
  double coi[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) coi[s] = roi[s] / M.wm[s];
  int Ti = static_cast<int>(temp);
  double Td = temp - Ti;
#pragma unroll
  for (int s = 0; s < NS; ++s) om[s] = 0.0;
  const int nr = M.na + M.nt + M.nl;
  for (int r = 0; r < nr; ++r) {
    const int cls = r < M.na ? 0 : (r < M.na + M.nt ? 1 : 2);
    const double* n1 = M.ni1 + (size_t)r * W;
    const double* n2 = M.ni2 + (size_t)r * W;
    const double* ep = M.eps + (size_t)r * W;
    double coM;
    if (n1[NS] > 0.0) {
      coM = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) coM = coM + coi[s] * ep[s];
    } else {
      coM = 1.0;
    }
    double pf = 1.0, pr = 1.0, rf, rr;
    if (cls == 0) {
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        double a = n1[s], b = n2[s];
        if (a != 0) pf = pf * chem_ipow(coi[s], (int)rint(a));
        if (b != 0) pr = pr * chem_ipow(coi[s], (int)rint(b));
      }
      rf = ctab(M, M.kf + (size_t)r * M.rows, Ti, Td) * pf * coM;
      rr = ctab(M, M.kb + (size_t)r * M.rows, Ti, Td) * pr * coM;
    } else {
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        double a = n1[s], b = n2[s];
        if (a != 0) pf = pf * rpow(coi[s], a);
        if (b != 0) pr = pr * rpow(coi[s], b);
      }
      const double* tb = cls == 1 ? M.troe + (size_t)(r - M.na) * 4 * M.rows : M.lind + (size_t)(r - M.na - M.nt) * 3 * M.rows;
      double kinf = ctab(M, tb, Ti, Td), k0 = ctab(M, tb + M.rows, Ti, Td), kc = ctab(M, tb + 2 * M.rows, Ti, Td);
      double Pr = k0 * coM / kinf;
      double k1 = kinf * (Pr / (1.0 + Pr));
      if (cls == 1) k1 = k1 * troe_F(Pr, ctab(M, tb + 3 * M.rows, Ti, Td));
      double k2 = k1 / kc;
      rf = k1 * pf;
      rr = k2 * pr;
    }
    double net = rf - rr;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double dn = n2[s] - n1[s];
      om[s] = om[s] + M.wm[s] * dn * net;
    }
  }
}

template <int NS>
__device__ __noinline__ void chem_rhs(const ChemDev& M, const double* Z, double* F) {
  double T = Z[NS];
  if (T < (double)M.tmin || T >= (double)(M.tmin + M.rows - 1) || T != T) {
#pragma unroll
    for (int q = 0; q <= NS; ++q) F[q] = -1.0;
    return;
  }
  int Ti = static_cast<int>(T);
  double Tf = T - (double)Ti;
  double roi[NS], dr[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) roi[s] = Z[s] > 0.0 ? Z[s] : 0.0;
  wdot<NS>(M, roi, T, dr);
  double ei = 0.0, rcv = 0.0;
  int r = Ti - M.tmin;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double* ht = M.h + (size_t)s * M.rows;
    const double* ct = M.cp + (size_t)s * M.rows;
    double h0 = ht[r], h1 = ht[r + 1], c0 = ct[r], c1 = ct[r + 1];
    double hv = h0 + Tf * (h1 - h0);
    double cv = c0 + Tf * (c1 - c0);
    ei = ei + (hv - M.Ri[s] * T) * dr[s];
    rcv = rcv + roi[s] * (cv - M.Ri[s]);
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) F[s] = dr[s];
  F[NS] = -ei / rcv;
}

// DEC (Moler ACM 423) on a register matrix a[i][j]; returns ier (0: regular).  ip[k] = pivot row of column k.
template <int N>
__device__ __forceinline__ int lu_dec(double (&a)[N][N], int (&ip)[N]) {
  ip[N - 1] = 1;
#pragma unroll
  for (int k = 0; k < N - 1; ++k) {
    int m = k;
    double am = fabs(a[k][k]);
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      double v = fabs(a[i][k]);
      if (v > am) { m = i; am = v; }
    }
    ip[k] = m;
    double t = a[k][k];
#pragma unroll
    for (int i = k + 1; i < N; ++i) t = (m == i) ? a[i][k] : t;
    if (m != k) {
      ip[N - 1] = -ip[N - 1];
#pragma unroll
      for (int i = k + 1; i < N; ++i) a[i][k] = (m == i) ? a[k][k] : a[i][k];
      a[k][k] = t;
    }
    if (t == 0.0) return k + 1;
    t = 1.0 / t;
#pragma unroll
    for (int i = k + 1; i < N; ++i) a[i][k] = -a[i][k] * t;
#pragma unroll
    for (int j = k + 1; j < N; ++j) {
      double tj = a[k][j];
#pragma unroll
      for (int i = k + 1; i < N; ++i) tj = (m == i) ? a[i][j] : tj;
      double akj = a[k][j];
#pragma unroll
      for (int i = k + 1; i < N; ++i) a[i][j] = (m == i) ? akj : a[i][j];
      a[k][j] = tj;
      if (tj != 0.0) {
#pragma unroll
        for (int i = k + 1; i < N; ++i) a[i][j] = a[i][j] + a[i][k] * tj;
      }
    }
  }
  if (a[N - 1][N - 1] == 0.0) return N;
  return 0;
}

template <int N>
__device__ __forceinline__ void lu_sol(const double (&a)[N][N], const int (&ip)[N], double (&b)[N]) {
#pragma unroll
  for (int k = 0; k < N - 1; ++k) {
    int m = ip[k];
    double t = b[k];
#pragma unroll
    for (int i = k + 1; i < N; ++i) t = (m == i) ? b[i] : t;
    double bk = b[k];
#pragma unroll
    for (int i = k + 1; i < N; ++i) b[i] = (m == i) ? bk : b[i];
    b[k] = t;
#pragma unroll
    for (int i = k + 1; i < N; ++i) b[i] = b[i] + a[i][k] * t;
  }
#pragma unroll
  for (int kb = 1; kb <= N - 1; ++kb) {
    const int k = N - kb;
    b[k] = b[k] / a[k][k];
    double t = -b[k];
#pragma unroll
    for (int i = 0; i < k; ++i) b[i] = b[i] + a[i][k] * t;
  }
  b[0] = b[0] / a[0][0];
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

// One lane per cell; persistent over the cell queue.  Each loop iteration is one step attempt of the lane's cell.
template <int NS>
__global__ void __launch_bounds__(64) k_chem(ChemDev M, ChemArgs A) {
  constexpr int NZ = NS + 1;
  using R = RodasC;
  const double uround = 1.1e-19, safe = 0.9, fac1 = 5.0, fac2 = 1.0 / 6.0;
  const int nmax = 100000;
  const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double* fj = A.fjac + lane;   // element q at fj[q * nlanes]
  const long long L = A.nlanes;

  // lane state
  long long cell = -1;
  bool active = false, newstep = true, reject = false, last = false;
  int nsing = 0, nst = 0, nacc = 0, nrej = 0, nfcn = 0;
  double x = 0.0, h = 0.0, hopt = 0.0, hacc = 0.0, erracc = 1e-2, hmaxn = 0.0, posneg = 1.0;
  double rho0 = 0.0, Tb = 0.0;
  double y[NZ], dy1[NZ];
  int idid = 0;
  bool done_all = false;

  while (true) {
    if (!active) {
      if (done_all) break;
      unsigned long long c = atomicAdd(A.queue, 1ull);
      if ((long long)c >= A.ncell) { done_all = true; break; }
      cell = (long long)c;
      // Chemistry_Newstate_blk: Z = (rho_i, T = p / (rho R))
      double rho = 0.0, Rm = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) { y[s] = A.state[(long long)s * A.stride + cell]; rho = rho + y[s]; }
#pragma unroll
      for (int s = 0; s < NS; ++s) Rm = Rm + M.Ri[s] * y[s] / rho;
      y[NS] = A.state[(long long)NS * A.stride + cell] / (rho * Rm);
      rho0 = rho;
      Tb = y[NS];
      // ROSCOR initial preparations (x = 0, xend = dt, H = 0 -> 1e-6)
      x = 0.0;
      double hmax = A.dt - x;
      posneg = copysign(1.0, A.dt - x);
      hmaxn = fmin(fabs(hmax), fabs(A.dt - x));
      h = 0.0;
      if (fabs(h) <= 10.0 * uround) h = 1.0e-6;
      h = fmin(fabs(h), hmaxn);
      h = copysign(h, posneg);
      reject = false; last = false; nsing = 0; nst = 0; nacc = 0; nrej = 0; nfcn = 0;
      hopt = h; hacc = 0.0; erracc = 1e-2;
      newstep = true;
      active = true;
      idid = 0;
    }
    bool finish = false;
    if (newstep) {
      // label 1: basic integration step
      if (nst > nmax) { idid = -2; finish = true; }
      else if (0.1 * fabs(h) <= fabs(x) * uround) { idid = -3; finish = true; }
      else if (last) { h = hopt; idid = 1; finish = true; }
      else {
        hopt = h;
        if ((x + h * 1.0001 - A.dt) * posneg >= 0.0) { h = A.dt - x; last = true; }
        chem_rhs<NS>(M, y, dy1);
        nfcn++;
        // numerical Jacobian, full (IJAC = 0): column i from y(i) + sqrt(uround max(1e-5, |y(i)|))
#pragma unroll
        for (int i = 0; i < NZ; ++i) {
          double ysafe = y[i];
          double delt = sqrt(uround * fmax(1.0e-5, fabs(ysafe)));
          y[i] = ysafe + delt;
          double f[NZ];
          chem_rhs<NS>(M, y, f);
#pragma unroll
          for (int j = 0; j < NZ; ++j) fj[(long long)(j + NZ * i) * L] = (f[j] - dy1[j]) / delt;
          y[i] = ysafe;
        }
        nfcn += NZ;
        newstep = false;
      }
    }
    if (!finish) {
      // label 2: stages
      double e[NZ][NZ];
      int ip[NZ];
      double fac = 1.0 / (h * R::GAMMA);
#pragma unroll
      for (int j = 0; j < NZ; ++j) {
#pragma unroll
        for (int i = 0; i < NZ; ++i) e[i][j] = -fj[(long long)(i + NZ * j) * L];
        e[j][j] = e[j][j] + fac;
      }
      int ier = lu_dec<NZ>(e, ip);
      if (ier != 0) {
        nsing++;
        if (nsing >= 5) { idid = -4; finish = true; }
        else { h = h * 0.5; reject = true; last = false; }
      } else {
        double hc21 = R::C21 / h, hc31 = R::C31 / h, hc32 = R::C32 / h, hc41 = R::C41 / h, hc42 = R::C42 / h;
        double hc43 = R::C43 / h, hc51 = R::C51 / h, hc52 = R::C52 / h, hc53 = R::C53 / h, hc54 = R::C54 / h;
        double hc61 = R::C61 / h, hc62 = R::C62 / h, hc63 = R::C63 / h, hc64 = R::C64 / h, hc65 = R::C65 / h;
        double ak1[NZ], ak2[NZ], ak3[NZ], ak4[NZ], ak5[NZ], ak6[NZ], yn[NZ], dy[NZ];
#pragma unroll
        for (int i = 0; i < NZ; ++i) ak1[i] = dy1[i];
        lu_sol<NZ>(e, ip, ak1);
#pragma unroll
        for (int i = 0; i < NZ; ++i) yn[i] = y[i] + R::A21 * ak1[i];
        chem_rhs<NS>(M, yn, dy);
#pragma unroll
        for (int i = 0; i < NZ; ++i) ak2[i] = dy[i] + hc21 * ak1[i];
        lu_sol<NZ>(e, ip, ak2);
#pragma unroll
        for (int i = 0; i < NZ; ++i) yn[i] = y[i] + R::A31 * ak1[i] + R::A32 * ak2[i];
        chem_rhs<NS>(M, yn, dy);
#pragma unroll
        for (int i = 0; i < NZ; ++i) ak3[i] = dy[i] + (hc31 * ak1[i] + hc32 * ak2[i]);
        lu_sol<NZ>(e, ip, ak3);
#pragma unroll
        for (int i = 0; i < NZ; ++i) yn[i] = y[i] + R::A41 * ak1[i] + R::A42 * ak2[i] + R::A43 * ak3[i];
        chem_rhs<NS>(M, yn, dy);
#pragma unroll
        for (int i = 0; i < NZ; ++i) ak4[i] = dy[i] + (hc41 * ak1[i] + hc42 * ak2[i] + hc43 * ak3[i]);
        lu_sol<NZ>(e, ip, ak4);
#pragma unroll
        for (int i = 0; i < NZ; ++i) yn[i] = y[i] + R::A51 * ak1[i] + R::A52 * ak2[i] + R::A53 * ak3[i] + R::A54 * ak4[i];
        chem_rhs<NS>(M, yn, dy);
#pragma unroll
        for (int i = 0; i < NZ; ++i) ak5[i] = dy[i] + (hc52 * ak2[i] + hc54 * ak4[i] + hc51 * ak1[i] + hc53 * ak3[i]);
        lu_sol<NZ>(e, ip, ak5);
#pragma unroll
        for (int i = 0; i < NZ; ++i) yn[i] = yn[i] + ak5[i];
        chem_rhs<NS>(M, yn, dy);
#pragma unroll
        for (int i = 0; i < NZ; ++i)
          ak6[i] = dy[i] + (hc61 * ak1[i] + hc62 * ak2[i] + hc65 * ak5[i] + hc64 * ak4[i] + hc63 * ak3[i]);
        lu_sol<NZ>(e, ip, ak6);
#pragma unroll
        for (int i = 0; i < NZ; ++i) yn[i] = yn[i] + ak6[i];
        nfcn += 5;
        nst++;
        double err = 0.0;
#pragma unroll
        for (int i = 0; i < NZ; ++i) {
          double sk = A.atol[i] + A.rtol[i] * fmax(fabs(y[i]), fabs(yn[i]));
          double q = ak6[i] / sk;
          err = err + q * q;
        }
        err = sqrt(err / NZ);
        fac = fmax(fac2, fmin(fac1, sqrt(sqrt(err)) / safe));   // err**0.25, as the host side
        double hnew = h / fac;
        if (err <= 1.0) {
          nacc++;
          if (nacc > 1) {
            double facgus = (hacc / h) * sqrt(sqrt(err * err / erracc)) / safe;
            facgus = fmax(fac2, fmin(fac1, facgus));
            fac = fmax(fac, facgus);
            hnew = h / fac;
          }
          hacc = h;
          erracc = fmax(1.0e-2, err);
#pragma unroll
          for (int i = 0; i < NZ; ++i) y[i] = yn[i];
          x = x + h;
          if (fabs(hnew) > hmaxn) hnew = posneg * hmaxn;
          if (reject) hnew = posneg * fmin(fabs(hnew), fabs(h));
          reject = false;
          h = hnew;
          newstep = true;
        } else {
          reject = true;
          last = false;
          h = hnew;
          if (nacc >= 1) nrej++;
        }
      }
    }
    if (finish) {
      // floors, p = T rho R, hr (Chemistry_Newstate_blk)
      double rs = 0.0, R2 = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (y[s] <= 1e-20) y[s] = 1e-20;
        A.state[(long long)s * A.stride + cell] = y[s];
        rs = rs + y[s];
      }
#pragma unroll
      for (int s = 0; s < NS; ++s) R2 = R2 + M.Ri[s] * y[s] / rs;
      A.state[(long long)NS * A.stride + cell] = y[NS] * rho0 * R2;
      A.hr[cell] = fabs(y[NS] - Tb) / Tb;
      A.idid[cell] = idid;
      A.nstep[cell] = nst;
      atomicAdd(&A.stats[0], (unsigned long long)nst);
      atomicAdd(&A.stats[1], (unsigned long long)nrej);
      atomicAdd(&A.stats[2], (unsigned long long)nfcn);
      if (idid < 0) atomicAdd(&A.stats[3], 1ull);
      atomicMax(&A.stats[4], (unsigned long long)nst);
      active = false;
    }
  }
}

}  // namespace mose
