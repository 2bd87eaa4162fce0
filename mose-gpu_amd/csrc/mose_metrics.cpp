// Host-side geometry of a multi-block structured mesh: cell metric tensor, cell lengths, volumes,
// face unit normals and areas, and the ghost-cell metrics the boundary records need.
//
// Follows src/lib/numerics/space/Lib_Metrics.f90 (Setup_Metrics :15-136, Compute_Norm_Area :199-379,
// Compute_Metric_Tensor :382-510, BC_Extrapolate_Metrics :765-1100, BC_Connect_Metrics :1105-1373)
// in the reference's floating-point order, so that the kernels read the geometry the reference reads.
// Compiled with -ffp-contract=off.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <algorithm>
#include <vector>

#include "mose_gpu.h"
#include "mose_internal.h"

namespace {

constexpr int GC = MOSE_GC;

struct V3 {
  double c[3];
};

struct BlockGeo {
  int n1, n2, n3;
  int64_t boff, f1off, f2off, f3off;
  int pi, pj;
  const double* node;   // (n1+1)(n2+1)(n3+1) x 3, i fastest
  int64_t row(int i, int j, int k) const {
    return boff + (i + GC - 1) + (int64_t)(j + GC - 1) * pi + (int64_t)(k + GC - 1) * pi * pj;
  }
  V3 nd(int i, int j, int k) const {
    const double* p = node + 3 * ((int64_t)i + (int64_t)(n1 + 1) * (j + (int64_t)(n2 + 1) * k));
    return V3{{p[0], p[1], p[2]}};
  }
};

struct Out {
  double *dl, *m, *vol, *fm1, *fm2, *fm3;
  int64_t NC, NF1, NF2, NF3;
};

inline V3 sub(const V3& a, const V3& b) { return V3{{a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]}}; }
inline V3 add(const V3& a, const V3& b) { return V3{{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }

// 3x1 - 3x2 + x3 (cubic node extrapolation), evaluated left to right
inline V3 cubic(const V3& a, const V3& b, const V3& c) {
  V3 r;
  for (int q = 0; q < 3; ++q) r.c[q] = 3.0 * a.c[q] - 3.0 * b.c[q] + c.c[q];
  return r;
}
inline V3 lin(const V3& a, const V3& b) {
  V3 r;
  for (int q = 0; q < 3; ++q) r.c[q] = 2.0 * a.c[q] - b.c[q];
  return r;
}

double tvol(const double* vx, const double* vy, const double* vz, int i1, int i2, int i3, int i4) {
  --i1; --i2; --i3; --i4;
  return std::fabs(((vx[i2] - vx[i1]) * ((vy[i3] - vy[i1]) * (vz[i4] - vz[i1]) - (vy[i4] - vy[i1]) * (vz[i3] - vz[i1])) +
                    (vy[i2] - vy[i1]) * ((vx[i4] - vx[i1]) * (vz[i3] - vz[i1]) - (vx[i3] - vx[i1]) * (vz[i4] - vz[i1])) +
                    (vz[i2] - vz[i1]) * ((vx[i3] - vx[i1]) * (vy[i4] - vy[i1]) - (vx[i4] - vx[i1]) * (vy[i3] - vy[i1]))) /
                   6.0);
}

// Compute_Metric_Tensor (Lib_Metrics.f90:382-510).  Returns false on a degenerate cell.
bool metric_tensor(const V3& N1, const V3& N2, const V3& N3, const V3& N4, const V3& N5, const V3& N6, const V3& N7,
                   const V3& N8, const Out& o, int64_t r) {
  double A[3][3];
  for (int q = 0; q < 3; ++q) {
    A[0][q] = N5.c[q] - N1.c[q] + N6.c[q] - N2.c[q] + N7.c[q] - N3.c[q] + N8.c[q] - N4.c[q];
    A[1][q] = N3.c[q] - N1.c[q] + N4.c[q] - N2.c[q] + N7.c[q] - N5.c[q] + N8.c[q] - N6.c[q];
    A[2][q] = N2.c[q] - N1.c[q] + N4.c[q] - N3.c[q] + N6.c[q] - N5.c[q] + N8.c[q] - N7.c[q];
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) A[a][b] = 0.25 * A[a][b];
  double det = A[0][0] * A[1][1] * A[2][2] - A[0][0] * A[1][2] * A[2][1] - A[0][1] * A[1][0] * A[2][2] +
               A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1] - A[0][2] * A[1][1] * A[2][0];
  if (std::fabs(det) == 0.0) return false;
  double cf[3][3];
  cf[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
  cf[0][1] = -(A[1][0] * A[2][2] - A[1][2] * A[2][0]);
  cf[0][2] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  cf[1][0] = -(A[0][1] * A[2][2] - A[0][2] * A[2][1]);
  cf[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
  cf[1][2] = -(A[0][0] * A[2][1] - A[0][1] * A[2][0]);
  cf[2][0] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
  cf[2][1] = -(A[0][0] * A[1][2] - A[0][2] * A[1][0]);
  cf[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
  // M = cofactor / det, M(a,b) stored at m[(a+3b)*NC + r]
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) o.m[(int64_t)(a + 3 * b) * o.NC + r] = cf[a][b] / det;
  for (int h = 0; h < 3; ++h)
    o.dl[(int64_t)h * o.NC + r] = std::sqrt(A[h][0] * A[h][0] + A[h][1] * A[h][1] + A[h][2] * A[h][2]);
  const V3* v[8] = {&N1, &N5, &N3, &N7, &N2, &N6, &N4, &N8};
  double vx[8], vy[8], vz[8];
  for (int q = 0; q < 8; ++q) { vx[q] = v[q]->c[0]; vy[q] = v[q]->c[1]; vz[q] = v[q]->c[2]; }
  double vol = tvol(vx, vy, vz, 1, 2, 3, 5) + tvol(vx, vy, vz, 2, 4, 3, 8) + tvol(vx, vy, vz, 5, 8, 6, 2) +
               tvol(vx, vy, vz, 5, 7, 8, 3) + tvol(vx, vy, vz, 5, 8, 2, 3);
  o.vol[r] = vol;
  return vol > 0.0;
}

inline double dsign(double a, double b) { return std::copysign(std::fabs(a), b); }

// Compute_Norm_Area (Lib_Metrics.f90:199-379)
void norm_area(const BlockGeo& B, const Out& o) {
  auto cross = [](const V3& d1, const V3& d2, double s) {
    V3 r;
    r.c[0] = (d1.c[1] * d2.c[2] - d1.c[2] * d2.c[1]) * s;
    r.c[1] = (d1.c[2] * d2.c[0] - d1.c[0] * d2.c[2]) * s;
    r.c[2] = (d1.c[0] * d2.c[1] - d1.c[1] * d2.c[0]) * s;
    return r;
  };
  auto quarter_sum = [](const V3& a, const V3& b, const V3& c, const V3& d) {
    V3 r;
    for (int q = 0; q < 3; ++q) r.c[q] = 0.25 * (a.c[q] + b.c[q] + c.c[q] + d.c[q]);
    return r;
  };
  auto dot = [](const V3& a, const V3& b) { return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2]; };
  V3 sn, d3;
  sn = cross(sub(B.nd(1, 1, 1), B.nd(1, 0, 0)), sub(B.nd(1, 0, 1), B.nd(1, 1, 0)), 1.0);
  d3 = sub(quarter_sum(B.nd(1, 1, 1), B.nd(1, 0, 1), B.nd(1, 0, 0), B.nd(1, 1, 0)),
           quarter_sum(B.nd(0, 1, 1), B.nd(0, 0, 1), B.nd(0, 0, 0), B.nd(0, 1, 0)));
  double signi = dsign(1.0, dot(d3, sn));
  sn = cross(sub(B.nd(0, 1, 1), B.nd(1, 1, 0)), sub(B.nd(1, 1, 1), B.nd(0, 1, 0)), 1.0);
  d3 = sub(quarter_sum(B.nd(1, 1, 1), B.nd(0, 1, 1), B.nd(0, 1, 0), B.nd(1, 1, 0)),
           quarter_sum(B.nd(1, 0, 1), B.nd(0, 0, 1), B.nd(0, 0, 0), B.nd(1, 0, 0)));
  double signj = dsign(1.0, dot(d3, sn));
  sn = cross(sub(B.nd(0, 1, 1), B.nd(1, 0, 1)), sub(B.nd(0, 0, 1), B.nd(1, 1, 1)), 1.0);
  d3 = sub(quarter_sum(B.nd(1, 1, 1), B.nd(0, 1, 1), B.nd(0, 0, 1), B.nd(1, 0, 1)),
           quarter_sum(B.nd(1, 1, 0), B.nd(0, 1, 0), B.nd(0, 0, 0), B.nd(1, 0, 0)));
  double signk = dsign(1.0, dot(d3, sn));
  auto put = [&](double* fm, int64_t NF, int64_t f, const V3& d, double sg) {
    double A = std::sqrt(d.c[0] * d.c[0] + d.c[1] * d.c[1] + d.c[2] * d.c[2]);
    double nx = 0.0, ny = 0.0, nz = 0.0;
    if (A != 0.0) { nx = d.c[0] / A * sg; ny = d.c[1] / A * sg; nz = d.c[2] / A * sg; }
    fm[f] = nx; fm[NF + f] = ny; fm[2 * NF + f] = nz; fm[3 * NF + f] = A;
  };
  for (int k = 1; k <= B.n3; ++k)
    for (int j = 1; j <= B.n2; ++j)
      for (int i = 0; i <= B.n1; ++i) {
        V3 d = cross(sub(B.nd(i, j, k), B.nd(i, j - 1, k - 1)), sub(B.nd(i, j - 1, k), B.nd(i, j, k - 1)), 0.5);
        put(o.fm1, o.NF1, B.f1off + i + (int64_t)(B.n1 + 1) * ((j - 1) + (int64_t)B.n2 * (k - 1)), d, signi);
      }
  for (int k = 1; k <= B.n3; ++k)
    for (int j = 0; j <= B.n2; ++j)
      for (int i = 1; i <= B.n1; ++i) {
        V3 d = cross(sub(B.nd(i - 1, j, k), B.nd(i, j, k - 1)), sub(B.nd(i, j, k), B.nd(i - 1, j, k - 1)), 0.5);
        put(o.fm2, o.NF2, B.f2off + (i - 1) + (int64_t)B.n1 * (j + (int64_t)(B.n2 + 1) * (k - 1)), d, signj);
      }
  for (int k = 0; k <= B.n3; ++k)
    for (int j = 1; j <= B.n2; ++j)
      for (int i = 1; i <= B.n1; ++i) {
        V3 d = cross(sub(B.nd(i - 1, j, k), B.nd(i, j - 1, k)), sub(B.nd(i - 1, j - 1, k), B.nd(i, j, k)), 0.5);
        put(o.fm3, o.NF3, B.f3off + (i - 1) + (int64_t)B.n1 * ((j - 1) + (int64_t)B.n2 * k), d, signk);
      }
}

void copy_cell(const Out& o, int64_t dst, int64_t src) {
  for (int h = 0; h < 3; ++h) o.dl[(int64_t)h * o.NC + dst] = o.dl[(int64_t)h * o.NC + src];
  for (int q = 0; q < 9; ++q) o.m[(int64_t)q * o.NC + dst] = o.m[(int64_t)q * o.NC + src];
  o.vol[dst] = o.vol[src];
}

// BC_Extrapolate_Metrics (Lib_Metrics.f90:765-1100) for the non-axisymmetric meshes (delthe = 0)
bool extrapolate_metrics(const BlockGeo& B, int ndir, int Im, int Jm, int Km, int Fm, const Out& o) {
  V3 N[8][GC * 2 + 1];   // N[node][g + GC]
  auto at = [&](int nodeid, int g) -> V3& { return N[nodeid][g + GC]; };
  bool ok = true;
  switch (Fm) {
    case 1:
      for (int g = 0; g <= GC; ++g) {
        at(0, g) = B.nd(Im + g - 1, Jm - 1, Km - 1); at(1, g) = B.nd(Im + g - 1, Jm - 1, Km);
        at(2, g) = B.nd(Im + g - 1, Jm, Km - 1); at(3, g) = B.nd(Im + g - 1, Jm, Km);
      }
      for (int g = -1; g >= -GC; --g) {
        for (int q = 0; q < 4; ++q) at(q, g) = cubic(at(q, g + 1), at(q, g + 2), at(q, g + 3));
        ok &= metric_tensor(at(0, g), at(1, g), at(2, g), at(3, g), at(0, g + 1), at(1, g + 1), at(2, g + 1), at(3, g + 1),
                            o, B.row(Im + g, Jm, Km));
      }
      break;
    case 2:
      for (int g = 0; g >= -GC; --g) {
        at(4, g) = B.nd(Im + g, Jm - 1, Km - 1); at(5, g) = B.nd(Im + g, Jm - 1, Km);
        at(6, g) = B.nd(Im + g, Jm, Km - 1); at(7, g) = B.nd(Im + g, Jm, Km);
      }
      for (int g = 1; g <= GC; ++g) {
        for (int q = 4; q < 8; ++q) at(q, g) = cubic(at(q, g - 1), at(q, g - 2), at(q, g - 3));
        ok &= metric_tensor(at(4, g - 1), at(5, g - 1), at(6, g - 1), at(7, g - 1), at(4, g), at(5, g), at(6, g), at(7, g),
                            o, B.row(Im + g, Jm, Km));
      }
      break;
    case 3: {
      bool c = B.n2 > 2;
      int top = c ? GC : 1;
      for (int g = 0; g <= top; ++g) {
        at(0, g) = B.nd(Im - 1, Jm + g - 1, Km - 1); at(1, g) = B.nd(Im - 1, Jm + g - 1, Km);
        at(4, g) = B.nd(Im, Jm + g - 1, Km - 1); at(5, g) = B.nd(Im, Jm + g - 1, Km);
      }
      for (int g = -1; g >= -GC; --g) {
        for (int q : {0, 1, 4, 5})
          at(q, g) = c ? cubic(at(q, g + 1), at(q, g + 2), at(q, g + 3)) : lin(at(q, g + 1), at(q, g + 2));
        ok &= metric_tensor(at(0, g), at(1, g), at(0, g + 1), at(1, g + 1), at(4, g), at(5, g), at(4, g + 1), at(5, g + 1),
                            o, B.row(Im, Jm + g, Km));
      }
      break;
    }
    case 4: {
      bool c = B.n2 > 2;
      int bot = c ? -GC : -1;
      for (int g = 0; g >= bot; --g) {
        at(2, g) = B.nd(Im - 1, Jm + g, Km - 1); at(3, g) = B.nd(Im - 1, Jm + g, Km);
        at(6, g) = B.nd(Im, Jm + g, Km - 1); at(7, g) = B.nd(Im, Jm + g, Km);
      }
      for (int g = 1; g <= GC; ++g) {
        for (int q : {2, 3, 6, 7})
          at(q, g) = c ? cubic(at(q, g - 1), at(q, g - 2), at(q, g - 3)) : lin(at(q, g - 1), at(q, g - 2));
        ok &= metric_tensor(at(2, g - 1), at(3, g - 1), at(2, g), at(3, g), at(6, g - 1), at(7, g - 1), at(6, g), at(7, g),
                            o, B.row(Im, Jm + g, Km));
      }
      break;
    }
    case 5:
    case 6: {
      bool lo = (Fm == 5);
      if (ndir == 1) {
        for (int g = 1; g <= GC; ++g) copy_cell(o, B.row(Im, Jm, lo ? Km - g : Km + g), B.row(Im, Jm, Km));
        break;
      }
      bool c = (ndir == 3);
      if (lo) {
        int top = c ? GC : 1;
        for (int g = 0; g <= top; ++g) {
          at(0, g) = B.nd(Im - 1, Jm - 1, Km + g - 1); at(2, g) = B.nd(Im - 1, Jm, Km + g - 1);
          at(4, g) = B.nd(Im, Jm - 1, Km + g - 1); at(6, g) = B.nd(Im, Jm, Km + g - 1);
        }
        for (int g = -1; g >= -GC; --g) {
          for (int q : {0, 2, 4, 6})
            at(q, g) = c ? cubic(at(q, g + 1), at(q, g + 2), at(q, g + 3)) : lin(at(q, g + 1), at(q, g + 2));
          ok &= metric_tensor(at(0, g), at(0, g + 1), at(2, g), at(2, g + 1), at(4, g), at(4, g + 1), at(6, g), at(6, g + 1),
                              o, B.row(Im, Jm, Km + g));
        }
      } else {
        int bot = c ? -GC : -1;
        for (int g = 0; g >= bot; --g) {
          at(1, g) = B.nd(Im - 1, Jm - 1, Km + g); at(3, g) = B.nd(Im - 1, Jm, Km + g);
          at(5, g) = B.nd(Im, Jm - 1, Km + g); at(7, g) = B.nd(Im, Jm, Km + g);
        }
        for (int g = 1; g <= GC; ++g) {
          for (int q : {1, 3, 5, 7})
            at(q, g) = c ? cubic(at(q, g - 1), at(q, g - 2), at(q, g - 3)) : lin(at(q, g - 1), at(q, g - 2));
          ok &= metric_tensor(at(1, g - 1), at(1, g), at(3, g - 1), at(3, g), at(5, g - 1), at(5, g), at(7, g - 1), at(7, g),
                              o, B.row(Im, Jm, Km + g));
        }
      }
      break;
    }
  }
  return ok;
}

// BC_Connect_Metrics (Lib_Metrics.f90:1105-1373): the ghost nodes are the donor nodes, translated by one
// period for a periodic record (MOSE 201)
bool connect_metrics(const BlockGeo& Bm, const BlockGeo& Bs, const mose_bc_record& r, const Out& o) {
  static const int guides[6][3] = {{0, 0, 0}, {-1, 0, 0}, {0, 0, 0}, {0, -1, 0}, {0, 0, 0}, {0, 0, -1}};
  static const int guides2[6][3] = {{1, 0, 0}, {-2, 0, 0}, {0, 1, 0}, {0, -2, 0}, {0, 0, 1}, {0, 0, -2}};
  int Im = r.i, Jm = r.j, Km = r.k, Fm = r.f, Is = r.is, Js = r.js, Ks = r.ks, Fs = r.fs;
  int d11 = r.d11, d12 = r.d12, d21 = r.d21, d22 = r.d22;
  V3 N[8][GC + 1];
  switch (Fm) {
    case 1: N[0][0] = Bm.nd(Im - 1, Jm - 1, Km - 1); N[1][0] = Bm.nd(Im - 1, Jm - 1, Km); N[2][0] = Bm.nd(Im - 1, Jm, Km - 1); N[3][0] = Bm.nd(Im - 1, Jm, Km); break;
    case 2: N[4][0] = Bm.nd(Im, Jm - 1, Km - 1); N[5][0] = Bm.nd(Im, Jm - 1, Km); N[6][0] = Bm.nd(Im, Jm, Km - 1); N[7][0] = Bm.nd(Im, Jm, Km); break;
    case 3: N[0][0] = Bm.nd(Im - 1, Jm - 1, Km - 1); N[1][0] = Bm.nd(Im - 1, Jm - 1, Km); N[4][0] = Bm.nd(Im, Jm - 1, Km - 1); N[5][0] = Bm.nd(Im, Jm - 1, Km); break;
    case 4: N[2][0] = Bm.nd(Im - 1, Jm, Km - 1); N[3][0] = Bm.nd(Im - 1, Jm, Km); N[6][0] = Bm.nd(Im, Jm, Km - 1); N[7][0] = Bm.nd(Im, Jm, Km); break;
    case 5: N[0][0] = Bm.nd(Im - 1, Jm - 1, Km - 1); N[2][0] = Bm.nd(Im - 1, Jm, Km - 1); N[4][0] = Bm.nd(Im, Jm - 1, Km - 1); N[6][0] = Bm.nd(Im, Jm, Km - 1); break;
    default: N[1][0] = Bm.nd(Im - 1, Jm - 1, Km); N[3][0] = Bm.nd(Im - 1, Jm, Km); N[5][0] = Bm.nd(Im, Jm - 1, Km); N[7][0] = Bm.nd(Im, Jm, Km); break;
  }
  auto map3 = [&](int Is1, int Js1, int Ks1, int& II, int& JJ, int& KK) {
    if (Fs <= 2) { II = Is1; JJ = (2 * Js1 + d11 + d21) / 2; KK = (2 * Ks1 + d12 + d22) / 2; }
    else if (Fs <= 4) { II = (2 * Is1 + d11 + d21) / 2; JJ = Js1; KK = (2 * Ks1 + d12 + d22) / 2; }
    else { II = (2 * Is1 + d11 + d21) / 2; JJ = (2 * Js1 + d12 + d22) / 2; KK = Ks1; }
  };
  V3 T{{0.0, 0.0, 0.0}};
  if (r.periodic) {
    int Is1 = Is + GC * guides[Fs - 1][0] - guides2[Fs - 1][0];
    int Js1 = Js + GC * guides[Fs - 1][1] - guides2[Fs - 1][1];
    int Ks1 = Ks + GC * guides[Fs - 1][2] - guides2[Fs - 1][2];
    int II, JJ, KK;
    map3(Is1, Js1, Ks1, II, JJ, KK);
    static const int own[6] = {3, 7, 5, 7, 6, 7};   // N4, N8, N6, N8, N7, N8
    T = sub(N[own[Fm - 1]][0], Bs.nd(II, JJ, KK));
  }
  // node slots written per Fm: (main, i1-1, i2-1, both) in N-index (0-based)
  static const int slot[6][4] = {{3, 1, 2, 0}, {7, 5, 6, 4}, {5, 1, 4, 0}, {7, 3, 6, 2}, {6, 2, 4, 0}, {7, 3, 5, 1}};
  bool ok = true;
  for (int g = 1; g <= GC; ++g) {
    int Is1 = Is + (GC - g) * guides[Fs - 1][0] + (g - 1) * guides2[Fs - 1][0];
    int Js1 = Js + (GC - g) * guides[Fs - 1][1] + (g - 1) * guides2[Fs - 1][1];
    int Ks1 = Ks + (GC - g) * guides[Fs - 1][2] + (g - 1) * guides2[Fs - 1][2];
    int II, JJ, KK;
    map3(Is1, Js1, Ks1, II, JJ, KK);
    int o1[3], o2[3], o3[3];
    if (Fs <= 2) { int a[3] = {0, d11, d12}, b[3] = {0, d21, d22}; for (int q = 0; q < 3; ++q) { o1[q] = a[q]; o2[q] = b[q]; o3[q] = a[q] + b[q]; } }
    else if (Fs <= 4) { int a[3] = {d11, 0, d12}, b[3] = {d21, 0, d22}; for (int q = 0; q < 3; ++q) { o1[q] = a[q]; o2[q] = b[q]; o3[q] = a[q] + b[q]; } }
    else { int a[3] = {d11, d12, 0}, b[3] = {d21, d22, 0}; for (int q = 0; q < 3; ++q) { o1[q] = a[q]; o2[q] = b[q]; o3[q] = a[q] + b[q]; } }
    N[slot[Fm - 1][0]][g] = add(Bs.nd(II, JJ, KK), T);
    N[slot[Fm - 1][1]][g] = add(Bs.nd(II - o1[0], JJ - o1[1], KK - o1[2]), T);
    N[slot[Fm - 1][2]][g] = add(Bs.nd(II - o2[0], JJ - o2[1], KK - o2[2]), T);
    N[slot[Fm - 1][3]][g] = add(Bs.nd(II - o3[0], JJ - o3[1], KK - o3[2]), T);
    switch (Fm) {
      case 1: ok &= metric_tensor(N[0][g], N[1][g], N[2][g], N[3][g], N[0][g - 1], N[1][g - 1], N[2][g - 1], N[3][g - 1], o, Bm.row(Im - g, Jm, Km)); break;
      case 2: ok &= metric_tensor(N[4][g - 1], N[5][g - 1], N[6][g - 1], N[7][g - 1], N[4][g], N[5][g], N[6][g], N[7][g], o, Bm.row(Im + g, Jm, Km)); break;
      case 3: ok &= metric_tensor(N[0][g], N[1][g], N[0][g - 1], N[1][g - 1], N[4][g], N[5][g], N[4][g - 1], N[5][g - 1], o, Bm.row(Im, Jm - g, Km)); break;
      case 4: ok &= metric_tensor(N[2][g - 1], N[3][g - 1], N[2][g], N[3][g], N[6][g - 1], N[7][g - 1], N[6][g], N[7][g], o, Bm.row(Im, Jm + g, Km)); break;
      case 5: ok &= metric_tensor(N[0][g], N[0][g - 1], N[2][g], N[2][g - 1], N[4][g], N[4][g - 1], N[6][g], N[6][g - 1], o, Bm.row(Im, Jm, Km - g)); break;
      default: ok &= metric_tensor(N[1][g - 1], N[1][g], N[3][g - 1], N[3][g], N[5][g - 1], N[5][g], N[7][g - 1], N[7][g], o, Bm.row(Im, Jm, Km + g)); break;
    }
  }
  return ok;
}

// BC_Symmetry_Metrics (Lib_Metrics.f90:1044-1102): ghost g of a symmetry face takes the metric, the lengths and the
// volume of its mirror image, the interior cell g - 1 away from the face.
void symmetry_metrics(const BlockGeo& B, int Im, int Jm, int Km, int Fm, const Out& o) {
  for (int g = 1; g <= GC; ++g) {
    int gi = Im, gj = Jm, gk = Km, si = Im, sj = Jm, sk = Km;
    switch (Fm) {
      case 1: gi = Im - g; si = Im + g - 1; break;
      case 2: gi = Im + g; si = Im - g + 1; break;
      case 3: gj = Jm - g; sj = Jm + g - 1; break;
      case 4: gj = Jm + g; sj = Jm - g + 1; break;
      case 5: gk = Km - g; sk = Km + g - 1; break;
      default: gk = Km + g; sk = Km - g + 1; break;
    }
    const int64_t rg = B.row(gi, gj, gk), rs = B.row(si, sj, sk);
    for (int q = 0; q < 9; ++q) o.m[q * o.NC + rg] = o.m[q * o.NC + rs];
    for (int h = 0; h < 3; ++h) o.dl[h * o.NC + rg] = o.dl[h * o.NC + rs];
    o.vol[rg] = o.vol[rs];
  }
}

}  // namespace

extern "C" int mose_metrics_build(int nb, const int* dims, const double* nodes, int nrec, const mose_bc_record* rec,
                                  int ndir, double* dl, double* m, double* vol, double* yn, double* fm1, double* fm2,
                                  double* fm3) {
  if (nb <= 0 || !dims || !nodes || !dl || !m || !vol || !fm1 || !fm2 || !fm3) return MOSE_ERR_ARG;
  std::vector<BlockGeo> B(nb);
  int64_t NC = 0, NF1 = 0, NF2 = 0, NF3 = 0, NN = 0;
  for (int b = 0; b < nb; ++b) {
    int n1 = dims[3 * b], n2 = dims[3 * b + 1], n3 = dims[3 * b + 2];
    if (n1 < 1 || n2 < 1 || n3 < 1) return MOSE_ERR_ARG;
    B[b] = BlockGeo{n1, n2, n3, NC, NF1, NF2, NF3, n1 + 2 * GC, n2 + 2 * GC, nodes + NN};
    NC += (int64_t)(n1 + 2 * GC) * (n2 + 2 * GC) * (n3 + 2 * GC);
    NF1 += (int64_t)(n1 + 1) * n2 * n3;
    NF2 += (int64_t)n1 * (n2 + 1) * n3;
    NF3 += (int64_t)n1 * n2 * (n3 + 1);
    NN += 3 * (int64_t)(n1 + 1) * (n2 + 1) * (n3 + 1);
  }
  Out o{dl, m, vol, fm1, fm2, fm3, NC, NF1, NF2, NF3};
  // ghost rows nobody computes stay at unit metric, unit length and unit volume
  for (int64_t r = 0; r < NC; ++r) {
    for (int h = 0; h < 3; ++h) dl[h * NC + r] = 1.0;
    for (int q = 0; q < 9; ++q) m[q * NC + r] = (q % 4 == 0) ? 1.0 : 0.0;
    vol[r] = 1.0;
    if (yn) yn[r] = 1.0;
  }
  bool ok = true;
  for (int b = 0; b < nb; ++b) {
    const BlockGeo& G = B[b];
    for (int k = 1; k <= G.n3; ++k)
      for (int j = 1; j <= G.n2; ++j)
        for (int i = 1; i <= G.n1; ++i)
          ok &= metric_tensor(G.nd(i - 1, j - 1, k - 1), G.nd(i - 1, j - 1, k), G.nd(i - 1, j, k - 1), G.nd(i - 1, j, k),
                              G.nd(i, j - 1, k - 1), G.nd(i, j - 1, k), G.nd(i, j, k - 1), G.nd(i, j, k), o, G.row(i, j, k));
    norm_area(G, o);
  }
  if (!ok) return MOSE_ERR_MESH;
  for (int e = 0; e < nrec; ++e) {
    const mose_bc_record& r = rec[e];
    if (r.b < 0 || r.b >= nb) return MOSE_ERR_ARG;
    if (r.typ == MOSE_BC_CONNECTION) {
      if (r.bs < 0 || r.bs >= nb) return MOSE_ERR_ARG;
      ok &= connect_metrics(B[r.b], B[r.bs], r, o);
    } else if (r.typ == MOSE_BC_SYMMETRY) {
      if (r.f < 1 || r.f > 6) return MOSE_ERR_ARG;
      symmetry_metrics(B[r.b], r.i, r.j, r.k, r.f, o);
    } else {
      ok &= extrapolate_metrics(B[r.b], ndir, r.i, r.j, r.k, r.f, o);
    }
  }
  return ok ? MOSE_OK : MOSE_ERR_MESH;
}

// ------------------------------------------------------------------------------------------------------------
// Wall distance: Compute_Yn (Lib_Metrics.f90:513-744) and Yn_Connection (:747-763)
// ------------------------------------------------------------------------------------------------------------
namespace {

// A 3-d tree over the wall-face centroids: median splits along x, y, z in turn; the nearest-point query visits the
// far side of a split only while the squared distance to the splitting plane is below the best squared distance,
// so it returns the minimum of dx^2 + dy^2 + dz^2 over all points, the value a scan of every point gives.
struct WallTree {
  std::vector<double> x, y, z;
  void build(int lo, int hi, int axis) {   // [lo, hi)
    if (hi - lo <= 1) return;
    const int mid = (lo + hi) / 2;
    std::vector<int> idx((size_t)(hi - lo));
    for (int q = lo; q < hi; ++q) idx[(size_t)(q - lo)] = q;
    const std::vector<double>& key = axis == 0 ? x : (axis == 1 ? y : z);
    std::nth_element(idx.begin(), idx.begin() + (mid - lo), idx.end(),
                     [&](int a, int b) { return key[(size_t)a] < key[(size_t)b]; });
    std::vector<double> tx, ty, tz;
    tx.reserve(idx.size()); ty.reserve(idx.size()); tz.reserve(idx.size());
    for (int q : idx) { tx.push_back(x[(size_t)q]); ty.push_back(y[(size_t)q]); tz.push_back(z[(size_t)q]); }
    for (int q = lo; q < hi; ++q) {
      x[(size_t)q] = tx[(size_t)(q - lo)]; y[(size_t)q] = ty[(size_t)(q - lo)]; z[(size_t)q] = tz[(size_t)(q - lo)];
    }
    build(lo, mid, (axis + 1) % 3);
    build(mid + 1, hi, (axis + 1) % 3);
  }
  void nearest(int lo, int hi, int axis, double qx, double qy, double qz, double& best2) const {
    if (lo >= hi) return;
    const int mid = (lo + hi) / 2;
    const double dx = x[(size_t)mid] - qx, dy = y[(size_t)mid] - qy, dz = z[(size_t)mid] - qz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < best2) best2 = d2;
    const double diff = axis == 0 ? qx - x[(size_t)mid] : (axis == 1 ? qy - y[(size_t)mid] : qz - z[(size_t)mid]);
    const int nxt = (axis + 1) % 3;
    if (diff < 0.0) {
      nearest(lo, mid, nxt, qx, qy, qz, best2);
      if (diff * diff < best2) nearest(mid + 1, hi, nxt, qx, qy, qz, best2);
    } else {
      nearest(mid + 1, hi, nxt, qx, qy, qz, best2);
      if (diff * diff < best2) nearest(lo, mid, nxt, qx, qy, qz, best2);
    }
  }
};

}  // namespace

extern "C" int mose_wall_distance(int nb, const int* dims, const double* nodes, int nrec, const mose_bc_record* rec,
                                  double* yn) {
  if (nb <= 0 || !dims || !nodes || !yn || (nrec > 0 && !rec)) return MOSE_ERR_ARG;
  std::vector<BlockGeo> B(nb);
  int64_t NC = 0, NN = 0;
  for (int b = 0; b < nb; ++b) {
    int n1 = dims[3 * b], n2 = dims[3 * b + 1], n3 = dims[3 * b + 2];
    if (n1 < 1 || n2 < 1 || n3 < 1) return MOSE_ERR_ARG;
    B[b] = BlockGeo{n1, n2, n3, NC, 0, 0, 0, n1 + 2 * GC, n2 + 2 * GC, nodes + NN};
    NC += (int64_t)(n1 + 2 * GC) * (n2 + 2 * GC) * (n3 + 2 * GC);
    NN += 3 * (int64_t)(n1 + 1) * (n2 + 1) * (n3 + 1);
  }
  // the centroids of the wall faces (Is_Wall: the no-slip types 5 and 6 of this path), 0.25 (n1 + n2 + n3 + n4)
  WallTree t;
  for (int e = 0; e < nrec; ++e) {
    const mose_bc_record& r = rec[e];
    if (r.typ != MOSE_BC_WALL_ADIABATIC && r.typ != MOSE_BC_WALL_ISOTHERMAL) continue;
    if (r.b < 0 || r.b >= nb || r.f < 1 || r.f > 6) return MOSE_ERR_ARG;
    const BlockGeo& G = B[r.b];
    const int Fm = r.f, m2 = Fm % 2;
    int c[4][3];
    if (Fm <= 2) {
      const int i = r.i - m2;
      int q[4][3] = {{i, r.j - 1, r.k - 1}, {i, r.j, r.k - 1}, {i, r.j - 1, r.k}, {i, r.j, r.k}};
      std::memcpy(c, q, sizeof c);
    } else if (Fm <= 4) {
      const int j = r.j - m2;
      int q[4][3] = {{r.i - 1, j, r.k - 1}, {r.i, j, r.k - 1}, {r.i - 1, j, r.k}, {r.i, j, r.k}};
      std::memcpy(c, q, sizeof c);
    } else {
      const int k = r.k - m2;
      int q[4][3] = {{r.i - 1, r.j - 1, k}, {r.i, r.j - 1, k}, {r.i - 1, r.j, k}, {r.i, r.j, k}};
      std::memcpy(c, q, sizeof c);
    }
    V3 n[4];
    for (int q = 0; q < 4; ++q) n[q] = G.nd(c[q][0], c[q][1], c[q][2]);
    t.x.push_back(0.25 * (n[0].c[0] + n[1].c[0] + n[2].c[0] + n[3].c[0]));
    t.y.push_back(0.25 * (n[0].c[1] + n[1].c[1] + n[2].c[1] + n[3].c[1]));
    t.z.push_back(0.25 * (n[0].c[2] + n[1].c[2] + n[2].c[2] + n[3].c[2]));
  }
  const int nw = (int)t.x.size();
  t.build(0, nw, 0);
  // every interior cell: the nearest centroid to the cell centre 1/8 (sum of the eight nodes, in Wall_Distance_Blk's
  // order); no wall: huge (kdtree_nearest on an empty tree)
  for (int b = 0; b < nb; ++b) {
    const BlockGeo& G = B[b];
    for (int k = 1; k <= G.n3; ++k)
      for (int j = 1; j <= G.n2; ++j)
        for (int i = 1; i <= G.n1; ++i) {
          const V3 q[8] = {G.nd(i, j, k), G.nd(i, j - 1, k), G.nd(i, j - 1, k - 1), G.nd(i, j, k - 1),
                           G.nd(i - 1, j, k), G.nd(i - 1, j - 1, k), G.nd(i - 1, j - 1, k - 1), G.nd(i - 1, j, k - 1)};
          double ctr[3];
          for (int a = 0; a < 3; ++a) {
            double sum = q[0].c[a];
            for (int v = 1; v < 8; ++v) sum = sum + q[v].c[a];
            ctr[a] = 1.0 / 8.0 * sum;
          }
          double best2 = std::numeric_limits<double>::max();
          if (nw > 0) t.nearest(0, nw, 0, ctr[0], ctr[1], ctr[2], best2);
          yn[G.row(i, j, k)] = nw > 0 ? std::sqrt(best2) : std::numeric_limits<double>::max();
        }
  }
  // Yn_Connection: the first ghost layer of a connection face takes the donor cell's distance
  for (int e = 0; e < nrec; ++e) {
    const mose_bc_record& r = rec[e];
    if (r.typ != MOSE_BC_CONNECTION) continue;
    if (r.b < 0 || r.b >= nb || r.bs < 0 || r.bs >= nb || r.f < 1 || r.f > 6) return MOSE_ERR_ARG;
    static const int guide[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    const int* gd = guide[r.f - 1];
    yn[B[r.b].row(r.i - gd[0], r.j - gd[1], r.k - gd[2])] = yn[B[r.bs].row(r.is, r.js, r.ks)];
  }
  return MOSE_OK;
}

// The unit tangent of each far-field face: the diagonal node(Fi,Fj,Fk) - node(one back along both transverse
// directions), over its norm (Setup_BC_Device, Mod_BC.f90:228-240).  The reference takes norm2 of the diagonal;
// here it is sqrt of the sum of squares in index order.
extern "C" int mose_records_tangent(int nb, const int* dims, const double* nodes, int nrec, mose_bc_record* rec) {
  if (nb <= 0 || !dims || !nodes || (nrec > 0 && !rec)) return MOSE_ERR_ARG;
  std::vector<BlockGeo> B(nb);
  int64_t NN = 0;
  for (int b = 0; b < nb; ++b) {
    int n1 = dims[3 * b], n2 = dims[3 * b + 1], n3 = dims[3 * b + 2];
    if (n1 < 1 || n2 < 1 || n3 < 1) return MOSE_ERR_ARG;
    B[b] = BlockGeo{n1, n2, n3, 0, 0, 0, 0, n1 + 2 * GC, n2 + 2 * GC, nodes + NN};
    NN += 3 * (int64_t)(n1 + 1) * (n2 + 1) * (n3 + 1);
  }
  for (int e = 0; e < nrec; ++e) {
    mose_bc_record& r = rec[e];
    if (r.typ != MOSE_BC_FARFIELD) continue;
    if (r.b < 0 || r.b >= nb || r.f < 1 || r.f > 6) return MOSE_ERR_ARG;
    const BlockGeo& G = B[r.b];
    int Fi = r.i, Fj = r.j, Fk = r.k;
    if (r.f <= 2) Fi = r.i - r.f % 2;
    else if (r.f <= 4) Fj = r.j - r.f % 2;
    else Fk = r.k - r.f % 2;
    V3 a = G.nd(Fi, Fj, Fk), c;
    if (r.f <= 2) c = G.nd(Fi, Fj - 1, Fk - 1);
    else if (r.f <= 4) c = G.nd(Fi - 1, Fj, Fk - 1);
    else c = G.nd(Fi - 1, Fj - 1, Fk);
    V3 t = sub(a, c);
    double tm = sqrt(t.c[0] * t.c[0] + t.c[1] * t.c[1] + t.c[2] * t.c[2]);
    if (!(tm > 0.0)) return MOSE_ERR_MESH;
    for (int q = 0; q < 3; ++q) r.tvec[q] = t.c[q] / tm;
  }
  return MOSE_OK;
}
