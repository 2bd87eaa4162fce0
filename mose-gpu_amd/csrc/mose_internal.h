// Internal definitions shared by the host side of the MI355X library (not part of the C ABI).
#ifndef MOSE_INTERNAL_H
#define MOSE_INTERNAL_H

#include <cstdint>

#include "mose_gpu.h"

#define MOSE_NTMAX 10000      // table rows 0..10000 (Mod_Convective.f90:36 NTMAX)
#define MOSE_MAX_PRIM 24
#define MOSE_RUNIV 8314.46261815324   // FLINT Runiv (Lib_ThermoTransport.f90:13)

// Per-block constants as the kernels see them.
struct mose_blk {
  int n1, n2, n3, pi, pj;
  int pad;
  int64_t boff;          // first arena row
  int64_t f1off, f2off, f3off;   // face-geometry offsets
  int64_t ifoff[3];      // interior-face buffer offsets (per direction)
  int64_t bpoff[6];      // boundary-face plane offsets (per face 1..6)
  int64_t cell0;         // first interior cell of the block in the flat interior numbering
};

// Thermo tables resident on the device
struct mose_dtab {
  const double* cp;      // [ns][MOSE_NTAB]
  const double* h;
  const double* mi;
  const double* k;
  double wm[MOSE_MAX_SPECIES];
  double Ri[MOSE_MAX_SPECIES];
  double mm025[MOSE_MAX_SPECIES][MOSE_MAX_SPECIES];
  double is8[MOSE_MAX_SPECIES][MOSE_MAX_SPECIES];
};

#endif
