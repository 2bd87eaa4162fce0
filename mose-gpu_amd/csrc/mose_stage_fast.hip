// The fused stage kernel of the performance path (MOSE_PATH_FAST, mose_stage_fast.hpp), in a translation unit
// compiled with fused multiply-add contraction and reciprocal multiplication (the library's Makefile gives the flags).
// The parity kernels (mose_capi.hip) keep the reference's roundings.
#include <hip/hip_runtime.h>

#include "mose_stage_fast.hpp"

namespace mose {

template <int NS, int NR, int BS, int SH>
static void launch_fast_shape(const CaseDev& D, const StageArgs& A, int rs, bool visc, unsigned grid, size_t lds,
                              hipStream_t s) {
  if (visc) {
    if (rs == MOSE_RS_LMROE) k_stage_fast<NS, NR, true, MOSE_RS_LMROE, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else if (rs == MOSE_RS_AUSMP) k_stage_fast<NS, NR, true, MOSE_RS_AUSMP, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else if (rs == MOSE_RS_HLLE) k_stage_fast<NS, NR, true, MOSE_RS_HLLE, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else k_stage_fast<NS, NR, true, MOSE_RS_HLLC, BS, SH><<<grid, BS, lds, s>>>(D, A);
  } else {
    if (rs == MOSE_RS_LMROE) k_stage_fast<NS, NR, false, MOSE_RS_LMROE, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else if (rs == MOSE_RS_AUSMP) k_stage_fast<NS, NR, false, MOSE_RS_AUSMP, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else if (rs == MOSE_RS_HLLE) k_stage_fast<NS, NR, false, MOSE_RS_HLLE, BS, SH><<<grid, BS, lds, s>>>(D, A);
    else k_stage_fast<NS, NR, false, MOSE_RS_HLLC, BS, SH><<<grid, BS, lds, s>>>(D, A);
  }
}

// the shapes of launch_stage_tiles: 8 x 8 x 4 and 8 x 8 x 8 fixed at compile time (below nine species), the others
// read at run time; 128-cell tiles run 256 threads, the others 512
template <int NS, int NR>
void launch_stage_fast(const CaseDev& D, const StageArgs& A, int rs, bool visc, unsigned grid, size_t lds,
                       hipStream_t s) {
  const int TI = A.TI, TJ = A.TJ, TK = A.TK;
  if (TI * TJ * TK == 128) { launch_fast_shape<NS, NR, 256, 0>(D, A, rs, visc, grid, lds, s); return; }
  if constexpr (NS < 9) {
    if (TI == 8 && TJ == 8 && TK == 4) { launch_fast_shape<NS, NR, 512, 1>(D, A, rs, visc, grid, lds, s); return; }
    if (TI == 8 && TJ == 8 && TK == 8) { launch_fast_shape<NS, NR, 512, 2>(D, A, rs, visc, grid, lds, s); return; }
  }
  launch_fast_shape<NS, NR, 512, 0>(D, A, rs, visc, grid, lds, s);
}

template void launch_stage_fast<1, 0>(const CaseDev&, const StageArgs&, int, bool, unsigned, size_t, hipStream_t);
template void launch_stage_fast<1, 1>(const CaseDev&, const StageArgs&, int, bool, unsigned, size_t, hipStream_t);
template void launch_stage_fast<4, 0>(const CaseDev&, const StageArgs&, int, bool, unsigned, size_t, hipStream_t);
template void launch_stage_fast<9, 0>(const CaseDev&, const StageArgs&, int, bool, unsigned, size_t, hipStream_t);

}  // namespace mose
