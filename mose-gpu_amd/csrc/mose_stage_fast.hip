// The fused stage kernel of the performance path (MOSE_PATH_FAST): the stage of mose_stage.hpp with the face fluxes
// of mose_fast.hpp, in a translation unit compiled with fused multiply-add contraction and reciprocal multiplication
// (the library's Makefile gives the flags).  The parity kernels (mose_capi.hip) keep the reference's roundings.
#include <hip/hip_runtime.h>

#include "mose_stage.hpp"

namespace mose {

template <int NS, int NR>
void launch_stage_fast(const CaseDev& D, const StageArgs& A, int rs, bool visc, unsigned grid, size_t lds,
                       hipStream_t s) {
  launch_stage_tiles<NS, NR, true>(D, A, rs, visc, grid, lds, s);
}

template void launch_stage_fast<1, 0>(const CaseDev&, const StageArgs&, int, bool, unsigned, size_t, hipStream_t);
template void launch_stage_fast<1, 1>(const CaseDev&, const StageArgs&, int, bool, unsigned, size_t, hipStream_t);
template void launch_stage_fast<4, 0>(const CaseDev&, const StageArgs&, int, bool, unsigned, size_t, hipStream_t);
template void launch_stage_fast<9, 0>(const CaseDev&, const StageArgs&, int, bool, unsigned, size_t, hipStream_t);

}  // namespace mose
