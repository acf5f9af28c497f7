// Register-resident search instantiations for FP8-e4m3 candidate lists (top-m, FP32 rescore).
#include "emh_search_kernels.h"

namespace emh {

hipError_t launch_reg_fp8(const SearchGeom &g, int M, int fr, const SearchArgs &a, hipStream_t s) {
    if (fr == 1) return launch_reg_m<OP_FP8, EMH_ROWS_F32, 1>(g, M, a, s);
    return launch_reg_m<OP_FP8, EMH_ROWS_F32>(g, M, a, s);
}

}  // namespace emh
