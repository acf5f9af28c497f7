// Register-resident search instantiations for FP32 (the drop-in exact search).
#include "emh_search_kernels.h"

namespace emh {

hipError_t launch_reg_f32(const SearchGeom &g, int rows, int fr, const SearchArgs &a, hipStream_t s) {
    // 128-row workgroups for f32 rows only (the tuner's second shape)
    if (fr == 1 && rows == EMH_ROWS_F32) return launch_reg_ns<OP_F32, EMH_ROWS_F32, 1, 1>(g, a, s);
    if (rows == EMH_ROWS_F32) return launch_reg_ns<OP_F32, EMH_ROWS_F32, 1>(g, a, s);
    if (rows == EMH_ROWS_I8) return launch_reg_ns<OP_F32, EMH_ROWS_I8, 1>(g, a, s);
    return launch_reg_ns<OP_F32, EMH_ROWS_U8, 1>(g, a, s);
}

}  // namespace emh
