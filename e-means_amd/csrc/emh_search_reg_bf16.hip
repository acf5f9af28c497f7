// Register-resident search instantiations for BF16 candidate lists (top-m, FP32 rescore).
#include "emh_search_kernels.h"

namespace emh {

hipError_t launch_reg_bf16(const SearchGeom &g, int rows, int M, int fr, const SearchArgs &a, hipStream_t s) {
    if (fr == 1 && rows == EMH_ROWS_F32) return launch_reg_m<OP_BF16, EMH_ROWS_F32, 1>(g, M, a, s);
    if (rows == EMH_ROWS_F32) return launch_reg_m<OP_BF16, EMH_ROWS_F32>(g, M, a, s);
    if (rows == EMH_ROWS_I8) return launch_reg_m<OP_BF16, EMH_ROWS_I8>(g, M, a, s);
    return launch_reg_m<OP_BF16, EMH_ROWS_U8>(g, M, a, s);
}

}  // namespace emh
