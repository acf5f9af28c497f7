// The LDS search kernel's instantiations: rows resident in LDS, or streamed
// through it in d-chunks, for rows wider than the register-resident
// kernel takes (and the I8 search's i64 scores).
#include "emh_search_kernels.h"

namespace emh {

hipError_t launch_lds(const SearchGeom &g, int rows, int M, const SearchArgs &a, hipStream_t s) {
    switch (g.op) {
        case OP_F32:
            if (M != 1) return hipErrorInvalidValue;
            if (rows == EMH_ROWS_F32) return launch_res<OP_F32, EMH_ROWS_F32, 1>(g, a, s);
            if (rows == EMH_ROWS_I8) return launch_res<OP_F32, EMH_ROWS_I8, 1>(g, a, s);
            return launch_res<OP_F32, EMH_ROWS_U8, 1>(g, a, s);
        case OP_BF16:
            if (rows == EMH_ROWS_F32) return launch_m<OP_BF16, EMH_ROWS_F32>(g, M, a, s);
            if (rows == EMH_ROWS_I8) return launch_m<OP_BF16, EMH_ROWS_I8>(g, M, a, s);
            return launch_m<OP_BF16, EMH_ROWS_U8>(g, M, a, s);
        case OP_FP8:
            return launch_m<OP_FP8, EMH_ROWS_F32>(g, M, a, s);
        case OP_I8:
            if (M != 1 || rows != EMH_ROWS_I8) return hipErrorInvalidValue;
            return launch_res<OP_I8, EMH_ROWS_I8, 1>(g, a, s);
        case OP_I16:
            if (M != 1 || rows != EMH_ROWS_I8) return hipErrorInvalidValue;
            return launch_one<OP_I16, EMH_ROWS_I8, 1, false>(g, a, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace emh
