// Register-resident search instantiations for the exact integer searches (I8, I16).
#include "emh_search_kernels.h"

namespace emh {

hipError_t launch_reg_i8(const SearchGeom &g, int fr, const SearchArgs &a, hipStream_t s) {
    if (fr == 1) return launch_reg_ns<OP_I8, EMH_ROWS_I8, 1, 1>(g, a, s);
    return launch_reg_ns<OP_I8, EMH_ROWS_I8, 1>(g, a, s);
}

// I16 rows are multiples of 128 bytes (a tile of two planes is then a
// whole number of 4 KiB workgroup loads)
hipError_t launch_reg_i16(const SearchGeom &g, const SearchArgs &a, hipStream_t s) {
    switch (g.row_bytes / 32) {
        case 4: return launch_reg<OP_I16, EMH_ROWS_I8, 1, 4>(a, s);
        case 8: return launch_reg<OP_I16, EMH_ROWS_I8, 1, 8>(a, s);
        case 12: return launch_reg<OP_I16, EMH_ROWS_I8, 1, 12>(a, s);
        case 16: return launch_reg<OP_I16, EMH_ROWS_I8, 1, 16>(a, s);
        default: return hipErrorInvalidValue;
    }
}


}  // namespace emh
