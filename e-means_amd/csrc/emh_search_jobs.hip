// The batched search's instantiations: many independent problems in one
// launch (search_reg_jobs_kernel), FP32, BF16, FP8 and I8.
#include "emh_search_kernels.h"

namespace emh {

template <int OP, int M>
static hipError_t launch_jobs_search_m(const void *args, const uint2 *wg, size_t nwg, hipStream_t s) {
    hipLaunchKernelGGL((search_reg_jobs_kernel<OP, M>), dim3((unsigned)nwg), dim3(256), 0, s, (const SearchArgs *)args,
                       wg);
    return hipGetLastError();
}

hipError_t launch_jobs_search_op(int op, int M, const void *search_args, const uint2 *wg, size_t nwg, hipStream_t s) {
    if (op == OP_F32) {
        if (M != 1) return hipErrorInvalidValue;
        return launch_jobs_search_m<OP_F32, 1>(search_args, wg, nwg, s);
    }
    if (op == OP_I8) {
        if (M != 1) return hipErrorInvalidValue;
        return launch_jobs_search_m<OP_I8, 1>(search_args, wg, nwg, s);
    }
    if (op == OP_FP8) {
        switch (M) {
            case 1: return launch_jobs_search_m<OP_FP8, 1>(search_args, wg, nwg, s);
            case 2: return launch_jobs_search_m<OP_FP8, 2>(search_args, wg, nwg, s);
            case 4: return launch_jobs_search_m<OP_FP8, 4>(search_args, wg, nwg, s);
            case 8: return launch_jobs_search_m<OP_FP8, 8>(search_args, wg, nwg, s);
            default: return hipErrorInvalidValue;
        }
    }
    if (op != OP_BF16) return hipErrorInvalidValue;
    switch (M) {
        case 1: return launch_jobs_search_m<OP_BF16, 1>(search_args, wg, nwg, s);
        case 2: return launch_jobs_search_m<OP_BF16, 2>(search_args, wg, nwg, s);
        case 4: return launch_jobs_search_m<OP_BF16, 4>(search_args, wg, nwg, s);
        case 8: return launch_jobs_search_m<OP_BF16, 8>(search_args, wg, nwg, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace emh
