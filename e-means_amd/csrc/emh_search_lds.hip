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

template <int OP, int M>
static hipError_t launch_lds_jobs_m(const void *args, const uint2 *wg, size_t nwg, size_t lds, hipStream_t s) {
    auto kern = search_lds_jobs_kernel<OP, M>;
    static size_t attr_bytes = 0;
    if (lds > 64 * 1024 && lds > attr_bytes) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr_bytes = lds;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(256), lds, s, (const SearchArgs *)args, wg);
    return hipGetLastError();
}

hipError_t launch_jobs_search_lds_op(int op, int M, const void *args, const uint2 *wg, size_t nwg, size_t lds,
                                     hipStream_t s) {
    if (op == OP_F32 || op == OP_I8 || op == OP_I16) {
        if (M != 1) return hipErrorInvalidValue;
        if (op == OP_F32) return launch_lds_jobs_m<OP_F32, 1>(args, wg, nwg, lds, s);
        if (op == OP_I8) return launch_lds_jobs_m<OP_I8, 1>(args, wg, nwg, lds, s);
        return launch_lds_jobs_m<OP_I16, 1>(args, wg, nwg, lds, s);
    }
    if (op != OP_BF16 && op != OP_FP8) return hipErrorInvalidValue;
    switch (M) {
        case 1: return op == OP_BF16 ? launch_lds_jobs_m<OP_BF16, 1>(args, wg, nwg, lds, s)
                                     : launch_lds_jobs_m<OP_FP8, 1>(args, wg, nwg, lds, s);
        case 2: return op == OP_BF16 ? launch_lds_jobs_m<OP_BF16, 2>(args, wg, nwg, lds, s)
                                     : launch_lds_jobs_m<OP_FP8, 2>(args, wg, nwg, lds, s);
        case 4: return op == OP_BF16 ? launch_lds_jobs_m<OP_BF16, 4>(args, wg, nwg, lds, s)
                                     : launch_lds_jobs_m<OP_FP8, 4>(args, wg, nwg, lds, s);
        case 8: return op == OP_BF16 ? launch_lds_jobs_m<OP_BF16, 8>(args, wg, nwg, lds, s)
                                     : launch_lds_jobs_m<OP_FP8, 8>(args, wg, nwg, lds, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace emh
