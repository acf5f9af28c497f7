// Device-side per-cluster counts and sums of a labelled batch.
//
// The reference folds each batch on the host, row by row in batch order:
// counts[l] += 1; sums[l] += row (f32), src/lib.rs:431-441, and keeps the
// last pass's totals for the closing mean, src/lib.rs:446-454, 481-489.
// Here the batch is grouped by label with a stable radix sort (row positions
// keep their batch order inside each label), and one lane per (cluster,
// dimension) folds that cluster's rows in that order, so every f32 sum is
// bit-identical to the host loop. The last pass's totals accumulate in FP64.
#include <rocprim/device/device_radix_sort.hpp>

#include "emh_internal.h"

namespace emh {

__global__ void fill_u32_kernel(uint32_t *p, size_t count, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void fill_f64_kernel(double *p, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void iota_hist_kernel(const uint32_t *labels, size_t m, size_t k, uint32_t *iota, uint32_t *counts) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
        iota[i] = (uint32_t)i;
        uint32_t l = labels[i];
        if (l < k) atomicAdd(&counts[l], 1u);
    }
}

// start[label] = first sorted position holding that label.
__global__ void segment_start_kernel(const uint32_t *sorted_labels, size_t m, size_t k, uint32_t *start) {
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < m; p += (size_t)gridDim.x * blockDim.x) {
        uint32_t l = sorted_labels[p];
        if (l < k && (p == 0 || sorted_labels[p - 1] != l)) start[l] = (uint32_t)p;
    }
}

template <int ROWS> __device__ inline float row_elem(const void *x, size_t idx) {
    if constexpr (ROWS == EMH_ROWS_F32) return ((const float *)x)[idx];
    else if constexpr (ROWS == EMH_ROWS_I8) return (float)((const int8_t *)x)[idx];
    else return (float)((const uint8_t *)x)[idx];
}

// One wave per cluster, lanes over dimensions; rows folded in batch order.
template <int ROWS>
__global__ __launch_bounds__(256) void segment_sum_kernel(const void *x, size_t d, const uint32_t *ids,
                                                          const uint32_t *sorted_pos, const uint32_t *start,
                                                          const uint32_t *counts, size_t k, float *sums) {
    const size_t jc = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (jc >= k) return;
    const uint32_t n = counts[jc];
    const uint32_t s0 = n ? start[jc] : 0;
    for (size_t t = lane; t < d; t += 64) {
        float acc = 0.0f;
        for (uint32_t q = 0; q < n; q++) {
            size_t r = sorted_pos[s0 + q];
            size_t src = ids ? (size_t)ids[r] : r;
            acc = __fadd_rn(acc, row_elem<ROWS>(x, src * d + t));
        }
        sums[jc * d + t] = acc;
    }
}

__global__ void pass_accumulate_kernel(const float *sums, const uint32_t *counts, size_t k, size_t d,
                                       double *pass_sums, uint64_t *pass_counts) {
    const size_t jc = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (jc >= k) return;
    const uint32_t n = counts[jc];
    if (n == 0) return;
    for (size_t t = lane; t < d; t += 64) pass_sums[jc * d + t] += (double)sums[jc * d + t];
    if (lane == 0) pass_counts[jc] += n;
}

// c_j = f32(total_j) * (1 / count_j as f32) for clusters the last pass used;
// with integer rows the f32 total is exact, as in src/lib.rs:485-487.
__global__ void closing_mean_kernel(const double *pass_sums, const uint64_t *pass_counts, size_t k, size_t d,
                                    float *centroids) {
    const size_t jc = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (jc >= k) return;
    const uint64_t n = pass_counts[jc];
    if (n == 0) return;
    const float inv = __fdiv_rn(1.0f, (float)n);
    for (size_t t = lane; t < d; t += 64) centroids[jc * d + t] = __fmul_rn((float)pass_sums[jc * d + t], inv);
}

static int key_bits(size_t k) {
    int b = 1;
    while (((size_t)1 << b) < k) b++;
    return b;
}

hipError_t accum_reserve(AccumScratch &a, size_t m, size_t k) {
    hipError_t e = hipSuccess;
    if (m > a.cap_rows) {
        hipFree(a.keys_alt);
        hipFree(a.vals);
        hipFree(a.vals_alt);
        hipFree(a.sort_tmp);
        a.keys_alt = a.vals = a.vals_alt = nullptr;
        a.sort_tmp = nullptr;
        if ((e = hipMalloc(&a.keys_alt, m * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&a.vals, m * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&a.vals_alt, m * 4)) != hipSuccess) return e;
        size_t bytes = 0;
        e = rocprim::radix_sort_pairs((void *)nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                      (const uint32_t *)nullptr, (uint32_t *)nullptr, m, 0, 32);
        if (e != hipSuccess) return e;
        a.sort_tmp_bytes = bytes;
        if ((e = hipMalloc(&a.sort_tmp, bytes ? bytes : 16)) != hipSuccess) return e;
        a.cap_rows = m;
    }
    if (k > a.cap_k) {
        hipFree(a.offsets);
        a.offsets = nullptr;
        if ((e = hipMalloc(&a.offsets, k * 4)) != hipSuccess) return e;
        a.cap_k = k;
    }
    return e;
}

void accum_free(AccumScratch &a) {
    hipFree(a.keys_alt);
    hipFree(a.vals);
    hipFree(a.vals_alt);
    hipFree(a.sort_tmp);
    hipFree(a.offsets);
    a = AccumScratch();
}

hipError_t launch_fill_u32(uint32_t *p, size_t count, uint32_t v, hipStream_t s) {
    if (!count) return hipSuccess;
    unsigned blocks = (unsigned)std::min<size_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, s, p, count, v);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, size_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    unsigned blocks = (unsigned)std::min<size_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_f64_kernel, dim3(blocks), dim3(256), 0, s, p, count);
    return hipGetLastError();
}

hipError_t launch_accumulate(AccumScratch &a, const void *x, int rows, size_t m, size_t d, const uint32_t *ids,
                             const uint32_t *labels, size_t k, float *sums, uint32_t *counts, hipStream_t s) {
    hipError_t e;
    if ((e = accum_reserve(a, m > 0 ? m : 1, k)) != hipSuccess) return e;
    if ((e = launch_fill_u32(counts, k, 0u, s)) != hipSuccess) return e;
    if (m > 0) {
        unsigned blocks = (unsigned)std::min<size_t>((m + 255) / 256, 4096);
        hipLaunchKernelGGL(iota_hist_kernel, dim3(blocks), dim3(256), 0, s, labels, m, k, a.vals, counts);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        size_t bytes = a.sort_tmp_bytes;
        e = rocprim::radix_sort_pairs(a.sort_tmp, bytes, labels, a.keys_alt, (const uint32_t *)a.vals, a.vals_alt, m,
                                      0, key_bits(k), s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(segment_start_kernel, dim3(blocks), dim3(256), 0, s, (const uint32_t *)a.keys_alt, m, k,
                           a.offsets);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    dim3 grid((unsigned)((k + 3) / 4));
    if (rows == EMH_ROWS_F32)
        hipLaunchKernelGGL(segment_sum_kernel<EMH_ROWS_F32>, grid, dim3(256), 0, s, x, d, ids, a.vals_alt, a.offsets, counts, k, sums);
    else if (rows == EMH_ROWS_I8)
        hipLaunchKernelGGL(segment_sum_kernel<EMH_ROWS_I8>, grid, dim3(256), 0, s, x, d, ids, a.vals_alt, a.offsets, counts, k, sums);
    else
        hipLaunchKernelGGL(segment_sum_kernel<EMH_ROWS_U8>, grid, dim3(256), 0, s, x, d, ids, a.vals_alt, a.offsets, counts, k, sums);
    return hipGetLastError();
}

hipError_t launch_pass_accumulate(const float *sums, const uint32_t *counts, size_t k, size_t d, double *pass_sums,
                                  uint64_t *pass_counts, hipStream_t s) {
    dim3 grid((unsigned)((k + 3) / 4));
    hipLaunchKernelGGL(pass_accumulate_kernel, grid, dim3(256), 0, s, sums, counts, k, d, pass_sums, pass_counts);
    return hipGetLastError();
}

hipError_t launch_closing_mean(const double *pass_sums, const uint64_t *pass_counts, size_t k, size_t d,
                               float *centroids, hipStream_t s) {
    dim3 grid((unsigned)((k + 3) / 4));
    hipLaunchKernelGGL(closing_mean_kernel, grid, dim3(256), 0, s, pass_sums, pass_counts, k, d, centroids);
    return hipGetLastError();
}

}  // namespace emh
