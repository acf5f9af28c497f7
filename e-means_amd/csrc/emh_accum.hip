// Device-side per-cluster counts and sums of a labelled batch.
//
// The reference folds each batch on the host, row by row in batch order:
// counts[l] += 1; sums[l] += row (f32), src/lib.rs:431-441, and keeps the
// last pass's totals for the closing mean, src/lib.rs:446-454, 481-489.
// Here the batch is grouped by label with a stable LSD radix sort (row
// positions keep their batch order inside each label), and the sorted rows
// are folded as one stream, one lane per dimension, each cluster's rows in
// that order, so every f32 sum is bit-identical to the host loop. The last
// pass's totals accumulate in FP64.
//
// The sort is this file's own (8-bit digits: block histograms, one-workgroup
// scan, stable scatter with wave ballots). It issues kernels only, no memset
// or allocation, so a captured graph can replay it.
#include <algorithm>

#include "emh_internal.h"

namespace emh {

__global__ void fill_u32_kernel(uint32_t *p, size_t count, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Zeroes three word ranges in one launch.
__global__ void zero3_kernel(uint32_t *a, size_t na, uint32_t *b, size_t nb, uint32_t *c, size_t nc) {
    const size_t total = na + nb + nc;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        if (i < na) a[i] = 0u;
        else if (i < na + nb) b[i - na] = 0u;
        else c[i - na - nb] = 0u;
    }
}

__global__ void fill_f64_kernel(double *p, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

template <int ROWS> __device__ inline float row_elem(const void *x, size_t idx) {
    if constexpr (ROWS == EMH_ROWS_F32) return ((const float *)x)[idx];
    else if constexpr (ROWS == EMH_ROWS_I8) return (float)((const int8_t *)x)[idx];
    else return (float)((const uint8_t *)x)[idx];
}

// The in-order fold of one cluster's n member rows into acc[c] (dimension
// tg + 64 c + lane, c < CC), by one wave: member row indices come 64 at a
// time (one per lane, src_of), and U rows' values are loaded ahead of their
// in-order adds. CC x U = 64 values per lane in flight: the fewer dimensions
// left, the more rows at once.
template <int ROWS, int CC, int U, class SrcOf>
__device__ __forceinline__ void fold_members(const void *x, size_t d, size_t tg, uint32_t n, int lane, SrcOf src_of,
                                             float (&acc)[4]) {
    for (uint32_t q0 = 0; q0 < n; q0 += 64) {
        const uint32_t cnt = n - q0 < 64 ? n - q0 : 64;
        uint32_t my_src = 0;  // row indices are below 2^32
        if ((uint32_t)lane < cnt) my_src = (uint32_t)src_of(q0 + lane);
        for (uint32_t u0 = 0; u0 < cnt; u0 += U) {
            float v[CC][U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                // the member's row start is uniform (scalar registers); each
                // lane adds its dimension
                const uint32_t src = (uint32_t)__builtin_amdgcn_readlane((int)my_src, (int)(u0 + u) & 63);
                const size_t row0 = (size_t)src * d + tg;
#pragma unroll
                for (int c = 0; c < CC; c++) {
                    const uint32_t t = 64u * c + (uint32_t)lane;
                    v[c][u] = (tg + t < d && u0 + u < cnt) ? row_elem<ROWS>(x, row0 + t) : 0.0f;
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (u0 + u < cnt)
#pragma unroll
                    for (int c = 0; c < CC; c++) acc[c] = __fadd_rn(acc[c], v[c][u]);
        }
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

// ---- stable LSD radix sort of (key, position) pairs, 8-bit digits ----------
constexpr int SORT_BLOCK = 1024;  // elements per workgroup (256 threads x 4 rounds)
// Batched sums: clusters with more members than this in one step are folded
// one wave per 64 dimensions (jobs_big_sum_kernel), the rest as a stream.
constexpr uint32_t SUM_BIG = 128;

// hist[digit * nblocks + block] = how many of the block's keys have the digit
__global__ __launch_bounds__(256) void digit_hist_kernel(const uint32_t *keys, size_t m, int shift, uint32_t *hist,
                                                         size_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t b0 = (size_t)blockIdx.x * SORT_BLOCK;
    for (int r = 0; r < SORT_BLOCK / 256; r++) {
        const size_t i = b0 + (size_t)r * 256 + threadIdx.x;
        if (i < m) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of the digit-major histogram (hist[digit * nblocks + block]):
// one launch of 256 workgroups scans each digit's row over the blocks and
// records the digit's total; the scatter adds the total of the smaller
// digits.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *sm, uint32_t *total) {
    // 256 threads: wave prefix by shuffles, then the 4 wave totals
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t inc = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)inc, off, 64);
        if (lane >= off) inc += o;
    }
    if (lane == 63) sm[w] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (int q = 0; q < w; q++) base += sm[q];
    *total = sm[0] + sm[1] + sm[2] + sm[3];
    __syncthreads();
    return base + inc - v;
}

__global__ __launch_bounds__(256) void digit_row_scan_kernel(uint32_t *hist, size_t nblocks, uint32_t *totals) {
    __shared__ uint32_t sm[4];
    uint32_t *row = hist + (size_t)blockIdx.x * nblocks;
    uint32_t carry = 0;
    for (size_t b0 = 0; b0 < nblocks; b0 += 256) {
        const size_t b = b0 + threadIdx.x;
        const uint32_t v = b < nblocks ? row[b] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, sm, &tot);
        if (b < nblocks) row[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}


// Stable scatter: an element's destination is the start of its digit in its
// block (scanned histogram) plus the number of equal digits before it in the
// block. Inside a wave the equal digits are found with one ballot per digit
// bit; waves and rounds are ordered through LDS counters, so the block keeps
// its input order.
// offs holds each digit's row scanned over the blocks (digit_row_scan_kernel);
// the digit's base (the total of the smaller digits) is added here from the
// digit totals.
// starts (or null): block 0 writes each digit's first sorted position.
__global__ __launch_bounds__(256) void digit_scatter_kernel(const uint32_t *keys, const uint32_t *vals, size_t m,
                                                            int shift, const uint32_t *offs, size_t nblocks,
                                                            const uint32_t *totals, uint32_t *keys_out,
                                                            uint32_t *vals_out, uint32_t *starts) {
    __shared__ uint32_t run[256];      // equal digits in the rounds before this one
    __shared__ uint32_t wcnt[4][256];  // per wave, this round
    __shared__ uint32_t sbase[256];    // total of the smaller digits
    __shared__ uint32_t ssm[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    {
        uint32_t all;
        sbase[tid] = block_exclusive_scan(totals[tid], ssm, &all);
        if (starts && blockIdx.x == 0) starts[tid] = sbase[tid];
    }
    run[tid] = 0;
    const size_t b0 = (size_t)blockIdx.x * SORT_BLOCK;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int r = 0; r < SORT_BLOCK / 256; r++) {
        for (int q = 0; q < 4; q++) wcnt[q][tid] = 0;
        __syncthreads();
        const size_t i = b0 + (size_t)r * 256 + tid;
        const bool live = i < m;
        const uint32_t key = live ? keys[i] : 0u;
        const uint32_t dg = (key >> shift) & 255u;
        uint64_t same = __ballot(live);
        for (int bit = 0; bit < 8; bit++) {
            const uint64_t bl = __ballot(live && ((dg >> bit) & 1u));
            same &= ((dg >> bit) & 1u) ? bl : ~bl;
        }
        const uint32_t before = (uint32_t)__popcll(same & lt);
        if (live && before == 0) wcnt[w][dg] = (uint32_t)__popcll(same);
        __syncthreads();
        if (live) {
            uint32_t pos = sbase[dg] + offs[(size_t)dg * nblocks + blockIdx.x] + run[dg] + before;
            for (int q = 0; q < w; q++) pos += wcnt[q][dg];
            keys_out[pos] = key;
            vals_out[pos] = vals ? vals[i] : (uint32_t)i;
        }
        __syncthreads();
        run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
    }
}

// The same sort in one pass over 11- or 12-bit digits (keys below BINS =
// 2048 or 4096), for the cluster labels of problems with k <= 4096: two or
// three launches instead of six.
constexpr int WIDE_BINS = 4096;  // the widest digit: scratch sizes

// BM: the histogram block-major (hist[block * BINS + digit], for the scatter
// that sums the earlier blocks' counts itself), else digit-major.
template <int BINS, bool BM>
__global__ __launch_bounds__(256) void digit_hist_wide_kernel(const uint32_t *keys, size_t m, uint32_t *hist,
                                                              size_t nblocks) {
    __shared__ uint32_t h[BINS];
    for (int i = threadIdx.x; i < BINS; i += 256) h[i] = 0;
    __syncthreads();
    const size_t b0 = (size_t)blockIdx.x * SORT_BLOCK;
    for (int r = 0; r < SORT_BLOCK / 256; r++) {
        const size_t i = b0 + (size_t)r * 256 + threadIdx.x;
        if (i < m) atomicAdd(&h[keys[i] & (BINS - 1)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < BINS; i += 256) {
        if constexpr (BM) hist[(size_t)blockIdx.x * BINS + i] = h[i];
        else hist[(size_t)i * nblocks + blockIdx.x] = h[i];
    }
}

// As digit_scatter_kernel; with BINS digits the per-wave counts (at most 64,
// one byte each) are not cleared wholesale each round: the leader of a digit
// group adds its group to the digit's running count and clears its own
// entry after the round.
// RAW: offs is the unscanned block-major histogram (digit_hist_wide_kernel)
// and each block sums the earlier blocks' counts of its digits itself (a few
// blocks: no scan launch), every load of it in flight at once; block 0 also
// writes the digit totals to totals_out.
// Otherwise offs and totals come from digit_row_scan_kernel.
template <bool RAW, int BINS>
__global__ __launch_bounds__(256) void digit_scatter_wide_kernel(const uint32_t *keys, const uint32_t *vals, size_t m,
                                                                 const uint32_t *offs, size_t nblocks,
                                                                 const uint32_t *totals, uint32_t *keys_out,
                                                                 uint32_t *vals_out, uint32_t *starts,
                                                                 uint32_t *totals_out) {
    constexpr int PER = BINS / 256, DB = BINS == 4096 ? 12 : 11;
    static_assert(BINS == 2048 || BINS == 4096, "digit width");
    static_assert(!RAW || BINS == 2048, "the block-prefix form keeps 2048 digits (LDS)");
    __shared__ uint32_t run[BINS];
    __shared__ uint8_t wcnt[4][BINS];
    __shared__ uint32_t sbase[BINS];
    __shared__ uint32_t pre[RAW ? BINS : 1];
    __shared__ uint32_t ssm[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    {
        // exclusive scan of the digit totals, PER consecutive digits a thread
        uint32_t loc[PER], tot[PER], tsum = 0;
        if constexpr (RAW) {
            // this thread's PER digits of every block: PER / 4 16-byte loads
            // a block, eight blocks' loads issued before any is added
            uint32_t pr[PER];
#pragma unroll
            for (int i = 0; i < PER; i++) pr[i] = tot[i] = 0u;
            const uint32_t *col = offs + (size_t)tid * PER;
            constexpr int BB = 8;
            for (size_t b0 = 0; b0 < nblocks; b0 += BB) {
                uint4 v[BB][PER / 4];
#pragma unroll
                for (int u = 0; u < BB; u++)
#pragma unroll
                    for (int q = 0; q < PER / 4; q++)
                        v[u][q] = b0 + u < nblocks ? *(const uint4 *)(col + (b0 + u) * BINS + 4 * q)
                                                   : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
                for (int u = 0; u < BB; u++) {
                    const bool before = b0 + u < blockIdx.x;
#pragma unroll
                    for (int q = 0; q < PER / 4; q++) {
                        const uint32_t c4[4] = {v[u][q].x, v[u][q].y, v[u][q].z, v[u][q].w};
#pragma unroll
                        for (int c = 0; c < 4; c++) {
                            tot[4 * q + c] += c4[c];
                            if (before) pr[4 * q + c] += c4[c];
                        }
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < PER; i++) {
                pre[tid * PER + i] = pr[i];
                if (blockIdx.x == 0) totals_out[tid * PER + i] = tot[i];
            }
        } else {
#pragma unroll
            for (int i = 0; i < PER; i++) tot[i] = totals[tid * PER + i];
        }
#pragma unroll
        for (int i = 0; i < PER; i++) {
            loc[i] = tsum;
            tsum += tot[i];
        }
        uint32_t all;
        const uint32_t base = block_exclusive_scan(tsum, ssm, &all);
#pragma unroll
        for (int i = 0; i < PER; i++) {
            sbase[tid * PER + i] = base + loc[i];
            if (starts && blockIdx.x == 0) starts[tid * PER + i] = base + loc[i];
            run[tid * PER + i] = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) wcnt[q][tid * PER + i] = 0;
        }
    }
    __syncthreads();
    const size_t b0 = (size_t)blockIdx.x * SORT_BLOCK;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int r = 0; r < SORT_BLOCK / 256; r++) {
        const size_t i = b0 + (size_t)r * 256 + tid;
        const bool live = i < m;
        const uint32_t key = live ? keys[i] : 0u;
        const uint32_t dg = key & (BINS - 1);
        uint64_t same = __ballot(live);
        for (int bit = 0; bit < DB; bit++) {
            const uint64_t bl = __ballot(live && ((dg >> bit) & 1u));
            same &= ((dg >> bit) & 1u) ? bl : ~bl;
        }
        const uint32_t before = (uint32_t)__popcll(same & lt), cnt = (uint32_t)__popcll(same);
        const bool leader = live && before == 0;
        if (leader) wcnt[w][dg] = (uint8_t)cnt;
        __syncthreads();
        if (live) {
            uint32_t pos = sbase[dg] + run[dg] + before;
            if constexpr (RAW) pos += pre[dg];
            else pos += offs[(size_t)dg * nblocks + blockIdx.x];
            for (int q = 0; q < w; q++) pos += wcnt[q][dg];
            keys_out[pos] = key;
            vals_out[pos] = vals ? vals[i] : (uint32_t)i;
        }
        __syncthreads();
        if (leader) {
            atomicAdd(&run[dg], cnt);
            wcnt[w][dg] = 0u;
        }
        __syncthreads();
    }
}

hipError_t accum_reserve(AccumScratch &a, size_t m, size_t k) {
    (void)k;  // the scratch depends on the batch only
    hipError_t e = hipSuccess;
    if (m > a.cap_rows) {
        hipFree(a.keys);
        hipFree(a.keys_alt);
        hipFree(a.vals);
        hipFree(a.vals_alt);
        hipFree(a.sort_tmp);
        a.keys = a.keys_alt = a.vals = a.vals_alt = nullptr;
        a.sort_tmp = nullptr;
        if ((e = hipMalloc(&a.keys, m * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&a.keys_alt, m * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&a.vals, m * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&a.vals_alt, m * 4)) != hipSuccess) return e;
        hipFree(a.big);
        a.big = nullptr;
        if ((e = hipMalloc(&a.big, (2 + 3 * (m / SUM_BIG + 1)) * 4)) != hipSuccess) return e;
        const size_t nblocks = (m + SORT_BLOCK - 1) / SORT_BLOCK;
        a.sort_tmp_bytes = WIDE_BINS * nblocks * 4 + WIDE_BINS * 4;  // histogram + digit totals
        if ((e = hipMalloc(&a.sort_tmp, a.sort_tmp_bytes)) != hipSuccess) return e;
        a.cap_rows = m;
    }
    if (!a.starts && (e = hipMalloc(&a.starts, WIDE_BINS * 4)) != hipSuccess) return e;
    return e;
}

void accum_free(AccumScratch &a) {
    hipFree(a.keys);
    hipFree(a.keys_alt);
    hipFree(a.vals);
    hipFree(a.vals_alt);
    hipFree(a.sort_tmp);
    hipFree(a.big);
    hipFree(a.starts);
    a = AccumScratch();
}

// vals_in: the value carried with each key (nullptr: its position).
// counts_out (or null): for a one-pass sort (bits <= 11) the number of each
// key (the pass's digit totals) and a.starts each key's first position.
hipError_t launch_sort_pairs(AccumScratch &a, const uint32_t *keys_in, const uint32_t *vals_in, size_t m, int bits,
                             const uint32_t **keys_out, const uint32_t **vals_out, hipStream_t s,
                             const uint32_t **counts_out = nullptr) {
    const size_t nblocks = (m + SORT_BLOCK - 1) / SORT_BLOCK;
    uint32_t *hist = (uint32_t *)a.sort_tmp;
    // ping-pong: pass 0 reads the input keys and values
    const uint32_t *kin = keys_in, *vin = vals_in;
    uint32_t *kb[2] = {a.keys, a.keys_alt}, *vb[2] = {a.vals, a.vals_alt};
    int cur = 0;
    if (bits > 8 && bits <= 11) {
        uint32_t *totals = hist + (size_t)2048 * nblocks;
        if (nblocks <= 64) {
            hipLaunchKernelGGL((digit_hist_wide_kernel<2048, true>), dim3((unsigned)nblocks), dim3(256), 0, s, kin, m,
                               hist, nblocks);
            // up to 64 blocks (64 Ki keys): each scatter block sums its digits'
            // counts over the blocks itself, two launches instead of three
            hipLaunchKernelGGL((digit_scatter_wide_kernel<true, 2048>), dim3((unsigned)nblocks), dim3(256), 0, s, kin,
                               vin, m, (const uint32_t *)hist, nblocks, (const uint32_t *)nullptr, kb[0], vb[0],
                               counts_out ? a.starts : nullptr, totals);
        } else {
            hipLaunchKernelGGL((digit_hist_wide_kernel<2048, false>), dim3((unsigned)nblocks), dim3(256), 0, s, kin, m,
                               hist, nblocks);
            hipLaunchKernelGGL(digit_row_scan_kernel, dim3(2048), dim3(256), 0, s, hist, nblocks, totals);
            hipLaunchKernelGGL((digit_scatter_wide_kernel<false, 2048>), dim3((unsigned)nblocks), dim3(256), 0, s, kin,
                               vin, m, (const uint32_t *)hist, nblocks, (const uint32_t *)totals, kb[0], vb[0],
                               counts_out ? a.starts : nullptr, (uint32_t *)nullptr);
        }
        *keys_out = kb[0];
        *vals_out = vb[0];
        if (counts_out) *counts_out = hist + (size_t)2048 * nblocks;
        return hipGetLastError();
    }
    if (bits == 12) {
        // 12-bit keys in one pass: histogram, the digits' rows scanned over
        // the blocks, scatter (three launches instead of two passes' six)
        uint32_t *totals = hist + (size_t)4096 * nblocks;
        hipLaunchKernelGGL((digit_hist_wide_kernel<4096, false>), dim3((unsigned)nblocks), dim3(256), 0, s, kin, m,
                           hist, nblocks);
        hipLaunchKernelGGL(digit_row_scan_kernel, dim3(4096), dim3(256), 0, s, hist, nblocks, totals);
        hipLaunchKernelGGL((digit_scatter_wide_kernel<false, 4096>), dim3((unsigned)nblocks), dim3(256), 0, s, kin,
                           vin, m, (const uint32_t *)hist, nblocks, (const uint32_t *)totals, kb[0], vb[0],
                           counts_out ? a.starts : nullptr, (uint32_t *)nullptr);
        *keys_out = kb[0];
        *vals_out = vb[0];
        if (counts_out) *counts_out = totals;
        return hipGetLastError();
    }
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(digit_hist_kernel, dim3((unsigned)nblocks), dim3(256), 0, s, kin, m, shift, hist, nblocks);
        hipLaunchKernelGGL(digit_row_scan_kernel, dim3(256), dim3(256), 0, s, hist, nblocks, hist + 256 * nblocks);
        const bool one_pass = counts_out && bits <= 8;
        hipLaunchKernelGGL(digit_scatter_kernel, dim3((unsigned)nblocks), dim3(256), 0, s, kin, vin, m, shift,
                           (const uint32_t *)hist, nblocks, (const uint32_t *)(hist + 256 * nblocks), kb[cur],
                           vb[cur], one_pass ? a.starts : nullptr);
        if (one_pass) *counts_out = hist + 256 * nblocks;
        kin = kb[cur];
        vin = vb[cur];
        cur ^= 1;
    }
    *keys_out = kin;
    *vals_out = vin;
    return hipGetLastError();
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

// ---- many jobs per launch ----

// key of batch slot i (slot_key); the first thread clears the big-cluster
// count for the sums.
__global__ void jobs_keys_kernel(const JobDev *jobs, const uint2 *rmap, size_t slots, uint32_t nclusters,
                                 uint32_t *keys, uint32_t *big_count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *big_count = 0;  // read by the sums after the sort
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += (size_t)gridDim.x * blockDim.x) {
        const uint2 r = rmap[i];
        keys[i] = slot_key(job_at<false>(jobs, r.x), r.y, nclusters);
    }
}

// ---- batched per-cluster sums: the sorted members as one stream ----------
//
// A step's clusters are small (a batch spread over k clusters), so one wave
// per cluster would spend its time on the prologue. Instead wave w owns the
// clusters whose first sorted member lies in [w SUM_SPAN, (w + 1) SUM_SPAN)
// and streams their members in sorted order: a cluster's members are its
// batch rows in batch order (stable sort), folded one after the other, and
// its sums written when the next cluster's first member comes up. Rows are
// loaded U at a time across cluster boundaries, lanes over dimensions
// (64 c + lane, c < CC).
constexpr uint32_t SUM_SPAN = 16;

// First position q in [p, lim) with keys[q] != prev (lim if none).
__device__ inline size_t first_change(const uint32_t *keys, size_t p, size_t lim, uint32_t prev, int lane) {
    for (size_t q0 = p; q0 < lim; q0 += 64) {
        const size_t q = q0 + lane;
        const uint64_t ch = __ballot(q < lim && keys[q] != prev);
        if (ch) return q0 + (size_t)(__ffsll((long long)ch) - 1);
    }
    return lim;
}

// One problem's view for the streamed fold: its rows, where its clusters'
// sums and counts go, and its global cluster range [base, end).
struct SumView {
    const void *x;             // n x d rows (f32, i8 or u8)
    size_t d;
    float *sums;               // (cluster - base) x d
    uint32_t *counts;          // (cluster - base)
    double *pass_sums;         // nullptr: no FP64 pass totals this step
    uint64_t *pass_counts;
    uint32_t base, end;
};

__device__ inline SumView job_view(const JobDev &j) {
    return SumView{j.x, j.d, j.sums, j.counts, j.last_epoch ? j.pass_sums : nullptr, j.pass_counts, j.cluster_base,
                   j.cluster_base + j.k};
}

// Element t of a row that starts at byte xr (uniform, scalar registers).
template <int ROWS> __device__ inline float row_at(const char *xr, uint32_t t) {
    if constexpr (ROWS == EMH_ROWS_F32) return *(const float *)(xr + 4u * t);
    else if constexpr (ROWS == EMH_ROWS_I8) return (float)*(const int8_t *)(xr + t);
    else return (float)*(const uint8_t *)(xr + t);
}

// Writes a finished cluster's dimensions [tg, tg + 64 CC): its f32 sums, and
// (the first group) its member count; with pass totals, those too.
template <int CC>
__device__ __forceinline__ void flush_cluster(const SumView &v, uint32_t key, size_t tg, const float (&acc)[CC],
                                              uint32_t cnt, int lane) {
    const size_t c = key - v.base, d = v.d;
#pragma unroll
    for (int cc = 0; cc < CC; cc++) {
        const size_t t = tg + 64 * cc + lane;
        if (t < d) {
            v.sums[c * d + t] = acc[cc];
            if (v.pass_sums) v.pass_sums[c * d + t] += (double)acc[cc];
        }
    }
    if (lane == 0 && tg == 0) {
        v.counts[c] = cnt;
        if (v.pass_sums) v.pass_counts[c] += cnt;
    }
}

// Streams the problem's clusters from position p (a cluster's first member)
// while they start before `stop` and belong to it, over dimensions
// [tg, tg + 64 CC); returns the position after the last member folded.
template <int ROWS, int CC, int U>
__device__ __forceinline__ size_t stream_clusters(const SumView &v, const uint32_t *keys, const uint32_t *src,
                                                  size_t slots, size_t p, size_t stop, size_t tg, int lane) {
    constexpr size_t ES = ROWS == EMH_ROWS_F32 ? 4 : 1;
    const size_t d = v.d;
    float acc[CC];
#pragma unroll
    for (int cc = 0; cc < CC; cc++) acc[cc] = 0.0f;
    uint32_t cur = keys[p], cnt = 0;
    for (;;) {
        // the next 64 positions: a member is taken when it continues the
        // running cluster, or starts one of the problem's clusters before
        // `stop` that also ends in this window (so it is not a big one; the
        // caller measures a cluster that runs past the window)
        const size_t q = p + lane;
        const uint32_t mk = q < slots ? keys[q] : 0xFFFFFFFFu;
        uint32_t before = (uint32_t)__shfl_up((int)mk, 1, 64);
        if (lane == 0) before = cur;
        const bool start = mk != before;
        const uint64_t starts = __ballot(start);
        const uint64_t later = lane < 63 ? starts >> (lane + 1) : 0ull;
        const bool take = !start || (q < stop && mk < v.end && later != 0ull);
        const uint64_t miss = __ballot(!take);
        const int ntake = miss ? __ffsll((long long)miss) - 1 : 64;
        const uint32_t ms = lane < ntake ? src[q] : 0u;
        for (int u0 = 0; u0 < ntake; u0 += U) {
            float val[CC][U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                // the member's row address is uniform: scalar registers
                // hold it, each lane adds its dimension
                const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)ms, (u0 + u) & 63);
                const char *xr = (const char *)v.x + ((size_t)r * d + tg) * ES;
#pragma unroll
                for (int cc = 0; cc < CC; cc++) {
                    const uint32_t t = 64u * cc + (uint32_t)lane;
                    val[cc][u] = (tg + t < d && u0 + u < ntake) ? row_at<ROWS>(xr, t) : 0.0f;
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (u0 + u < ntake) {
                    const uint32_t ku = (uint32_t)__builtin_amdgcn_readlane((int)mk, (u0 + u) & 63);
                    if (ku != cur) {
                        flush_cluster<CC>(v, cur, tg, acc, cnt, lane);
#pragma unroll
                        for (int cc = 0; cc < CC; cc++) acc[cc] = 0.0f;
                        cnt = 0;
                        cur = ku;
                    }
#pragma unroll
                    for (int cc = 0; cc < CC; cc++) acc[cc] = __fadd_rn(acc[cc], val[cc][u]);
                    cnt++;
                }
            }
        }
        p += (size_t)ntake;
        if (ntake < 64) break;
    }
    flush_cluster<CC>(v, cur, tg, acc, cnt, lane);
    return p;
}

// stream_clusters with CC picked from the dimensions left at tg
template <int ROWS>
__device__ __forceinline__ size_t stream_dims(const SumView &v, const uint32_t *keys, const uint32_t *src,
                                              size_t slots, size_t p, size_t stop, size_t tg, int lane) {
    const size_t left = v.d - tg;
    if (left > 128) return stream_clusters<ROWS, 4, 16>(v, keys, src, slots, p, stop, tg, lane);
    if (left > 64) return stream_clusters<ROWS, 2, 32>(v, keys, src, slots, p, stop, tg, lane);
    return stream_clusters<ROWS, 1, 64>(v, keys, src, slots, p, stop, tg, lane);
}

// A big cluster's entry in the list (after the count word): its key, first
// sorted position and members; its count is written here.
__device__ inline void list_big(uint32_t *big, uint32_t *count_slot, uint32_t key, size_t p, size_t n) {
    const uint32_t at = atomicAdd(&big[0], 1u);
    big[2 + 3 * at] = key;
    big[3 + 3 * at] = (uint32_t)p;
    big[4 + 3 * at] = (uint32_t)n;
    *count_slot = (uint32_t)n;
}

// keys / src: the sorted slots' global clusters (nclusters for spare slots,
// last) and job-relative rows. Each cluster's member count goes to counts
// (a global cluster's entry); big clusters are listed in big (after the
// count word) for jobs_big_sum_kernel. A run of clusters streamed together
// belongs to one job (stream_clusters stops at the job's last cluster), so
// its dimensions are taken 256 at a time over the same run.
template <int ROWS>
__global__ __launch_bounds__(256) void jobs_stream_sum_kernel(const JobDev *jobs, const uint32_t *cjob,
                                                              uint32_t nclusters, const uint32_t *keys,
                                                              const uint32_t *src, uint32_t *counts, size_t slots,
                                                              uint32_t *big) {
    const int lane = threadIdx.x & 63;
    size_t p = ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * SUM_SPAN;
    if (p >= slots) return;
    const size_t stop = p + SUM_SPAN < slots ? p + SUM_SPAN : slots;
    // the members before the first cluster that starts in this span belong
    // to an earlier wave
    if (p > 0) p = first_change(keys, p, stop, keys[p - 1], lane);
    while (p < stop) {
        const uint32_t key = keys[p];
        if (key >= nclusters) break;
        const size_t n = first_change(keys, p, slots, key, lane) - p;  // the cluster's members
        if (n > SUM_BIG) {
            if (lane == 0) list_big(big, &counts[key], key, p, n);
            p += n;
            continue;
        }
        const JobDev j = job_at<true>(jobs, cjob[key]);
        const SumView v = job_view(j);
        size_t next = p;
        for (size_t tg = 0; tg < v.d; tg += 256) next = stream_dims<ROWS>(v, keys, src, slots, p, stop, tg, lane);
        p = next;
    }
}

// The big clusters of this step: wave t of the grid-stride loop folds
// dimensions [64 (t % G), + 64) of listed cluster t / G (G = ceil(d / 64)
// of the widest job in the batched case, of the problem otherwise), its
// members in batch order, 64 rows in flight.
constexpr unsigned BIG_SUM_BLOCKS = 512;

template <int ROWS>
__device__ __forceinline__ void fold_big(const SumView &v, const uint32_t *src, uint32_t key, uint32_t p0, uint32_t n,
                                         size_t tg, int lane) {
    auto src_of = [&](uint32_t q) -> size_t { return (size_t)src[p0 + q]; };
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    fold_members<ROWS, 1, 64>(v.x, v.d, tg, n, lane, src_of, acc);
    const size_t c = key - v.base, dd = tg + lane;
    if (dd < v.d) {
        v.sums[c * v.d + dd] = acc[0];
        if (v.pass_sums) v.pass_sums[c * v.d + dd] += (double)acc[0];
    }
    if (lane == 0 && tg == 0 && v.pass_sums) v.pass_counts[c] += n;
}

template <int ROWS>
__global__ __launch_bounds__(256) void jobs_big_sum_kernel(const JobDev *jobs, const uint32_t *cjob,
                                                           const uint32_t *src, const uint32_t *big, uint32_t groups) {
    const int lane = threadIdx.x & 63;
    const size_t total = (size_t)big[0] * groups, waves = (size_t)gridDim.x * 4;
    for (size_t t = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < total; t += waves) {
        const size_t e = t / groups;
        const uint32_t key = big[2 + 3 * e], p0 = big[3 + 3 * e], n = big[4 + 3 * e];
        const size_t tg = (t - e * groups) * 64;
        const JobDev j = job_at<true>(jobs, cjob[key]);
        if (tg >= j.d) continue;
        fold_big<ROWS>(job_view(j), src, key, p0, n, tg, lane);
    }
}

// One problem (any row width): keys are the sorted labels (< k), src the
// rows; wave w streams the clusters starting in its span, 256 dimensions at
// a time; big clusters go to the list for big_sum_kernel.
template <int ROWS>
__global__ __launch_bounds__(256) void stream_sum_kernel(const void *x, size_t d, const uint32_t *keys,
                                                         const uint32_t *src, size_t m, size_t k, float *sums,
                                                         uint32_t *counts, uint32_t *big) {
    const int lane = threadIdx.x & 63;
    size_t p0 = ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * SUM_SPAN;
    if (p0 >= m) return;
    const size_t stop = p0 + SUM_SPAN < m ? p0 + SUM_SPAN : m;
    if (p0 > 0) p0 = first_change(keys, p0, stop, keys[p0 - 1], lane);
    const SumView v{x, d, sums, counts, nullptr, nullptr, 0u, (uint32_t)k};
    for (size_t tg = 0; tg < d; tg += 256) {
        size_t p = p0;
        while (p < stop) {
            const uint32_t key = keys[p];
            if (key >= k) break;
            const size_t n = first_change(keys, p, m, key, lane) - p;
            if (n > SUM_BIG) {
                if (lane == 0 && tg == 0) list_big(big, &counts[key], key, p, n);
                p += n;
                continue;
            }
            p = stream_dims<ROWS>(v, keys, src, m, p, stop, tg, lane);
        }
    }
}

// Integer rows (i8, u8), clusters of at most SUM_BIG members: every partial
// sum of the reference's f32 fold in batch order (src/lib.rs:431-441) is an
// integer of magnitude at most 128 x 255 < 2^24, so each f32 add is exact
// and the fold's result is the exact integer sum in any order. One wave
// adds a cluster's member rows in int32, 16 bytes a lane: lane (slot, chunk)
// = (lane >> L, lane & (2^L - 1)) takes chunk `chunk` of every 2^(6-L)-th
// member from member `slot` on (2^L = d / 16 chunks a row), the slots are
// summed by shuffles, and slot 0 writes the f32 sums; the cluster's count
// goes to counts.
template <int ROWS>
__device__ __forceinline__ void int_cluster_sum(const uint8_t *x, size_t d, int L, const uint32_t *src, size_t p,
                                                uint32_t n, uint32_t key, float *sums, uint32_t *counts, int lane) {
    const int nch = 1 << L, R = 64 >> L;
    const int slot = lane >> L, ch = lane & (nch - 1);
    int32_t acc[16];
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0;
    auto add = [&](const uint4 w) {
        const uint32_t v[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                if constexpr (ROWS == EMH_ROWS_I8) acc[4 * q + b] += (int32_t)(int8_t)(v[q] >> (8 * b));
                else acc[4 * q + b] += (int32_t)((v[q] >> (8 * b)) & 255u);
            }
    };
    // four members a lane in flight
    for (uint32_t it = 0; it < n; it += 4 * (uint32_t)R) {
        uint4 w[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t q = it + (uint32_t)(u * R + slot);
            w[u] = make_uint4(0u, 0u, 0u, 0u);
            if (q < n) w[u] = *(const uint4 *)(x + (size_t)src[p + q] * d + (size_t)ch * 16);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) add(w[u]);
    }
    for (int off = 32; off >= nch; off >>= 1)
#pragma unroll
        for (int e = 0; e < 16; e++) acc[e] += __shfl_xor(acc[e], off, 64);
    if (slot == 0) {
        float4 *o = (float4 *)(sums + (size_t)key * d + (size_t)ch * 16);
#pragma unroll
        for (int q = 0; q < 4; q++)
            o[q] = make_float4((float)acc[4 * q], (float)acc[4 * q + 1], (float)acc[4 * q + 2], (float)acc[4 * q + 3]);
    }
    if (lane == 0) counts[key] = n;
}

// stream_sum_kernel for integer rows whose d is 16 times a power of two (at
// most 1024) and whose rows are 16-byte aligned: the same clusters per wave
// (those whose first sorted member lies in the wave's SUM_SPAN positions),
// each summed by int_cluster_sum over all its dimensions at once; big
// clusters go to the list for big_sum_kernel.
template <int ROWS>
__global__ __launch_bounds__(256) void stream_sum_int_kernel(const void *x, size_t d, int L, const uint32_t *keys,
                                                             const uint32_t *src, size_t m, size_t k, float *sums,
                                                             uint32_t *counts, uint32_t *big) {
    const int lane = threadIdx.x & 63;
    size_t p = ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * SUM_SPAN;
    if (p >= m) return;
    const size_t stop = p + SUM_SPAN < m ? p + SUM_SPAN : m;
    if (p > 0) p = first_change(keys, p, stop, keys[p - 1], lane);
    while (p < stop) {
        const uint32_t key = keys[p];
        if (key >= k) break;
        const size_t n = first_change(keys, p, m, key, lane) - p;
        if (n > SUM_BIG) {
            if (lane == 0) list_big(big, &counts[key], key, p, n);
        } else {
            int_cluster_sum<ROWS>((const uint8_t *)x, d, L, src, p, (uint32_t)n, key, sums, counts, lane);
        }
        p += n;
    }
}

template <int ROWS>
__global__ __launch_bounds__(256) void big_sum_kernel(const void *x, size_t d, const uint32_t *src, size_t k,
                                                      float *sums, uint32_t *counts, const uint32_t *big) {
    const int lane = threadIdx.x & 63;
    const size_t groups = (d + 63) / 64;
    const size_t total = (size_t)big[0] * groups, waves = (size_t)gridDim.x * 4;
    const SumView v{x, d, sums, counts, nullptr, nullptr, 0u, (uint32_t)k};
    for (size_t t = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < total; t += waves) {
        const size_t e = t / groups;
        fold_big<ROWS>(v, src, big[2 + 3 * e], big[3 + 3 * e], big[4 + 3 * e], (t - e * groups) * 64, lane);
    }
}

// One problem whose labels sort in one pass (k <= 2048): the sort's key
// counts and first positions say where each cluster's members lie, so wave
// w folds dimensions [64 (w % G), + 64) of cluster w / G (G = ceil(d / 64))
// over its members in batch order, with no search for cluster boundaries;
// it writes the cluster's count (also 0) and, for clusters with members,
// its sums.
template <int ROWS>
__global__ __launch_bounds__(256) void cluster_sum_kernel(const void *x, size_t d, const uint32_t *src,
                                                          const uint32_t *starts, const uint32_t *nkey, size_t k,
                                                          uint32_t groups, float *sums, uint32_t *counts) {
    const int lane = threadIdx.x & 63;
    const size_t w = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const size_t c = w / groups;
    if (c >= k) return;
    const size_t tg = (w - c * groups) * 64;
    const uint32_t n = nkey[c], p0 = starts[c];
    if (tg == 0 && lane == 0) counts[c] = n;
    if (n == 0) return;
    auto src_of = [&](uint32_t q) -> size_t { return (size_t)src[p0 + q]; };
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    fold_members<ROWS, 1, 64>(x, d, tg, n, lane, src_of, acc);
    if (tg + lane < d) sums[c * d + tg + lane] = acc[0];
}

// Per-cluster counts and f32 sums of one batch, each cluster's rows folded
// in batch order (src/lib.rs:431-441): labels sorted stably, carrying their
// rows (ids[i], or i), then per cluster (k <= 2048) or the sorted stream;
// empty clusters keep zeros.
hipError_t launch_accumulate(AccumScratch &a, const void *x, int rows, size_t m, size_t d, const uint32_t *ids,
                             const uint32_t *labels, size_t k, float *sums, uint32_t *counts, hipStream_t s,
                             bool cleared) {
    hipError_t e;
    uint32_t *big_before = a.big;
    if ((e = accum_reserve(a, m > 0 ? m : 1, k)) != hipSuccess) return e;
    if (!cleared || a.big != big_before) {
        // counts, sums (+0.0f) and the big-cluster count start at zero
        const size_t zn = k + k * d + 1;
        hipLaunchKernelGGL(zero3_kernel, dim3((unsigned)std::min<size_t>((zn + 255) / 256, 4096)), dim3(256), 0, s,
                           counts, k, (uint32_t *)sums, k * d, a.big, (size_t)1);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (m == 0) return hipSuccess;
    const uint32_t *sorted_keys, *sorted_src, *nkey = nullptr;
    const int bits = key_bits(k);
    if ((e = launch_sort_pairs(a, labels, ids, m, bits, &sorted_keys, &sorted_src, s, bits <= 11 ? &nkey : nullptr)) !=
        hipSuccess)
        return e;
    if (nkey) {
        const uint32_t groups = (uint32_t)((d + 63) / 64);
        const dim3 cgrid((unsigned)((k * groups + 3) / 4));
        if (rows == EMH_ROWS_F32)
            hipLaunchKernelGGL(cluster_sum_kernel<EMH_ROWS_F32>, cgrid, dim3(256), 0, s, x, d, sorted_src, a.starts,
                               nkey, k, groups, sums, counts);
        else if (rows == EMH_ROWS_I8)
            hipLaunchKernelGGL(cluster_sum_kernel<EMH_ROWS_I8>, cgrid, dim3(256), 0, s, x, d, sorted_src, a.starts,
                               nkey, k, groups, sums, counts);
        else
            hipLaunchKernelGGL(cluster_sum_kernel<EMH_ROWS_U8>, cgrid, dim3(256), 0, s, x, d, sorted_src, a.starts,
                               nkey, k, groups, sums, counts);
        return hipGetLastError();
    }
    const dim3 grid((unsigned)(((m + SUM_SPAN - 1) / SUM_SPAN + 3) / 4)), bgrid(BIG_SUM_BLOCKS);
    // integer rows of 16 x 2^L bytes (L <= 6), 16-byte aligned: whole rows
    // at once, int32 sums
    int L = -1;
    if (rows != EMH_ROWS_F32 && d % 16 == 0 && ((uintptr_t)x & 15) == 0)
        for (int l = 0; l <= 6; l++)
            if (d == ((size_t)16 << l)) L = l;
    if (L >= 0) {
        if (rows == EMH_ROWS_I8) {
            hipLaunchKernelGGL(stream_sum_int_kernel<EMH_ROWS_I8>, grid, dim3(256), 0, s, x, d, L, sorted_keys,
                               sorted_src, m, k, sums, counts, a.big);
            hipLaunchKernelGGL(big_sum_kernel<EMH_ROWS_I8>, bgrid, dim3(256), 0, s, x, d, sorted_src, k, sums, counts,
                               (const uint32_t *)a.big);
        } else {
            hipLaunchKernelGGL(stream_sum_int_kernel<EMH_ROWS_U8>, grid, dim3(256), 0, s, x, d, L, sorted_keys,
                               sorted_src, m, k, sums, counts, a.big);
            hipLaunchKernelGGL(big_sum_kernel<EMH_ROWS_U8>, bgrid, dim3(256), 0, s, x, d, sorted_src, k, sums, counts,
                               (const uint32_t *)a.big);
        }
        return hipGetLastError();
    }
    switch (rows) {
        case EMH_ROWS_F32:
            hipLaunchKernelGGL(stream_sum_kernel<EMH_ROWS_F32>, grid, dim3(256), 0, s, x, d, sorted_keys, sorted_src, m,
                               k, sums, counts, a.big);
            hipLaunchKernelGGL(big_sum_kernel<EMH_ROWS_F32>, bgrid, dim3(256), 0, s, x, d, sorted_src, k, sums, counts,
                               (const uint32_t *)a.big);
            break;
        case EMH_ROWS_I8:
            hipLaunchKernelGGL(stream_sum_kernel<EMH_ROWS_I8>, grid, dim3(256), 0, s, x, d, sorted_keys, sorted_src, m, k,
                               sums, counts, a.big);
            hipLaunchKernelGGL(big_sum_kernel<EMH_ROWS_I8>, bgrid, dim3(256), 0, s, x, d, sorted_src, k, sums, counts,
                               (const uint32_t *)a.big);
            break;
        default:
            hipLaunchKernelGGL(stream_sum_kernel<EMH_ROWS_U8>, grid, dim3(256), 0, s, x, d, sorted_keys, sorted_src, m, k,
                               sums, counts, a.big);
            hipLaunchKernelGGL(big_sum_kernel<EMH_ROWS_U8>, bgrid, dim3(256), 0, s, x, d, sorted_src, k, sums, counts,
                               (const uint32_t *)a.big);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_jobs_keys(AccumScratch &a, const JobDev *jobs, const uint2 *rmap, size_t slots,
                            bool keys_written, size_t nclusters, uint32_t *keys, uint32_t *counts_all,
                            hipStream_t s) {
    hipError_t e;
    if ((e = accum_reserve(a, slots > 0 ? slots : 1, nclusters)) != hipSuccess) return e;
    (void)counts_all;  // zero here: set up so, and left so by jobs_ema_kernel
    if (slots == 0 || keys_written) return hipSuccess;
    unsigned blocks = (unsigned)std::min<size_t>((slots + 255) / 256, 4096);
    hipLaunchKernelGGL(jobs_keys_kernel, dim3(blocks), dim3(256), 0, s, jobs, rmap, slots, (uint32_t)nclusters, keys,
                       a.big);
    return hipGetLastError();
}

hipError_t launch_jobs_sums(AccumScratch &a, const JobDev *jobs, int rows, size_t d_max, size_t slots,
                            const uint32_t *ids, const uint32_t *cjob, size_t nclusters, const uint32_t *keys,
                            uint32_t *counts_all, hipStream_t s) {
    if (slots == 0) return hipSuccess;
    if (rows != EMH_ROWS_F32 && rows != EMH_ROWS_I8) return hipErrorInvalidValue;
    hipError_t e;
    // slot i's job-relative row is ids[i] (the jobs' id ranges are laid out
    // in slot order)
    const uint32_t *sorted_keys, *sorted_src;
    if ((e = launch_sort_pairs(a, keys, ids, slots, key_bits(nclusters + 1), &sorted_keys, &sorted_src, s)) !=
        hipSuccess)
        return e;
    const size_t waves = (slots + SUM_SPAN - 1) / SUM_SPAN;
    const dim3 grid((unsigned)((waves + 3) / 4)), bgrid(BIG_SUM_BLOCKS);
    const uint32_t groups = (uint32_t)((d_max + 63) / 64);
    if (rows == EMH_ROWS_F32) {
        hipLaunchKernelGGL(jobs_stream_sum_kernel<EMH_ROWS_F32>, grid, dim3(256), 0, s, jobs, cjob,
                           (uint32_t)nclusters, sorted_keys, sorted_src, counts_all, slots, a.big);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(jobs_big_sum_kernel<EMH_ROWS_F32>, bgrid, dim3(256), 0, s, jobs, cjob, sorted_src,
                           (const uint32_t *)a.big, groups);
    } else {
        hipLaunchKernelGGL(jobs_stream_sum_kernel<EMH_ROWS_I8>, grid, dim3(256), 0, s, jobs, cjob,
                           (uint32_t)nclusters, sorted_keys, sorted_src, counts_all, slots, a.big);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(jobs_big_sum_kernel<EMH_ROWS_I8>, bgrid, dim3(256), 0, s, jobs, cjob, sorted_src,
                           (const uint32_t *)a.big, groups);
    }
    return hipGetLastError();
}

// ---- the EMA step and the revival split on device-resident centroids ----

// N' = (1 - a) N + a n; c' = c + (a n / N') (s / n - c), each f32 operation
// rounded to nearest in the order of src/lib.rs:287-309, so the centroids are
// bit-identical to the host loop. One wave per cluster.
__device__ __forceinline__ void ema_cluster(float *cen, float *usage, const float *sums, const uint32_t *counts,
                                            size_t j, size_t d, float alpha, float decay, int replace, int lane) {
    float u = usage[j];
    if (!replace) u = __fmul_rn(u, decay);
    const uint32_t cnt = counts[j];
    if (cnt) {
        const float n = (float)cnt;
        const float retained = replace ? 0.0f : u;
        const float an = __fmul_rn(alpha, n);
        const float mass = __fadd_rn(retained, an);
        const float weight = __fdiv_rn(an, mass);
        float *c = cen + j * d;
        const float *sm = sums + j * d;
        for (size_t t = lane; t < d; t += 64) {
            const float mean = __fdiv_rn(sm[t], n);
            c[t] = weight == 1.0f ? mean : __fadd_rn(c[t], __fmul_rn(weight, __fsub_rn(mean, c[t])));
        }
        u = mass;
    }
    if (lane == 0) usage[j] = u;
}

// clear_counts (the counts array itself, or null): each count is zeroed once
// its wave has read it, and workgroup 0 zeroes *clear_word, so the next
// step's accumulate needs no clearing launch.
// pass_sums / pass_counts (the last epoch, else null): each cluster's f32
// sums and count are first added to the FP64 pass totals (what
// pass_accumulate_kernel does, before the EMA consumes them).
__global__ __launch_bounds__(256) void ema_apply_kernel(float *cen, float *usage, const float *sums,
                                                        const uint32_t *counts, size_t k, size_t d, float alpha,
                                                        float decay, int replace, uint32_t *clear_counts,
                                                        uint32_t *clear_word, double *pass_sums,
                                                        uint64_t *pass_counts) {
    if (clear_word && blockIdx.x == 0 && threadIdx.x == 0) *clear_word = 0u;
    const size_t j = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= k) return;
    const int lane = threadIdx.x & 63;
    if (pass_sums) {
        const uint32_t n = counts[j];
        if (n) {
            for (size_t t = lane; t < d; t += 64) pass_sums[j * d + t] += (double)sums[j * d + t];
            if (lane == 0) pass_counts[j] += n;
        }
    }
    ema_cluster(cen, usage, sums, counts, j, d, alpha, decay, replace, lane);
    if (clear_counts && (threadIdx.x & 63) == 0) clear_counts[j] = 0u;
}

// The step's last kernel also leaves the next step's accumulators at zero:
// each cluster's count once the EMA has read it, and (workgroup 0) the
// jobs' stats words, which the search has read.
__global__ __launch_bounds__(256) void jobs_ema_kernel(const JobDev *jobs, const uint32_t *cjob, size_t nclusters,
                                                       unsigned *stats, size_t nstats) {
    if (blockIdx.x == 0)
        for (size_t i = threadIdx.x; i < nstats; i += 256) stats[i] = 0u;
    const size_t g = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= nclusters) return;
    const JobDev j = job_at<true>(jobs, cjob[g]);
    if (j.m == 0) return;
    const int lane = threadIdx.x & 63;
    const size_t c = g - j.cluster_base;
    ema_cluster(j.cen, j.usage, j.sums, j.counts, c, j.d, j.alpha, j.decay, (int)j.replace, lane);
    if (lane == 0) j.counts[c] = 0u;
}

hipError_t launch_jobs_ema(const JobDev *jobs, const uint32_t *cjob, size_t nclusters, unsigned *stats, size_t nstats,
                           hipStream_t s) {
    hipLaunchKernelGGL(jobs_ema_kernel, dim3((unsigned)((nclusters + 3) / 4)), dim3(256), 0, s, jobs, cjob, nclusters,
                       stats, nstats);
    return hipGetLastError();
}

// Revival (src/lib.rs:330-340): donor row c -> c (1 - eps_i), copy ->
// c (1 + eps_i), eps_i = +1e-4 on even and -1e-4 on odd dimensions; if the
// two rows came out equal the copy's first value moves up one ulp. One wave
// per (donor, unused) pair; pairs never share a row.
// out[j] = row ids[j] of x widened to f32 (the initial centroids of a run
// whose rows are already on the device, src/lib.rs:374-381). One wave per row.
__global__ __launch_bounds__(256) void gather_rows_kernel(const void *x, int rows, size_t d, const uint32_t *ids,
                                                          size_t m, float *out) {
    const size_t j = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= m) return;
    const size_t src = (size_t)ids[j] * d;
    for (size_t i = threadIdx.x & 63; i < d; i += 64) {
        float v;
        if (rows == EMH_ROWS_I8) v = (float)((const int8_t *)x)[src + i];
        else if (rows == EMH_ROWS_U8) v = (float)((const uint8_t *)x)[src + i];
        else v = ((const float *)x)[src + i];
        out[j * d + i] = v;
    }
}

hipError_t launch_gather_rows(const void *x, int rows, size_t d, const uint32_t *ids, size_t m, float *out,
                              hipStream_t s) {
    if (!m) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, s, x, rows, d, ids, m, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(64) void revive_split_kernel(float *cen, const uint32_t *pairs, size_t npairs, size_t d,
                                                          float *usage) {
    const size_t t = blockIdx.x;
    const int lane = threadIdx.x;
    if (t >= npairs) return;
    const uint32_t src = pairs[2 * t], dst = pairs[2 * t + 1];
    // the pairs' clusters are distinct (a donor is used once, and is never
    // an unused cluster), so each pair's masses are its own
    if (usage && lane == 0) {
        const float half = __fdiv_rn(usage[src], 2.0f);
        usage[src] = half;
        usage[dst] = half;
    }
    float *cs = cen + (size_t)src * d, *cd = cen + (size_t)dst * d;
    const float lo = 1.0f - 1e-4f, hi = 1.0f + 1e-4f;
    bool same = true;
    for (size_t i = lane; i < d; i += 64) {
        const float c = cs[i];
        const float a = __fmul_rn(c, (i % 2 == 0) ? lo : hi), b = __fmul_rn(c, (i % 2 == 0) ? hi : lo);
        cs[i] = a;
        cd[i] = b;
        same = same && a == b;
    }
    const bool all_same = __ballot(!same) == 0;
    if (all_same && lane == 0) cd[0] = nextafterf(cd[0], INFINITY);
}

hipError_t launch_ema_apply(float *cen, float *usage, const float *sums, const uint32_t *counts, size_t k, size_t d,
                            float alpha, hipStream_t s, uint32_t *clear_counts, uint32_t *clear_word,
                            double *pass_sums, uint64_t *pass_counts) {
    const int replace = alpha == 1.0f;
    const float decay = 1.0f - alpha;
    hipLaunchKernelGGL(ema_apply_kernel, dim3((unsigned)((k + 3) / 4)), dim3(256), 0, s, cen, usage, sums, counts, k, d,
                       alpha, decay, replace, clear_counts, clear_word, pass_sums, pass_counts);
    return hipGetLastError();
}

// Revival splits of many jobs: triple t = (job, donor, unused).
__global__ __launch_bounds__(64) void jobs_revive_split_kernel(const JobDev *jobs, const uint32_t *triples,
                                                              size_t ntriples) {
    const size_t t = blockIdx.x;
    const int lane = threadIdx.x;
    if (t >= ntriples) return;
    const JobDev j = job_at<true>(jobs, triples[3 * t]);
    const size_t d = j.d;
    float *cs = j.cen + (size_t)triples[3 * t + 1] * d, *cd = j.cen + (size_t)triples[3 * t + 2] * d;
    const float lo = 1.0f - 1e-4f, hi = 1.0f + 1e-4f;
    bool same = true;
    for (size_t i = lane; i < d; i += 64) {
        const float c = cs[i];
        const float a = __fmul_rn(c, (i % 2 == 0) ? lo : hi), b = __fmul_rn(c, (i % 2 == 0) ? hi : lo);
        cs[i] = a;
        cd[i] = b;
        same = same && a == b;
    }
    const bool all_same = __ballot(!same) == 0;
    if (all_same && lane == 0) cd[0] = nextafterf(cd[0], INFINITY);
}

hipError_t launch_jobs_revive(const JobDev *jobs, const uint32_t *triples, size_t ntriples, hipStream_t s) {
    if (!ntriples) return hipSuccess;
    hipLaunchKernelGGL(jobs_revive_split_kernel, dim3((unsigned)ntriples), dim3(64), 0, s, jobs, triples, ntriples);
    return hipGetLastError();
}

hipError_t launch_revive_split(float *cen, const uint32_t *pairs, size_t npairs, size_t d, hipStream_t s,
                               float *usage) {
    if (!npairs) return hipSuccess;
    hipLaunchKernelGGL(revive_split_kernel, dim3((unsigned)npairs), dim3(64), 0, s, cen, pairs, npairs, d, usage);
    return hipGetLastError();
}

// ---- one problem's training step replayed as a graph ----
// The step's kernels take everything that changes from step to step from
// the device: ctr[0] is the step's index (advanced by the step's last
// kernel), ctr[1] a copy the stage kernel makes for the later kernels, and
// params[step] the step's batch position, EMA rate and flags.

// Copies the step's batch ids (its slice of the epoch order) to a fixed
// array the step's kernels read, one per thread.
__global__ __launch_bounds__(256) void step_stage_kernel(const uint32_t *order, const StepParam *params,
                                                         uint32_t *ctr, uint32_t *step_ids, size_t nb) {
    const uint32_t c = ctr[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[1] = c;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) step_ids[i] = order[params[c].b0 + i];
}

// A step that ran without the graph still takes its index.
__global__ void step_advance_kernel(uint32_t *ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctr[0] += 1u;
}

// The step's last kernel: the EMA step with the running step's rate and
// last-epoch flag (ema_apply_kernel's, one wave per cluster) and the step's
// labels (last epoch only) and distances into pinned host slot
// (step % nslot): labels at [0, nb), distances at [batch, batch + nb) of
// the slot, and before a revival check each cluster's new mass at
// [2 batch + j].
__global__ __launch_bounds__(256) void step_finish_kernel(float *cen, float *usage, const float *sums,
                                                          uint32_t *counts, size_t k, size_t d,
                                                          const StepParam *params, uint32_t *ctr,
                                                          uint32_t *clear_word, double *pass_sums,
                                                          uint64_t *pass_counts, const uint32_t *labels,
                                                          const float *dist, size_t nb, size_t batch,
                                                          StepSlots slots) {
    const uint32_t c = ctr[1];
    const StepParam p = params[c];
    uint32_t *dst = slots.p[c % (uint32_t)slots.n];
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[0] = c + 1u;  // the next step's index
    // the host reads a step's labels only in the last epoch (the last
    // pass's labels), its distances always (the objective)
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += stride) {
        if (p.is_last) dst[i] = labels[i];
        dst[batch + i] = __float_as_uint(dist[i]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *clear_word = 0u;
    const size_t j = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= k) return;
    const int lane = threadIdx.x & 63;
    if (p.is_last) {
        const uint32_t n = counts[j];
        if (n) {
            for (size_t t = lane; t < d; t += 64) pass_sums[j * d + t] += (double)sums[j * d + t];
            if (lane == 0) pass_counts[j] += n;
        }
    }
    ema_cluster(cen, usage, sums, counts, j, d, p.alpha, p.decay, (int)p.replace, lane);
    if (lane == 0) {
        counts[j] = 0u;
        if (p.revive) dst[2 * batch + j] = __float_as_uint(usage[j]);
    }
}

hipError_t launch_step_stage(const uint32_t *order, const StepParam *params, uint32_t *ctr, uint32_t *step_ids,
                             size_t nb, hipStream_t s) {
    hipLaunchKernelGGL(step_stage_kernel, dim3((unsigned)std::max<size_t>((nb + 255) / 256, 1)), dim3(256), 0, s,
                       order, params, ctr, step_ids, nb);
    return hipGetLastError();
}

hipError_t launch_step_advance(uint32_t *ctr, hipStream_t s) {
    hipLaunchKernelGGL(step_advance_kernel, dim3(1), dim3(64), 0, s, ctr);
    return hipGetLastError();
}

hipError_t launch_step_finish(float *cen, float *usage, const float *sums, uint32_t *counts, size_t k, size_t d,
                              const StepParam *params, uint32_t *ctr, uint32_t *clear_word, double *pass_sums,
                              uint64_t *pass_counts, const uint32_t *labels, const float *dist, size_t nb,
                              size_t batch, const StepSlots &slots, hipStream_t s) {
    if (slots.n < 1 || slots.n > 8) return hipErrorInvalidValue;
    const size_t blocks = std::max<size_t>((k + 3) / 4, std::min<size_t>((nb + 255) / 256, 256));
    hipLaunchKernelGGL(step_finish_kernel, dim3((unsigned)blocks), dim3(256), 0, s, cen, usage, sums, counts, k, d,
                       params, ctr, clear_word, pass_sums, pass_counts, labels, dist, nb, batch, slots);
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
