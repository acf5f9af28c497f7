// The report's nearest-centroid step on the device (src/report.rs:169-183).
//
// For every centroid i: the live centroid j != i (cluster size > 0) with the
// smallest sqdist (src/report.rs:104-114), the lowest j among equal
// distances, and sqrt of that distance; (i, +inf) when there is none. The
// distance is computed in the reference's own f32 order, so indices and
// distances are bit-identical to it: sixteen running sums, lane i taking
// elements i, i+16, ..., folded left from 0, plus the remainder past the last
// full 16-chunk folded left on its own and added last. That order is not a
// dot-product expansion, so the step runs on the vector ALUs, not MFMA.
//
// One workgroup per pair of 32-centroid tiles (I <= J): sqdist is symmetric
// bit for bit ((a - b)^2 == (b - a)^2 in IEEE), so each pair is computed once
// and offered to both rows. The 256 lanes form 16 groups of 16; lane l of a
// group owns running sum l of an 8 x 8 block of pairs. The tiles are staged in
// LDS 128 elements at a time; afterwards the sixteen sums of each pair are
// transposed through LDS and folded in order by one lane. A candidate is the
// key (distance bits << 32 | index): distances are >= +0, so unsigned order of
// the key is (distance, index) order, and an atomic minimum per row gives the
// reference's strict-< scan result whatever order the tiles finish in.
#include <math.h>

#include <algorithm>

#include "emh_assigner.h"

namespace emh {

constexpr int NT = 32;          // centroids per tile side
constexpr int NSTAGE = 128;     // elements staged per pass (eight 16-chunks)
constexpr int NSTRIDE = 132;    // padded LDS row stride (floats)
constexpr int RED_STRIDE = 17;  // padded stride of one pair's sixteen sums
constexpr int LDS_FLOATS = std::max(2 * NT * NSTRIDE, 512 * RED_STRIDE);

constexpr uint64_t KEY_NONE = (uint64_t)0x7F800000u << 32;  // +inf

__global__ __launch_bounds__(256) void nearest_init_kernel(uint64_t *keys, size_t k) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < k; i += (size_t)gridDim.x * 256)
        keys[i] = KEY_NONE | i;
}

__global__ __launch_bounds__(256) void nearest_pairs_kernel(const float *__restrict__ c, size_t k, size_t d,
                                                            const uint32_t *__restrict__ sizes,
                                                            unsigned long long *keys) {
    const unsigned I = blockIdx.y, J = blockIdx.x;
    if (J < I) return;
    __shared__ float lds[LDS_FLOATS];
    __shared__ unsigned long long row_min[NT], col_min[NT];
    const int tid = threadIdx.x, g = tid >> 4, l = tid & 15, gr = g >> 2, gc = g & 3;
    const size_t i0 = (size_t)I * NT, j0 = (size_t)J * NT;
    const size_t full = d / 16 * 16;

    float acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
        for (int b = 0; b < 8; b++) acc[a][b] = 0.0f;

    float *R = lds, *Cc = lds + NT * NSTRIDE;
    for (size_t s0 = 0; s0 < full; s0 += NSTAGE) {
        const int len = (int)std::min<size_t>(NSTAGE, full - s0);
        for (int e = tid; e < NT * NSTAGE; e += 256) {
            const int r = e / NSTAGE, q = e % NSTAGE;
            float vr = 0.0f, vc = 0.0f;
            if (q < len) {
                if (i0 + r < k) vr = c[(i0 + r) * d + s0 + q];
                if (j0 + r < k) vc = c[(j0 + r) * d + s0 + q];
            }
            R[r * NSTRIDE + q] = vr;
            Cc[r * NSTRIDE + q] = vc;
        }
        __syncthreads();
        for (int ch = 0; ch < len; ch += 16) {
            float xr[8], yc[8];
#pragma unroll
            for (int a = 0; a < 8; a++) xr[a] = R[(gr * 8 + a) * NSTRIDE + ch + l];
#pragma unroll
            for (int b = 0; b < 8; b++) yc[b] = Cc[(gc * 8 + b) * NSTRIDE + ch + l];
#pragma unroll
            for (int a = 0; a < 8; a++)
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const float t = __fsub_rn(xr[a], yc[b]);
                    acc[a][b] = __fadd_rn(acc[a][b], __fmul_rn(t, t));
                }
        }
        __syncthreads();
    }

    if (tid < NT) {
        row_min[tid] = ~0ull;
        col_min[tid] = ~0ull;
    }
    float *red = lds;  // the staging buffer is free once the last pass is done
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 8; b++) red[(g * 32 + a * 8 + b) * RED_STRIDE + l] = acc[h * 4 + a][b];
        __syncthreads();
        for (int p = tid; p < 512; p += 256) {
            const int pg = p >> 5, a = (p >> 3) & 3, b = p & 7;
            const int li = (pg >> 2) * 8 + h * 4 + a, lj = (pg & 3) * 8 + b;
            const size_t i = i0 + li, j = j0 + lj;
            if (i >= k || j >= k || (I == J && li >= lj)) continue;
            float s = 0.0f;
#pragma unroll
            for (int q = 0; q < 16; q++) s = __fadd_rn(s, red[p * RED_STRIDE + q]);
            float tail = 0.0f;
            for (size_t e = full; e < d; e++) {
                const float t = __fsub_rn(c[i * d + e], c[j * d + e]);
                tail = __fadd_rn(tail, __fmul_rn(t, t));
            }
            s = __fadd_rn(s, tail);
            if (!(s < INFINITY)) continue;  // NaN and +inf never replace the starting (i, +inf)
            const unsigned long long hi = (unsigned long long)__float_as_uint(s) << 32;
            if (!sizes || sizes[j]) atomicMin(&row_min[li], hi | j);
            if (!sizes || sizes[i]) atomicMin(&col_min[lj], hi | i);
        }
        __syncthreads();
    }
    if (tid < NT) {
        if (i0 + tid < k && row_min[tid] != ~0ull) atomicMin(&keys[i0 + tid], row_min[tid]);
    } else if (tid < 2 * NT) {
        const int t = tid - NT;
        if (j0 + t < k && col_min[t] != ~0ull) atomicMin(&keys[j0 + t], col_min[t]);
    }
}

__global__ __launch_bounds__(256) void nearest_finish_kernel(const uint64_t *keys, size_t k, uint32_t *nearest,
                                                             float *dist) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < k; i += (size_t)gridDim.x * 256) {
        const uint64_t key = keys[i];
        nearest[i] = (uint32_t)key;
        // sqrtf is the correctly rounded square root under the library's build
        // flags (__fsqrt_rn is the hardware approximation unless rounded
        // operations are requested), as f32::sqrt
        dist[i] = sqrtf(__uint_as_float((uint32_t)(key >> 32)));
    }
}

static int nearest_check(uint64_t k, uint64_t d) {
    if (k == 0 || d == 0) return fail(EMH_E_ARG, "report_nearest: need k > 0 centroids of d > 0 dimensions");
    if (k > 0xFFFFFFFFull) return fail(EMH_E_ARG, "report_nearest: k must fit u32 labels");
    if ((k + NT - 1) / NT > 65535) return fail(EMH_E_ARG, "report_nearest: k above 2,097,120 centroids");
    return EMH_OK;
}

// keys: k u64 of workspace
static int launch_nearest(const float *c, size_t k, size_t d, const uint32_t *sizes, uint64_t *keys,
                          uint32_t *nearest, float *dist, hipStream_t s) {
    const unsigned lin = (unsigned)std::min<size_t>((k + 255) / 256, 4096);
    const unsigned nt = (unsigned)((k + NT - 1) / NT);
    hipLaunchKernelGGL(nearest_init_kernel, dim3(lin), dim3(256), 0, s, keys, k);
    hipLaunchKernelGGL(nearest_pairs_kernel, dim3(nt, nt), dim3(256), 0, s, c, k, d, sizes,
                       (unsigned long long *)keys);
    hipLaunchKernelGGL(nearest_finish_kernel, dim3(lin), dim3(256), 0, s, (const uint64_t *)keys, k, nearest, dist);
    return check(hipGetLastError(), "report_nearest");
}

}  // namespace emh

using namespace emh;

extern "C" {

int emh_report_nearest(int device, const float *centroids, uint64_t k, uint64_t d, const uint32_t *sizes,
                       uint32_t *nearest, float *nearest_dist) {
    int rc;
    if (!centroids || !nearest || !nearest_dist) return fail(EMH_E_ARG, "report_nearest: NULL argument");
    if ((rc = nearest_check(k, d))) return rc;
    return with_device(device, [&](hipStream_t s) {
        DevBuf cd, sd, kd, nd, dd;
        if ((rc = check(cd.ensure(k * d * 4), "centroids")) || (rc = check(kd.ensure(k * 8), "keys")) ||
            (rc = check(nd.ensure(k * 4), "nearest")) || (rc = check(dd.ensure(k * 4), "distances")) ||
            (sizes && (rc = check(sd.ensure(k * 4), "sizes"))) ||
            (rc = check(hipMemcpyAsync(cd.p, centroids, k * d * 4, hipMemcpyHostToDevice, s), "upload centroids")) ||
            (sizes && (rc = check(hipMemcpyAsync(sd.p, sizes, k * 4, hipMemcpyHostToDevice, s), "upload sizes"))) ||
            (rc = launch_nearest(cd.as<float>(), k, d, sizes ? sd.as<uint32_t>() : nullptr, kd.as<uint64_t>(),
                                 nd.as<uint32_t>(), dd.as<float>(), s)) ||
            (rc = check(hipMemcpyAsync(nearest, nd.p, k * 4, hipMemcpyDeviceToHost, s), "download nearest")) ||
            (rc = check(hipMemcpyAsync(nearest_dist, dd.p, k * 4, hipMemcpyDeviceToHost, s), "download distances")) ||
            (rc = check(hipStreamSynchronize(s), "report_nearest")))
            return rc;
        return (int)EMH_OK;
    });
}

int emh_report_nearest_device(const float *centroids_dev, uint64_t k, uint64_t d, const uint32_t *sizes_dev,
                              uint64_t *keys_dev, uint32_t *nearest_dev, float *dist_dev, void *stream) {
    int rc;
    if (!centroids_dev || !keys_dev || !nearest_dev || !dist_dev)
        return fail(EMH_E_ARG, "report_nearest_device: NULL argument");
    if ((rc = nearest_check(k, d))) return rc;
    return launch_nearest(centroids_dev, k, d, sizes_dev, keys_dev, nearest_dev, dist_dev, (hipStream_t)stream);
}

}  // extern "C"
