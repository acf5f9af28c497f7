// Calibrated f32 -> i8 codec on the device (src/quantize.rs:43-96).
//
// fit: per-dimension shift = the sample mean, summed in f64 row by row in
// row order (one thread per dimension, so the f64 bits are the reference's),
// and one global scale from the quantile's order statistic of |x - shift|
// under f32::total_cmp, found exactly by a radix select over total-order
// keys (a max-reduction when the quantile is 1, faiss's calibration). The
// scalar tail (target / peak, power-of-two rounding) runs on the host with
// the same libm calls as the reference. encode / decode / clipped_fraction
// are elementwise and HBM-bound: encode reads 4 bytes and writes 1 per value.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "emh_assigner.h"

namespace emh {

// f32::total_cmp order as an unsigned key
__device__ __host__ inline uint32_t total_key(uint32_t u) { return (u & 0x80000000u) ? ~u : (u | 0x80000000u); }
__host__ inline uint32_t total_key_inverse(uint32_t k) { return (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k; }

// mean_i = sum_r x[r][i] in f64, rows in order; 8 rows of loads ahead of the adds
__global__ __launch_bounds__(64) void col_mean_kernel(const float *x, size_t n, size_t d, float *shift) {
    const size_t t = (size_t)blockIdx.x * 64 + threadIdx.x;
    if (t >= d) return;
    double acc = 0.0;
    size_t r = 0;
    for (; r + 8 <= n; r += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = x[(r + u) * d + t];
#pragma unroll
        for (int u = 0; u < 8; u++) acc = __dadd_rn(acc, (double)v[u]);
    }
    for (; r < n; r++) acc = __dadd_rn(acc, (double)x[r * d + t]);
    shift[t] = (float)__ddiv_rn(acc, (double)n);
}

// The same sums with the loads taken off the add chain: a workgroup owns 32
// dimensions, all 256 lanes stage MEAN_TR-row tiles through LDS (16-byte
// loads, double buffered, the next tile's loads in flight during the adds),
// and lanes 0..31 of wave 0 add the staged rows in row order, so each
// dimension's f64 chain is the one above, bit for bit. Needs d % 4 == 0 and a
// 16-byte aligned x.
constexpr int MEAN_TR = 256;
constexpr int MEAN_DIMS = 32;
__global__ __launch_bounds__(256) void col_mean_tiled_kernel(const float4 *x, size_t n, size_t d, float *shift) {
    __shared__ float4 tile[2][MEAN_TR][MEAN_DIMS / 4];
    const size_t dq = d / 4, j0 = (size_t)blockIdx.x * MEAN_DIMS;
    const unsigned quad = threadIdx.x % 8, row_in = threadIdx.x / 8;  // 32 rows per pass, 8 passes
    const bool quad_ok = j0 + 4 * quad < d;
    const size_t ntiles = (n + MEAN_TR - 1) / MEAN_TR;
    float4 v[MEAN_TR / 32];
    auto load = [&](size_t r0) {
#pragma unroll
        for (int u = 0; u < MEAN_TR / 32; u++) {
            const size_t r = r0 + row_in + 32 * u;
            v[u] = quad_ok && r < n ? x[r * dq + j0 / 4 + quad] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int u = 0; u < MEAN_TR / 32; u++) tile[b][row_in + 32 * u][quad] = v[u];
    };
    load(0);
    store(0);
    __syncthreads();
    double acc = 0.0;
    const float *lane_col = (const float *)&tile[0][0][0] + threadIdx.x;
    for (size_t t = 0; t < ntiles; t++) {
        const int b = (int)(t & 1);
        if (t + 1 < ntiles) load((t + 1) * MEAN_TR);
        if (threadIdx.x < MEAN_DIMS) {
            const size_t rows = std::min<size_t>(MEAN_TR, n - t * MEAN_TR);
            const float *col = lane_col + (size_t)b * MEAN_TR * MEAN_DIMS;
#pragma unroll 8
            for (size_t i = 0; i < rows; i++) acc = __dadd_rn(acc, (double)col[i * MEAN_DIMS]);
        }
        if (t + 1 < ntiles) store(b ^ 1);
        __syncthreads();
    }
    if (threadIdx.x < MEAN_DIMS && j0 + threadIdx.x < d) shift[j0 + threadIdx.x] = (float)__ddiv_rn(acc, (double)n);
}

// Elementwise kernels walk rows by workgroup and dimensions by lane.
#define EMH_FOR_ELEMS(n, d)                                                        \
    for (size_t r = blockIdx.x; r < (n); r += gridDim.x)                          \
        for (size_t t = threadIdx.x, i = r * (d) + t; t < (d); t += blockDim.x, i += blockDim.x)

// largest total-order key of |x - shift|
__global__ __launch_bounds__(256) void mag_max_kernel(const float *x, size_t n, size_t d, const float *shift,
                                                      unsigned *out) {
    uint32_t best = 0;
    EMH_FOR_ELEMS(n, d) best = max(best, total_key(__float_as_uint(fabsf(__fsub_rn(x[i], shift[t])))));
    for (int off = 32; off > 0; off >>= 1) best = max(best, (uint32_t)__shfl_xor((int)best, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

// histogram of key digit (key >> bit) & 255 among keys whose bits above `bit`
// equal `prefix`
__global__ __launch_bounds__(256) void mag_hist_kernel(const float *x, size_t n, size_t d, const float *shift,
                                                       uint32_t prefix, int bit, unsigned long long *hist) {
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t hi_mask = bit >= 24 ? 0u : (0xFFFFFFFFu << (bit + 8));
    EMH_FOR_ELEMS(n, d) {
        const uint32_t k = total_key(__float_as_uint(fabsf(__fsub_rn(x[i], shift[t]))));
        if ((k & hi_mask) == (prefix & hi_mask)) atomicAdd(&h[(k >> bit) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// code = round((x - s) * scale) clamped to [-128, 127], NaN -> 0 (Rust's `as`)
__global__ __launch_bounds__(256) void encode_kernel(const float *x, size_t n, size_t d, const float *shift,
                                                     float scale, int8_t *codes) {
    EMH_FOR_ELEMS(n, d) {
        float r = roundf(__fmul_rn(__fsub_rn(x[i], shift[t]), scale));
        r = r != r ? 0.0f : fminf(fmaxf(r, -128.0f), 127.0f);
        codes[i] = (int8_t)r;
    }
}

__device__ inline uint32_t encode_one(float x, float s, float scale) {
    float r = roundf(__fmul_rn(__fsub_rn(x, s), scale));
    r = r != r ? 0.0f : fminf(fmaxf(r, -128.0f), 127.0f);
    return (uint32_t)(uint8_t)(int8_t)r;
}

// The same codes four values at a time: 16-byte loads, 4-byte stores. A
// workgroup covers `rows_per` rows of dq = d/4 quads per pass and keeps
// ENC_U passes of loads in flight before it stores.
constexpr int ENC_U = 4;
template <class F>
__device__ inline void for_quads(const float4 *x, size_t n, unsigned dq, unsigned rows_per, F f) {
    const unsigned span = rows_per * dq;  // <= 256 when dq <= 256, else dq
    const size_t step = (size_t)gridDim.x * rows_per * ENC_U;
    for (size_t r0 = (size_t)blockIdx.x * rows_per * ENC_U; r0 < n; r0 += step) {
        for (unsigned e = threadIdx.x; e < span; e += 256) {
            const unsigned rr = e / dq, q = e - rr * dq;
            float4 v[ENC_U];
#pragma unroll
            for (int u = 0; u < ENC_U; u++) {
                const size_t r = r0 + (size_t)u * rows_per + rr;
                if (r < n) v[u] = x[r * dq + q];
            }
#pragma unroll
            for (int u = 0; u < ENC_U; u++) {
                const size_t r = r0 + (size_t)u * rows_per + rr;
                if (r < n) f(r * dq + q, q, v[u]);
            }
        }
    }
}

__global__ __launch_bounds__(256) void encode4_kernel(const float4 *x, size_t n, unsigned dq, unsigned rows_per,
                                                      const float4 *shift, float scale, uint32_t *codes) {
    for_quads(x, n, dq, rows_per, [&](size_t i, unsigned q, float4 v) {
        const float4 sh = shift[q];
        codes[i] = encode_one(v.x, sh.x, scale) | encode_one(v.y, sh.y, scale) << 8 |
                   encode_one(v.z, sh.z, scale) << 16 | encode_one(v.w, sh.w, scale) << 24;
    });
}

__device__ inline uint32_t mag_key(float x, float s) { return total_key(__float_as_uint(fabsf(__fsub_rn(x, s)))); }

// mag_max_kernel four values at a time
__global__ __launch_bounds__(256) void mag_max4_kernel(const float4 *x, size_t n, unsigned dq, unsigned rows_per,
                                                       const float4 *shift, unsigned *out) {
    uint32_t best = 0;
    for_quads(x, n, dq, rows_per, [&](size_t, unsigned q, float4 v) {
        const float4 sh = shift[q];
        best = max(best, max(max(mag_key(v.x, sh.x), mag_key(v.y, sh.y)), max(mag_key(v.z, sh.z), mag_key(v.w, sh.w))));
    });
    for (int off = 32; off > 0; off >>= 1) best = max(best, (uint32_t)__shfl_xor((int)best, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

// v * (1 / scale) + s, without contraction
__global__ __launch_bounds__(256) void decode_kernel(const float *codes, size_t n, size_t d, const float *shift,
                                                     float inv, float *out) {
    EMH_FOR_ELEMS(n, d) out[i] = __fadd_rn(__fmul_rn(codes[i], inv), shift[t]);
}

// values whose scaled offset is outside [-128.5, 127.5) (NaN included)
__global__ __launch_bounds__(256) void clipped_kernel(const float *x, size_t n, size_t d, const float *shift,
                                                      float scale, unsigned long long *out) {
    unsigned long long c = 0;
    EMH_FOR_ELEMS(n, d) {
        const float v = __fmul_rn(__fsub_rn(x[i], shift[t]), scale);
        c += !(v >= -128.5f && v < 127.5f);
    }
    for (int off = 32; off > 0; off >>= 1) c += (unsigned long long)__shfl_xor((long long)c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

static dim3 row_grid(size_t n) { return dim3((unsigned)std::max<size_t>(1, std::min<size_t>(n, 16384))); }

// encode on device buffers: the 4-wide kernel when d % 4 == 0 and the
// pointers allow 16-byte loads and 4-byte stores, else the scalar one
struct QuadGrid {
    unsigned dq, rows_per, grid;
};
static QuadGrid quad_grid(size_t n, size_t d) {
    QuadGrid g;
    g.dq = (unsigned)(d / 4);
    g.rows_per = g.dq >= 256 ? 1u : 256u / g.dq;
    const size_t groups = (n + (size_t)g.rows_per * ENC_U - 1) / ((size_t)g.rows_per * ENC_U);
    g.grid = (unsigned)std::max<size_t>(1, std::min<size_t>(groups, 8192));
    return g;
}
static bool quads_ok(size_t d, const void *x, const void *shift) {
    return d % 4 == 0 && d / 4 <= 0xFFFFFFFFu && (uintptr_t)x % 16 == 0 && (uintptr_t)shift % 16 == 0;
}

static int launch_encode(const float *x, size_t n, size_t d, const float *shift, float scale, int8_t *codes,
                         hipStream_t s) {
    if (quads_ok(d, x, shift) && (uintptr_t)codes % 4 == 0) {
        const QuadGrid g = quad_grid(n, d);
        hipLaunchKernelGGL(encode4_kernel, dim3(g.grid), dim3(256), 0, s, (const float4 *)x, n, g.dq, g.rows_per,
                           (const float4 *)shift, scale, (uint32_t *)codes);
    } else {
        hipLaunchKernelGGL(encode_kernel, row_grid(n), dim3(256), 0, s, x, n, d, shift, scale, codes);
    }
    return check(hipGetLastError(), "encode");
}

// fit on device rows; shift_dev (d floats) written, scale returned
static int quant_fit_dev(const float *x, size_t n, size_t d, float target_peak, double quantile, int power_of_two,
                         float *shift_dev, float *scale_out, hipStream_t s) {
    int rc;
    if (d % 4 == 0 && (uintptr_t)x % 16 == 0)
        hipLaunchKernelGGL(col_mean_tiled_kernel, dim3((unsigned)((d + MEAN_DIMS - 1) / MEAN_DIMS)), dim3(256), 0, s,
                           (const float4 *)x, n, d, shift_dev);
    else
        hipLaunchKernelGGL(col_mean_kernel, dim3((unsigned)((d + 63) / 64)), dim3(64), 0, s, x, n, d, shift_dev);
    if ((rc = check(hipGetLastError(), "quantizer mean"))) return rc;
    const size_t count = n * d;
    DevBuf scratch;
    if ((rc = check(scratch.ensure(256 * 8), "quantizer scratch"))) return rc;
    // src/quantize.rs:57: idx = round((len - 1) * clamp(quantile, 0, 1))
    const double q = std::min(std::max(quantile, 0.0), 1.0);
    uint64_t idx = (uint64_t)llround((double)(count - 1) * q);
    uint32_t key = 0;
    if (idx == count - 1) {
        if ((rc = check(hipMemsetAsync(scratch.p, 0, 4, s), "quantizer scratch"))) return rc;
        if (quads_ok(d, x, shift_dev)) {
            const QuadGrid g = quad_grid(n, d);
            hipLaunchKernelGGL(mag_max4_kernel, dim3(g.grid), dim3(256), 0, s, (const float4 *)x, n, g.dq, g.rows_per,
                               (const float4 *)shift_dev, scratch.as<unsigned>());
        } else {
            hipLaunchKernelGGL(mag_max_kernel, row_grid(n), dim3(256), 0, s, x, n, d, (const float *)shift_dev,
                               scratch.as<unsigned>());
        }
        if ((rc = check(hipGetLastError(), "quantizer peak")) ||
            (rc = check(hipMemcpyAsync(&key, scratch.p, 4, hipMemcpyDeviceToHost, s), "quantizer peak")) ||
            (rc = check(hipStreamSynchronize(s), "quantizer peak")))
            return rc;
    } else {
        // radix select: the idx-th smallest key, 8 bits at a time from the top
        uint32_t prefix = 0;
        unsigned long long hist[256];
        for (int bit = 24; bit >= 0; bit -= 8) {
            if ((rc = check(hipMemsetAsync(scratch.p, 0, 256 * 8, s), "quantizer scratch"))) return rc;
            hipLaunchKernelGGL(mag_hist_kernel, row_grid(n), dim3(256), 0, s, x, n, d, (const float *)shift_dev,
                               prefix, bit, scratch.as<unsigned long long>());
            if ((rc = check(hipGetLastError(), "quantizer select")) ||
                (rc = check(hipMemcpyAsync(hist, scratch.p, 256 * 8, hipMemcpyDeviceToHost, s), "quantizer select")) ||
                (rc = check(hipStreamSynchronize(s), "quantizer select")))
                return rc;
            uint32_t digit = 0;
            while (digit < 255 && idx >= hist[digit]) idx -= hist[digit++];
            prefix |= digit << bit;
        }
        key = prefix;
    }
    float peak;
    const uint32_t bits = total_key_inverse(key);
    memcpy(&peak, &bits, 4);
    // src/quantize.rs:59-63
    float scale = peak > 0.0f ? target_peak / peak : 1.0f;
    if (power_of_two) {
        // 2^(floor(log2 scale) as i32): Rust's cast saturates and maps NaN to 0
        const float l = floorf(log2f(scale));
        const int e = l != l ? 0 : (l > 1000.0f ? 1000 : (l < -1000.0f ? -1000 : (int)l));
        scale = ldexpf(1.0f, e);
    }
    *scale_out = scale;
    return EMH_OK;
}

}  // namespace emh

using namespace emh;

extern "C" {

int emh_quant_fit(int device, const float *sample, uint64_t n, uint64_t d, float target_peak, double quantile,
                  int power_of_two, float *shift_out, float *scale_out) {
    if (!sample || !shift_out || !scale_out || n == 0 || d == 0)
        return fail(EMH_E_ARG, "quant_fit: need a non-empty n x d sample and outputs");
    return with_device(device, [&](hipStream_t s) {
        DevBuf x, sh;
        int rc;
        if ((rc = check(x.ensure(n * d * 4), "quantizer sample")) || (rc = check(sh.ensure(d * 4), "quantizer shift")) ||
            (rc = check(hipMemcpyAsync(x.p, sample, n * d * 4, hipMemcpyHostToDevice, s), "upload sample")) ||
            (rc = quant_fit_dev(x.as<float>(), n, d, target_peak, quantile, power_of_two, sh.as<float>(), scale_out, s)) ||
            (rc = check(hipMemcpyAsync(shift_out, sh.p, d * 4, hipMemcpyDeviceToHost, s), "download shift")) ||
            (rc = check(hipStreamSynchronize(s), "quant_fit")))
            return rc;
        return (int)EMH_OK;
    });
}

int emh_quant_encode(int device, const float *x, uint64_t n, uint64_t d, const float *shift, float scale,
                     int8_t *codes) {
    if (!x || !shift || !codes || d == 0) return fail(EMH_E_ARG, "quant_encode: bad arguments");
    if (n == 0) return EMH_OK;
    return with_device(device, [&](hipStream_t s) {
        DevBuf xd, sh, cd;
        int rc;
        const size_t count = n * d;
        if ((rc = check(xd.ensure(count * 4), "quantizer rows")) || (rc = check(sh.ensure(d * 4), "quantizer shift")) ||
            (rc = check(cd.ensure(count), "quantizer codes")) ||
            (rc = check(hipMemcpyAsync(xd.p, x, count * 4, hipMemcpyHostToDevice, s), "upload rows")) ||
            (rc = check(hipMemcpyAsync(sh.p, shift, d * 4, hipMemcpyHostToDevice, s), "upload shift")))
            return rc;
        if ((rc = launch_encode(xd.as<float>(), n, d, sh.as<float>(), scale, cd.as<int8_t>(), s)) ||
            (rc = check(hipMemcpyAsync(codes, cd.p, count, hipMemcpyDeviceToHost, s), "download codes")) ||
            (rc = check(hipStreamSynchronize(s), "quant_encode")))
            return rc;
        return (int)EMH_OK;
    });
}

int emh_quant_decode(int device, const float *codes, uint64_t n, uint64_t d, const float *shift, float scale,
                     float *out) {
    if (!codes || !shift || !out || d == 0) return fail(EMH_E_ARG, "quant_decode: bad arguments");
    if (n == 0) return EMH_OK;
    return with_device(device, [&](hipStream_t s) {
        DevBuf cd, sh, od;
        int rc;
        const size_t count = n * d;
        if ((rc = check(cd.ensure(count * 4), "quantizer codes")) || (rc = check(sh.ensure(d * 4), "quantizer shift")) ||
            (rc = check(od.ensure(count * 4), "quantizer rows")) ||
            (rc = check(hipMemcpyAsync(cd.p, codes, count * 4, hipMemcpyHostToDevice, s), "upload codes")) ||
            (rc = check(hipMemcpyAsync(sh.p, shift, d * 4, hipMemcpyHostToDevice, s), "upload shift")))
            return rc;
        const float inv = 1.0f / scale;  // src/quantize.rs:82
        hipLaunchKernelGGL(decode_kernel, row_grid(n), dim3(256), 0, s, cd.as<float>(), (size_t)n, (size_t)d,
                           sh.as<float>(), inv, od.as<float>());
        if ((rc = check(hipGetLastError(), "decode")) ||
            (rc = check(hipMemcpyAsync(out, od.p, count * 4, hipMemcpyDeviceToHost, s), "download rows")) ||
            (rc = check(hipStreamSynchronize(s), "quant_decode")))
            return rc;
        return (int)EMH_OK;
    });
}

int emh_quant_clipped_fraction(int device, const float *x, uint64_t n, uint64_t d, const float *shift, float scale,
                               double *out) {
    if (!x || !shift || !out || d == 0 || n == 0) return fail(EMH_E_ARG, "quant_clipped_fraction: bad arguments");
    return with_device(device, [&](hipStream_t s) {
        DevBuf xd, sh, cnt;
        int rc;
        const size_t count = n * d;
        unsigned long long c = 0;
        if ((rc = check(xd.ensure(count * 4), "quantizer rows")) || (rc = check(sh.ensure(d * 4), "quantizer shift")) ||
            (rc = check(cnt.ensure(8), "quantizer count")) ||
            (rc = check(hipMemcpyAsync(xd.p, x, count * 4, hipMemcpyHostToDevice, s), "upload rows")) ||
            (rc = check(hipMemcpyAsync(sh.p, shift, d * 4, hipMemcpyHostToDevice, s), "upload shift")) ||
            (rc = check(hipMemsetAsync(cnt.p, 0, 8, s), "clear count")))
            return rc;
        hipLaunchKernelGGL(clipped_kernel, row_grid(n), dim3(256), 0, s, xd.as<float>(), (size_t)n, (size_t)d,
                           sh.as<float>(), scale, cnt.as<unsigned long long>());
        if ((rc = check(hipGetLastError(), "clipped")) ||
            (rc = check(hipMemcpyAsync(&c, cnt.p, 8, hipMemcpyDeviceToHost, s), "download count")) ||
            (rc = check(hipStreamSynchronize(s), "quant_clipped_fraction")))
            return rc;
        *out = (double)c / (double)count;
        return (int)EMH_OK;
    });
}

// Device-pointer variants on the caller's stream (rows already in HBM).
int emh_quant_fit_device(const float *sample_dev, uint64_t n, uint64_t d, float target_peak, double quantile,
                         int power_of_two, float *shift_dev, float *scale_out, void *stream) {
    if (!sample_dev || !shift_dev || !scale_out || n == 0 || d == 0)
        return fail(EMH_E_ARG, "quant_fit_device: need a non-empty n x d sample and outputs");
    return quant_fit_dev(sample_dev, n, d, target_peak, quantile, power_of_two, shift_dev, scale_out,
                         (hipStream_t)stream);
}

int emh_quant_encode_device(const float *x_dev, uint64_t n, uint64_t d, const float *shift_dev, float scale,
                            int8_t *codes_dev, void *stream) {
    if (!x_dev || !shift_dev || !codes_dev || d == 0) return fail(EMH_E_ARG, "quant_encode_device: bad arguments");
    if (n == 0) return EMH_OK;
    return launch_encode(x_dev, n, d, shift_dev, scale, codes_dev, (hipStream_t)stream);
}

}  // extern "C"
