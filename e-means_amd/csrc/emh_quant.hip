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

// fit on device rows; shift_dev (d floats) written, scale returned
static int quant_fit_dev(const float *x, size_t n, size_t d, float target_peak, double quantile, int power_of_two,
                         float *shift_dev, float *scale_out, hipStream_t s) {
    int rc;
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
        hipLaunchKernelGGL(mag_max_kernel, row_grid(n), dim3(256), 0, s, x, n, d, (const float *)shift_dev,
                           scratch.as<unsigned>());
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

template <class F> static int with_device(int device, F &&body) {
    int rc;
    if ((rc = check(hipSetDevice(device), "hipSetDevice"))) return rc;
    hipStream_t s;
    if ((rc = check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"))) return rc;
    rc = body(s);
    (void)hipStreamDestroy(s);
    return rc;
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
        hipLaunchKernelGGL(encode_kernel, row_grid(n), dim3(256), 0, s, xd.as<float>(), (size_t)n, (size_t)d,
                           sh.as<float>(), scale, cd.as<int8_t>());
        if ((rc = check(hipGetLastError(), "encode")) ||
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
    hipLaunchKernelGGL(encode_kernel, row_grid(n), dim3(256), 0, (hipStream_t)stream, x_dev, (size_t)n, (size_t)d,
                       shift_dev, scale, codes_dev);
    return check(hipGetLastError(), "encode");
}

}  // extern "C"
