// The report's O(k^2 d) steps on the device (src/report.rs:169-183, 208).
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
#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "emh_assigner.h"

namespace emh {

constexpr int NT = 32;          // centroids per tile side
constexpr int NSTAGE = 128;     // elements staged per pass (eight 16-chunks)
constexpr int NSTRIDE = 132;    // padded LDS row stride (floats)
constexpr int RED_STRIDE = 17;  // padded stride of one pair's sixteen sums
constexpr int LDS_FLOATS = std::max(2 * NT * NSTRIDE, 512 * RED_STRIDE);

constexpr uint64_t KEY_NONE = (uint64_t)0x7F800000u << 32;  // +inf
constexpr uint64_t REPORT_MAX_K = 131072;  // k x k bits = 2 GiB for the family step

__global__ __launch_bounds__(256) void nearest_init_kernel(uint64_t *keys, size_t k) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < k; i += (size_t)gridDim.x * 256)
        keys[i] = KEY_NONE | i;
}

// What a pair tile produces: the nearest-centroid keys, or the family step's
// "within radius" relation (sqrt(sqdist) <= radius, src/report.rs:208) as a
// k x ceil(k/32) bit matrix, word (i, T) holding bit (j - 32T) for the
// centroids j of tile T. Each word is written by exactly one workgroup.
enum PairMode { PAIR_NEAREST = 0, PAIR_RADIUS = 1 };

template <int MODE>
__global__ __launch_bounds__(256) void pairs_kernel(const float *__restrict__ c, size_t k, size_t d,
                                                    const uint32_t *__restrict__ sizes, unsigned long long *keys,
                                                    float radius, uint32_t *adj) {
    const unsigned I = blockIdx.y, J = blockIdx.x;
    if (J < I) return;
    __shared__ float lds[LDS_FLOATS];
    __shared__ unsigned long long row_min[NT], col_min[NT];
    __shared__ uint32_t row_bits[NT], col_bits[NT];
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
        row_bits[tid] = 0;
        col_bits[tid] = 0;
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
            if (MODE == PAIR_RADIUS) {
                if (sqrtf(s) <= radius) {  // NaN: false, as the reference's comparison
                    atomicOr(&row_bits[li], 1u << lj);
                    atomicOr(&col_bits[lj], 1u << li);
                }
                continue;
            }
            if (!(s < INFINITY)) continue;  // NaN and +inf never replace the starting (i, +inf)
            const unsigned long long hi = (unsigned long long)__float_as_uint(s) << 32;
            if (!sizes || sizes[j]) atomicMin(&row_min[li], hi | j);
            if (!sizes || sizes[i]) atomicMin(&col_min[lj], hi | i);
        }
        __syncthreads();
    }
    if (MODE == PAIR_RADIUS) {
        const size_t words = (k + NT - 1) / NT;
        if (tid < NT) {
            if (i0 + tid < k) adj[(i0 + tid) * words + J] = I == J ? row_bits[tid] | col_bits[tid] : row_bits[tid];
        } else if (tid < 2 * NT && I != J) {
            const int t = tid - NT;
            if (j0 + t < k) adj[(j0 + t) * words + I] = col_bits[t];
        }
        return;
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
    hipLaunchKernelGGL(pairs_kernel<PAIR_NEAREST>, dim3(nt, nt), dim3(256), 0, s, c, k, d, sizes,
                       (unsigned long long *)keys, 0.0f, (uint32_t *)nullptr);
    hipLaunchKernelGGL(nearest_finish_kernel, dim3(lin), dim3(256), 0, s, (const uint64_t *)keys, k, nearest, dist);
    return check(hipGetLastError(), "report_nearest");
}

// adj: k x ceil(k/32) u32, every word written
static int launch_radius(const float *c, size_t k, size_t d, float radius, uint32_t *adj, hipStream_t s) {
    const unsigned nt = (unsigned)((k + NT - 1) / NT);
    hipLaunchKernelGGL(pairs_kernel<PAIR_RADIUS>, dim3(nt, nt), dim3(256), 0, s, c, k, d, (const uint32_t *)nullptr,
                       (unsigned long long *)nullptr, radius, adj);
    return check(hipGetLastError(), "report radius");
}

// ---- report::report on the host around the two device steps --------------
// (src/report.rs:119-250). Sorts are over f32::total_cmp keys; quantile
// indices round half away from zero (f64::round); sums and divisions are the
// reference's, in its order.

static inline uint32_t total_key(float v) {
    uint32_t u;
    memcpy(&u, &v, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
static void sort_total(std::vector<float> &v) {
    std::sort(v.begin(), v.end(), [](float a, float b) { return total_key(a) < total_key(b); });
}
// src/report.rs:83-88
static float quantile(const std::vector<float> &sorted, size_t lo, size_t len, double q) {
    if (len == 0) return 0.0f;
    return sorted[lo + (size_t)llround((double)(len - 1) * q)];
}
// src/report.rs:91-96: (median, MAD) of an ascending range
static void median_mad(const std::vector<float> &sorted, size_t lo, size_t len, float &med, float &mad) {
    med = quantile(sorted, lo, len, 0.5);
    std::vector<float> dev(len);
    for (size_t t = 0; t < len; t++) dev[t] = fabsf(sorted[lo + t] - med);
    sort_total(dev);
    mad = quantile(dev, 0, len, 0.5);
}
static float median_of(std::vector<float> v) {
    sort_total(v);
    return quantile(v, 0, v.size(), 0.5);
}

struct ReportOut {
    emh_cluster_stats *clusters;
    float *local_z, *global_z;
    uint32_t *family_leader;
    uint64_t *family_size;
    double *family_share;
    uint32_t *family_start, *family_members;
    emh_report_summary *summary;
};

static int report_impl(int device, const uint32_t *labels, const float *sqdist, size_t n, const float *centroids,
                       size_t k, size_t d, const emh_report_config &cfg, const ReportOut &o) {
    int rc;
    for (size_t r = 0; r < n; r++)
        if (labels[r] >= k) return fail(EMH_E_ARG, "report: a label is not below k");
    std::vector<float> dist(n);
    for (size_t r = 0; r < n; r++) dist[r] = sqrtf(fmaxf(sqdist[r], 0.0f));  // v.max(0.0).sqrt()

    // counting sort of member distances by cluster (src/report.rs:125-137)
    std::vector<size_t> start(k + 1, 0);
    for (size_t r = 0; r < n; r++) start[labels[r] + 1]++;
    for (size_t j = 0; j < k; j++) start[j + 1] += start[j];
    std::vector<float> grouped(n);
    {
        std::vector<size_t> cur(start.begin(), start.end() - 1);
        for (size_t r = 0; r < n; r++) grouped[cur[labels[r]]++] = dist[r];
    }
    // corpus median and MAD (src/report.rs:139-142): order statistics, so a
    // selection gives the elements a full sort would put at those indices
    float g_med = 0.0f, g_mad = 0.0f;
    if (n) {
        auto less = [](float a, float b) { return total_key(a) < total_key(b); };
        const size_t mi = (size_t)llround((double)(n - 1) * 0.5);
        std::vector<float> all(dist);
        std::nth_element(all.begin(), all.begin() + mi, all.end(), less);
        g_med = all[mi];
        for (float &v : all) v = fabsf(v - g_med);
        std::nth_element(all.begin(), all.begin() + mi, all.end(), less);
        g_mad = all[mi];
    }
    const float g_scale = fmaxf(1.4826f * g_mad, FLT_MIN);

    // per-cluster statistics (src/report.rs:145-167), clusters spread over
    // host threads (each sorts and folds its own range of `grouped`)
    std::vector<uint32_t> sizes(k);
    auto cluster = [&](size_t j) {
        emh_cluster_stats &c = o.clusters[j];
        memset(&c, 0, sizeof c);
        const size_t lo = start[j], len = start[j + 1] - lo;
        sizes[j] = (uint32_t)len;
        if (!len) return;
        std::sort(grouped.begin() + lo, grouped.begin() + lo + len,
                  [](float a, float b) { return total_key(a) < total_key(b); });
        c.size = (uint32_t)len;
        median_mad(grouped, lo, len, c.median, c.mad);
        double sq = 0.0;  // f64, in the sorted order (src/report.rs:157)
        for (size_t t = 0; t < len; t++) sq += (double)grouped[lo + t] * (double)grouped[lo + t];
        c.rms = (float)sqrt(sq / (double)len);
        c.p90 = quantile(grouped, lo, len, 0.9);
        c.p99 = quantile(grouped, lo, len, 0.99);
        c.max = grouped[lo + len - 1];
    };
    {
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<size_t> next{0};
        auto worker = [&]() {
            for (size_t j0; (j0 = next.fetch_add(64)) < k;)
                for (size_t j = j0; j < std::min(k, j0 + 64); j++) cluster(j);
        };
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < nt && (size_t)t * 64 < k; t++) pool.emplace_back(worker);
        worker();
        for (auto &t : pool) t.join();
    }

    // nearest live other centroid on the device (src/report.rs:169-183),
    // then the "within radius" relation of the family step (src/report.rs:208)
    std::vector<uint32_t> nidx(k);
    std::vector<float> ndist(k);
    const size_t words = (k + NT - 1) / NT;
    std::vector<uint32_t> adj;
    float median_rms = 0.0f, median_nearest = 0.0f;
    rc = with_device(device, [&](hipStream_t s) {
        DevBuf cd, sd, kd, nd, dd, ad;
        int r;
        if ((r = check(cd.ensure(k * d * 4), "centroids")) || (r = check(sd.ensure(k * 4), "sizes")) ||
            (r = check(kd.ensure(k * 8), "keys")) || (r = check(nd.ensure(k * 4), "nearest")) ||
            (r = check(dd.ensure(k * 4), "distances")) ||
            (r = check(hipMemcpyAsync(cd.p, centroids, k * d * 4, hipMemcpyHostToDevice, s), "upload centroids")) ||
            (r = check(hipMemcpyAsync(sd.p, sizes.data(), k * 4, hipMemcpyHostToDevice, s), "upload sizes")) ||
            (r = launch_nearest(cd.as<float>(), k, d, sd.as<uint32_t>(), kd.as<uint64_t>(), nd.as<uint32_t>(),
                                dd.as<float>(), s)) ||
            (r = check(hipMemcpyAsync(nidx.data(), nd.p, k * 4, hipMemcpyDeviceToHost, s), "download nearest")) ||
            (r = check(hipMemcpyAsync(ndist.data(), dd.p, k * 4, hipMemcpyDeviceToHost, s), "download distances")) ||
            (r = check(hipStreamSynchronize(s), "report nearest")))
            return r;
        std::vector<float> live_rms, live_nd;
        for (size_t j = 0; j < k; j++)
            if (sizes[j]) {
                live_rms.push_back(o.clusters[j].rms);
                live_nd.push_back(ndist[j]);
            }
        median_rms = median_of(live_rms);
        median_nearest = median_of(live_nd);
        const float radius = cfg.family_radius * median_nearest;
        adj.resize(k * words);
        if ((r = check(ad.ensure(k * words * 4), "radius bits")) ||
            (r = launch_radius(cd.as<float>(), k, d, radius, ad.as<uint32_t>(), s)) ||
            (r = check(hipMemcpyAsync(adj.data(), ad.p, k * words * 4, hipMemcpyDeviceToHost, s), "download bits")) ||
            (r = check(hipStreamSynchronize(s), "report radius")))
            return r;
        return (int)EMH_OK;
    });
    if (rc) return rc;
    for (size_t j = 0; j < k; j++) {
        emh_cluster_stats &c = o.clusters[j];
        c.nearest_centroid = nidx[j];
        c.nearest_dist = ndist[j];
        c.looseness = c.rms / fmaxf(median_rms, FLT_MIN);
        c.isolation = ndist[j] / fmaxf(median_nearest, FLT_MIN);
    }

    // families (src/report.rs:190-223): live clusters by decreasing size (equal
    // sizes by index), each unassigned one leads the unassigned live clusters
    // within the radius, members in that same order
    std::vector<uint32_t> order;
    for (size_t j = 0; j < k; j++)
        if (sizes[j]) order.push_back((uint32_t)j);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return sizes[a] > sizes[b]; });
    std::vector<uint32_t> pos(k, UINT32_MAX), family_of(k, UINT32_MAX);
    for (size_t t = 0; t < order.size(); t++) pos[order[t]] = (uint32_t)t;
    struct Fam {
        uint32_t leader;
        std::vector<uint32_t> members;
        uint64_t size;
    };
    std::vector<Fam> fams;
    std::vector<uint32_t> joins;
    for (uint32_t lead : order) {
        if (family_of[lead] != UINT32_MAX) continue;
        joins.clear();
        joins.push_back(lead);
        const uint32_t *row = adj.data() + (size_t)lead * words;
        for (size_t w = 0; w < words; w++)
            for (uint32_t bits = row[w]; bits; bits &= bits - 1) {
                const size_t j = w * NT + (size_t)__builtin_ctz(bits);
                if (j != lead && sizes[j] && family_of[j] == UINT32_MAX) joins.push_back((uint32_t)j);
            }
        std::sort(joins.begin(), joins.end(), [&](uint32_t a, uint32_t b) { return pos[a] < pos[b]; });
        Fam f{lead, joins, 0};
        for (uint32_t j : joins) {
            family_of[j] = (uint32_t)fams.size();
            f.size += sizes[j];
        }
        fams.push_back(std::move(f));
    }
    std::vector<uint32_t> rank(fams.size());
    for (size_t f = 0; f < fams.size(); f++) rank[f] = (uint32_t)f;
    std::stable_sort(rank.begin(), rank.end(), [&](uint32_t a, uint32_t b) { return fams[a].size > fams[b].size; });
    size_t m = 0;
    for (size_t f = 0; f < rank.size(); f++) {
        const Fam &F = fams[rank[f]];
        o.family_leader[f] = F.leader;
        o.family_size[f] = F.size;
        o.family_share[f] = (double)F.size / (double)n;
        o.family_start[f] = (uint32_t)m;
        for (uint32_t j : F.members) {
            o.family_members[m++] = j;
            o.clusters[j].family = (uint32_t)f;
        }
    }
    o.family_start[rank.size()] = (uint32_t)m;

    // z-scores (src/report.rs:226-239)
    for (size_t r = 0; r < n; r++) {
        const float gz = (dist[r] - g_med) / g_scale;
        const emh_cluster_stats &c = o.clusters[labels[r]];
        o.global_z[r] = gz;
        o.local_z[r] = ((size_t)c.size < cfg.min_local_size || c.mad <= 0.0f)
                           ? gz
                           : (dist[r] - c.median) / (1.4826f * c.mad);
    }
    o.summary->median_dist = g_med;
    o.summary->median_rms = median_rms;
    o.summary->median_nearest_dist = median_nearest;
    o.summary->n_families = (uint32_t)fams.size();
    return EMH_OK;
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

void emh_report_config_default(emh_report_config *cfg) {
    if (!cfg) return;
    cfg->family_radius = 1.25f;  // ReportConfig::default (src/report.rs:30-34)
    cfg->min_local_size = 30;
}

int emh_report(int device, const uint32_t *labels, const float *sqdist, uint64_t n, const float *centroids,
               uint64_t k, uint64_t d, const emh_report_config *cfg, emh_cluster_stats *clusters, float *local_z,
               float *global_z, uint32_t *family_leader, uint64_t *family_size, double *family_share,
               uint32_t *family_start, uint32_t *family_members, emh_report_summary *summary) {
    int rc;
    if (!centroids || !cfg || !clusters || !family_leader || !family_size || !family_share || !family_start ||
        !family_members || !summary || (n && (!labels || !sqdist || !local_z || !global_z)))
        return fail(EMH_E_ARG, "report: NULL argument");
    if ((rc = nearest_check(k, d))) return rc;
    if (n > 0xFFFFFFFFull) return fail(EMH_E_ARG, "report: n must fit u32 offsets");
    if (k > REPORT_MAX_K)
        return fail(EMH_E_UNSUPPORTED, "report: the family step's k x k bit relation is offered up to k = 131072");
    ReportOut o{clusters, local_z, global_z, family_leader, family_size, family_share, family_start, family_members,
                summary};
    return report_impl(device, labels, sqdist, n, centroids, k, d, *cfg, o);
}

}  // extern "C"
