// The assigner object behind the C ABI handle (shared by the API and the
// training driver).
#pragma once

#include <string>
#include <vector>

#include "emh_internal.h"

namespace emh {

// Grow-only device allocation, freed with its owner.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), cap(o.cap) {
        o.p = nullptr;
        o.cap = 0;
    }
    ~DevBuf() { release(); }
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return (T *)p; }
};

// Tile of a register-resident search launch: fr 32-row fragments per wave
// (128 fr rows per workgroup), ns centroid slices per row block; ms: the
// tuner's timing of it.
struct Tile {
    int fr = 2, ns = 1;
    float ms = 0.0f;
};

}  // namespace emh

struct emh_assigner {
    int device = 0;
    int precision = EMH_FP32;
    int op = emh::OP_F32;
    int M = 1;  // candidates per row (1 = plain argmin)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string device_name;
    std::string arch;  // gcnArchName / CU count: the tile cache's device key

    emh::SearchGeom g{};
    bool has_centroids = false;
    bool has_bias = false;

    // centroid side (set_centroids / set_bias)
    emh::DevBuf cen_f32, cen_op, norm32, norm64, hn, cn, bias, cexp, stats;
    // the f32 centroid table the search reads: cen_f32, or (training's
    // per-step snapshot of a search without a rescore) the caller's device
    // table itself, valid until the next set_centroids
    const float *cen_ref = nullptr;
    const float *cen() const { return cen_ref ? cen_ref : cen_f32.as<float>(); }
    // row side (per assign call)
    emh::DevBuf x_stage, labels_stage, dist_stage, cand, x8, row_scale, row_norm;
    // tile of the register-resident search (autotuned): centroid-range
    // split results, the last launch's tile, forced values (0: tuned)
    emh::DevBuf part;
    int cu_count = 256;
    emh::Tile last_tile;
    int forced_slices = 0;
    int forced_fr = 0;
    // per-cluster accumulation
    emh::AccumScratch acc;

    // optional per-kernel timing: 4 events per search call (row prep, search, rescore)
    bool timing = false;
    std::vector<hipEvent_t> events;
    size_t events_used = 0;
    hipEvent_t timing_event();

    ~emh_assigner();
};

namespace emh {
// Search over device rows (x_dev of `rows` type, optionally gathered by ids).
int run_search(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev, uint32_t *labels_dev,
               float *dist_dev);
int fail(int code, const std::string &msg);
// set_centroids from host (synchronous) or device (stream-ordered) memory.
// borrow (device source, FP32 / I8 / I16 only): read the caller's table in
// place instead of copying it; it must stay unchanged until the searches
// queued before the next set_centroids have run.
int set_centroids_impl(emh_assigner *h, const float *centroids, size_t k, size_t d, bool device_src,
                       bool borrow = false);
// set_bias from host memory; `sync` false leaves it stream-ordered (the
// caller keeps `bias`, pinned, unchanged until the stream has passed it),
// except for the I8 search, which reads its range flag back.
int set_bias_impl(emh_assigner *h, const float *bias, size_t k, bool sync);
int check(hipError_t e, const char *what);
// Run body(stream) on `device` with a stream of its own, destroyed afterwards
// (the host-pointer entry points that take a device ordinal).
template <class F> inline int with_device(int device, F &&body) {
    int rc;
    if ((rc = check(hipSetDevice(device), "hipSetDevice"))) return rc;
    hipStream_t s;
    if ((rc = check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"))) return rc;
    rc = body(s);
    (void)hipStreamDestroy(s);
    return rc;
}
}  // namespace emh
