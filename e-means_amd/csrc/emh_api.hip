// C ABI of the e-means nearest-centroid search (include/emeans_hip.h).
//
// Mirrors the reference's plugin traits Assigner / AssignRows<T>
// (src/lib.rs:75-91) with the CUDA backend's behaviour (src/cuda.rs) as the
// contract: set_centroids copies the table and computes norms, set_bias adds
// to the scores until the next set_centroids, assign is synchronous and ties
// go to the lowest index. Errors come back as codes plus a message instead of
// the reference's panics.
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <new>

#include <map>
#include <mutex>
#include <tuple>

#include "emh_assigner.h"

namespace emh {

static thread_local std::string g_last_error = "no error";

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

int check(hipError_t e, const char *what) {
    if (e == hipSuccess) return EMH_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(EMH_E_MEMORY, std::string("not enough GPU memory: ") + what);
    return fail(EMH_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

static int op_of(int precision) {
    switch (precision) {
        case EMH_FP32: return OP_F32;
        case EMH_BF16_RESCORE: return OP_BF16;
        case EMH_FP8_RESCORE: return OP_FP8;
        case EMH_I8: return OP_I8;
        case EMH_I16: return OP_I16;
        default: return -1;
    }
}

static size_t row_bytes_of(int rows) { return rows == EMH_ROWS_F32 ? 4 : 1; }

// Records a timing event on the stream when timing is on.
static int mark(emh_assigner *h, hipEvent_t *slot) {
    if (!h->timing) return EMH_OK;
    *slot = h->timing_event();
    if (!*slot) return fail(EMH_E_DEVICE, "timing event");
    return check(hipEventRecord(*slot, h->stream), "hipEventRecord");
}

// ---- tile autotuner ----
// A register-resident search launch sweeps 128 fr rows per workgroup (fr
// 32-row fragments per wave, 1 or 2) and may split every row block's
// centroid tiles over nslice workgroups (merge_slices_kernel merges them).
// Neither changes the results, so for launches that leave the chip short of
// two workgroups per CU the tuner times the candidate (fr, nslice) pairs on
// the call itself, the first time a shape (op, row type, m, row bytes,
// tiles, row blocks) is seen, keeps the fastest, and records it in a JSON
// file that later processes read: EMH_TUNE_CACHE names the file ("" keeps
// the choices in memory only), by default
// $XDG_CACHE_HOME/emeans_hip/tiles.json or ~/.cache/emeans_hip/tiles.json.
// EMH_AUTOTUNE=0 replaces the timing by a rule: fr = 2 and the smallest
// split that puts two workgroups on every CU.
namespace {
std::mutex g_tune_mu;
std::map<std::string, Tile> g_tiles;
bool g_tiles_loaded = false;

bool autotune_enabled() {
    const char *e = getenv("EMH_AUTOTUNE");
    return !(e && e[0] == '0');
}

std::string tune_cache_path() {
    const char *e = getenv("EMH_TUNE_CACHE");
    if (e) return std::string(e);
    const char *x = getenv("XDG_CACHE_HOME");
    if (x && x[0]) return std::string(x) + "/emeans_hip/tiles.json";
    const char *home = getenv("HOME");
    if (home && home[0]) return std::string(home) + "/.cache/emeans_hip/tiles.json";
    return std::string();
}

// The file is one JSON object, one entry a line:
//   "<arch>|<op>|<rows>|<M>|<row bytes>|<tiles>|<row blocks>": {"fr": 2, "nslice": 4, "ms": 0.1234},
// (called with g_tune_mu held)
void load_tiles() {
    if (g_tiles_loaded) return;
    g_tiles_loaded = true;
    const std::string path = tune_cache_path();
    if (path.empty()) return;
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return;
    char line[512], key[400];
    while (fgets(line, sizeof line, f)) {
        int fr = 0, ns = 0;
        if (sscanf(line, " \"%399[^\"]\": {\"fr\": %d, \"nslice\": %d", key, &fr, &ns) == 3 && (fr == 1 || fr == 2) &&
            ns >= 1 && ns <= 4096)
            g_tiles[key] = Tile{fr, ns, 0.0f};
    }
    fclose(f);
}

// (called with g_tune_mu held); best effort: a cache that cannot be
// written leaves the choices in memory
void save_tiles() {
    const std::string path = tune_cache_path();
    if (path.empty()) return;
    for (size_t p = path.find('/', 1); p != std::string::npos; p = path.find('/', p + 1))
        (void)mkdir(path.substr(0, p).c_str(), 0755);
    const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return;
    fprintf(f, "{\n");
    size_t i = 0;
    for (const auto &kv : g_tiles)
        fprintf(f, " \"%s\": {\"fr\": %d, \"nslice\": %d, \"ms\": %.4f}%s\n", kv.first.c_str(), kv.second.fr,
                kv.second.ns, kv.second.ms, ++i < g_tiles.size() ? "," : "");
    fprintf(f, "}\n");
    const bool ok = fclose(f) == 0;
    if (!ok || rename(tmp.c_str(), path.c_str()) != 0) (void)remove(tmp.c_str());
}
}  // namespace

// Search launch with tile t; `part` holds slice_part_bytes of per-slice
// results when t.ns > 1.
static int search_launch(emh_assigner *h, int M, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                         const uint8_t *x8, const float *rs, const float *rn, uint32_t *labels_dev, float *dist_dev,
                         uint32_t *cand, Tile t, uint32_t *part) {
    const SearchGeom &g = h->g;
    h->last_tile = t;
    return check(launch_search(g, M, x_dev, rows, m, ids_dev, h->cen_op.p, h->hn.as<float>(), h->cn.as<int64_t>(), x8,
                               rs, rn, h->cexp.as<uint8_t>(), h->stats.as<unsigned>(), labels_dev, dist_dev, cand,
                               h->stream, t.ns, t.ns > 1 ? part : nullptr, t.fr),
                 "search kernel");
}

// fr = 1 is instantiated for f32 rows of the FP32 and BF16 searches, and
// for the FP8 and I8 searches
static bool fr1_available(const emh_assigner *h, int rows) {
    if (!h->g.reg) return false;
    if (h->op == OP_F32 || h->op == OP_BF16) return rows == EMH_ROWS_F32;
    return h->op == OP_FP8 || (h->op == OP_I8 && h->g.i8_packed);
}

// The rule without timing: fr = 2 and the smallest power-of-two split that
// puts two workgroups on every CU (1 when the launch fills the chip already).
static Tile split_rule(const emh_assigner *h, size_t m) {
    const SearchGeom &g = h->g;
    const size_t rowblocks = (m + 255) / 256;
    const int nt = reg_tiles(g);
    const size_t target = 2 * (size_t)h->cu_count;
    Tile t;
    if (!g.reg || rowblocks >= target || nt < 4) return t;
    for (int s = 1; s <= nt / 2 && rowblocks * (size_t)s <= 4 * target; s *= 2) {
        t.ns = s;
        if (rowblocks * (size_t)s >= target) break;
    }
    return t;
}

// a forced tile (emh_set_tile / emh_set_split), if any
static bool forced_tile(const emh_assigner *h, int rows, Tile *t) {
    if (h->forced_fr <= 0 && h->forced_slices <= 0) return false;
    t->fr = h->forced_fr == 1 && fr1_available(h, rows) ? 1 : 2;
    t->ns = h->forced_slices > 0 ? h->forced_slices : 1;
    return true;
}

static int choose_tile(emh_assigner *h, int M, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                       const uint8_t *x8, const float *rs, const float *rn, uint32_t *labels_dev, float *dist_dev,
                       uint32_t *cand, Tile *out) {
    *out = Tile{};
    const SearchGeom &g = h->g;
    if (forced_tile(h, rows, out)) return EMH_OK;
    if (!g.reg || h->timing) return EMH_OK;
    const size_t rowblocks = (m + 255) / 256;
    const int nt = reg_tiles(g);
    const size_t target = 2 * (size_t)h->cu_count;  // the kernel runs two workgroups per CU
    if (rowblocks >= target || nt < 4) return EMH_OK;
    std::vector<Tile> cands;
    for (int fr = 2; fr >= 1; fr--) {
        if (fr == 1 && !fr1_available(h, rows)) break;
        const size_t rb = (m + 128 * fr - 1) / (128 * fr);
        for (int s = 1; s <= nt / 2 && rb * (size_t)s <= 4 * target; s *= 2) cands.push_back(Tile{fr, s, 0.0f});
    }
    if (cands.size() < 2) return EMH_OK;
    const std::string key = h->arch + "|" + std::to_string(h->op) + "|" + std::to_string(rows) + "|" +
                            std::to_string(M) + "|" + std::to_string(g.row_bytes) + "|" + std::to_string(nt) + "|" +
                            std::to_string(rowblocks);
    if (!autotune_enabled()) {
        *out = split_rule(h, m);
        return EMH_OK;
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        load_tiles();
        auto it = g_tiles.find(key);
        if (it != g_tiles.end()) {
            *out = it->second;
            if (out->fr == 1 && !fr1_available(h, rows)) out->fr = 2;
            return EMH_OK;
        }
    }
    int rc;
    size_t part_bytes = 0;
    for (const Tile &t : cands) part_bytes = std::max(part_bytes, slice_part_bytes(M, m, t.ns, h->op));
    if ((rc = check(h->part.ensure(part_bytes), "split buffer"))) return rc;
    hipEvent_t e0, e1;
    if ((rc = check(hipEventCreate(&e0), "tune event")) || (rc = check(hipEventCreate(&e1), "tune event"))) return rc;
    uint32_t *part = h->part.as<uint32_t>();
    // Launches of a few tens of microseconds time noisily and the clock
    // ramps while the first ones run: every candidate gets a warm-up
    // launch, then ROUNDS passes over all candidates time one launch each,
    // and a candidate's time is its fastest.
    constexpr int ROUNDS = 3;
    for (size_t i = 0; i < cands.size() && !rc; i++) {
        cands[i].ms = INFINITY;
        rc = search_launch(h, M, x_dev, rows, m, ids_dev, x8, rs, rn, labels_dev, dist_dev, cand, cands[i], part);
    }
    for (int round = 0; round < ROUNDS && !rc; round++)
        for (size_t i = 0; i < cands.size(); i++) {
            float ms = 0.0f;
            if ((rc = check(hipEventRecord(e0, h->stream), "tune event")) ||
                (rc = search_launch(h, M, x_dev, rows, m, ids_dev, x8, rs, rn, labels_dev, dist_dev, cand, cands[i],
                                    part)) ||
                (rc = check(hipEventRecord(e1, h->stream), "tune event")) ||
                (rc = check(hipEventSynchronize(e1), "tune sync")))
                break;
            (void)hipEventElapsedTime(&ms, e0, e1);
            cands[i].ms = std::min(cands[i].ms, ms);
        }
    Tile best = cands.front();
    for (const Tile &t : cands)
        if (t.ms < best.ms) best = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tiles[key] = best;
    save_tiles();
    *out = best;
    return EMH_OK;
}

// The row-side buffers of one search call.
struct RowScratch {
    uint32_t *cand = nullptr;  // m x M candidate ids (BF16 / FP8)
    uint8_t *x8 = nullptr;     // FP8: m x d_pad encoded rows
    float *rs = nullptr;       // FP8: m row scales
    float *rn = nullptr;       // FP8: m squared row norms
    uint32_t *part = nullptr;  // nslice > 1: per-slice results
};

// Encoding (FP8), search and rescore (BF16 / FP8) of one call on the
// buffers of `rs_`, with `ns` centroid slices; allocates nothing.
static int search_stages(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                         uint32_t *labels_dev, float *dist_dev, const RowScratch &b, Tile t, hipEvent_t *ev,
                         bool timed) {
    auto mk = [&](int i) { return timed ? mark(h, &ev[i]) : EMH_OK; };
    const SearchGeom &g = h->g;
    int rc;
    if (h->op == OP_F32 || h->op == OP_I8 || h->op == OP_I16) {
        if ((rc = mk(1))) return rc;
        if ((rc = search_launch(h, 1, x_dev, rows, m, ids_dev, nullptr, nullptr, nullptr, labels_dev, dist_dev,
                                nullptr, t, b.part)))
            return rc;
        if ((rc = mk(2))) return rc;
        return mk(3);
    }
    if ((rc = mk(1))) return rc;
    if ((rc = search_launch(h, h->M, x_dev, rows, m, ids_dev, b.x8, b.rs, b.rn, nullptr, nullptr, b.cand, t, b.part)))
        return rc;
    if ((rc = mk(2))) return rc;
    if ((rc = check(launch_rescore(h->M, x_dev, rows, m, g.d, ids_dev, h->cen(),
                                   h->has_bias ? h->bias.as<float>() : nullptr, g.k, b.cand, labels_dev, dist_dev,
                                   h->stream),
                    "rescore kernel")))
        return rc;
    return mk(3);
}

static int encode_rows(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                       const RowScratch &b) {
    const SearchGeom &g = h->g;
    return check(launch_rows_fp8(x_dev, rows, m, g.d, g.d_pad, ids_dev, b.x8, b.rs, b.rn, h->stream),
                 "fp8 row encoding");
}

int run_search(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev, uint32_t *labels_dev,
               float *dist_dev) {
    if (m == 0) return EMH_OK;
    const SearchGeom &g = h->g;
    int rc;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    if ((rc = mark(h, &ev[0]))) return rc;
    RowScratch b;
    const bool cand_search = h->op == OP_BF16 || h->op == OP_FP8;
    if (cand_search) {
        if ((rc = check(h->cand.ensure(m * h->M * 4), "candidate buffer"))) return rc;
        b.cand = h->cand.as<uint32_t>();
    }
    // FP8: the encoded rows, unless the register-resident kernel encodes them itself
    if (h->op == OP_FP8 && !fp8_encodes_in_search(g, rows)) {
        if ((rc = check(h->x8.ensure(m * g.d_pad), "fp8 row buffer"))) return rc;
        if ((rc = check(h->row_scale.ensure(m * 4), "row scale buffer"))) return rc;
        if ((rc = check(h->row_norm.ensure(m * 4), "row norm buffer"))) return rc;
        b.x8 = h->x8.as<uint8_t>();
        b.rs = h->row_scale.as<float>();
        b.rn = h->row_norm.as<float>();
        if ((rc = encode_rows(h, x_dev, rows, m, ids_dev, b))) return rc;
    }
    const int M = cand_search ? h->M : 1;
    Tile t;
    if ((rc = choose_tile(h, M, x_dev, rows, m, ids_dev, b.x8, b.rs, b.rn, cand_search ? nullptr : labels_dev,
                          cand_search ? nullptr : dist_dev, b.cand, &t)))
        return rc;
    if (t.ns > 1) {
        if ((rc = check(h->part.ensure(slice_part_bytes(M, m, t.ns, h->op)), "split buffer"))) return rc;
        b.part = h->part.as<uint32_t>();
    }
    return search_stages(h, x_dev, rows, m, ids_dev, labels_dev, dist_dev, b, t, ev, true);
}

// ---- the search without allocations ----
// The row-side buffers of emh_assign_device_noalloc live in this array (one
// copy per GPU, held until the process ends). Its device address on each
// GPU is taken once, by emh_create on that GPU, and kept in g_scratch.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch[MAX_DEVICES];

// Called with `device` current.
static int scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return fail(EMH_E_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return EMH_OK;
    void *ptr = nullptr;
    int rc = check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    g_scratch[device] = ptr;
    return EMH_OK;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Scratch layout for m rows and ns slices: cand | x8 | row scales | row
// norms | slice results; returns the bytes used.
static size_t noalloc_layout(const emh_assigner *h, int rows, size_t m, int ns, uint8_t *base, RowScratch *b) {
    const SearchGeom &g = h->g;
    size_t off = 0;
    auto take = [&](size_t bytes) -> uint8_t * {
        uint8_t *p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    if (h->op == OP_BF16 || h->op == OP_FP8) b->cand = (uint32_t *)take(m * h->M * 4);
    if (h->op == OP_FP8 && !fp8_encodes_in_search(g, rows)) {  // else the kernel encodes its rows
        b->x8 = take(m * g.d_pad);
        b->rs = (float *)take(m * 4);
        b->rn = (float *)take(m * 4);
    }
    if (ns > 1) b->part = (uint32_t *)take(slice_part_bytes(h->op == OP_BF16 || h->op == OP_FP8 ? h->M : 1, m, ns, h->op));
    return off;
}

int run_search_noalloc(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                       uint32_t *labels_dev, float *dist_dev) {
    if (m == 0) return EMH_OK;
    uint8_t *base = h->device >= 0 && h->device < MAX_DEVICES ? (uint8_t *)g_scratch[h->device] : nullptr;
    if (!base) return fail(EMH_E_STATE, "no-allocation search: the scratch array of this device is not set up");
    RowScratch probe;
    const size_t need1 = noalloc_layout(h, rows, m, 1, nullptr, &probe);
    if (need1 > NOALLOC_SCRATCH_BYTES) {
        const size_t per_row = need1 / m + 1;
        return fail(EMH_E_ARG, "no-allocation search: " + std::to_string(m) + " rows need " + std::to_string(need1) +
                                   " bytes of scratch, above the 67108864-byte maximum (at most about " +
                                   std::to_string(NOALLOC_SCRATCH_BYTES / per_row) + " rows per call at this d and m)");
    }
    // a forced tile, else the rule's; its split when the slice results fit
    Tile t;
    if (!forced_tile(h, rows, &t)) t = split_rule(h, m);
    if (t.ns > 1) {
        const int nt = reg_tiles(h->g);
        const int per = (nt + t.ns - 1) / t.ns;
        t.ns = per > 0 ? (nt + per - 1) / per : 1;  // the slices launch_search forms
        RowScratch p2;
        if (!h->g.reg || noalloc_layout(h, rows, m, t.ns, nullptr, &p2) > NOALLOC_SCRATCH_BYTES) t.ns = 1;
    }
    RowScratch b;
    noalloc_layout(h, rows, m, t.ns, base, &b);
    int rc;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    if (h->op == OP_FP8 && !fp8_encodes_in_search(h->g, rows) && (rc = encode_rows(h, x_dev, rows, m, ids_dev, b)))
        return rc;
    return search_stages(h, x_dev, rows, m, ids_dev, labels_dev, dist_dev, b, t, ev, false);
}

}  // namespace emh

using namespace emh;

hipEvent_t emh_assigner::timing_event() {
    if (events_used == events.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        events.push_back(e);
    }
    return events[events_used++];
}

emh_assigner::~emh_assigner() {
    for (DevBuf *b : {&cen_f32, &cen_op, &norm32, &norm64, &hn, &cn, &bias, &cexp, &stats, &x_stage, &labels_stage,
                      &dist_stage, &cand, &x8, &row_scale, &row_norm, &part})
        b->release();
    accum_free(acc);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}

// I8: whether every score fits the packed int32 key of the register-resident
// kernel; I16: whether every offset fits its exact f64 keys (flags written by
// the bias kernel). Otherwise the LDS kernel's i64 scores run.
static int refresh_i8_range(emh_assigner *h, bool with_bias) {
    h->g.i8_packed = false;
    h->g.i16_f64 = false;
    if (h->op == OP_I16 && h->g.reg) {
        // I16 on the register kernel: without a bias cn = sum c16^2 <= d 2^30
        // < 2^46 (d <= 512 there); with one, the bias kernel's flag says
        if (!with_bias) {
            h->g.i16_f64 = true;
            return EMH_OK;
        }
        uint32_t flag = 1;
        int rc = check(hipMemcpy(&flag, h->stats.as<uint32_t>() + 2, 4, hipMemcpyDeviceToHost), "read range flag");
        if (rc) return rc;
        h->g.i16_f64 = flag == 0;
        return EMH_OK;
    }
    if (h->op != OP_I8 || !h->g.reg || (int64_t)h->g.d * 32768 >= ((int64_t)1 << 26)) return EMH_OK;
    // without a bias cn = sum c8^2 <= d 2^14, inside the limit 2^26 - d 2^15
    // whenever d 3 2^14 < 2^26: no need to read the flag back (no sync)
    if (!with_bias && (int64_t)h->g.d * 49152 < ((int64_t)1 << 26)) {
        h->g.i8_packed = true;
        return EMH_OK;
    }
    uint32_t flag = 1;
    int rc = check(hipMemcpy(&flag, h->stats.as<uint32_t>() + 2, 4, hipMemcpyDeviceToHost), "read range flag");
    if (rc) return rc;
    h->g.i8_packed = flag == 0;
    return EMH_OK;
}

namespace emh {
// set_centroids from host memory (synchronous: the caller may reuse the
// table) or from device memory (stream-ordered).
int set_centroids_impl(emh_assigner *h, const float *centroids, size_t k, size_t d, bool device_src, bool borrow) {
    if (!h) return fail(EMH_E_ARG, "emh_set_centroids: NULL handle");
    if (k == 0 || d == 0 || !centroids) return fail(EMH_E_ARG, "set_centroids: need k > 0, d > 0 and a k x d table");
    if ((h->op == OP_I8 || h->op == OP_I16) && d > 32768)
        return fail(EMH_E_ARG, "integer assignment supports d <= 32768");
    if (k > 0xFFFFFFFEull) return fail(EMH_E_ARG, "set_centroids: k must fit in 32 bits");
    int rc;
    if ((rc = check(hipSetDevice(h->device), "hipSetDevice"))) return rc;
    SearchGeom g = make_geom(h->op, k, d);
    // the rescoring searches read the f32 table after later model updates
    // may have changed the caller's: they always keep a copy
    borrow = borrow && device_src && h->op != OP_BF16 && h->op != OP_FP8;
    if (!borrow && (rc = check(h->cen_f32.ensure(k * d * 4), "centroid buffer"))) return rc;
    if ((rc = check(h->cen_op.ensure(operand_bytes(g)), "operand buffer"))) return rc;
    if ((rc = check(h->norm32.ensure(g.k_pad * 4), "norm buffer"))) return rc;
    if ((rc = check(h->norm64.ensure(g.k_pad * 8), "norm buffer"))) return rc;
    if ((rc = check(h->hn.ensure(g.k_pad * 4), "score buffer"))) return rc;
    if ((rc = check(h->cn.ensure(g.k_pad * 8), "score buffer"))) return rc;
    if ((rc = check(h->cexp.ensure(g.k_pad), "scale buffer"))) return rc;
    if ((rc = check(h->stats.ensure(12), "stats buffer"))) return rc;
    h->g = g;
    h->has_centroids = false;
    h->has_bias = false;
    h->cen_ref = borrow ? centroids : nullptr;
    if (!borrow && (rc = check(hipMemcpyAsync(h->cen_f32.p, centroids, k * d * 4,
                                              device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->stream),
                               "copy centroids")))
        return rc;
    // FP32, and I8 where its range flag is not read (no bias, d below 1366),
    // take their unbiased score offsets from the prep kernel itself: no
    // stats words to clear, no separate offsets pass
    const bool fused = h->op == OP_F32 || (h->op == OP_I8 && (int64_t)d * 49152 < ((int64_t)1 << 26));
    if (!fused && (rc = check(launch_fill_u32(h->stats.as<uint32_t>(), 3, 0u, h->stream), "fill"))) return rc;
    if ((rc = check(launch_prep_centroids(g, h->cen(), h->cen_op.p, h->norm32.as<float>(), h->norm64.as<int64_t>(),
                                          h->cexp.as<uint8_t>(), h->stats.as<unsigned>(), h->stream,
                                          fused && h->op == OP_F32 ? h->hn.as<float>() : nullptr,
                                          fused && h->op == OP_I8 ? h->cn.as<int64_t>() : nullptr),
                    "centroid prep")))
        return rc;
    if (!fused && (rc = check(launch_set_bias(g, h->norm32.as<float>(), h->norm64.as<int64_t>(), nullptr,
                                              h->hn.as<float>(), h->cn.as<int64_t>(), h->stats.as<unsigned>(),
                                              h->stream),
                              "score offsets")))
        return rc;
    // The caller may reuse `centroids` on return (pageable source).
    if (!device_src && (rc = check(hipStreamSynchronize(h->stream), "set_centroids"))) return rc;
    if ((rc = refresh_i8_range(h, false))) return rc;
    h->has_centroids = true;
    return EMH_OK;
}

int set_bias_impl(emh_assigner *h, const float *bias, size_t k, bool sync) {
    if (!h) return fail(EMH_E_ARG, "emh_set_bias: NULL handle");
    if (!h->has_centroids) return fail(EMH_E_STATE, "set_centroids must be called first");
    if (!bias || k != h->g.k) return fail(EMH_E_ARG, "set_bias: bias must have k entries");
    int rc;
    if ((rc = check(hipSetDevice(h->device), "hipSetDevice"))) return rc;
    if ((rc = check(h->bias.ensure(k * 4), "bias buffer"))) return rc;
    if ((rc = check(hipMemcpyAsync(h->bias.p, bias, k * 4, hipMemcpyHostToDevice, h->stream), "copy bias"))) return rc;
    if ((rc = check(launch_fill_u32(h->stats.as<uint32_t>() + 1, 2, 0u, h->stream), "fill"))) return rc;
    if ((rc = check(launch_set_bias(h->g, h->norm32.as<float>(), h->norm64.as<int64_t>(), h->bias.as<float>(),
                                    h->hn.as<float>(), h->cn.as<int64_t>(), h->stats.as<unsigned>(), h->stream),
                    "score offsets")))
        return rc;
    if ((sync || h->op == OP_I8) && (rc = check(hipStreamSynchronize(h->stream), "set_bias"))) return rc;
    if ((rc = refresh_i8_range(h, true))) return rc;
    h->has_bias = true;
    return EMH_OK;
}
}  // namespace emh

extern "C" {

const char *emh_last_error(void) { return g_last_error.c_str(); }

int emh_create(int device, int precision, int rescore_m, emh_assigner **out) {
    if (!out) return fail(EMH_E_ARG, "emh_create: out is NULL");
    *out = nullptr;
    int op = op_of(precision);
    if (op < 0) return fail(EMH_E_ARG, "emh_create: unknown precision");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(EMH_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= count) return fail(EMH_E_ARG, "emh_create: device ordinal out of range");
    int rc;
    if ((rc = check(hipSetDevice(device), "hipSetDevice"))) return rc;
    hipDeviceProp_t prop;
    if ((rc = check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties"))) return rc;
    emh_assigner *h = new (std::nothrow) emh_assigner();
    if (!h) return fail(EMH_E_MEMORY, "host allocation failed");
    h->device = device;
    h->precision = precision;
    h->op = op;
    if (op == OP_BF16 || op == OP_FP8) {
        int m = rescore_m <= 0 ? 4 : rescore_m;
        if (m > MAX_M) m = MAX_M;
        int p = 1;
        while (p < m) p <<= 1;
        h->M = p;
    }
    h->device_name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
    h->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    h->arch = std::string(prop.gcnArchName) + "/" + std::to_string(h->cu_count);
    // the device address of this GPU's copy of the no-allocation scratch
    if ((rc = scratch_setup(device))) {
        delete h;
        return rc;
    }
    if ((rc = check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate"))) {
        delete h;
        return rc;
    }
    h->own_stream = true;
    *out = h;
    return EMH_OK;
}

void emh_destroy(emh_assigner *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
}

int emh_set_stream(emh_assigner *h, void *stream) {
    if (!h) return fail(EMH_E_ARG, "emh_set_stream: NULL handle");
    if (h->own_stream && h->stream) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipStreamDestroy(h->stream);
    }
    h->own_stream = false;
    h->stream = (hipStream_t)stream;
    return EMH_OK;
}

int emh_synchronize(emh_assigner *h) {
    if (!h) return fail(EMH_E_ARG, "emh_synchronize: NULL handle");
    return check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
}

int emh_set_timing(emh_assigner *h, int on) {
    if (!h) return fail(EMH_E_ARG, "emh_set_timing: NULL handle");
    int rc;
    if ((rc = check(hipStreamSynchronize(h->stream), "hipStreamSynchronize"))) return rc;
    h->timing = on != 0;
    h->events_used = 0;
    return EMH_OK;
}

int emh_get_timing(emh_assigner *h, double *ms_out, uint64_t *calls) {
    if (!h || !ms_out || !calls) return fail(EMH_E_ARG, "emh_get_timing: NULL argument");
    int rc;
    if ((rc = check(hipStreamSynchronize(h->stream), "hipStreamSynchronize"))) return rc;
    ms_out[0] = ms_out[1] = ms_out[2] = 0.0;
    const size_t n = h->events_used / 4;
    for (size_t c = 0; c < n; c++) {
        for (int s = 0; s < 3; s++) {
            float ms = 0.0f;
            if ((rc = check(hipEventElapsedTime(&ms, h->events[4 * c + s], h->events[4 * c + s + 1]),
                            "hipEventElapsedTime")))
                return rc;
            ms_out[s] += ms;
        }
    }
    *calls = n;
    h->events_used = 0;
    return EMH_OK;
}


int emh_set_centroids(emh_assigner *h, const float *centroids, size_t k, size_t d) {
    return set_centroids_impl(h, centroids, k, d, false);
}

int emh_set_bias(emh_assigner *h, const float *bias, size_t k) { return emh::set_bias_impl(h, bias, k, true); }

int emh_set_split(emh_assigner *h, int nslice) {
    if (!h) return fail(EMH_E_ARG, "emh_set_split: NULL handle");
    if (nslice < 0 || nslice > 4096) return fail(EMH_E_ARG, "emh_set_split: nslice must be in [0, 4096]");
    h->forced_slices = nslice;
    return EMH_OK;
}

int emh_set_tile(emh_assigner *h, int fragments) {
    if (!h) return fail(EMH_E_ARG, "emh_set_tile: NULL handle");
    if (fragments < 0 || fragments > 2) return fail(EMH_E_ARG, "emh_set_tile: fragments must be 0 (tuned), 1 or 2");
    h->forced_fr = fragments;
    return EMH_OK;
}

int emh_last_tile(const emh_assigner *h, int *fragments, int *nslice) {
    if (!h || !fragments || !nslice) return fail(EMH_E_ARG, "emh_last_tile: NULL argument");
    *fragments = h->last_tile.fr;
    *nslice = h->last_tile.ns;
    return EMH_OK;
}

int emh_name(const emh_assigner *h, char *buf, size_t cap) {
    if (!h || !buf || cap == 0) return fail(EMH_E_ARG, "emh_name: bad arguments");
    static const char *kind[] = {"FP32 fused MFMA argmin", "BF16 MFMA + FP32 rescore", "FP8-e4m3 MFMA + FP32 rescore",
                                 "I8 MFMA exact", "I16 (two i8 MFMA planes) exact"};
    char tmp[512];
    int n;
    if (h->op == OP_BF16 || h->op == OP_FP8)
        n = snprintf(tmp, sizeof tmp, "HIP %s (top-%d), %s", kind[h->precision], h->M, h->device_name.c_str());
    else
        n = snprintf(tmp, sizeof tmp, "HIP %s, %s", kind[h->precision], h->device_name.c_str());
    if (h->has_centroids)
        snprintf(tmp + n, sizeof tmp - n, ", k=%zu d=%zu (%s rows%s)", h->g.k, h->g.d,
                 h->g.reg ? "register-resident" : h->g.resident ? "LDS-resident" : "chunked",
                 h->last_tile.ns > 1 ? ", centroids split" : "");
    strncpy(buf, tmp, cap - 1);
    buf[cap - 1] = 0;
    return EMH_OK;
}

static int assign_host(emh_assigner *h, const void *x, int rows, size_t m, uint32_t *labels, float *dist) {
    if (!h) return fail(EMH_E_ARG, "assign: NULL handle");
    if (!h->has_centroids) return fail(EMH_E_STATE, "set_centroids must be called first");
    if (m == 0) return EMH_OK;
    if (!x || !labels || !dist) return fail(EMH_E_ARG, "assign: NULL buffer");
    if ((h->op == OP_I8 || h->op == OP_I16) && rows != EMH_ROWS_I8)
        return fail(EMH_E_UNSUPPORTED, "the integer search takes i8 rows only (AssignRows<i8>)");
    int rc;
    if ((rc = check(hipSetDevice(h->device), "hipSetDevice"))) return rc;
    const size_t d = h->g.d, rb = row_bytes_of(rows);
    // Rows per upload: up to 1 Mi rows or 1 GiB of staging, whichever is smaller.
    size_t cap = std::min<size_t>(m, std::max<size_t>(1, std::min<size_t>(1u << 20, (1ull << 30) / (d * rb))));
    if ((rc = check(h->x_stage.ensure(cap * d * rb), "row staging"))) return rc;
    if ((rc = check(h->labels_stage.ensure(cap * 4), "label staging"))) return rc;
    if ((rc = check(h->dist_stage.ensure(cap * 4), "distance staging"))) return rc;
    for (size_t r0 = 0; r0 < m; r0 += cap) {
        size_t mc = std::min(cap, m - r0);
        if ((rc = check(hipMemcpyAsync(h->x_stage.p, (const char *)x + r0 * d * rb, mc * d * rb, hipMemcpyHostToDevice,
                                       h->stream),
                        "copy rows")))
            return rc;
        if ((rc = run_search(h, h->x_stage.p, rows, mc, nullptr, h->labels_stage.as<uint32_t>(),
                             h->dist_stage.as<float>())))
            return rc;
        if ((rc = check(hipMemcpyAsync(labels + r0, h->labels_stage.p, mc * 4, hipMemcpyDeviceToHost, h->stream),
                        "copy labels")))
            return rc;
        if ((rc = check(hipMemcpyAsync(dist + r0, h->dist_stage.p, mc * 4, hipMemcpyDeviceToHost, h->stream),
                        "copy distances")))
            return rc;
        if ((rc = check(hipStreamSynchronize(h->stream), "assign"))) return rc;
    }
    return EMH_OK;
}

int emh_assign_f32(emh_assigner *h, const float *x, size_t m, uint32_t *labels, float *dist) {
    return assign_host(h, x, EMH_ROWS_F32, m, labels, dist);
}
int emh_assign_i8(emh_assigner *h, const int8_t *x, size_t m, uint32_t *labels, float *dist) {
    return assign_host(h, x, EMH_ROWS_I8, m, labels, dist);
}
int emh_assign_u8(emh_assigner *h, const uint8_t *x, size_t m, uint32_t *labels, float *dist) {
    return assign_host(h, x, EMH_ROWS_U8, m, labels, dist);
}

int emh_assign_device(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                      uint32_t *labels_dev, float *dist_dev) {
    if (!h) return fail(EMH_E_ARG, "assign: NULL handle");
    if (!h->has_centroids) return fail(EMH_E_STATE, "set_centroids must be called first");
    if (rows < EMH_ROWS_F32 || rows > EMH_ROWS_U8) return fail(EMH_E_ARG, "assign: unknown row type");
    if (m == 0) return EMH_OK;
    if (!x_dev || !labels_dev || !dist_dev) return fail(EMH_E_ARG, "assign: NULL buffer");
    if ((h->op == OP_I8 || h->op == OP_I16) && rows != EMH_ROWS_I8)
        return fail(EMH_E_UNSUPPORTED, "the integer search takes i8 rows only (AssignRows<i8>)");
    int rc;
    if ((rc = check(hipSetDevice(h->device), "hipSetDevice"))) return rc;
    return run_search(h, x_dev, rows, m, ids_dev, labels_dev, dist_dev);
}

int emh_assign_device_noalloc(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                              uint32_t *labels_dev, float *dist_dev) {
    if (!h) return fail(EMH_E_ARG, "assign: NULL handle");
    if (!h->has_centroids) return fail(EMH_E_STATE, "set_centroids must be called first");
    if (rows < EMH_ROWS_F32 || rows > EMH_ROWS_U8) return fail(EMH_E_ARG, "assign: unknown row type");
    if (m == 0) return EMH_OK;
    if (!x_dev || !labels_dev || !dist_dev) return fail(EMH_E_ARG, "assign: NULL buffer");
    if ((h->op == OP_I8 || h->op == OP_I16) && rows != EMH_ROWS_I8)
        return fail(EMH_E_UNSUPPORTED, "the integer search takes i8 rows only (AssignRows<i8>)");
    int rc;
    if ((rc = check(hipSetDevice(h->device), "hipSetDevice"))) return rc;
    return run_search_noalloc(h, x_dev, rows, m, ids_dev, labels_dev, dist_dev);
}

int emh_accumulate_device(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                          const uint32_t *labels_dev, float *sums_dev, uint32_t *counts_dev) {
    if (!h) return fail(EMH_E_ARG, "accumulate: NULL handle");
    if (!h->has_centroids) return fail(EMH_E_STATE, "set_centroids must be called first");
    if (rows < EMH_ROWS_F32 || rows > EMH_ROWS_U8) return fail(EMH_E_ARG, "accumulate: unknown row type");
    if (!sums_dev || !counts_dev || (m && (!x_dev || !labels_dev))) return fail(EMH_E_ARG, "accumulate: NULL buffer");
    int rc;
    if ((rc = check(hipSetDevice(h->device), "hipSetDevice"))) return rc;
    return check(launch_accumulate(h->acc, x_dev, rows, m, h->g.d, ids_dev, labels_dev, h->g.k, sums_dev, counts_dev,
                                   h->stream),
                 "accumulate");
}

}  // extern "C"
