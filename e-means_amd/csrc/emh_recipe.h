// Host side of the e-means recipe, shared by the single-problem and the
// batched training drivers: constants, RNG, recipe laws, model state
// (src/lib.rs:31-43, 165-345) and the per-batch loop of src/lib.rs:400-466
// as a resumable runner.
#pragma once

#include <math.h>
#include <string.h>

#include <algorithm>
#include <future>
#include <thread>
#include <string>
#include <vector>

#include "emh_assigner.h"

namespace emh {


// ---- constants of the reference (src/lib.rs:31-43) ----------------------
static const double ALPHA_MIN_FACTOR = 0.1;
static const size_t CHECK_UNUSED_EVERY = 5;
static const float INIT_USAGE = 1.0f / 65536.0f;
static const float UNUSED_THRESHOLD = 0.1f;
static const size_t NO_REVIVAL_MAX_D = 2;
static const size_t AUTO_QUALITY_MAX_D = 16;
static const double BATCH_THROUGHPUT[4] = {0.035134, 0.134749, 0.876473, 0.813344};
static const double BATCH_QUALITY[4] = {0.00439175, 0.134749, 0.876473, 0.813344};
static const double ALPHA_LAW[5] = {0.8244, -0.4617, 0.9003, -0.4498, -0.1057};
static const size_t ASSIGN_CHUNK = 65536;

// ---- xoshiro256++ seeded by splitmix64 (src/lib.rs:227-274) -------------
class Rng {
  public:
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto &w : s_) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            w = x ^ (x >> 31);
        }
    }
    uint64_t next() {
        const uint64_t r = rotl(s_[0] + s_[3], 23) + s_[0];
        const uint64_t t = s_[1] << 17;
        s_[2] ^= s_[0];
        s_[3] ^= s_[1];
        s_[1] ^= s_[2];
        s_[0] ^= s_[3];
        s_[2] ^= t;
        s_[3] = rotl(s_[3], 45);
        return r;
    }
    uint64_t below(uint64_t bound) {  // Lemire
        // accept x when the low word of x * bound is >= (2^64 - bound) % bound;
        // a low word >= bound always is, so the division waits for the rare
        // low word below bound (the same draws, the same results)
        unsigned __int128 m = (unsigned __int128)next() * bound;
        if ((uint64_t)m < bound) {
            const uint64_t t = (0 - bound) % bound;
            while ((uint64_t)m < t) m = (unsigned __int128)next() * bound;
        }
        return (uint64_t)(m >> 64);
    }
    double open01() { return ((double)(next() >> 11) + 0.5) / 9007199254740992.0; }
    // Fisher-Yates from the top: for i = n - 1 down to 1, swap p[i] with
    // p[below(i + 1)]. The draws do not depend on the array, so they are
    // taken LA steps ahead (in the same order) and the random slot is
    // prefetched.
    void permutation(uint32_t n, std::vector<uint32_t> &p) {
        p.resize(n);
        for (uint32_t i = 0; i < n; i++) p[i] = i;
        if (n < 2) return;
        constexpr uint32_t LA = 16;
        uint32_t ahead[LA];
        uint32_t next_i = n - 1;  // the next step whose draw is taken
        auto draw = [&]() {
            const uint32_t j = (uint32_t)below((uint64_t)next_i + 1);
            ahead[next_i % LA] = j;
            __builtin_prefetch(&p[j], 1);
            next_i--;
        };
        for (uint32_t t = 0; t < LA && next_i >= 1; t++) draw();
        for (uint32_t i = n - 1; i >= 1; i--) {
            const uint32_t j = ahead[i % LA];
            if (next_i >= 1) draw();  // step i - LA, into the slot just read
            std::swap(p[i], p[j]);
        }
    }

  private:
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t s_[4];
};

// ---- recipe (src/lib.rs:165-225) -----------------------------------------
static uint64_t round_pow2(double v) {
    if (v <= 1.0) return 1;
    const double low = ldexp(1.0, (int)floor(log2(v)));
    const double high = 2.0 * low;
    return (uint64_t)(v / low < high / v ? low : high);
}

static int resolve(uint64_t n, uint64_t d, uint64_t k, const emh_config &cfg, emh_recipe &r) {
    if (!(cfg.nepochs > 0.0) || cfg.nepochs >= (double)(EMH_MAX_EPOCHS - 1)) return -1;
    int law = cfg.batch_law;
    if (law == 0) law = d <= AUTO_QUALITY_MAX_D ? 2 : 1;
    uint64_t batch;
    if (cfg.has_batch_size) {
        batch = std::min<uint64_t>(std::max<uint64_t>(cfg.batch_size, 1), n);
    } else {
        const double *c = law == 1 ? BATCH_THROUGHPUT : BATCH_QUALITY;
        const double raw = c[0] * pow((double)k, c[1]) * pow((double)n, c[2]) * pow(cfg.nepochs, c[3]);
        const uint64_t cap = law == 1 ? round_pow2((double)n / 2.0) : n / 2;
        batch = std::max<uint64_t>(std::min(std::min(round_pow2(raw), cap), n), 1);
    }
    r.batch_size = batch;
    if (cfg.has_alpha0) {
        r.alpha0 = cfg.alpha0;
    } else {
        const double a = ALPHA_LAW[0] * pow((double)k, ALPHA_LAW[1]) * pow((double)batch, ALPHA_LAW[2]) *
                         pow((double)n, ALPHA_LAW[3]) * pow(cfg.nepochs, ALPHA_LAW[4]);
        r.alpha0 = std::min(std::max(a, 1e-4), 1.0);
    }
    const uint64_t whole = (uint64_t)floor(cfg.nepochs);
    const uint64_t rem = (uint64_t)nearbyint((cfg.nepochs - (double)whole) * (double)n);
    r.n_epochs = 0;
    if (whole == 0) {
        r.epoch_rows[r.n_epochs++] = rem;
    } else {
        for (uint64_t i = 0; i + 1 < whole; i++) r.epoch_rows[r.n_epochs++] = n;
        if (rem > 0) r.epoch_rows[r.n_epochs++] = rem;
        r.epoch_rows[r.n_epochs++] = n;
    }
    r.total_steps = 0;
    for (uint32_t e = 0; e < r.n_epochs; e++) r.total_steps += (r.epoch_rows[e] + batch - 1) / batch;
    r.unused_threshold = cfg.has_unused_threshold ? cfg.unused_threshold
                                                  : (d <= NO_REVIVAL_MAX_D ? 0.0f : UNUSED_THRESHOLD);
    return 0;
}

static double scheduled_alpha(double alpha0, uint64_t step, uint64_t total) {
    const double frac = (double)std::min(step, total) / (double)std::max<uint64_t>(total, 1);
    const double cosine = 0.5 * (1.0 + cos(M_PI * frac));
    return alpha0 * (ALPHA_MIN_FACTOR + (1.0 - ALPHA_MIN_FACTOR) * cosine);
}

// ---- model state (src/lib.rs:276-345) -------------------------------------
struct State {
    size_t k, d;
    std::vector<float> centroids, usage;
    size_t n_splits = 0;

    // Gumbel-top-k donor split of clusters below tau * mean mass
    // (src/lib.rs:314-344), planned on `usage` (updated here): the (donor,
    // unused) pairs whose centroid rows the device then splits
    // (revive_split_kernel). The EMA step itself runs on the device
    // (ema_apply_kernel); the model's centroids live there.
    void revive_plan(float tau, Rng &rng, std::vector<std::pair<size_t, size_t>> &pairs) {
        pairs.clear();
        float total = 0.0f;
        for (float u : usage) total += u;
        const float threshold = total * tau / (float)k;
        std::vector<size_t> unused;
        for (size_t j = 0; j < k; j++)
            if (usage[j] < threshold) unused.push_back(j);
        if (unused.empty()) return;
        std::vector<std::pair<double, size_t>> donors;
        for (size_t j = 0; j < k; j++)
            if (usage[j] >= threshold && usage[j] > 0.0f)
                donors.emplace_back(log((double)usage[j]) - log(-log(rng.open01())), j);
        const size_t nsplit = std::min(unused.size(), donors.size());
        std::sort(donors.begin(), donors.end(),
                  [](const std::pair<double, size_t> &a, const std::pair<double, size_t> &b) { return a.first > b.first; });
        for (size_t t = 0; t < nsplit; t++) {
            const size_t src = donors[t].second, dst = unused[t];
            const float half = usage[src] / 2.0f;
            usage[src] = half;
            usage[dst] = half;
            pairs.emplace_back(src, dst);
        }
        n_splits += nsplit;
    }
};

static float widen(const void *x, int rows, size_t idx) {
    if (rows == EMH_ROWS_I8) return (float)((const int8_t *)x)[idx];
    if (rows == EMH_ROWS_U8) return (float)((const uint8_t *)x)[idx];
    return ((const float *)x)[idx];
}

// Byte layout of a block made of 256-byte aligned segments.
struct Layout {
    size_t bytes = 0;
    size_t add(size_t n) {
        const size_t at = (bytes + 255) / 256 * 256;
        bytes = at + (n ? n : 1);
        return at;
    }
};

// A segment of such a block.
struct Seg {
    void *p = nullptr;
    template <class T> T *as() const { return (T *)p; }
};

// Pinned host buffer (hipHostMalloc).
struct HostBuf {
    void *p = nullptr;
    hipError_t alloc(size_t bytes) { return hipHostMalloc(&p, bytes ? bytes : 16, hipHostMallocDefault); }
    // host-coherent: kernels read it in place, seeing the host's latest writes
    hipError_t alloc_coherent(size_t bytes) {
        return hipHostMalloc(&p, bytes ? bytes : 16, hipHostMallocCoherent | hipHostMallocMapped);
    }
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
    template <class T> T *as() const { return (T *)p; }
};


// One training run's host state, advanced one batch at a time: next()
// stages the batch (src/lib.rs:400-417), the caller runs the device half
// (gather, search, per-cluster sums, EMA step, revival split) and folds the
// batch's labels and distances in with finish_stats (src/lib.rs:419-466).
// p[0..n) = v, on a few threads for large n (the first touch of a fresh
// output array is page faults, which spread across cores).
inline void fill_u32_host(uint32_t *p, size_t n, uint32_t v) {
    const size_t nt = n < (size_t)1 << 20 ? 1 : std::min<size_t>(8, std::max(1u, std::thread::hardware_concurrency()));
    if (nt == 1) {
        std::fill(p, p + n, v);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; t++)
        th.emplace_back([=] { std::fill(p + n * t / nt, p + n * (t + 1) / nt, v); });
    for (auto &x : th) x.join();
}

struct JobRun {
    const void *x = nullptr;
    int rows = EMH_ROWS_F32;
    size_t n = 0, d = 0, k = 0;
    emh_config cfg{};
    float *centroids_out = nullptr;
    uint32_t *labels_out = nullptr;
    emh_train_out *out = nullptr;
    emh_recipe R{};
    State st{};
    Rng split_rng{0};
    std::vector<uint32_t> order;
    std::vector<float> bias;
    double mean_sqdist = 0.0, epoch_obj = 0.0;
    uint64_t step = 0;
    uint32_t epoch = 0;
    size_t b0 = 0, chunk = 0;
    bool done = false;
    // the staged batch
    size_t nb = 0;
    const uint32_t *ids = nullptr;
    bool penalized = false, is_last = false, last_epoch_start = false;
    // deferred statistics: a batch's objective and labels are folded in
    // (finish_stats_deferred, in batch order) after later batches have been
    // staged; the epoch objective is then recorded by the batch closing it.
    // Only for runs without the balance penalty, whose next batch does not
    // depend on the previous one's statistics.
    bool deferred = false;
    // the next epoch's batch order, drawn on a background thread while the
    // current epoch runs (started when it starts); next() takes it instead
    // of drawing it then
    std::vector<uint32_t> order_next;
    uint32_t order_next_epoch = UINT32_MAX;
    std::future<void> order_job;
    struct StepMeta {
        size_t nb;
        const uint32_t *ids;
        bool is_last, closes_epoch;
    };

    // input checks (src/lib.rs:349-361), recipe, initial centroids
    // (random data rows, src/lib.rs:374-381)
    // rows on the device: the caller gathers the initial centroids (rows
    // init_rows of x, widened to f32) on the GPU; st.centroids stays unset
    bool rows_on_device = false;
    std::vector<uint32_t> init_rows;

    int init(const void *x_, int rows_, size_t n_, size_t d_, size_t k_, const emh_config &cfg_, float *cout,
             uint32_t *lout, emh_train_out *out_, bool on_device = false) {
        if (d_ == 0 || k_ == 0 || !x_) return fail(EMH_E_ARG, "x must be a non-empty row-major n x d array");
        if (n_ < k_) return fail(EMH_E_ARG, "need at least k = " + std::to_string(k_) + " rows, got " + std::to_string(n_));
        if (n_ > 0xFFFFFFFFull) return fail(EMH_E_ARG, "more than 2^32 - 1 rows: subsample or shard the input");
        if (!(cfg_.nepochs > 0.0)) return fail(EMH_E_ARG, "nepochs must be positive");
        if (rows_ < EMH_ROWS_F32 || rows_ > EMH_ROWS_U8) return fail(EMH_E_ARG, "unknown row type");
        if (!cout || !lout || !out_) return fail(EMH_E_ARG, "train: NULL output");
        x = x_;
        rows = rows_;
        n = n_;
        d = d_;
        k = k_;
        cfg = cfg_;
        centroids_out = cout;
        labels_out = lout;
        out = out_;
        memset(&R, 0, sizeof R);
        if (resolve(n, d, k, cfg, R)) return fail(EMH_E_ARG, "nepochs out of range");
        rows_on_device = on_device;
        st = State{k, d, std::vector<float>(on_device ? 0 : k * d), std::vector<float>(k, INIT_USAGE)};
        // the first epoch's order (next() takes it) drawn beside the initial
        // rows' permutation: two independent streams of the same length
        if (R.n_epochs > 0)
            order_job = std::async(std::launch::async, [this] {
                Rng(cfg.seed * 1000003ull + 1).permutation((uint32_t)n, order_next);
                order_next_epoch = 0;
            });
        {
            std::vector<uint32_t> init;
            Rng(cfg.seed).permutation((uint32_t)n, init);
            if (on_device) {
                init_rows.assign(init.begin(), init.begin() + k);
            } else {
                for (size_t j = 0; j < k; j++)
                    for (size_t i = 0; i < d; i++) st.centroids[j * d + i] = widen(x, rows, (size_t)init[j] * d + i);
            }
        }
        split_rng = Rng(cfg.seed ^ 0x5EED5EED5EED5EEDull);
        chunk = std::min(ASSIGN_CHUNK, (size_t)R.batch_size);
        bias.assign(cfg.balance > 0.0f ? k : 0, 0.0f);
        fill_u32_host(labels_out, n, 0xFFFFFFFFu);
        out->n_objective = 0;
        mean_sqdist = 0.0;
        step = 0;
        epoch = 0;
        b0 = 0;
        done = R.n_epochs == 0;
        return EMH_OK;
    }

    // Stages the next batch: its rows (ids), and the balance bias against the
    // fixed centroid snapshot (src/lib.rs:408-417). False once done.
    bool next() {
        if (done) return false;
        if (b0 == 0) {
            if (order_job.valid()) order_job.get();  // this epoch's order, drawn in the background
            if (order_next_epoch == epoch) {
                order.swap(order_next);
                order_next_epoch = UINT32_MAX;
            } else {
                Rng(cfg.seed * 1000003ull + (uint64_t)epoch + 1).permutation((uint32_t)n, order);
            }
            // the next epoch's order (the same permutation next() would draw
            // then) while this epoch runs; order_next no longer holds ids a
            // pending batch reads: only the last epoch's labels use them
            if (epoch + 1 < R.n_epochs) {
                const uint32_t e1 = epoch + 1;
                order_job = std::async(std::launch::async, [this, e1] {
                    Rng(cfg.seed * 1000003ull + (uint64_t)e1 + 1).permutation((uint32_t)n, order_next);
                    order_next_epoch = e1;
                });
            }
            if (!deferred) epoch_obj = 0.0;
        }
        is_last = epoch == R.n_epochs - 1;
        last_epoch_start = is_last && b0 == 0;
        const size_t epoch_n = R.epoch_rows[epoch];
        nb = std::min((size_t)R.batch_size, epoch_n - b0);
        ids = order.data() + b0;
        penalized = cfg.balance > 0.0f && mean_sqdist > 0.0;
        if (penalized) {
            double mm = 0.0;
            for (float u : st.usage) mm += (double)u;
            const double mean_mass = mm / (double)k;
            const float scale = (float)((double)cfg.balance * mean_sqdist / mean_mass);
            for (size_t j = 0; j < k; j++) bias[j] = scale * st.usage[j];
        }
        return true;
    }

    // Objective, true distances, last-pass labels (src/lib.rs:425-440).
    void finish_stats(const uint32_t *lab, float *dist) {
        double batch_obj = 0.0;
        for (size_t p0 = 0; p0 < nb; p0 += chunk) {
            const size_t m = std::min(chunk, nb - p0);
            if (penalized) {
                // report true squared distances (src/lib.rs:425-430)
                for (size_t r = p0; r < p0 + m; r++) {
                    const float v = dist[r] - bias[lab[r]];
                    dist[r] = v > 0.0f ? v : 0.0f;
                }
            }
            double cs = 0.0;
            for (size_t r = p0; r < p0 + m; r++) cs += (double)dist[r];
            batch_obj += cs;
        }
        if (is_last)
            for (size_t r = 0; r < nb; r++) labels_out[ids[r]] = lab[r];
        epoch_obj += batch_obj;
        mean_sqdist = batch_obj / (double)nb;
    }

    // finish_stats for a batch whose true distances (bias taken off) the
    // device has made and whose last-pass labels it keeps.
    void finish_stats_device(const float *dist) {
        double batch_obj = 0.0;
        for (size_t p0 = 0; p0 < nb; p0 += chunk) {
            const size_t m = std::min(chunk, nb - p0);
            double cs = 0.0;
            for (size_t r = p0; r < p0 + m; r++) cs += (double)dist[r];
            batch_obj += cs;
        }
        epoch_obj += batch_obj;
        mean_sqdist = batch_obj / (double)nb;
    }

    // The staged batch, for finish_stats_deferred.
    StepMeta meta() const { return StepMeta{nb, ids, is_last, b0 + nb >= R.epoch_rows[epoch]}; }

    // finish_stats for a batch staged earlier (no balance penalty), then the
    // epoch objective when the batch closes its epoch. The ids of the last
    // epoch stay valid: no permutation is drawn after it.
    void finish_stats_deferred(const StepMeta &m, const uint32_t *lab, const float *dist) {
        double batch_obj = 0.0;
        for (size_t p0 = 0; p0 < m.nb; p0 += chunk) {
            const size_t cm = std::min(chunk, m.nb - p0);
            double cs = 0.0;
            for (size_t r = p0; r < p0 + cm; r++) cs += (double)dist[r];
            batch_obj += cs;
        }
        if (m.is_last)
            for (size_t r = 0; r < m.nb; r++) labels_out[m.ids[r]] = lab[r];
        epoch_obj += batch_obj;
        mean_sqdist = batch_obj / (double)m.nb;
        if (m.closes_epoch) {
            if (out->n_objective < EMH_MAX_EPOCHS) out->objective[out->n_objective++] = epoch_obj;
            epoch_obj = 0.0;
        }
    }

    // This step's EMA rate (src/lib.rs:455) and whether revival follows it
    // (src/lib.rs:463-465).
    float alpha() const { return (float)scheduled_alpha(R.alpha0, step, R.total_steps); }
    bool revive_due() const { return (step + 1) % CHECK_UNUSED_EVERY == 0 && R.unused_threshold > 0.0f; }

    // Step, batch position and epoch bookkeeping after a batch.
    void advance() {
        step++;
        b0 += nb;
        if (b0 >= R.epoch_rows[epoch]) {
            if (!deferred && out->n_objective < EMH_MAX_EPOCHS) out->objective[out->n_objective++] = epoch_obj;
            epoch++;
            b0 = 0;
            if (epoch == R.n_epochs) done = true;
        }
    }
};

}  // namespace emh
