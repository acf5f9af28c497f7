// The e-means training driver over the HIP assigner.
//
// The recipe stays on the host exactly as the reference crate has it
// (src/lib.rs:165-491): recipe laws, xoshiro256++ batch order, step
// schedule, balance bias and the revival plan. What moves to the device is
// the data path of each batch and the model it updates: the rows are
// uploaded once and per epoch only its batch order (4 bytes per row); the
// gather, the search, the per-cluster sums/counts and the EMA step (f32
// operations in the host loop's order, so bit-identical) run on the GPU
// against device-resident centroids; the last pass's totals accumulate in
// FP64 on the GPU for the closing mean. Each batch's ids are a slice of
// the uploaded order; labels and distances come down per batch in one copy
// (objective, last-pass labels), the usage vector only when the bias or a
// revival needs it; without the balance penalty the host folds them in a
// few batches later, so staging and launching overlap the GPU's work.
#include <atomic>
#include <chrono>
#include <thread>

#include "emh_recipe.h"

namespace emh {

// x: host rows (uploaded once here) or, with rows_on_device, rows already in
// HBM (no copy; the initial centroids are gathered on the device).
static int train_impl(emh_assigner *h, const void *x, int rows, size_t n, size_t d, size_t k, const emh_config &cfg,
                      float *centroids_out, uint32_t *labels_out, emh_train_out *out, bool rows_on_device) {
    JobRun J;
    int rc;
    // the steps' searches read the device model in place (FP32 / I8 / I16):
    // the assigner must not keep that pointer once the model is freed
    struct Unborrow {
        emh_assigner *h;
        ~Unborrow() { h->cen_ref = nullptr; }
    } unborrow{h};
    if ((rc = J.init(x, rows, n, d, k, cfg, centroids_out, labels_out, out, rows_on_device))) return rc;
    if ((h->op == OP_I8 || h->op == OP_I16) && rows != EMH_ROWS_I8)
        return fail(EMH_E_UNSUPPORTED, "the integer search takes i8 rows only (AssignRows<i8>)");
    const size_t batch = J.R.batch_size;
    const size_t rb = rows == EMH_ROWS_F32 ? 4 : 1;
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = check(hipSetDevice(h->device), "hipSetDevice"))) return rc;

    // Device-resident rows, model (centroids k x d, usage k) and per-batch
    // buffers. The EMA step runs on the device; the host keeps the recipe
    // and pulls the usage vector down when the balance bias or a revival
    // needs it.
    DevBuf data, ids_d, out_d, sums_d, counts_d, pass_d, passn_d, cent_d, usage_d, pairs_d;
    if (!rows_on_device &&
        ((rc = check(data.ensure(n * d * rb), "resident rows")) ||
         (rc = check(hipMemcpyAsync(data.p, x, n * d * rb, hipMemcpyHostToDevice, h->stream), "upload rows"))))
        return rc;
    const void *rows_dev = rows_on_device ? x : data.p;
    size_t epoch_max = batch;
    for (size_t e : J.R.epoch_rows) epoch_max = std::max(epoch_max, e);
    if ((rc = check(ids_d.ensure(epoch_max * 4), "epoch order"))) return rc;
    // labels [0, batch) then distances [batch, 2 batch): one download a batch
    if ((rc = check(out_d.ensure(2 * batch * 4), "labels and distances"))) return rc;
    uint32_t *lab_dev = out_d.as<uint32_t>();
    float *dist_dev = out_d.as<float>() + batch;
    if ((rc = check(sums_d.ensure(k * d * 4), "sums"))) return rc;
    if ((rc = check(counts_d.ensure(k * 4), "counts"))) return rc;
    if ((rc = check(pass_d.ensure(k * d * 8), "pass sums"))) return rc;
    if ((rc = check(passn_d.ensure(k * 8), "pass counts"))) return rc;
    if ((rc = check(cent_d.ensure(k * d * 4), "centroids"))) return rc;
    if ((rc = check(usage_d.ensure(k * 4), "usage"))) return rc;
    if ((rc = check(pairs_d.ensure(k * 8), "revival pairs"))) return rc;
    HostBuf ids_h, out_h, usage_h, pairs_h, bias_h;
    if ((rc = check(bias_h.alloc(k * 4), "pinned bias"))) return rc;
    if ((rc = check(ids_h.alloc(epoch_max * 4), "pinned epoch order"))) return rc;
    if ((rc = check(out_h.alloc(2 * batch * 4), "pinned labels and distances"))) return rc;
    if ((rc = check(usage_h.alloc(k * 4), "pinned usage"))) return rc;
    if ((rc = check(pairs_h.alloc(k * 8), "pinned pairs"))) return rc;
    memcpy(usage_h.p, J.st.usage.data(), k * 4);
    DevBuf init_d;
    if (rows_on_device) {
        // the initial centroids: rows J.init_rows of the device table
        HostBuf init_h;
        if ((rc = check(init_h.alloc(k * 4), "pinned initial rows")) ||
            (rc = check(init_d.ensure(k * 4), "initial rows")))
            return rc;
        memcpy(init_h.p, J.init_rows.data(), k * 4);
        if ((rc = check(hipMemcpyAsync(init_d.p, init_h.p, k * 4, hipMemcpyHostToDevice, h->stream), "initial rows")) ||
            (rc = check(launch_gather_rows(x, rows, d, init_d.as<uint32_t>(), k, cent_d.as<float>(), h->stream),
                        "initial centroids")) ||
            (rc = check(hipStreamSynchronize(h->stream), "initial centroids")))  // before init_h is freed
            return rc;
    } else if ((rc = check(hipMemcpyAsync(cent_d.p, J.st.centroids.data(), k * d * 4, hipMemcpyHostToDevice,
                                          h->stream),
                           "upload centroids"))) {
        return rc;
    }
    if ((rc = check(hipMemcpyAsync(usage_d.p, usage_h.p, k * 4, hipMemcpyHostToDevice, h->stream), "upload usage")))
        return rc;
    std::vector<std::pair<size_t, size_t>> pairs;

    // Without the balance penalty no batch waits for the previous one's
    // labels: each batch's labels and distances come down into one of NSLOT
    // pinned slots behind an event, and a helper thread folds them into the
    // statistics in batch order (the objective's f64 chain and the last
    // pass's labels) while this thread stages and launches the next
    // batches; a slot is reused once its batch is folded.
    const bool async = cfg.balance <= 0.0f;
    J.deferred = async;
    constexpr int NSLOT = 4;
    HostBuf out_s[NSLOT];
    struct Events {
        hipEvent_t e[NSLOT] = {}, order = nullptr;  // order: the epoch order's upload
        ~Events() {
            for (hipEvent_t v : e)
                if (v) (void)hipEventDestroy(v);
            if (order) (void)hipEventDestroy(order);
        }
    } ev;
    JobRun::StepMeta meta[NSLOT];
    if (async)
        for (int i = 0; i < NSLOT; i++)
            if ((rc = check(out_s[i].alloc(2 * batch * 4), "pinned labels and distances")) ||
                (rc = check(hipEventCreateWithFlags(&ev.e[i], hipEventDisableTiming), "event")))
                return rc;
    if ((rc = check(hipEventCreateWithFlags(&ev.order, hipEventDisableTiming), "event"))) return rc;
    bool order_pending = false, counts_cleared = false;
    // posted: batches whose slot is queued; folded: batches the helper has
    // taken into the statistics (both count from the first batch)
    struct Folder {
        std::atomic<uint64_t> posted{0}, folded{0};
        std::atomic<bool> stop{false};
        std::atomic<int> err{0};
        std::thread th;
        ~Folder() {
            stop.store(true, std::memory_order_release);
            if (th.joinable()) th.join();
        }
    } fold;
    if (async) {
        const int dev = h->device;
        fold.th = std::thread([&, dev] {
            (void)hipSetDevice(dev);
            uint64_t next = 0;
            for (;;) {
                while (fold.posted.load(std::memory_order_acquire) == next) {
                    if (fold.stop.load(std::memory_order_acquire) &&
                        fold.posted.load(std::memory_order_acquire) == next)
                        return;
                    std::this_thread::yield();
                }
                const int sl = (int)(next % NSLOT);
                if (hipEventSynchronize(ev.e[sl]) != hipSuccess) fold.err.store(1, std::memory_order_release);
                J.finish_stats_deferred(meta[sl], out_s[sl].as<uint32_t>(), out_s[sl].as<float>() + batch);
                fold.folded.store(++next, std::memory_order_release);
            }
        });
    }
    // waits until the helper has folded the first `count` batches
    auto folded_up_to = [&](uint64_t count) -> int {
        while (fold.folded.load(std::memory_order_acquire) < count) std::this_thread::yield();
        return fold.err.load(std::memory_order_acquire) ? fail(EMH_E_DEVICE, "batch statistics: a step failed")
                                                        : (int)EMH_OK;
    };

    while (J.next()) {
        const size_t nb = J.nb;
        const int sl = (int)(J.step % NSLOT);
        // slot sl held batch step - NSLOT: reused once that batch is folded
        if (async && J.step >= (uint64_t)NSLOT && (rc = folded_up_to(J.step - NSLOT + 1))) return rc;
        if (J.last_epoch_start) {
            if ((rc = check(launch_fill_f64(pass_d.as<double>(), k * d, h->stream), "clear pass sums"))) return rc;
            if ((rc = check(launch_fill_f64((double *)passn_d.p, k, h->stream), "clear pass counts"))) return rc;
        }
        // One batch against a fixed centroid snapshot (src/lib.rs:408-417):
        // the searches without a rescore read the model in place (the EMA
        // changes it only after they have run, in stream order)
        if ((rc = set_centroids_impl(h, cent_d.as<float>(), k, d, true, true))) return rc;
        if (J.penalized) {
            // pinned and stream-ordered: the step ends with a stream sync
            // before the next bias is staged (penalized runs are not async)
            memcpy(bias_h.p, J.bias.data(), k * 4);
            if ((rc = set_bias_impl(h, bias_h.as<float>(), k, false))) return rc;
        }
        if (J.b0 == 0) {
            // The epoch's batch order goes up once; batch ids are slices of
            // it. The device copy is stream-ordered behind the last epoch's
            // readers; the pinned staging waits for its previous upload.
            const size_t epoch_n = J.R.epoch_rows[J.epoch];
            if (order_pending && (rc = check(hipEventSynchronize(ev.order), "epoch order"))) return rc;
            memcpy(ids_h.p, J.order.data(), epoch_n * 4);
            if ((rc = check(hipMemcpyAsync(ids_d.p, ids_h.p, epoch_n * 4, hipMemcpyHostToDevice, h->stream),
                            "copy epoch order")) ||
                (rc = check(hipEventRecord(ev.order, h->stream), "event")))
                return rc;
            order_pending = true;
        }
        const uint32_t *ids_dev = ids_d.as<uint32_t>() + J.b0;
        // gather + search on the device (src/lib.rs:419-424)
        if ((rc = run_search(h, rows_dev, rows, nb, ids_dev, lab_dev, dist_dev)))
            return rc;
        // per-cluster counts and sums in batch order (src/lib.rs:431-441)
        // after the first step the EMA step leaves the counts and the
        // big-cluster count at zero (no clearing launch)
        if ((rc = check(launch_accumulate(h->acc, rows_dev, rows, nb, d, ids_dev, lab_dev, k,
                                          sums_d.as<float>(), counts_d.as<uint32_t>(), h->stream, counts_cleared),
                        "accumulate")))
            return rc;
        // the EMA step (src/lib.rs:455-456); in the last epoch the same
        // launch first adds the batch's sums and counts to the FP64 pass
        // totals (src/lib.rs:446-454)
        if ((rc = check(launch_ema_apply(cent_d.as<float>(), usage_d.as<float>(), sums_d.as<float>(),
                                         counts_d.as<uint32_t>(), k, d, J.alpha(), h->stream,
                                         counts_d.as<uint32_t>(), h->acc.big,
                                         J.is_last ? pass_d.as<double>() : nullptr,
                                         J.is_last ? passn_d.as<uint64_t>() : nullptr),
                        "EMA step")))
            return rc;
        counts_cleared = true;
        const bool revive = J.revive_due();
        const bool need_usage = revive || cfg.balance > 0.0f;
        if (async) {
            if ((rc = check(hipMemcpyAsync(out_s[sl].p, out_d.p, (batch + nb) * 4, hipMemcpyDeviceToHost, h->stream),
                            "labels and distances")) ||
                (rc = check(hipEventRecord(ev.e[sl], h->stream), "event")))
                return rc;
            meta[sl] = J.meta();
            fold.posted.store(J.step + 1, std::memory_order_release);
            // the revival plan needs this step's masses, nothing of the
            // batches still pending: those fold in later as usual
            if (revive &&
                ((rc = check(hipMemcpyAsync(usage_h.p, usage_d.p, k * 4, hipMemcpyDeviceToHost, h->stream), "usage")) ||
                 (rc = check(hipStreamSynchronize(h->stream), "batch"))))
                return rc;
            if (revive) memcpy(J.st.usage.data(), usage_h.p, k * 4);
        } else {
            if ((rc = check(hipMemcpyAsync(out_h.p, out_d.p, (batch + nb) * 4, hipMemcpyDeviceToHost, h->stream),
                            "labels and distances")) ||
                (need_usage &&
                 (rc = check(hipMemcpyAsync(usage_h.p, usage_d.p, k * 4, hipMemcpyDeviceToHost, h->stream), "usage"))) ||
                (rc = check(hipStreamSynchronize(h->stream), "batch")))
                return rc;
            if (need_usage) memcpy(J.st.usage.data(), usage_h.p, k * 4);
            J.finish_stats(out_h.as<uint32_t>(), out_h.as<float>() + batch);
        }
        if (revive) {
            // src/lib.rs:463-465: planned on the host (its RNG), rows split on the device
            J.st.revive_plan(J.R.unused_threshold, J.split_rng, pairs);
            if (!pairs.empty()) {
                uint32_t *pp = pairs_h.as<uint32_t>();
                for (size_t t = 0; t < pairs.size(); t++) {
                    pp[2 * t] = (uint32_t)pairs[t].first;
                    pp[2 * t + 1] = (uint32_t)pairs[t].second;
                }
                memcpy(usage_h.p, J.st.usage.data(), k * 4);
                // stream-ordered: the pinned staging is next written after
                // the next revival's synchronisation, when these copies are
                // long done
                if ((rc = check(hipMemcpyAsync(pairs_d.p, pairs_h.p, pairs.size() * 8, hipMemcpyHostToDevice, h->stream),
                                "revival pairs")) ||
                    (rc = check(hipMemcpyAsync(usage_d.p, usage_h.p, k * 4, hipMemcpyHostToDevice, h->stream), "usage")) ||
                    (rc = check(launch_revive_split(cent_d.as<float>(), pairs_d.as<uint32_t>(), pairs.size(), d, h->stream),
                                "revival")))
                    return rc;
            }
        }
        J.advance();
    }
    if (async && (rc = folded_up_to(fold.posted.load(std::memory_order_acquire)))) return rc;

    // Closing mean on the device from the FP64 pass totals (src/lib.rs:481-489).
    if ((rc = check(launch_closing_mean(pass_d.as<double>(), passn_d.as<uint64_t>(), k, d, cent_d.as<float>(),
                                        h->stream),
                    "closing mean")) ||
        (rc = check(hipMemcpyAsync(centroids_out, cent_d.p, k * d * 4, hipMemcpyDeviceToHost, h->stream),
                    "copy centroids")) ||
        (rc = check(hipStreamSynchronize(h->stream), "closing mean")))
        return rc;
    out->n_splits = J.st.n_splits;
    out->recipe = J.R;
    out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return EMH_OK;
}

}  // namespace emh

using namespace emh;

extern "C" {

void emh_config_default(emh_config *cfg) {
    if (!cfg) return;
    memset(cfg, 0, sizeof *cfg);
    cfg->nepochs = 5.0;
    cfg->seed = 1234;
}

int emh_resolve_recipe(uint64_t n, uint64_t d, uint64_t k, const emh_config *cfg, emh_recipe *out) {
    if (!cfg || !out) return fail(EMH_E_ARG, "resolve_recipe: NULL argument");
    memset(out, 0, sizeof *out);
    if (resolve(n, d, k, *cfg, *out)) return fail(EMH_E_ARG, "nepochs must be in (0, 1023)");
    return EMH_OK;
}

int emh_train(emh_assigner *h, const void *x, int rows, size_t n, size_t d, size_t k, const emh_config *cfg,
              float *centroids_out, uint32_t *labels_out, emh_train_out *out) {
    if (!h || !cfg) return fail(EMH_E_ARG, "train: NULL argument");
    return train_impl(h, x, rows, n, d, k, *cfg, centroids_out, labels_out, out, false);
}

int emh_train_device(emh_assigner *h, const void *x_dev, int rows, size_t n, size_t d, size_t k,
                     const emh_config *cfg, float *centroids_out, uint32_t *labels_out, emh_train_out *out) {
    if (!h || !cfg) return fail(EMH_E_ARG, "train_device: NULL argument");
    return train_impl(h, x_dev, rows, n, d, k, *cfg, centroids_out, labels_out, out, true);
}

int emh_assign_all(emh_assigner *h, const void *x, int rows, size_t n, size_t d, const float *centroids, size_t k,
                   uint32_t *labels, float *dist) {
    // assign (src/lib.rs:512-520): one set_centroids, then the rows.
    if (!h) return fail(EMH_E_ARG, "assign: NULL handle");
    int rc;
    if ((rc = emh_set_centroids(h, centroids, k, d))) return rc;
    if (rows == EMH_ROWS_F32) return emh_assign_f32(h, (const float *)x, n, labels, dist);
    if (rows == EMH_ROWS_I8) return emh_assign_i8(h, (const int8_t *)x, n, labels, dist);
    if (rows == EMH_ROWS_U8) return emh_assign_u8(h, (const uint8_t *)x, n, labels, dist);
    return fail(EMH_E_ARG, "assign: unknown row type");
}

}  // extern "C"
