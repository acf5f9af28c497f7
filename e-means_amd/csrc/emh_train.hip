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
#include <map>
#include <set>
#include <thread>

#include "emh_recipe.h"

namespace emh {

// The step graph takes a search whose launches read nothing back to the
// host: the register-resident kernel, and for I8 offsets inside the packed
// key without the range flag (d 3 2^14 < 2^26). EMH_TRAIN_GRAPH=0 runs
// every step without it.
static bool step_graph_fits(int op, size_t k, size_t d) {
    const char *e = getenv("EMH_TRAIN_GRAPH");
    if (e && e[0] == '0') return false;
    if (!make_geom(op, k, d).reg) return false;
    return op != OP_I8 || (int64_t)d * 49152 < ((int64_t)1 << 26);
}

// Every step's batch position, EMA rate and last-epoch flag, in the order
// JobRun::next and JobRun::advance step through them (src/lib.rs:400-456).
static std::vector<StepParam> step_params(const emh_recipe &R) {
    std::vector<StepParam> P;
    uint64_t step = 0;
    for (uint32_t e = 0; e < R.n_epochs; e++) {
        size_t b0 = 0;
        do {
            StepParam p{};
            p.b0 = (uint32_t)b0;
            p.nb = (uint32_t)std::min<uint64_t>(R.batch_size, R.epoch_rows[e] - b0);
            p.alpha = (float)scheduled_alpha(R.alpha0, step, R.total_steps);
            p.decay = 1.0f - p.alpha;
            p.replace = p.alpha == 1.0f;
            p.is_last = e + 1 == R.n_epochs;
            p.revive = (step + 1) % CHECK_UNUSED_EVERY == 0 && R.unused_threshold > 0.0f;  // JobRun::revive_due
            P.push_back(p);
            b0 += p.nb;
            step++;
        } while (b0 < R.epoch_rows[e]);
    }
    if (P.empty()) P.push_back(StepParam{});
    return P;
}

// The assigner's device buffers a captured step reads or writes (operand
// tables, search and sort scratch). A buffer that grows is freed and
// allocated anew (DevBuf::ensure, accum_reserve), so a graph captured before
// holds a freed address: it is captured again once any of these has moved.
static std::vector<const void *> step_buffers(const emh_assigner *h) {
    return {h->cen_f32.p, h->cen_op.p,    h->norm32.p, h->norm64.p,    h->hn.p,       h->cn.p,
            h->cexp.p,    h->stats.p,     h->bias.p,   h->cand.p,      h->x8.p,       h->row_scale.p,
            h->row_norm.p, h->part.p,     h->acc.keys, h->acc.keys_alt, h->acc.vals,  h->acc.vals_alt,
            h->acc.big,   h->acc.sort_tmp, h->acc.starts};
}

struct StepGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const void *> bufs;  // step_buffers() when captured
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~StepGraph() { reset(); }
};

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

    // Without the balance penalty no batch waits for the previous one's
    // labels: each batch's labels and distances come down into one of NSLOT
    // pinned slots behind an event, and a helper thread folds them into the
    // statistics in batch order (the objective's f64 chain and the last
    // pass's labels) while this thread stages and launches the next
    // batches; a slot is reused once its batch is folded.
    const bool async = cfg.balance <= 0.0f;
    J.deferred = async;
    constexpr int NSLOT = 4;
    // Replayed steps. Without the balance penalty every batch size, once one
    // step of it has run kernel by kernel (choosing its search tiles), runs
    // as one captured graph: the host's cost of a step is then one graph
    // launch instead of a dozen kernel launches and a copy (each of those
    // costs the host about as long as the step's short kernels run). The
    // graph reads what changes between steps from the device: a step index
    // that its first kernel takes and advances, and per-step values (batch
    // position, EMA rate, last-epoch and revival flags) the recipe fixes
    // before the first step. Steps run without it advance the index too, so
    // the index is J.step.
    const bool graph_mode = async && !h->timing && step_graph_fits(h->op, k, d);
    const std::vector<StepParam> P = graph_mode ? step_params(J.R) : std::vector<StepParam>();

    // Device-resident rows, model (centroids k x d, usage k) and per-batch
    // buffers. The EMA step runs on the device; the host keeps the recipe
    // and pulls the usage vector down when the balance bias or a revival
    // needs it. The per-run buffers are segments of one device block and
    // one pinned block (an allocation costs the host about as long as a
    // step's kernels run).
    DevBuf data;
    if (!rows_on_device &&
        ((rc = check(data.ensure(n * d * rb), "resident rows")) ||
         (rc = check(hipMemcpyAsync(data.p, x, n * d * rb, hipMemcpyHostToDevice, h->stream), "upload rows"))))
        return rc;
    const void *rows_dev = rows_on_device ? x : data.p;
    size_t epoch_max = batch;
    for (size_t e : J.R.epoch_rows) epoch_max = std::max(epoch_max, e);
    Layout dl, hl;
    const size_t o_ids = dl.add(epoch_max * 4), o_out = dl.add(2 * batch * 4), o_sums = dl.add(k * d * 4),
                 o_counts = dl.add(k * 4), o_pass = dl.add(k * d * 8), o_passn = dl.add(k * 8),
                 o_cent = dl.add(k * d * 4), o_usage = dl.add(k * 4), o_init = dl.add(k * 4),
                 o_params = dl.add(P.size() * sizeof(StepParam)), o_ctr = dl.add(8), o_sids = dl.add(batch * 4);
    const size_t h_bias = hl.add(k * 4), h_ids = hl.add(epoch_max * 4), h_out = hl.add(2 * batch * 4),
                 h_usage = hl.add(k * 4), h_init = hl.add(k * 4), h_params = hl.add(P.size() * sizeof(StepParam));
    size_t h_slot[NSLOT];
    for (int i = 0; i < NSLOT; i++) h_slot[i] = async ? hl.add((2 * batch + k) * 4) : 0;
    DevBuf dblock;
    HostBuf hblock, pairs_h;
    if ((rc = check(dblock.ensure(dl.bytes), "training buffers")) ||
        (rc = check(hblock.alloc(hl.bytes), "pinned training buffers")))
        return rc;
    auto dseg = [&](size_t o) { return Seg{(uint8_t *)dblock.p + o}; };
    auto hseg = [&](size_t o) { return Seg{(uint8_t *)hblock.p + o}; };
    // labels [0, batch) then distances [batch, 2 batch): one download a batch
    const Seg ids_d = dseg(o_ids), out_d = dseg(o_out), sums_d = dseg(o_sums), counts_d = dseg(o_counts),
              pass_d = dseg(o_pass), passn_d = dseg(o_passn), cent_d = dseg(o_cent), usage_d = dseg(o_usage),
              init_d = dseg(o_init), params_d = dseg(o_params), ctr_d = dseg(o_ctr), step_ids_d = dseg(o_sids);
    const Seg bias_h = hseg(h_bias), ids_h = hseg(h_ids), out_h = hseg(h_out), usage_h = hseg(h_usage),
              init_h = hseg(h_init), params_h = hseg(h_params);
    Seg out_s[NSLOT];
    for (int i = 0; i < NSLOT; i++) out_s[i] = hseg(h_slot[i]);
    uint32_t *lab_dev = out_d.as<uint32_t>();
    float *dist_dev = out_d.as<float>() + batch;
    // revival pairs: read in place by the split kernel (no upload)
    uint32_t *pairs_dev = nullptr;
    if ((rc = check(pairs_h.alloc_coherent(k * 8), "pinned pairs")) ||
        (rc = check(hipHostGetDevicePointer((void **)&pairs_dev, pairs_h.p, 0), "pinned pairs")))
        return rc;
    memcpy(usage_h.p, J.st.usage.data(), k * 4);
    if (rows_on_device) {
        // the initial centroids: rows J.init_rows of the device table
        memcpy(init_h.p, J.init_rows.data(), k * 4);
        if ((rc = check(hipMemcpyAsync(init_d.p, init_h.p, k * 4, hipMemcpyHostToDevice, h->stream), "initial rows")) ||
            (rc = check(launch_gather_rows(x, rows, d, init_d.as<uint32_t>(), k, cent_d.as<float>(), h->stream),
                        "initial centroids")))
            return rc;
    } else if ((rc = check(hipMemcpyAsync(cent_d.p, J.st.centroids.data(), k * d * 4, hipMemcpyHostToDevice,
                                          h->stream),
                           "upload centroids"))) {
        return rc;
    }
    if ((rc = check(hipMemcpyAsync(usage_d.p, usage_h.p, k * 4, hipMemcpyHostToDevice, h->stream), "upload usage")))
        return rc;
    std::vector<std::pair<size_t, size_t>> pairs;
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
            if ((rc = check(hipEventCreateWithFlags(&ev.e[i], hipEventDisableTiming), "event")))
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

    std::map<size_t, StepGraph> graphs;  // by batch rows: the full batch, the short one closing an epoch
    std::set<size_t> ran;                // batch sizes one step has run without a graph (tiles chosen)
    StepSlots slots_dev{};
    if (graph_mode) {
        memcpy(params_h.p, P.data(), P.size() * sizeof(StepParam));
        if ((rc = check(hipMemcpyAsync(params_d.p, params_h.p, P.size() * sizeof(StepParam), hipMemcpyHostToDevice,
                                       h->stream),
                        "step table")) ||
            (rc = check(launch_fill_u32(ctr_d.as<uint32_t>(), 2, 0u, h->stream), "step index")))
            return rc;
        slots_dev.n = NSLOT;
        for (int i = 0; i < NSLOT; i++) slots_dev.p[i] = out_s[i].as<uint32_t>();
    }

    while (J.next()) {
        const size_t nb = J.nb;
        const int sl = (int)(J.step % NSLOT);
        const bool replay = graph_mode && ran.count(nb) != 0;
        // slot sl held batch step - NSLOT: reused once that batch is folded
        if (async && J.step >= (uint64_t)NSLOT && (rc = folded_up_to(J.step - NSLOT + 1))) return rc;
        if (J.last_epoch_start) {
            if ((rc = check(launch_fill_f64(pass_d.as<double>(), k * d, h->stream), "clear pass sums"))) return rc;
            if ((rc = check(launch_fill_f64((double *)passn_d.p, k, h->stream), "clear pass counts"))) return rc;
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
        if (replay) {
            StepGraph &sg = graphs[nb];
            if (sg.exec && sg.bufs != step_buffers(h)) {
                // a buffer it addresses has moved since the capture (another
                // batch size chose a wider search split, say): every posted
                // step, replays of it included, is done once folded
                if ((rc = folded_up_to(fold.posted.load(std::memory_order_acquire)))) return rc;
                sg.reset();
            }
            if (!sg.exec) {
                // the step as one graph: stage (step index, batch ids),
                // centroid prep, search, per-cluster sums, EMA, labels and
                // distances into the step's pinned slot. The helper thread's
                // waits on events of this stream would void the capture:
                // every posted batch is folded first (once per run).
                if ((rc = folded_up_to(fold.posted.load(std::memory_order_acquire))) ||
                    (rc = check(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "begin capture")))
                    return rc;
                uint32_t *sids = step_ids_d.as<uint32_t>();
                uint32_t *ctr = ctr_d.as<uint32_t>();
                int crc = check(launch_step_stage(ids_d.as<uint32_t>(), params_d.as<StepParam>(), ctr, sids, nb,
                                                  h->stream),
                                "step stage");
                if (!crc) crc = set_centroids_impl(h, cent_d.as<float>(), k, d, true, true);
                if (!crc) crc = run_search(h, rows_dev, rows, nb, sids, lab_dev, dist_dev);
                if (!crc)
                    crc = check(launch_accumulate(h->acc, rows_dev, rows, nb, d, sids, lab_dev, k, sums_d.as<float>(),
                                                  counts_d.as<uint32_t>(), h->stream, true),
                                "accumulate");
                if (!crc)
                    crc = check(launch_step_finish(cent_d.as<float>(), usage_d.as<float>(), sums_d.as<float>(),
                                                   counts_d.as<uint32_t>(), k, d, params_d.as<StepParam>(), ctr,
                                                   h->acc.big, pass_d.as<double>(), passn_d.as<uint64_t>(), lab_dev,
                                                   dist_dev, nb, batch, slots_dev, h->stream),
                                "EMA step");
                const hipError_t ee = hipStreamEndCapture(h->stream, &sg.graph);
                if (crc) return crc;
                if ((rc = check(ee, "end capture")) ||
                    (rc = check(hipGraphInstantiate(&sg.exec, sg.graph, nullptr, nullptr, 0), "instantiate step")))
                    return rc;
                sg.bufs = step_buffers(h);
            }
            if ((rc = check(hipGraphLaunch(sg.exec, h->stream), "step launch"))) return rc;
        } else {
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
        if (graph_mode && (rc = check(launch_step_advance(ctr_d.as<uint32_t>(), h->stream), "step index"))) return rc;
        if (graph_mode) ran.insert(nb);
        }
        counts_cleared = true;
        const bool revive = J.revive_due();
        const bool need_usage = revive || cfg.balance > 0.0f;
        if (async) {
            // (a replayed step's own last kernel has written its slot)
            if ((!replay && (rc = check(hipMemcpyAsync(out_s[sl].p, out_d.p, (batch + nb) * 4, hipMemcpyDeviceToHost,
                                                      h->stream),
                                        "labels and distances"))) ||
                (rc = check(hipEventRecord(ev.e[sl], h->stream), "event")))
                return rc;
            meta[sl] = J.meta();
            fold.posted.store(J.step + 1, std::memory_order_release);
            // the revival plan needs this step's masses, nothing of the
            // batches still pending: those fold in later as usual (a
            // replayed step has put them in its slot)
            if (revive && replay) {
                if ((rc = check(hipEventSynchronize(ev.e[sl]), "batch"))) return rc;
                memcpy(J.st.usage.data(), out_s[sl].as<float>() + 2 * batch, k * 4);
            } else if (revive) {
                if ((rc = check(hipMemcpyAsync(usage_h.p, usage_d.p, k * 4, hipMemcpyDeviceToHost, h->stream),
                                "usage")) ||
                    (rc = check(hipStreamSynchronize(h->stream), "batch")))
                    return rc;
                memcpy(J.st.usage.data(), usage_h.p, k * 4);
            }
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
                // the split kernel reads the pairs in place and halves the
                // donors' masses on the device as the plan did on the host
                // (the same f32 division of the same values). The pinned
                // pairs are next written after the next revival check's
                // synchronisation, when this kernel is long done.
                if ((rc = check(launch_revive_split(cent_d.as<float>(), pairs_dev, pairs.size(), d, h->stream,
                                                    usage_d.as<float>()),
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
