// Batched training: many independent e-means problems in lockstep.
//
// Each job keeps the host recipe of src/lib.rs:348-491 exactly (JobRun);
// what is shared is the device step. Every job's model (centroids, cluster
// masses) stays on the device, every job's batch of one training step is a
// slot range of one concatenated batch and its clusters a range of one
// concatenated cluster list, and the step (upload of the job table, batch
// ids and biases; prep of every job's centroid operand; the search per
// row-width class; the FP32 rescore (BF16) or the merge of a split search
// (FP32, I8); the per-cluster sums with one stable
// sort; the EMA step; download of labels, distances and masses) is captured
// once as a hipGraph and replayed once per step. The kernels read the step's
// row counts, rates and flags from the device job table that the replay's
// first copy refreshes, so the graph never changes; jobs that have finished
// keep their slots and idle. Revival splits run between replays.
#include <atomic>
#include <chrono>
#include <functional>
#include <future>
#include <thread>

#include "emh_recipe.h"

namespace emh {

namespace {

// Runs fn(0..n-1) on a few persistent host threads plus the caller; each
// index runs exactly once, and run() returns when all have. Between runs the
// workers poll (yielding the core) instead of sleeping: a step's host halves
// come tens of microseconds apart, less than a sleeping thread takes to wake.
// A run is handed to worker w by storing its number in w's `go` (release)
// after the run's state is written; w reports it with `done` (release), and
// run() returns only when every worker has, so no worker is still inside a
// run when the next one rewrites the state.
class HostPool {
  public:
    explicit HostPool(unsigned nthreads) : w_(nthreads) {
        for (unsigned t = 0; t < nthreads; t++) th_.emplace_back([this, t] { work(w_[t]); });
    }
    ~HostPool() {
        stop_.store(true, std::memory_order_relaxed);
        gen_++;
        for (auto &w : w_) w.go.store(gen_, std::memory_order_release);
        for (auto &t : th_) t.join();
    }
    void run(size_t n, const std::function<void(size_t)> &fn) {
        fn_ = &fn;
        n_ = n;
        next_.store(0, std::memory_order_relaxed);
        gen_++;
        for (auto &w : w_) w.go.store(gen_, std::memory_order_release);
        drain();
        for (auto &w : w_)
            while (w.done.load(std::memory_order_acquire) != gen_) std::this_thread::yield();
    }

  private:
    struct alignas(64) Slot {
        std::atomic<uint64_t> go{0}, done{0};
    };
    void drain() {
        for (;;) {
            const size_t i = next_.fetch_add(1, std::memory_order_relaxed);
            if (i >= n_) return;
            (*fn_)(i);
        }
    }
    void work(Slot &w) {
        uint64_t seen = 0;
        for (;;) {
            uint64_t g;
            while ((g = w.go.load(std::memory_order_acquire)) == seen) std::this_thread::yield();
            seen = g;
            if (stop_.load(std::memory_order_relaxed)) return;
            drain();
            w.done.store(g, std::memory_order_release);
        }
    }
    std::vector<Slot> w_;
    std::vector<std::thread> th_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0;
    std::atomic<size_t> next_{0};
    uint64_t gen_ = 0;
    std::atomic<bool> stop_{false};
};

// A second stream and the events that fork and join it inside a capture.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t e[4] = {};
    hipError_t init() {
        hipError_t r = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        for (int i = 0; i < 4 && r == hipSuccess; i++) r = hipEventCreateWithFlags(&e[i], hipEventDisableTiming);
        return r;
    }
    ~SideStream() {
        for (hipEvent_t v : e)
            if (v) (void)hipEventDestroy(v);
        if (s) (void)hipStreamDestroy(s);
    }
};

struct GraphGuard {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphGuard() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

}  // namespace

static int batch_impl(emh_assigner *h, emh_job *jobs, size_t njobs, emh_batch_stats *stats) {
    if (!jobs || njobs == 0) return fail(EMH_E_ARG, "train_batch: need at least one job");
    if (njobs > 0xFFFFu) return fail(EMH_E_ARG, "train_batch: at most 65535 jobs per call");
    if (h->op != OP_F32 && h->op != OP_BF16 && h->op != OP_FP8 && h->op != OP_I8 && h->op != OP_I16)
        return fail(EMH_E_UNSUPPORTED, "train_batch: unknown search precision");
    const int op = h->op, M = (op == OP_BF16 || op == OP_FP8) ? h->M : 1;
    const bool cand_search = op == OP_BF16 || op == OP_FP8;
    // the I8 search takes i8 rows (AssignRows<i8>), the float searches f32
    const bool int_search = op == OP_I8 || op == OP_I16;
    const int rows = int_search ? EMH_ROWS_I8 : EMH_ROWS_F32;
    const size_t rb = int_search ? 1 : 4;
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    if ((rc = check(hipSetDevice(h->device), "hipSetDevice"))) return rc;

    // Host threads for the jobs' host halves (each job's recipe state is its
    // own): the caller plus up to 15 workers.
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    HostPool pool(std::min<unsigned>({hw, 16u, (unsigned)njobs}) - 1);

    // Host runners (set up in parallel: the initial rows come from a
    // permutation of each job's rows), geometry and slot ranges.
    std::vector<JobRun> run(njobs);
    std::vector<int> init_rc(njobs, EMH_OK);
    std::vector<std::string> init_err(njobs);
    const std::function<void(size_t)> init_job = [&](size_t j) {
        const emh_job &jb = jobs[j];
        if (jb.rows != rows) {
            init_rc[j] = EMH_E_UNSUPPORTED;
            init_err[j] = int_search ? "train_batch: the integer searches take i8 rows only"
                                     : "train_batch: the float searches take f32 rows only";
            return;
        }
        init_rc[j] = run[j].init(jb.x, jb.rows, jb.n, jb.d, jb.k, jb.cfg, jb.centroids_out, jb.labels_out, jb.out,
                                 jb.x_on_device != 0);
        if (init_rc[j] != EMH_OK) init_err[j] = emh_last_error();  // this thread's message
    };
    pool.run(njobs, init_job);
    for (size_t j = 0; j < njobs; j++)
        if (init_rc[j] != EMH_OK) return fail(init_rc[j], init_err[j]);
    std::vector<SearchGeom> geo(njobs);
    std::vector<size_t> row_base(njobs), cl_base(njobs), cen_base(njobs), cpad_base(njobs), op_base(njobs),
        part_base(njobs), nrow_base(njobs), x8_base(njobs);
    std::vector<JobSlices> slices(njobs);
    size_t slots = 0, nclusters = 0, cen_floats = 0, cpad = 0, op_bytes = 0, part_words = 0, rows_total = 0, d_max = 0,
           x8_bytes = 0;
    for (size_t j = 0; j < njobs; j++) {
        const emh_job &jb = jobs[j];
        geo[j] = make_geom(op, jb.k, jb.d);
        if (op == OP_I16) {
            // I16 jobs run the LDS kernel's exact i64 scores, d streamed in
            // chunks (as a single problem's rows past 512 bytes)
            geo[j].reg = false;
            geo[j].resident = false;
            geo[j].row_bytes = CHUNK_BYTES;
        }
        d_max = std::max(d_max, (size_t)jb.d);
        row_base[j] = slots;
        slots += run[j].R.batch_size;
        nrow_base[j] = rows_total;
        rows_total += jb.n;
        cl_base[j] = nclusters;
        nclusters += jb.k;
        cen_base[j] = cen_floats;
        cen_floats += jb.k * jb.d;
        cpad_base[j] = cpad;
        cpad += geo[j].k_pad;
        op_base[j] = op_bytes;
        op_bytes += geo[j].reg ? geo[j].k_pad * (size_t)geo[j].row_bytes : operand_table_bytes(geo[j]);
        // jobs with operand rows past 512 bytes run the LDS kernel, unsplit
        slices[j] = geo[j].reg ? job_slices(geo[j]) : JobSlices{1u, 1u};
        part_base[j] = part_words;
        part_words += run[j].R.batch_size * job_part_words(op, M, slices[j].nslice);
        // FP8: the step's encoded batch rows, d_pad bytes a slot
        x8_base[j] = x8_bytes;
        if (op == OP_FP8) x8_bytes += run[j].R.batch_size * geo[j].d_pad;
    }
    if (slots > 0xFFFFFFFFull || nclusters >= 0xFFFFFFFFull)
        return fail(EMH_E_ARG, "train_batch: more than 2^32 - 1 batch rows or clusters in one call");

    // Resident rows, one table per job (the caller's own when on the device).
    std::vector<DevBuf> rows_d(njobs);
    std::vector<const void *> rows_ptr(njobs);
    for (size_t j = 0; j < njobs; j++) {
        if (jobs[j].x_on_device) {
            rows_ptr[j] = jobs[j].x;
            continue;
        }
        const size_t bytes = jobs[j].n * jobs[j].d * rb;
        if ((rc = check(rows_d[j].ensure(bytes), "resident rows"))) return rc;
        if ((rc = check(hipMemcpyAsync(rows_d[j].p, jobs[j].x, bytes, hipMemcpyHostToDevice, h->stream), "upload rows")))
            return rc;
        rows_ptr[j] = rows_d[j].p;
    }

    // Staged per-step input (host pinned -> device, one copy) and output
    // (device -> host pinned, one copy), same layout on both sides. The
    // output block's mass region is also the device-resident masses.
    const size_t sa_bytes = search_args_bytes();
    Layout in;
    const size_t in_job = in.add(njobs * sizeof(JobDev)), in_sa = in.add(njobs * sa_bytes),
                 in_bias = in.add(nclusters * 4), in_ids = in.add(slots * 4);
    Layout outl;
    const size_t out_dist = outl.add(slots * 4), out_use = outl.add(nclusters * 4);
    DevBuf in_d, out_d, cen_model, sums_d, counts_d, cen_op, norm32, hn, cn, stats_d, cand, keys, pass_sums,
        pass_counts, maps, pairs_d, part_d, lab_d, labels_all, x8_d, rscale_d, rnorm_d, cexp_d;
    HostBuf in_h, out_h, pairs_h;
    if ((rc = check(in_d.ensure(in.bytes), "step input")) || (rc = check(out_d.ensure(outl.bytes), "step output")) ||
        (rc = check(in_h.alloc(in.bytes), "pinned step input")) ||
        (rc = check(out_h.alloc(outl.bytes), "pinned step output")) ||
        (rc = check(pairs_h.alloc(nclusters * 12), "pinned revival triples")) ||
        (rc = check(pairs_d.ensure(nclusters * 12), "revival triples")) ||
        (rc = check(cen_model.ensure(cen_floats * 4), "centroids")) ||
        (rc = check(sums_d.ensure(cen_floats * 4), "sums")) || (rc = check(counts_d.ensure(nclusters * 4), "counts")) ||
        (rc = check(cen_op.ensure(op_bytes), "operand tables")) || (rc = check(norm32.ensure(cpad * 4), "norms")) ||
        (rc = check(hn.ensure(cpad * 4), "offsets")) ||
        (int_search && (rc = check(cn.ensure(cpad * 8), "integer offsets"))) ||
        (op == OP_FP8 && ((rc = check(x8_d.ensure(x8_bytes), "fp8 rows")) ||
                          (rc = check(cexp_d.ensure(cpad), "fp8 centroid scales")) ||
                          (rc = check(rscale_d.ensure(slots * 4), "fp8 row scales")) ||
                          (rc = check(rnorm_d.ensure(slots * 4), "fp8 row norms")))) ||
        (rc = check(stats_d.ensure(njobs * 16), "stats")) ||
        (rc = check(cand.ensure(slots * 4 * (size_t)M), "candidates")) ||
        (rc = check(keys.ensure(slots * 4), "keys")) || (rc = check(pass_sums.ensure(cen_floats * 8), "pass sums")) ||
        (rc = check(pass_counts.ensure(nclusters * 8), "pass counts")) ||
        (part_words && (rc = check(part_d.ensure(part_words * 4), "slice results"))) ||
        (rc = check(lab_d.ensure(slots * 4), "labels")) ||
        (rc = check(labels_all.ensure(rows_total * 4), "last-pass labels")) ||
        (rc = check(launch_fill_u32(labels_all.as<uint32_t>(), rows_total, 0xFFFFFFFFu, h->stream), "labels")) ||
        // the step graph expects these at zero; its last kernel leaves them so
        (rc = check(launch_fill_u32(stats_d.as<uint32_t>(), 4 * njobs, 0u, h->stream), "stats")) ||
        (rc = check(launch_fill_u32(counts_d.as<uint32_t>(), nclusters, 0u, h->stream), "counts")))
        return rc;
    memset(in_h.p, 0, in.bytes);
    uint8_t *in_dp = (uint8_t *)in_d.p, *out_dp = (uint8_t *)out_d.p;
    uint8_t *in_hp = (uint8_t *)in_h.p, *out_hp = (uint8_t *)out_h.p;
    float *use_dev = (float *)(out_dp + out_use);
    // initial models: random data rows (JobRun::init), masses 1/65536
    // (device rows: gathered on the device from the rows JobRun::init picked)
    size_t ninit = 0;
    for (size_t j = 0; j < njobs; j++) ninit += run[j].init_rows.size();
    DevBuf init_d;
    HostBuf init_h;
    if (ninit && ((rc = check(init_d.ensure(ninit * 4), "initial rows")) ||
                  (rc = check(init_h.alloc(ninit * 4), "pinned initial rows"))))
        return rc;
    size_t ioff = 0;
    for (size_t j = 0; j < njobs; j++) {
        memcpy(out_hp + out_use + cl_base[j] * 4, run[j].st.usage.data(), jobs[j].k * 4);
        if (jobs[j].x_on_device) {
            memcpy(init_h.as<uint32_t>() + ioff, run[j].init_rows.data(), jobs[j].k * 4);
            ioff += jobs[j].k;
        } else if ((rc = check(hipMemcpyAsync(cen_model.as<float>() + cen_base[j], run[j].st.centroids.data(),
                                              jobs[j].k * jobs[j].d * 4, hipMemcpyHostToDevice, h->stream),
                               "upload centroids"))) {
            return rc;
        }
    }
    if (ninit) {
        if ((rc = check(hipMemcpyAsync(init_d.p, init_h.p, ninit * 4, hipMemcpyHostToDevice, h->stream), "initial rows")))
            return rc;
        ioff = 0;
        for (size_t j = 0; j < njobs; j++) {
            if (!jobs[j].x_on_device) continue;
            if ((rc = check(launch_gather_rows(jobs[j].x, rows, jobs[j].d, init_d.as<uint32_t>() + ioff,
                                               jobs[j].k, cen_model.as<float>() + cen_base[j], h->stream),
                            "initial centroids")))
                return rc;
            ioff += jobs[j].k;
        }
    }
    if ((rc = check(hipMemcpyAsync(use_dev, out_hp + out_use, nclusters * 4, hipMemcpyHostToDevice, h->stream),
                    "upload masses")))
        return rc;

    // Static maps: search workgroups, prep entries, batch slots, clusters.
    // One workgroup per (job, slice, 256 batch rows), every row width in the
    // same launch; the jobs whose slices are the most work (tiles x tile
    // centroids x row length) come first, so the longest sweeps start first.
    // Workgroups go to the 8 XCDs in turn (dispatch i runs on XCD i % 8), so
    // the list is dealt out in runs of XCD_RUN consecutive entries, one XCD
    // per run: the row blocks of one slice then share that XCD's L2 for the
    // slice's centroid tiles. Padding entries (first row past every batch)
    // exit at once.
    constexpr size_t XCDS = 8, XCD_RUN = 8;
    std::vector<uint2> wg_all, wg_lds, cmap(cpad), rmap(slots);
    size_t lds_max = 0;  // the LDS kernel's dynamic LDS for the widest-row job
    for (size_t j = 0; j < njobs; j++) {
        if (geo[j].reg) continue;
        lds_max = std::max(lds_max, search_lds_bytes(geo[j]));
        for (size_t r0 = 0; r0 < run[j].R.batch_size; r0 += BM) wg_lds.push_back(make_uint2((unsigned)j, (unsigned)r0));
    }
    {
        std::vector<size_t> order;
        for (size_t j = 0; j < njobs; j++)
            if (geo[j].reg) order.push_back(j);
        auto work = [&](size_t j) { return (size_t)slices[j].tiles_per_slice * reg_tile(geo[j]) * geo[j].d_pad; };
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work(a) > work(b); });
        std::vector<uint2> list;
        for (size_t j : order)
            for (uint32_t sl = 0; sl < slices[j].nslice; sl++)
                for (size_t r0 = 0; r0 < run[j].R.batch_size; r0 += 256)
                    list.push_back(make_uint2((unsigned)j, (unsigned)r0 | sl));
        std::vector<std::vector<uint2>> per(XCDS);
        for (size_t i = 0; i < list.size(); i++) per[(i / XCD_RUN) % XCDS].push_back(list[i]);
        size_t rounds = 0;
        for (const auto &v : per) rounds = std::max(rounds, v.size());
        const uint2 pad = make_uint2(0u, 0xFFFFFF00u);
        for (size_t t = 0; t < rounds; t++)
            for (size_t x = 0; x < XCDS; x++) wg_all.push_back(t < per[x].size() ? per[x][t] : pad);
    }
    std::vector<uint32_t> cjob(nclusters);
    for (size_t j = 0; j < njobs; j++) {
        for (size_t c = 0; c < geo[j].k_pad; c++) cmap[cpad_base[j] + c] = make_uint2((unsigned)j, (unsigned)c);
        for (size_t r = 0; r < run[j].R.batch_size; r++) rmap[row_base[j] + r] = make_uint2((unsigned)j, (unsigned)r);
        for (size_t c = 0; c < jobs[j].k; c++) cjob[cl_base[j] + c] = (uint32_t)j;
    }
    Layout ml;
    const size_t m_wg = ml.add(wg_all.size() * 8), m_c = ml.add(cpad * 8), m_r = ml.add(slots * 8),
                 m_j = ml.add(nclusters * 4), m_lwg = ml.add(wg_lds.size() * 8);
    if ((rc = check(maps.ensure(ml.bytes), "maps"))) return rc;
    uint8_t *mp = (uint8_t *)maps.p;
    if ((rc = check(hipMemcpyAsync(mp + m_wg, wg_all.data(), wg_all.size() * 8, hipMemcpyHostToDevice, h->stream), "maps")) ||
        (rc = check(hipMemcpyAsync(mp + m_c, cmap.data(), cpad * 8, hipMemcpyHostToDevice, h->stream), "maps")) ||
        (rc = check(hipMemcpyAsync(mp + m_r, rmap.data(), slots * 8, hipMemcpyHostToDevice, h->stream), "maps")) ||
        (rc = check(hipMemcpyAsync(mp + m_j, cjob.data(), nclusters * 4, hipMemcpyHostToDevice, h->stream), "maps")) ||
        (!wg_lds.empty() &&
         (rc = check(hipMemcpyAsync(mp + m_lwg, wg_lds.data(), wg_lds.size() * 8, hipMemcpyHostToDevice, h->stream),
                     "maps"))) ||
        (rc = check(hipStreamSynchronize(h->stream), "upload")))
        return rc;

    // Static part of the job table (device addresses).
    std::vector<JobDev> jd(njobs);
    for (size_t j = 0; j < njobs; j++) {
        JobDev &t = jd[j];
        memset(&t, 0, sizeof t);
        t.x = rows_ptr[j];
        t.ids = (const uint32_t *)(in_dp + in_ids) + row_base[j];
        t.cen = cen_model.as<float>() + cen_base[j];
        t.bias = (const float *)(in_dp + in_bias) + cl_base[j];
        t.usage = use_dev + cl_base[j];
        t.cen_op = (uint8_t *)cen_op.p + op_base[j];
        t.norm32 = norm32.as<float>() + cpad_base[j];
        t.hn = hn.as<float>() + cpad_base[j];
        t.cn = int_search ? cn.as<int64_t>() + cpad_base[j] : nullptr;
        t.cexp = op == OP_FP8 ? cexp_d.as<uint8_t>() + cpad_base[j] : nullptr;
        t.shift = (uint32_t)geo[j].shift;
        t.stats = stats_d.as<unsigned>() + 4 * j;
        t.labels = lab_d.as<uint32_t>() + row_base[j];
        t.labels_all = labels_all.as<uint32_t>() + nrow_base[j];
        t.dist = (float *)(out_dp + out_dist) + row_base[j];
        t.cand = cand.as<uint32_t>() + row_base[j] * M;
        t.sums = sums_d.as<float>() + cen_base[j];
        t.counts = counts_d.as<uint32_t>() + cl_base[j];
        t.pass_sums = pass_sums.as<double>() + cen_base[j];
        t.pass_counts = pass_counts.as<uint64_t>() + cl_base[j];
        t.d = (uint32_t)jobs[j].d;
        t.k = (uint32_t)jobs[j].k;
        t.k_pad = (uint32_t)geo[j].k_pad;
        t.d_pad = (uint32_t)geo[j].d_pad;
        // the register-resident search reads a tile-major table, the LDS
        // kernel a row-major one (tiles of one centroid)
        t.row_bytes = (uint32_t)op_row_bytes(geo[j]);
        t.tb = geo[j].reg ? (uint32_t)reg_tile(geo[j]) : 1u;
        t.nslice = slices[j].nslice;
        t.tiles_per_slice = slices[j].tiles_per_slice;
        t.part = slices[j].nslice > 1 ? part_d.as<uint32_t>() + part_base[j] : nullptr;
        if (op == OP_FP8) {
            t.x8 = x8_d.as<uint8_t>() + x8_base[j];
            t.row_scale = rscale_d.as<float>() + row_base[j];
            t.row_norm = rnorm_d.as<float>() + row_base[j];
        }
        t.row_base = (uint32_t)row_base[j];
        t.cluster_base = (uint32_t)cl_base[j];
    }
    if ((rc = check(accum_reserve(h->acc, slots, nclusters), "sort scratch"))) return rc;

    // Capture the step.
    const JobDev *jobs_dev = (const JobDev *)(in_dp + in_job);
    const void *sa_dev = in_dp + in_sa;
    // The labels' and distances' download is not in the graph: the graph
    // records an external event once they are final (after the rescore) and
    // a copy stream downloads them while the per-cluster sums and the EMA
    // step (which read the labels and write neither) run.
    GraphGuard gg;
    SideStream side;
    if ((rc = check(side.init(), "copy stream"))) return rc;
    if ((rc = check(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "begin capture"))) return rc;
    // capture info: the nodes the next captured operation would follow
    auto last_nodes = [&](std::vector<hipGraphNode_t> &out) -> hipError_t {
        hipStreamCaptureStatus cs;
        const hipGraphNode_t *deps = nullptr;
        size_t ndeps = 0;
        const hipError_t r = hipStreamGetCaptureInfo_v2(h->stream, &cs, nullptr, nullptr, &deps, &ndeps);
        if (r == hipSuccess) out.assign(deps, deps + ndeps);
        return r;
    };
    // the batch ids are uploaded on the copy stream too, beside the job
    // tables' upload and the prep kernel; the first search waits for them
    std::vector<hipGraphNode_t> search_first;
    hipError_t ce = hipMemcpyAsync(in_dp, in_hp, in_ids, hipMemcpyHostToDevice, h->stream);
    if (ce == hipSuccess) ce = launch_jobs_prep(op, jobs_dev, (const uint2 *)(mp + m_c), cpad, h->stream);
    // FP8: the batch rows' encoding is the first reader of the batch ids
    if (ce == hipSuccess && op == OP_FP8) ce = launch_jobs_rows_fp8(jobs_dev, (const uint2 *)(mp + m_r), slots, h->stream);
    if (ce == hipSuccess && op == OP_FP8) ce = last_nodes(search_first);
    if (ce == hipSuccess) ce = launch_jobs_search(op, M, sa_dev, (const uint2 *)(mp + m_wg), wg_all.size(), h->stream);
    if (ce == hipSuccess && op != OP_FP8) ce = last_nodes(search_first);
    if (ce == hipSuccess)
        ce = launch_jobs_search_lds(op, M, sa_dev, (const uint2 *)(mp + m_lwg), wg_lds.size(), lds_max, h->stream);
    // the candidate searches' rescore writes the sort keys in the same pass
    if (ce == hipSuccess && cand_search)
        ce = launch_jobs_rescore(M, jobs_dev, (const uint2 *)(mp + m_r), slots, h->stream, (uint32_t)nclusters,
                                 keys.as<uint32_t>(), (uint32_t *)h->acc.big);
    if (ce == hipSuccess && !cand_search && part_words)
        ce = launch_jobs_merge(op, jobs_dev, (const uint2 *)(mp + m_r), slots, h->stream);
    if (ce == hipSuccess)
        ce = launch_jobs_keys(h->acc, jobs_dev, (const uint2 *)(mp + m_r), slots, cand_search, nclusters,
                              keys.as<uint32_t>(), counts_d.as<uint32_t>(), h->stream);
    // the distances are final here (true distances): an event record node
    // follows these nodes once the capture ends
    std::vector<hipGraphNode_t> dist_deps;
    if (ce == hipSuccess) ce = last_nodes(dist_deps);
    if (ce == hipSuccess)
        ce = launch_jobs_sums(h->acc, jobs_dev, rows, d_max, slots, (const uint32_t *)(in_dp + in_ids),
                              (const uint32_t *)(mp + m_j), nclusters, keys.as<uint32_t>(), counts_d.as<uint32_t>(),
                              h->stream);
    if (ce == hipSuccess)
        ce = launch_jobs_ema(jobs_dev, (const uint32_t *)(mp + m_j), nclusters, stats_d.as<unsigned>(), 4 * njobs,
                             h->stream);
    if (ce == hipSuccess)
        ce = hipMemcpyAsync(out_hp + out_use, out_dp + out_use, outl.bytes - out_use, hipMemcpyDeviceToHost, h->stream);
    hipError_t ee = hipStreamEndCapture(h->stream, &gg.graph);
    if ((rc = check(ce, "capture step")) || (rc = check(ee, "end capture"))) return rc;
    hipGraphNode_t dist_event, ids_wait;
    if (search_first.size() != 1) return fail(EMH_E_DEVICE, "train_batch: unexpected step graph");
    if ((rc = check(hipGraphAddEventRecordNode(&dist_event, gg.graph, dist_deps.data(), dist_deps.size(),
                                               side.e[0]),
                    "distances event")) ||
        (rc = check(hipGraphAddEventWaitNode(&ids_wait, gg.graph, nullptr, 0, side.e[1]), "ids event")) ||
        (rc = check(hipGraphAddDependencies(gg.graph, &ids_wait, search_first.data(), 1), "ids event")) ||
        (rc = check(hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0), "instantiate step")))
        return rc;
    size_t nodes = 0;
    (void)hipGraphGetNodes(gg.graph, nullptr, &nodes);

    // Lockstep steps. The host halves of the jobs are independent (own RNG,
    // own staging slots) and run on a few host threads; revival triples are
    // gathered in job order.
    uint64_t steps = 0;
    std::vector<char> live(njobs);
    std::vector<std::vector<std::pair<size_t, size_t>>> jpairs(njobs);
    JobDev *jd_h = (JobDev *)(in_hp + in_job);
    // I8 jobs run the packed-key search only, whose scores need
    // |sum c8^2 + round(bias)| < 2^26 - d 2^15 (sum c8^2 is in [0, d 2^14]);
    // a bias past that (the single-problem driver then takes the i64 search)
    // stops the batch with an error
    std::vector<char> out_of_range(njobs, 0);
    const auto i8_bias_fits = [&](const JobRun &J) {
        const double lim = (double)(((int64_t)1 << 26) - (int64_t)J.d * 32768), top = (double)J.d * 16384.0;
        for (float b : J.bias) {
            const double r = std::round((double)b);
            if (!(r + top < lim && r > -lim)) return false;
        }
        return true;
    };
    const std::function<void(size_t)> stage = [&](size_t j) {
        JobRun &J = run[j];
        JobDev t = jd[j];
        live[j] = J.next();
        if (live[j]) {
            if (J.penalized && op == OP_I8 && geo[j].reg && !i8_bias_fits(J)) out_of_range[j] = 1;
            if (J.penalized) memcpy(in_hp + in_bias + cl_base[j] * 4, J.bias.data(), J.k * 4);
            memcpy(in_hp + in_ids + row_base[j] * 4, J.ids, J.nb * 4);
            t.m = (uint32_t)J.nb;
            t.penalized = J.penalized;
            t.last_epoch = J.is_last;
            t.clear_pass = J.last_epoch_start;
            t.alpha = J.alpha();
            t.decay = 1.0f - t.alpha;
            t.replace = t.alpha == 1.0f;
        }
        jd_h[j] = t;
        fill_search_args(in_hp + in_sa + j * sa_bytes, geo[j], t);
    };
    const std::function<void(size_t)> finish = [&](size_t j) {
        jpairs[j].clear();
        if (!live[j]) return;
        JobRun &J = run[j];
        memcpy(J.st.usage.data(), out_hp + out_use + cl_base[j] * 4, J.k * 4);
        J.finish_stats_device((const float *)(out_hp + out_dist) + row_base[j]);
        if (J.revive_due()) {
            J.st.revive_plan(J.R.unused_threshold, J.split_rng, jpairs[j]);
            if (!jpairs[j].empty()) memcpy(out_hp + out_use + cl_base[j] * 4, J.st.usage.data(), J.k * 4);
        }
        J.advance();
    };
    // Each job's closing mean (src/lib.rs:481-489, from its FP64 pass totals
    // over its final EMA centroids) and the download of its centroids and
    // last-pass labels run on a closing stream, from a background thread, as
    // soon as its last step (and that step's revivals) is done: later steps
    // never touch a finished job's data, and the copies into the caller's
    // arrays block only that thread.
    float *cen_dev = cen_model.as<float>();
    SideStream closer;
    if ((rc = check(closer.init(), "closing stream"))) return rc;
    const auto close_job = [&](size_t j) -> std::pair<int, std::string> {
        int r;
        const size_t k = jobs[j].k, d = jobs[j].d;
        if ((r = check(hipSetDevice(h->device), "hipSetDevice")) ||
            (r = check(launch_closing_mean(jd[j].pass_sums, jd[j].pass_counts, k, d, cen_dev + cen_base[j], closer.s),
                       "closing mean")) ||
            (r = check(hipMemcpyAsync(jobs[j].centroids_out, cen_dev + cen_base[j], k * d * 4, hipMemcpyDeviceToHost,
                                      closer.s),
                       "copy centroids")) ||
            (r = check(hipMemcpyAsync(jobs[j].labels_out, labels_all.as<uint32_t>() + nrow_base[j], jobs[j].n * 4,
                                      hipMemcpyDeviceToHost, closer.s),
                       "copy labels")) ||
            (r = check(hipStreamSynchronize(closer.s), "closing mean")))
            return {r, emh_last_error()};
        return {(int)EMH_OK, std::string()};
    };
    std::vector<char> closed(njobs, 0);
    std::vector<std::future<std::pair<int, std::string>>> closing;
    const auto close_finished = [&]() {
        for (size_t j = 0; j < njobs; j++)
            if (!live[j] && !closed[j]) {
                closed[j] = 1;
                closing.push_back(std::async(std::launch::async, close_job, j));
            }
    };
    double host_s = 0.0, wait_s = 0.0, launch_s = 0.0;
    using clk = std::chrono::steady_clock;
    auto since = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
    // a job's host half of one step and the staging of its next batch run
    // together, one pool run per step
    const std::function<void(size_t)> finish_stage = [&](size_t j) {
        finish(j);
        stage(j);
    };
    const auto range_error = [&]() -> int {
        for (size_t j = 0; j < njobs; j++)
            if (out_of_range[j])
                return fail(EMH_E_UNSUPPORTED, "train_batch: job " + std::to_string(j) +
                                                   "'s balance bias leaves the packed I8 score range "
                                                   "(|sum c8^2 + round(bias)| < 2^26 - d 2^15); emh_train takes it");
        return EMH_OK;
    };
    auto th = clk::now();
    pool.run(njobs, stage);
    host_s += since(th);
    if ((rc = range_error())) return rc;
    for (;;) {
        if (std::find(live.begin(), live.end(), (char)1) == live.end()) break;
        th = clk::now();
        if ((rc = check(hipMemcpyAsync(in_dp + in_ids, in_hp + in_ids, in.bytes - in_ids, hipMemcpyHostToDevice,
                                       side.s),
                        "batch ids")) ||
            (rc = check(hipEventRecord(side.e[1], side.s), "batch ids")) ||
            (rc = check(hipGraphLaunch(gg.exec, h->stream), "step launch")) ||
            (rc = check(hipStreamWaitEvent(side.s, side.e[0], 0), "labels ready")) ||
            (rc = check(hipMemcpyAsync(out_hp, out_dp, out_use, hipMemcpyDeviceToHost, side.s), "distances")))
            return rc;
        launch_s += since(th);
        th = clk::now();
        if ((rc = check(hipStreamSynchronize(h->stream), "step")) ||
            (rc = check(hipStreamSynchronize(side.s), "distances")))
            return rc;
        wait_s += since(th);
        steps++;
        // host half per job, then its next batch; revival planned on the
        // downloaded masses, rows split on the device before the next replay
        // (one upload of all triples and masses, one kernel)
        th = clk::now();
        pool.run(njobs, finish_stage);
        host_s += since(th);
        if ((rc = range_error())) break;
        uint32_t *tp = pairs_h.as<uint32_t>();
        size_t ntrip = 0;
        for (size_t j = 0; j < njobs; j++)
            for (const auto &pr : jpairs[j]) {
                tp[3 * ntrip] = (uint32_t)j;
                tp[3 * ntrip + 1] = (uint32_t)pr.first;
                tp[3 * ntrip + 2] = (uint32_t)pr.second;
                ntrip++;
            }
        const bool revived = ntrip > 0;
        if (revived &&
            ((rc = check(hipMemcpyAsync(pairs_d.p, tp, ntrip * 12, hipMemcpyHostToDevice, h->stream), "revival")) ||
             (rc = check(hipMemcpyAsync(use_dev, out_hp + out_use, nclusters * 4, hipMemcpyHostToDevice, h->stream),
                         "masses")) ||
             (rc = check(launch_jobs_revive(jobs_dev, pairs_d.as<uint32_t>(), ntrip, h->stream), "revival"))))
            return rc;
        if (revived && (rc = check(hipStreamSynchronize(h->stream), "revival"))) return rc;
        close_finished();
    }

    close_finished();  // jobs that never stepped
    for (auto &f : closing) {
        const auto r = f.get();
        if (r.first != EMH_OK && rc == EMH_OK) rc = fail(r.first, r.second);
    }
    if (rc) return rc;
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (size_t j = 0; j < njobs; j++) {
        jobs[j].out->n_splits = run[j].st.n_splits;
        jobs[j].out->recipe = run[j].R;
        jobs[j].out->seconds = secs;
    }
    if (stats) {
        stats->steps = steps;
        stats->graph_nodes = nodes;
        stats->seconds = secs;
        stats->host_seconds = host_s;
        stats->wait_seconds = wait_s;
        stats->launch_seconds = launch_s;
    }
    return EMH_OK;
}

}  // namespace emh

using namespace emh;

extern "C" int emh_train_batch(emh_assigner *h, emh_job *jobs, size_t njobs, emh_batch_stats *stats) {
    if (!h) return fail(EMH_E_ARG, "train_batch: NULL handle");
    return batch_impl(h, jobs, njobs, stats);
}
