// Internal declarations shared by the e-means HIP library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <string>

#include "../../include/emeans_hip.h"

namespace emh {

// OP_I16: CentroidBits::I16, the centroid stored as two i8 planes (see
// search_lds_core).
enum Op { OP_F32 = 0, OP_BF16 = 1, OP_FP8 = 2, OP_I8 = 3, OP_I16 = 4 };

// Workgroup geometry of the search kernels: 4 waves, each owning 32 rows, and
// centroid tiles of BN centroids (two 32x32 MFMA blocks per wave).
constexpr int WAVES = 4;
constexpr int BM = 32 * WAVES;  // rows per workgroup
constexpr int BN = 64;          // centroids per tile
constexpr int CHUNK_BYTES = 128;  // d-chunk per row when rows are not resident
constexpr int ROW_LDS_BUDGET = 64 * 1024;  // bytes of LDS for a resident row block
constexpr int MAX_M = 8;        // top-m candidates kept by the rescoring searches

inline int op_elem_bytes(int op) { return op == OP_F32 ? 4 : (op == OP_BF16 ? 2 : 1); }
// Row padding unit in bytes: each lane half of a row must be a whole number
// of 16-byte quads (the MFMA loops read halves quad by quad), and the BF16 /
// FP8 searches pad to 64 bytes so a half is an even number of quads.
inline int op_granule(int op) { return (op == OP_BF16 || op == OP_FP8) ? 64 : 32; }

struct SearchGeom {
    int op;
    bool reg;           // rows fit the register-resident kernel (operand row <= 512 bytes)
    bool i8_packed = false;  // I8: every |2 x.c - cn| < 2^26, so scores fit the packed int32 key
    bool i16_f64 = false;    // I16: every |cn| < 2^46, so the register kernel's f64 keys are exact
    bool resident;      // the workgroup's rows stay in LDS for the whole centroid sweep
    size_t d, d_pad;    // elements; d_pad * elem_bytes is the operand row stride
    size_t k, k_pad;
    int row_bytes;      // bytes of one LDS row segment (d_pad*es when resident, else CHUNK_BYTES)
    int shift = 0;      // I16: centroids are round(c * 2^shift) (src/cpu_i8.rs:95-107)
};

SearchGeom make_geom(int op, size_t k, size_t d);

// I16 centroids as i8 planes: c16 = 2^8 hi + (lo + 128), hi the signed top
// byte, lo the low byte less 128, so x.c16 = 2^8 x.hi + x.lo + 128 sum(x).
constexpr int I16_PLANES = 2;
// Centroids per tile of the register-resident search (one or two 32-wide
// MFMA blocks).
// (I16: 32 centroids, whose two planes are the tile's two MFMA blocks.)
inline int reg_tile(const SearchGeom &g) { return g.op == OP_I16 ? 32 : g.row_bytes <= 256 ? 64 : 32; }
// Bytes of the operand centroid buffer: the row-major table (k_pad rows of
// d_pad elements; row_bytes is only the LDS segment when rows stream in
// chunks), followed on the register-resident path (where a row is
// row_bytes) by the same table in tile-major order (reg_operand).
inline size_t operand_table_bytes(const SearchGeom &g) {
    return g.k_pad * g.d_pad * (size_t)op_elem_bytes(g.op) * (g.op == OP_I16 ? I16_PLANES : 1);
}
inline size_t operand_bytes(const SearchGeom &g) {
    return g.reg ? 2 * operand_table_bytes(g) : operand_table_bytes(g);
}
// Bytes of one operand row in the table a search reads: the register-
// resident kernel's row, or the LDS kernel's whole padded row.
inline size_t op_row_bytes(const SearchGeom &g) {
    return g.reg ? (size_t)g.row_bytes : g.d_pad * (size_t)op_elem_bytes(g.op);
}
inline const void *reg_operand(const SearchGeom &g, const void *cen_op) {
    return (const uint8_t *)cen_op + operand_table_bytes(g);
}

// Launchers (stream-ordered, no synchronisation).
// stats (u32 words, starting at 0): [0] unused, [1] max(0, -min_j bias_j)
// as f32 bits (set by the bias kernel), [2] the I8 / I16 range flag.
// cen_op holds operand_bytes(g): the prep launcher fills both layouts.
// hn_out / cn_out: also write the unbiased score offsets (what
// launch_set_bias without a bias writes) in the same pass.
// cexp (FP8, k_pad bytes): each centroid's MFMA scale (E8M0 of 1 / s_c).
hipError_t launch_prep_centroids(const SearchGeom &g, const float *c, void *cen_op, float *norm32, int64_t *norm64,
                                 uint8_t *cexp, unsigned *stats, hipStream_t s, float *hn_out = nullptr,
                                 int64_t *cn_out = nullptr);
hipError_t launch_set_bias(const SearchGeom &g, const float *norm32, const int64_t *norm64, const float *bias,
                           float *hn, int64_t *cn, unsigned *stats, hipStream_t s);
hipError_t launch_rows_fp8(const void *x, int rows, size_t m, size_t d, size_t d_pad, const uint32_t *ids,
                           uint8_t *x8, float *row_scale, float *row_norm, hipStream_t s);
bool reg_path(const SearchGeom &g);
// FP8 on f32 rows through the register-resident kernel: the kernel encodes
// its rows itself (no separate encoding pass, x8 / row_scale / row_norm
// unused).
inline bool fp8_encodes_in_search(const SearchGeom &g, int rows) {
    return g.op == OP_FP8 && g.reg && rows == EMH_ROWS_F32;
}
// fr (register-resident path): 32-row fragments per wave, 1 or 2 (128 or 256
// rows per workgroup); the results do not depend on it.
// nslice > 1 (register-resident path only): the centroid tiles are split
// over nslice workgroups per row block (fewer when tiles run out), each
// writing partial results to `part` (slice_part_bytes), then merged into
// labels/dist or cand. The results do not depend on nslice.
hipError_t launch_search(const SearchGeom &g, int M, const void *x, int rows, size_t m, const uint32_t *ids,
                         const void *cen_op, const float *hn, const int64_t *cn, const uint8_t *x8,
                         const float *row_scale, const float *row_norm, const uint8_t *cexp,
                         const unsigned *stats, uint32_t *labels, float *dist, uint32_t *cand, hipStream_t s,
                         int nslice = 1, uint32_t *part = nullptr, int fr = 2);
int reg_tiles(const SearchGeom &g);
size_t slice_part_bytes(int M, size_t m, int nslice, int op);
hipError_t launch_rescore(int M, const void *x, int rows, size_t m, size_t d, const uint32_t *ids, const float *c,
                          const float *bias, size_t k, const uint32_t *cand, uint32_t *labels, float *dist,
                          hipStream_t s);

// ---- many independent problems per launch (batched training) -------------
// One problem's device view for one step; the host rewrites the per-step
// fields before each replay of the step graph. Pointers address this job's
// slot of the shared per-step arrays.
struct JobDev {
    const void *x;           // resident rows, n x d (f32; i8 for the I8 search)
    const uint32_t *ids;     // this step's batch row positions
    float *cen;              // the job's device-resident centroids, k x d
    const float *bias;       // this step's balance bias, k (read when penalized)
    float *usage;            // the job's device-resident cluster masses, k
    uint8_t *cen_op;         // tile-major operand table (register-resident search)
    float *norm32, *hn;      // k_pad each
    int64_t *cn;             // I8: k_pad integer score offsets sum c8^2 + round(bias)
    unsigned *stats;         // 4 words: [1] max -bias (f32 bits); the others unused
    uint8_t *cexp;           // FP8: k_pad centroid MFMA scales (E8M0 of 1 / s_c)
    uint32_t *labels;        // batch slot
    float *dist;             // batch slot (true distances once the keys kernel has run)
    uint32_t *labels_all;    // n: the last pass's label of every row (device-resident)
    uint32_t *cand;          // batch slot x M
    float *sums;             // k x d
    uint32_t *counts;        // k
    double *pass_sums;       // k x d
    uint64_t *pass_counts;   // k
    uint32_t *part;          // centroid-range split: per (row, slice) results, batch slot x nslice
    uint8_t *x8;             // FP8: this step's batch rows as e4m3, batch slot x d_pad
    float *row_scale;        // FP8: batch slot (power-of-two encoding scale)
    float *row_norm;         // FP8: batch slot (|x|^2 of the f32 row)
    uint32_t d, k, k_pad, d_pad, row_bytes, tb;
    uint32_t shift;          // I16: centroids are round(c 2^shift)
    uint32_t nslice, tiles_per_slice;  // the search's centroid-range split (1: none)
    uint32_t row_base, cluster_base;  // offsets in the concatenated batch / cluster ranges
    // per step
    uint32_t m;              // rows this step (0: the job is done)
    uint32_t penalized;      // balance bias applies this step
    uint32_t last_epoch;     // fold this step's sums into the FP64 pass totals
    uint32_t clear_pass;     // first batch of the last epoch: zero the pass totals first
    float alpha, decay;      // this step's EMA rate and 1 - rate
    uint32_t replace;        // alpha == 1
};

// A pointer read from a device table addresses global memory, which the
// compiler cannot see: its loads and stores would take the flat forms, whose
// completion LDS waits also hold for. Passing it through an empty asm as an
// address-space-1 pointer says so. global_ptr_s: the value is the same in
// every lane of the wave (kept in scalar registers); global_ptr_v: any.
template <class T> __device__ inline T *global_ptr_s(T *p) {
    const uint64_t v = (uint64_t)p;
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    __attribute__((address_space(1))) T *g = (__attribute__((address_space(1))) T *)((hi << 32) | lo);
    asm("" : "+s"(g));
    return (T *)g;
}
template <class T> __device__ inline T *global_ptr_v(T *p) {
    __attribute__((address_space(1))) T *g = (__attribute__((address_space(1))) T *)p;
    asm("" : "+v"(g));
    return (T *)g;
}

// jobs[i] with every pointer marked global (UNIFORM: i is the same in every
// lane of the wave).
template <bool UNIFORM> __device__ inline JobDev job_at(const JobDev *jobs, size_t i) {
    JobDev j = jobs[i];
#define EMH_G(f)                    \
    if constexpr (UNIFORM)          \
        j.f = global_ptr_s(j.f);    \
    else                            \
        j.f = global_ptr_v(j.f)
    EMH_G(x);
    EMH_G(ids);
    EMH_G(cen);
    EMH_G(bias);
    EMH_G(usage);
    EMH_G(cen_op);
    EMH_G(norm32);
    EMH_G(hn);
    EMH_G(cn);
    EMH_G(stats);
    EMH_G(labels);
    EMH_G(dist);
    EMH_G(labels_all);
    EMH_G(cand);
    EMH_G(sums);
    EMH_G(counts);
    EMH_G(pass_sums);
    EMH_G(pass_counts);
    EMH_G(part);
    EMH_G(x8);
    EMH_G(row_scale);
    EMH_G(row_norm);
    EMH_G(cexp);
#undef EMH_G
    return j;
}

// The last epoch's first batch zeroes its job's FP64 totals: centroid c's
// row, by the 16 lanes gl of the group that preps that centroid.
__device__ inline void clear_pass_row(const JobDev &j, size_t c, int gl) {
    if (!j.clear_pass || c >= j.k) return;
    for (size_t t = gl; t < j.d; t += 16) j.pass_sums[c * j.d + t] = 0.0;
    if (gl == 0) j.pass_counts[c] = 0;
}

// The sort key of a batch slot holding row `row` of job j's batch: the job's
// first global cluster + the row's label, or nclusters for a slot past the
// job's rows this step (sorts last, never folded). Once the labels are in:
// a penalized job's distance becomes the true one (the score less the
// bias), and in the last epoch the row's label is kept.
__device__ inline uint32_t slot_key(const JobDev &j, uint32_t row, uint32_t nclusters) {
    uint32_t key = nclusters;
    if (row < j.m) {
        const uint32_t l = j.labels[row];
        if (l < j.k) {
            key = j.cluster_base + l;
            if (j.penalized) {
                const float v = __fsub_rn(j.dist[row], j.bias[l]);
                j.dist[row] = v > 0.0f ? v : 0.0f;
            }
        }
        if (j.last_epoch) j.labels_all[j.ids[row]] = l;
    }
    return key;
}

// prep: every (job, centroid < k_pad) entry of cmap (cmap lists each job's
// entries together); operand, norms, offsets with this step's bias, stats
// (zeroed before), FP8 centroid scales; in the last epoch's first batch the
// centroid's FP64 pass totals are zeroed too (clear_pass_row).
hipError_t launch_jobs_prep(int op, const JobDev *jobs, const uint2 *cmap, size_t entries, hipStream_t s);
// FP8: every slot's batch row (rmap: (job, row) per slot) encoded as
// rows_fp8_kernel encodes a single problem's rows.
hipError_t launch_jobs_rows_fp8(const JobDev *jobs, const uint2 *rmap, size_t slots, hipStream_t s);
// search of every job in one launch: wg holds (job, first row | slice) per
// workgroup (first rows are multiples of 256, slices below 256).
hipError_t launch_jobs_search(int op, int M, const void *search_args, const uint2 *wg, size_t nwg, hipStream_t s);
// the same for the jobs whose operand rows are too wide for it, on the LDS
// kernel: wg holds (job, first row) per workgroup (first rows are multiples
// of BM); lds_bytes: the largest such job's search_lds_bytes.
hipError_t launch_jobs_search_lds(int op, int M, const void *search_args, const uint2 *wg, size_t nwg,
                                  size_t lds_bytes, hipStream_t s);
// dynamic LDS of the LDS kernel for a geometry
size_t search_lds_bytes(const SearchGeom &g);
// rescoring of every job's candidates (merging a split job's slice lists
// first): rmap holds (job, row) per batch slot. With keys given, each slot's
// sort key (slot_key) is written in the same pass and *big_count zeroed,
// which launch_jobs_keys then leaves out.
hipError_t launch_jobs_rescore(int M, const JobDev *jobs, const uint2 *rmap, size_t slots, hipStream_t s,
                               uint32_t nclusters = 0, uint32_t *keys = nullptr, uint32_t *big_count = nullptr);
// FP32 and I8 jobs with a split search: each slot's slice results merged
// into its label and distance.
hipError_t launch_jobs_merge(int op, const JobDev *jobs, const uint2 *rmap, size_t slots, hipStream_t s);
// Centroid-range split of one job's search: slices of consecutive tiles
// sized so that every workgroup sweeps about the same operand volume.
struct JobSlices {
    uint32_t nslice, tiles_per_slice;
};
JobSlices job_slices(const SearchGeom &g);
// Words of slice results per batch row of a job (M candidates, or one best).
size_t job_part_words(int op, int M, uint32_t nslice);
// Builds a SearchArgs table entry (opaque bytes) for one job's step.
size_t search_args_bytes();
void fill_search_args(void *dst, const SearchGeom &g, const JobDev &j);

// Device-side per-cluster accumulation (emh_accum.hip).
struct AccumScratch {
    uint32_t *keys = nullptr, *keys_alt = nullptr, *vals = nullptr, *vals_alt = nullptr;
    uint32_t *big = nullptr;  // batched sums: [0] count, then (cluster, first position, members) triples
    uint32_t *starts = nullptr;  // one-pass sorts (keys below 2048): each key's first sorted position
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    size_t cap_rows = 0;
};
hipError_t accum_reserve(AccumScratch &a, size_t m, size_t k);
void accum_free(AccumScratch &a);
// `cleared`: counts and the big-cluster count are already zero (an EMA step
// with clear_counts left them so) and the sums of clusters without members
// are not read; the clearing launch is skipped.
hipError_t launch_accumulate(AccumScratch &a, const void *x, int rows, size_t m, size_t d, const uint32_t *ids,
                             const uint32_t *labels, size_t k, float *sums, uint32_t *counts, hipStream_t s,
                             bool cleared = false);
hipError_t launch_pass_accumulate(const float *sums, const uint32_t *counts, size_t k, size_t d, double *pass_sums,
                                  uint64_t *pass_counts, hipStream_t s);
hipError_t launch_closing_mean(const double *pass_sums, const uint64_t *pass_counts, size_t k, size_t d,
                               float *centroids, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, size_t count, uint32_t v, hipStream_t s);
// The EMA step of src/lib.rs:287-309 on device centroids/usage (k x d, k).
// With clear_counts it also zeroes each count once read, and *clear_word.
// With pass_sums / pass_counts (the last epoch) it first adds each
// cluster's sums and count to the FP64 pass totals (launch_pass_accumulate).
hipError_t launch_ema_apply(float *cen, float *usage, const float *sums, const uint32_t *counts, size_t k, size_t d,
                            float alpha, hipStream_t s, uint32_t *clear_counts = nullptr,
                            uint32_t *clear_word = nullptr, double *pass_sums = nullptr,
                            uint64_t *pass_counts = nullptr);
// The revival row split of src/lib.rs:330-340 for (donor, unused) pairs;
// with `usage`, also the masses' halving (src/lib.rs:334-336: both take half
// the donor's mass). pairs may be host-coherent pinned memory.
hipError_t launch_revive_split(float *cen, const uint32_t *pairs, size_t npairs, size_t d, hipStream_t s,
                               float *usage = nullptr);
// out (m x d f32) = rows ids of x (f32 / i8 / u8), widened
hipError_t launch_gather_rows(const void *x, int rows, size_t d, const uint32_t *ids, size_t m, float *out,
                              hipStream_t s);
// cjob maps a global cluster to its job.
// Revival row splits of many jobs: triples (job, donor, unused).
hipError_t launch_jobs_revive(const JobDev *jobs, const uint32_t *triples, size_t ntriples, hipStream_t s);
// The EMA step of every job still running, one wave per global cluster.
// It also zeroes each cluster's count after reading it and (nstats words)
// the jobs' stats, for the next step.
hipError_t launch_jobs_ema(const JobDev *jobs, const uint32_t *cjob, size_t nclusters, unsigned *stats, size_t nstats,
                           hipStream_t s);
// Batched, first half: the last epoch's first batch zeroes its FP64 totals;
// keys = cluster_base + label for live slots (nclusters for spare slots),
// counts_all (all zero on entry); penalized jobs' distances lose their bias (the
// true squared distances, src/lib.rs:425-430); last-epoch labels are stored
// at their rows in labels_all.
// keys_written: the rescore pass has written the keys already (only the
// scratch is reserved here).
hipError_t launch_jobs_keys(AccumScratch &a, const JobDev *jobs, const uint2 *rmap, size_t slots,
                            bool keys_written, size_t nclusters, uint32_t *keys, uint32_t *counts_all,
                            hipStream_t s);
// Second half: one stable sort, then per-job member counts (counts_all,
// zeroed by the first half), f32 sums and FP64 pass totals. rows: the jobs'
// row type (f32 or i8, one for all); d_max: the widest job's row length.
hipError_t launch_jobs_sums(AccumScratch &a, const JobDev *jobs, int rows, size_t d_max, size_t slots,
                            const uint32_t *ids, const uint32_t *cjob, size_t nclusters, const uint32_t *keys,
                            uint32_t *counts_all, hipStream_t s);
hipError_t launch_fill_f64(double *p, size_t count, hipStream_t s);
// ---- one problem's training step as a replayed graph (emh_accum.hip) ----
// Per-step values of a run, fixed by the recipe before the first step.
struct StepParam {
    uint32_t b0, nb;     // the batch's position in the epoch order, its rows
    float alpha, decay;  // the EMA rate and 1 - rate (f32, as the host computes them)
    uint32_t replace;    // alpha == 1
    uint32_t is_last;    // last epoch: the step's sums go into the FP64 pass totals
    uint32_t revive;     // a revival check follows: the masses go to the host slot too
    uint32_t pad;
};
// Pinned host slots the labels and distances of step s go to (slot s % n).
struct StepSlots {
    uint32_t *p[8];
    int n;
};
// ctr[0]: the step's index, ctr[1]: its copy for the step's later kernels.
// The stage kernel copies ctr[0] to ctr[1] and the step's ids (order +
// params[step].b0, nb of them) to step_ids; the step's last kernel
// (launch_step_finish) advances ctr[0], and a step run without the graph
// advances it with launch_step_advance.
hipError_t launch_step_stage(const uint32_t *order, const StepParam *params, uint32_t *ctr, uint32_t *step_ids,
                             size_t nb, hipStream_t s);
hipError_t launch_step_advance(uint32_t *ctr, hipStream_t s);
// The step's last kernel: launch_ema_apply with the running step's rate
// and last-epoch flag (each count zeroed once read, and *clear_word), and
// the step's labels at [0, nb) and distances at [batch, batch + nb) of its
// slot; on steps flagged `revive` also the k new masses at [2 batch, + k).
hipError_t launch_step_finish(float *cen, float *usage, const float *sums, uint32_t *counts, size_t k, size_t d,
                              const StepParam *params, uint32_t *ctr, uint32_t *clear_word, double *pass_sums,
                              uint64_t *pass_counts, const uint32_t *labels, const float *dist, size_t nb,
                              size_t batch, const StepSlots &slots, hipStream_t s);


// Error reporting shared by the C ABI.
void set_error(const std::string &msg);

}  // namespace emh
