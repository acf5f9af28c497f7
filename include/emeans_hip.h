/* emeans_hip.h - C ABI of the MI355X (gfx950) e-means nearest-centroid search.
 *
 * The first block is the drop-in boundary: one function per method of the
 * reference crate's plugin traits (andrewluetgers/e-means)
 *
 *     trait Assigner            src/lib.rs:75-84
 *     trait AssignRows<T>       src/lib.rs:87-91
 *
 * so a Rust `impl Assigner for HipAssigner` wraps exactly these functions
 * (an extern "C" block plus the impl). The second block holds extensions
 * the reference has no counterpart for: device-resident rows, device-side
 * per-cluster sums/counts, the training driver of src/lib.rs:348-491
 * running over them, and many trainings in lockstep.
 *
 * Conventions: every function returns 0 on success and a nonzero EMH_E_*
 * code on failure; emh_last_error() then holds a message. Nothing panics,
 * throws or aborts across this boundary. All pointers are plain host
 * pointers unless the name says _dev. Calls on one handle must come from one
 * thread at a time (the trait methods take &mut self).
 */
#ifndef EMEANS_HIP_H
#define EMEANS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
#pragma GCC visibility push(default)

typedef struct emh_assigner emh_assigner;

/* Search precision of an assigner. */
typedef enum {
    EMH_FP32 = 0,          /* fused FP32 |x|^2 - 2x.c + |c|^2 argmin: the drop-in */
    EMH_BF16_RESCORE = 1,  /* BF16 MFMA, FP32 accumulate, top-m kept, FP32 rescore */
    EMH_FP8_RESCORE = 2,   /* FP8-e4m3 (per-row and per-centroid scaled) MFMA, top-m, FP32 rescore */
    EMH_I8 = 3,            /* i8 MFMA, int32 dots: bit-exact CpuI8Assigner(I8) */
    EMH_I16 = 4            /* i8 MFMA on split i16 centroids: CpuI8Assigner(I16) */
} emh_precision;

/* Row element types (the reference's Scalar impls, src/lib.rs:47-71). */
typedef enum { EMH_ROWS_F32 = 0, EMH_ROWS_I8 = 1, EMH_ROWS_U8 = 2 } emh_rows;

enum {
    EMH_OK = 0,
    EMH_E_ARG = 1,      /* bad argument or contract violation (the reference asserts) */
    EMH_E_STATE = 2,    /* e.g. assign before set_centroids */
    EMH_E_MEMORY = 3,   /* not enough device memory (src/cuda.rs:103-110) */
    EMH_E_DEVICE = 4,   /* HIP runtime error */
    EMH_E_UNSUPPORTED = 5 /* row type / precision combination not offered */
};

/* ---- Drop-in boundary: trait Assigner + AssignRows<T> ------------------- */

/* Replaces CudaAssigner::new(ordinal) (src/cuda.rs:69-91) and
 * CpuI8Assigner::new(bits) (src/cpu_i8.rs:61-69). rescore_m: candidates kept
 * per row by the BF16/FP8 searches (1..8; 0 picks 4); ignored otherwise. */
int emh_create(int device, int precision, int rescore_m, emh_assigner **out);
void emh_destroy(emh_assigner *h);

/* Assigner::set_centroids (src/lib.rs:77; src/cuda.rs:127-135). k x d
 * row-major f32. Resets any bias (src/lib.rs:78-81). */
int emh_set_centroids(emh_assigner *h, const float *centroids, size_t k, size_t d);

/* Assigner::set_bias (src/lib.rs:81; src/cpu.rs:95-99; src/cpu_i8.rs:112-117).
 * bias has k entries; rows then minimise |x - c_j|^2 + bias_j and the returned
 * distance includes the bias. */
int emh_set_bias(emh_assigner *h, const float *bias, size_t k);

/* Assigner::name (src/lib.rs:83). Writes a NUL-terminated string. */
int emh_name(const emh_assigner *h, char *buf, size_t cap);

/* AssignRows<T>::assign (src/lib.rs:90; src/cuda.rs:155-213): x is m x d
 * row-major; labels[m], dist[m]. Ties go to the lowest index. Synchronous:
 * outputs are valid on return. EMH_I8/EMH_I16 take i8 rows only, as
 * CpuI8Assigner implements AssignRows<i8> only (src/cpu_i8.rs:130). */
int emh_assign_f32(emh_assigner *h, const float *x, size_t m, uint32_t *labels, float *dist);
int emh_assign_i8(emh_assigner *h, const int8_t *x, size_t m, uint32_t *labels, float *dist);
int emh_assign_u8(emh_assigner *h, const uint8_t *x, size_t m, uint32_t *labels, float *dist);

/* Message of the last failure on this thread (never NULL). */
const char *emh_last_error(void);

/* ---- Extensions (no counterpart in the reference traits) ---------------- */

/* Run on this HIP stream (hipStream_t) instead of the assigner's own. */
int emh_set_stream(emh_assigner *h, void *stream);

/* Device-pointer search: x_dev (m x d, `rows` type) already in HBM; optional
 * ids_dev (m u32) gathers row ids_dev[r] of x_dev for row r (the gather of
 * src/lib.rs:421-423 done on device). Asynchronous on the assigner's stream;
 * emh_synchronize() waits. */
int emh_assign_device(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                      uint32_t *labels_dev, float *dist_dev);
int emh_synchronize(emh_assigner *h);

/* emh_assign_device without any allocation (no counterpart in the
 * reference: for callers that cannot allocate device memory, such as code
 * under HIP graph capture or tight serving loops). Same arguments, results
 * and stream as emh_assign_device; the call itself allocates nothing: the
 * row-side buffers (candidate lists, FP8 row encodings, centroid-split
 * results) live in one static device array of EMH_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends. A call whose rows need more than that
 * fails with EMH_E_ARG (BF16 top-4: 16 bytes per row; FP8 top-8: 32 bytes
 * per row, plus d_pad + 8 for rows over 512 bytes or i8/u8 rows, whose
 * encodings are staged; FP32 / I8 / I16 need none unless split). Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads, other handles) race, so order them. No timing
 * events are recorded. */
#define EMH_NOALLOC_SCRATCH_BYTES (64ull << 20)
int emh_assign_device_noalloc(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                              uint32_t *labels_dev, float *dist_dev);

/* Per-kernel timing with HIP events on the assigner's stream. While on, each
 * search call records events around its three stages; emh_get_timing waits,
 * returns the summed milliseconds of {row encoding, search, rescore} in
 * ms_out[0..2] and the number of search calls, and resets the totals. */
int emh_set_timing(emh_assigner *h, int on);
int emh_get_timing(emh_assigner *h, double *ms_out, uint64_t *calls);

/* Per-cluster counts and f32 sums of a labelled batch, each sum folded in
 * batch row order (bit-identical to the host loop of src/lib.rs:431-441).
 * Device pointers; sums_dev is k x d f32, counts_dev k u32. */
int emh_accumulate_device(emh_assigner *h, const void *x_dev, int rows, size_t m, const uint32_t *ids_dev,
                          const uint32_t *labels_dev, float *sums_dev, uint32_t *counts_dev);

/* Training configuration: mirror of Config (src/lib.rs:103-139). */
typedef struct {
    double nepochs;
    int32_t batch_law;            /* 0 auto, 1 throughput, 2 quality */
    int32_t has_batch_size;
    uint64_t batch_size;
    int32_t has_alpha0;
    int32_t has_unused_threshold;
    double alpha0;
    float unused_threshold;
    float balance;
    uint64_t seed;
} emh_config;

/* Entries of the per-epoch tables below: nepochs must be below
 * EMH_MAX_EPOCHS - 1 (1023 epochs; resolve/train return EMH_E_ARG beyond).
 * The reference's Vec has no bound (src/lib.rs:159, 180-192). */
#define EMH_MAX_EPOCHS 1024

/* Recipe (src/lib.rs:142-149). */
typedef struct {
    uint64_t batch_size;
    double alpha0;
    uint64_t total_steps;
    float unused_threshold;
    uint32_t n_epochs;
    uint64_t epoch_rows[EMH_MAX_EPOCHS];
} emh_recipe;

typedef struct {
    uint64_t n_splits;
    uint32_t n_objective;
    double objective[EMH_MAX_EPOCHS];
    emh_recipe recipe;
    double seconds;               /* wall time of the training loop */
} emh_train_out;

void emh_config_default(emh_config *cfg);
/* resolve_recipe (src/lib.rs:201-225). */
int emh_resolve_recipe(uint64_t n, uint64_t d, uint64_t k, const emh_config *cfg, emh_recipe *out);

/* train (src/lib.rs:348-491) with this assigner: rows are uploaded once and
 * stay in HBM; each batch's gather, search and per-cluster sums/counts run on
 * the device; the EMA update, revival and schedule run on the host exactly
 * as the reference. The last pass's per-cluster sums accumulate in FP64 on
 * the device for the closing mean. centroids_out: k x d; labels_out: n. */
int emh_train(emh_assigner *h, const void *x, int rows, size_t n, size_t d, size_t k, const emh_config *cfg,
              float *centroids_out, uint32_t *labels_out, emh_train_out *out);
/* emh_train on rows already in device memory (x_dev, n x d, row-major): no
 * copy of the rows; the initial centroids are gathered on the device. The
 * outputs are host arrays as for emh_train, and the results are the same. */
int emh_train_device(emh_assigner *h, const void *x_dev, int rows, size_t n, size_t d, size_t k,
                     const emh_config *cfg, float *centroids_out, uint32_t *labels_out, emh_train_out *out);

/* assign (src/lib.rs:512-520): exact labels and distances for all n rows. */
int emh_assign_all(emh_assigner *h, const void *x, int rows, size_t n, size_t d, const float *centroids, size_t k,
                   uint32_t *labels, float *dist);

/* Int8Quantizer (src/quantize.rs:43-96) on the device: the calibrated
 * f32 -> i8 codec that produces the int8 rows. fit = Int8Quantizer::fit with
 * Calibration {target_peak, quantile, power_of_two} (faiss: 112, 1.0, 1):
 * shift_out gets d per-dimension means (f64 sums in row order), scale_out
 * the global scale. encode/decode/clipped_fraction = the methods of the same
 * names. Results equal the reference's bit for bit. The host-pointer forms
 * upload, compute and download on `device`; the _device forms take device
 * pointers and a hipStream_t (as void *), leaving everything on the GPU. */
int emh_quant_fit(int device, const float *sample, uint64_t n, uint64_t d, float target_peak, double quantile,
                  int power_of_two, float *shift_out, float *scale_out);
int emh_quant_encode(int device, const float *x, uint64_t n, uint64_t d, const float *shift, float scale,
                     int8_t *codes);
int emh_quant_decode(int device, const float *codes, uint64_t n, uint64_t d, const float *shift, float scale,
                     float *out);
int emh_quant_clipped_fraction(int device, const float *x, uint64_t n, uint64_t d, const float *shift, float scale,
                               double *out);
int emh_quant_fit_device(const float *sample_dev, uint64_t n, uint64_t d, float target_peak, double quantile,
                         int power_of_two, float *shift_dev, float *scale_out, void *stream);
int emh_quant_encode_device(const float *x_dev, uint64_t n, uint64_t d, const float *shift_dev, float scale,
                            int8_t *codes_dev, void *stream);

/* The report's nearest-centroid step (src/report.rs:169-183, the O(k^2 d) part
 * of report::report): for each of the k centroids (k x d, row-major) the
 * nearest other centroid among live clusters (sizes[j] > 0; sizes NULL: all
 * live) and the plain L2 distance to it, (i, +inf) when there is none. The
 * distance is the reference's sqdist (src/report.rs:104-114) in its own f32
 * order, so indices and distances equal the reference's bit for bit. The
 * _device form takes device pointers, k u64 of workspace (keys_dev) and a
 * hipStream_t (as void *). */
int emh_report_nearest(int device, const float *centroids, uint64_t k, uint64_t d, const uint32_t *sizes,
                       uint32_t *nearest, float *nearest_dist);
int emh_report_nearest_device(const float *centroids_dev, uint64_t k, uint64_t d, const uint32_t *sizes_dev,
                              uint64_t *keys_dev, uint32_t *nearest_dev, float *dist_dev, void *stream);

/* report::report (src/report.rs:119-250) with its two O(k^2 d) steps on the
 * device (the nearest-centroid step above, and the family step's "within
 * radius" test, src/report.rs:208) and the per-cluster statistics, families
 * and z-scores on the host in the reference's arithmetic. Inputs are what
 * assign returns (labels < k, squared distances) and the k x d centroids.
 * Outputs: clusters[k] (ClusterStats field by field), local_z[n],
 * global_z[n], and the families sorted by decreasing size as parallel arrays
 * of capacity k: family_leader, family_size, family_share, and member lists
 * family_members[family_start[f] .. family_start[f + 1]) (family_start has
 * k + 1 entries). Clusters of equal size are taken in index order where the
 * reference's unstable sort leaves the order open. k <= 131072. */
typedef struct {
    float family_radius;     /* ReportConfig::family_radius (1.25) */
    uint64_t min_local_size; /* ReportConfig::min_local_size (30) */
} emh_report_config;
typedef struct {
    uint32_t size;
    float rms, median, mad, p90, p99, max, looseness;
    uint32_t nearest_centroid;
    float nearest_dist, isolation;
    uint32_t family;
} emh_cluster_stats;
typedef struct {
    float median_dist, median_rms, median_nearest_dist;
    uint32_t n_families;
} emh_report_summary;
void emh_report_config_default(emh_report_config *cfg);
int emh_report(int device, const uint32_t *labels, const float *sqdist, uint64_t n, const float *centroids,
               uint64_t k, uint64_t d, const emh_report_config *cfg, emh_cluster_stats *clusters, float *local_z,
               float *global_z, uint32_t *family_leader, uint64_t *family_size, double *family_share,
               uint32_t *family_start, uint32_t *family_members, emh_report_summary *summary);

/* Centroid-range split of the register-resident search (no counterpart in
 * the reference). A launch with few rows fills the chip by splitting every
 * row block's centroid tiles over several workgroups and merging their
 * per-row results; the results are the same for every split. nslice = 0
 * (the default) lets the built-in autotuner time the candidate splits the
 * first time a shape is seen and keep the fastest for the process
 * (EMH_AUTOTUNE=0 in the environment replaces the timing by a rule);
 * nslice >= 1 forces that many slices (fewer when the tiles run out). */
int emh_set_split(emh_assigner *h, int nslice);

/* Tile of the register-resident search (no counterpart in the reference):
 * a workgroup sweeps 128 x fragments rows (fragments = 1 or 2 per wave)
 * against its centroid slice. 0 (the default) lets the autotuner time the
 * candidate (fragments, split) pairs the first time a launch shape that
 * leaves the chip short of work is seen; its choices persist in a JSON file
 * (EMH_TUNE_CACHE, default ~/.cache/emeans_hip/tiles.json; "" keeps them in
 * memory) that later processes read. The results are the same for every
 * tile. emh_last_tile reports the tile of the assigner's last search launch
 * (fragments, centroid slices). */
int emh_set_tile(emh_assigner *h, int fragments);
int emh_last_tile(const emh_assigner *h, int *fragments, int *nslice);

/* Many independent problems, trained in lockstep (no counterpart in the
 * reference, which trains one problem per train() call, src/lib.rs:348).
 * Each job is one train() call's arguments; every job ends with the results
 * emh_train gives it on the same assigner (same precision and top-m). One
 * training step of all jobs still running is one replay of a captured
 * hipGraph: upload of every job's centroid snapshot, bias and batch ids,
 * then the gather + search (+ FP32 rescore) + per-cluster sums of all jobs,
 * then one download. Precision FP32, BF16_RESCORE or FP8_RESCORE on f32
 * rows, or I8 or I16 on i8 rows, any d: jobs whose operand rows pass 512
 * bytes (FP32: d > 128, BF16: d > 256, FP8 and I8: d > 512), and I16 jobs,
 * run the LDS search kernel in the same step. I8 jobs of at most 512-byte rows run the
 * packed-key search only: a balance bias with |sum c8^2 + round(bias)| >=
 * 2^26 - d 2^15 stops the call with EMH_E_UNSUPPORTED (emh_train then takes
 * the i64 search). */
typedef struct emh_job {
    const void *x; /* n x d rows, row-major */
    int rows;      /* enum emh_rows: EMH_ROWS_F32 (EMH_ROWS_I8 for the I8 search) */
    uint64_t n, d, k;
    emh_config cfg;
    float *centroids_out; /* k x d */
    uint32_t *labels_out; /* n */
    emh_train_out *out;
    int x_on_device;      /* nonzero: x points to device memory (rows already in HBM, not copied) */
} emh_job;

typedef struct emh_batch_stats {
    uint64_t steps;       /* graph replays: one per lockstep training step */
    uint64_t graph_nodes; /* nodes of the per-step graph */
    double seconds;       /* wall time of the call (row uploads included) */
    double host_seconds;  /* of which the host halves of the steps (staging, statistics, revival plans) */
    double wait_seconds;  /* and waiting for the replays to finish */
    double launch_seconds; /* and in the graph launch calls */
} emh_batch_stats;

int emh_train_batch(emh_assigner *h, emh_job *jobs, size_t njobs, emh_batch_stats *stats);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif
#endif
