/* CPU oracle for the e-means nearest-centroid search path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a plain-C restatement of the
 * reference crate (andrewluetgers/e-means, Rust) and is used as the checker
 * by tests/, by __graft_entry__.smoke() and as the CPU baseline leg of
 * bench.py. The product library (e-means_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it follows. Pinning of each
 * function (what checks it against known answers) is listed in
 * emeans_oracle.c's header comment.
 */
#ifndef EMEANS_ORACLE_H
#define EMEANS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG (src/lib.rs:227-274) ------------------------------------------ */
typedef struct { uint64_t s[4]; } or_rng;
void or_rng_init(or_rng *r, uint64_t seed);
uint64_t or_rng_next(or_rng *r);
uint64_t or_rng_below(or_rng *r, uint64_t bound);
double or_rng_open01(or_rng *r);
void or_rng_permutation(or_rng *r, uint32_t n, uint32_t *p);
void or_rng_fill_open01(or_rng *r, size_t count, double *out);
void or_splitmix64_stream(uint64_t seed, size_t count, uint64_t *out);

/* ---- Recipe (src/lib.rs:93-225) ---------------------------------------- */
enum { OR_LAW_AUTO = 0, OR_LAW_THROUGHPUT = 1, OR_LAW_QUALITY = 2 };
typedef struct {
    double nepochs;
    int32_t batch_law;
    int32_t has_batch_size;
    uint64_t batch_size;
    int32_t has_alpha0;
    int32_t has_unused_threshold;
    double alpha0;
    float unused_threshold;
    float balance;
    uint64_t seed;
} or_config;
typedef struct {
    uint64_t batch_size;
    double alpha0;
    uint64_t total_steps;
    float unused_threshold;
    uint32_t n_epochs;          /* entries in epoch_rows */
    uint64_t epoch_rows[1024];
} or_recipe;
void or_config_default(or_config *cfg);
int or_resolve_recipe(uint64_t n, uint64_t d, uint64_t k, const or_config *cfg, or_recipe *out);
double or_scheduled_alpha(double alpha0, uint64_t step, uint64_t total);

/* ---- State (src/lib.rs:276-345) ---------------------------------------- */
void or_state_apply(size_t k, size_t d, float *centroids, float *usage, const uint32_t *touched,
                    size_t ntouched, const uint32_t *counts, const float *sums, float alpha);
size_t or_state_revive(size_t k, size_t d, float *centroids, float *usage, float tau, or_rng *rng);

/* ---- CPU f32 search (src/cpu.rs) --------------------------------------- */
enum { OR_SIMD_PORTABLE = 0, OR_SIMD_AVX512 = 2 };
enum { OR_ROWS_F32 = 0, OR_ROWS_I8 = 1, OR_ROWS_U8 = 2 };
int or_simd_best(void);
float or_sqnorm(const float *v, size_t d);
/* norms[j] = sqnorm(c_j) (+ bias[j] when bias != NULL): set_centroids + set_bias */
void or_cpu_norms(const float *c, size_t k, size_t d, const float *bias, float *norms);
int or_cpu_assign(int simd, int nthreads, const void *x, int rows, size_t n, size_t d, const float *c,
                  const float *norms, size_t k, uint32_t *labels, float *dist);

/* ---- CPU int8 search (src/cpu_i8.rs) ----------------------------------- */
enum { OR_BITS_I8 = 0, OR_BITS_I16 = 1 };
#define OR_I8_MAX_D 32768
typedef struct {
    int bits, vnni;
    size_t k, d;
    uint32_t shift;
    int8_t *c8;
    int16_t *c16;
    int64_t *offset;
    int64_t mult;
} or_i8;
int or_i8_has_vnni(void);
void or_i8_init(or_i8 *a, int bits, int vnni);
void or_i8_free(or_i8 *a);
int or_i8_set_centroids(or_i8 *a, const float *c, size_t k, size_t d);
void or_i8_set_bias(or_i8 *a, const float *bias);
int or_i8_assign(const or_i8 *a, int nthreads, const int8_t *x, size_t n, uint32_t *labels, float *dist);
/* flat accessors for ctypes */
int or_i8_prepare(int bits, const float *c, size_t k, size_t d, const float *bias, int8_t *c8, int16_t *c16,
                  int64_t *offset, int64_t *mult, uint32_t *shift);
int or_i8_search(int bits, int vnni, int nthreads, const int8_t *x, size_t n, size_t d, const float *c,
                 size_t k, const float *bias, uint32_t *labels, float *dist);

/* ---- Quantizer (src/quantize.rs) --------------------------------------- */
void or_quant_fit(const float *sample, size_t n, size_t d, float target_peak, double quantile, int power_of_two,
                  float *shift, float *scale);
void or_quant_encode(const float *x, size_t n, size_t d, const float *shift, float scale, int8_t *out);
/* the same with the reference's parallel loops on nthreads pinned threads */
void or_quant_fit_mt(int nthreads, const float *sample, size_t n, size_t d, float target_peak, double quantile,
                     int power_of_two, float *shift, float *scale);
void or_quant_encode_mt(int nthreads, const float *x, size_t n, size_t d, const float *shift, float scale,
                        int8_t *out);
/* nonzero: pool threads are pinned one per CPU of the affinity mask */
int or_pool_pins_threads(void);
void or_quant_decode(const float *codes, size_t n, size_t d, const float *shift, float scale, float *out);
double or_quant_clipped_fraction(const float *x, size_t n, size_t d, const float *shift, float scale);

/* ---- Report's nearest-centroid step (src/report.rs:104-114, 169-183) ---- */
float or_report_sqdist(const float *a, const float *b, size_t d);
/* rows [row0, row1): nearest live other centroid and its L2 distance; sizes
 * NULL means every cluster is live. Outputs are indexed from row0. */
void or_report_nearest(int nthreads, const float *c, size_t k, size_t d, const uint32_t *sizes, size_t row0,
                       size_t row1, uint32_t *nearest, float *nearest_dist);

/* ---- Brute-force FP64 / exact references -------------------------------- */
void or_brute_f64(int nthreads, const void *x, int rows, size_t n, size_t d, const float *c, size_t k,
                  const float *bias, uint32_t *labels, double *dist);
void or_dist_f64(int nthreads, const void *x, int rows, size_t n, size_t d, const float *c, const uint32_t *labels,
                 double *dist);
void or_brute_i64(const int8_t *x, size_t n, size_t d, const int8_t *c, size_t k, uint32_t *labels, int64_t *dist);

/* ---- Training (src/lib.rs:347-520) ------------------------------------- */
enum { OR_ASSIGNER_F32 = 0, OR_ASSIGNER_I8 = 1, OR_ASSIGNER_I16 = 2 };
typedef struct {
    uint64_t n_splits;
    uint32_t n_objective;
    double objective[1024];
    or_recipe recipe;
} or_train_out;
int or_train(const void *x, int rows, size_t n, size_t d, size_t k, const or_config *cfg, int assigner, int simd,
             int nthreads, float *centroids, uint32_t *labels, or_train_out *out);
int or_assign(const void *x, int rows, size_t n, size_t d, const float *c, size_t k, int simd, int nthreads,
              uint32_t *labels, float *dist);
void or_balance_stats(const uint32_t *labels, size_t n, size_t k, double *imbalance, double *max_over_mean,
                      uint64_t *empty);

#ifdef __cplusplus
}
#endif
#endif
