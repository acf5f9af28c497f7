/* CPU oracle for the e-means nearest-centroid search path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference crate
 * (andrewluetgers/e-means, Rust). Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg load it; the product never does.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off, no fast-math: Rust never
 * contracts a*b+c, so neither may this file except where the reference calls
 * an FMA intrinsic). SIMD kernels carry per-function target attributes and are
 * chosen at run time, so the library runs on hosts without AVX-512.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   RNG            splitmix64 published first outputs for seed 0; a big-integer
 *                  re-derivation of xoshiro256++ from Vigna's published code;
 *                  permutation validity and uniformity.
 *   recipe         faiss 1.16.0 golden values held by src/lib.rs:526-540;
 *                  schedule endpoints src/lib.rs:542-546.
 *   apply/revive   Lloyd closed form src/lib.rs:548-562; revival invariants
 *                  src/lib.rs:564-573.
 *   f32 search     FP64 brute force with the src/cpu.rs:282-309 bound on its
 *                  seven shapes; portable and AVX-512 kernels agree.
 *   i8/i16 search  exact i64 brute force (src/cpu_i8.rs:294-324) and the
 *                  fractional-centroid case (src/cpu_i8.rs:326-336).
 *   quantizer      known-answer test src/quantize.rs:102-116.
 *   report nearest the report test's 1-D case (src/report.rs:283-312);
 *                  hand-derived f32 values that only the sixteen-lane order
 *                  with a separate remainder fold produces; exact integer
 *                  brute force; ties, self and empty-cluster exclusion.
 *   train          closing-mean invariant, loss bound and monotone objective
 *                  (src/lib.rs:575-614).
 * Parity unpinned: the exact f32 bits of the AVX-512 horizontal reduction
 * (src/cpu.rs:228-230 calls _mm512_reduce_add_ps, whose tree the Rust source
 * does not fix). This file reduces 512 -> 256 -> 128 -> pairs, the tree of the
 * published stdarch and GCC implementations.
 */
#define _GNU_SOURCE
#include "emeans_oracle.h"

#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ========================================================================= */
/* RNG: xoshiro256++ seeded by splitmix64 (src/lib.rs:227-274)               */
/* ========================================================================= */

static uint64_t splitmix_next(uint64_t *z) {
    /* src/lib.rs:233-239 */
    *z += 0x9E3779B97F4A7C15ull;
    uint64_t x = *z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

void or_splitmix64_stream(uint64_t seed, size_t count, uint64_t *out) {
    uint64_t z = seed;
    for (size_t i = 0; i < count; i++) out[i] = splitmix_next(&z);
}

void or_rng_init(or_rng *r, uint64_t seed) {
    /* src/lib.rs:231-241 */
    uint64_t z = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix_next(&z);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t or_rng_next(or_rng *r) {
    /* src/lib.rs:242-253 */
    uint64_t *s = r->s;
    uint64_t res = rotl64(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return res;
}

uint64_t or_rng_below(or_rng *r, uint64_t bound) {
    /* Lemire's unbiased multiply-shift, src/lib.rs:255-262 */
    for (;;) {
        unsigned __int128 m = (unsigned __int128)or_rng_next(r) * (unsigned __int128)bound;
        if ((uint64_t)m >= (0 - bound) % bound) return (uint64_t)(m >> 64);
    }
}

double or_rng_open01(or_rng *r) {
    /* src/lib.rs:264-266 */
    return ((double)(or_rng_next(r) >> 11) + 0.5) / (double)(1ull << 53);
}

void or_rng_fill_open01(or_rng *r, size_t count, double *out) {
    for (size_t i = 0; i < count; i++) out[i] = or_rng_open01(r);
}

void or_rng_permutation(or_rng *r, uint32_t n, uint32_t *p) {
    /* Fisher-Yates from the top, src/lib.rs:267-273 */
    for (uint32_t i = 0; i < n; i++) p[i] = i;
    for (uint32_t i = n; i-- > 1;) {
        uint64_t j = or_rng_below(r, (uint64_t)i + 1);
        uint32_t t = p[i];
        p[i] = p[j];
        p[j] = t;
    }
}

/* ========================================================================= */
/* Recipe (src/lib.rs:31-43, 126-225)                                        */
/* ========================================================================= */

static const double ALPHA_MIN_FACTOR = 0.1;
static const float INIT_USAGE = 1.0f / 65536.0f;
static const float SPLIT_EPSILON = 1e-4f;
static const float UNUSED_THRESHOLD = 0.1f;
static const size_t NO_REVIVAL_MAX_D = 2;
static const size_t AUTO_QUALITY_MAX_D = 16;
static const size_t CHECK_UNUSED_EVERY = 5;
static const double BATCH_THROUGHPUT[4] = {0.035134, 0.134749, 0.876473, 0.813344};
static const double BATCH_QUALITY[4] = {0.00439175, 0.134749, 0.876473, 0.813344};
static const double ALPHA_LAW[5] = {0.8244, -0.4617, 0.9003, -0.4498, -0.1057};
#define ASSIGN_CHUNK 65536u

void or_config_default(or_config *cfg) {
    /* src/lib.rs:126-139 */
    memset(cfg, 0, sizeof *cfg);
    cfg->nepochs = 5.0;
    cfg->batch_law = OR_LAW_AUTO;
    cfg->balance = 0.0f;
    cfg->seed = 1234;
}

static uint64_t round_power_of_two(double v) {
    /* src/lib.rs:165-172 */
    if (v <= 1.0) return 1;
    double low = ldexp(1.0, (int)floor(log2(v)));
    double high = 2.0 * low;
    return (uint64_t)((v / low < high / v) ? low : high);
}

static double step_law(uint64_t n, uint64_t k, double nepochs, uint64_t batch) {
    /* src/lib.rs:174-178 */
    double a = ALPHA_LAW[0] * pow((double)k, ALPHA_LAW[1]) * pow((double)batch, ALPHA_LAW[2]) *
               pow((double)n, ALPHA_LAW[3]) * pow(nepochs, ALPHA_LAW[4]);
    return a < 1e-4 ? 1e-4 : (a > 1.0 ? 1.0 : a);
}

static uint32_t epoch_rows(uint64_t n, double nepochs, uint64_t *out) {
    /* src/lib.rs:180-192 (round_ties_even == nearbyint in the default mode) */
    uint64_t whole = (uint64_t)floor(nepochs);
    uint64_t rem = (uint64_t)nearbyint((nepochs - (double)whole) * (double)n);
    uint32_t m = 0;
    if (whole == 0) {
        out[m++] = rem;
        return m;
    }
    for (uint64_t i = 0; i + 1 < whole && m < 1021; i++) out[m++] = n;
    if (rem > 0) out[m++] = rem;
    out[m++] = n;
    return m;
}

double or_scheduled_alpha(double alpha0, uint64_t step, uint64_t total) {
    /* src/lib.rs:194-198 */
    double frac = (double)(step < total ? step : total) / (double)(total > 1 ? total : 1);
    double cosine = 0.5 * (1.0 + cos(M_PI * frac));
    return alpha0 * (ALPHA_MIN_FACTOR + (1.0 - ALPHA_MIN_FACTOR) * cosine);
}

int or_resolve_recipe(uint64_t n, uint64_t d, uint64_t k, const or_config *cfg, or_recipe *out) {
    /* src/lib.rs:201-225 */
    if (cfg->nepochs >= 1023.0) return -1; /* epoch_rows table bound of this oracle */
    int law = cfg->batch_law;
    if (law == OR_LAW_AUTO) law = d <= AUTO_QUALITY_MAX_D ? OR_LAW_QUALITY : OR_LAW_THROUGHPUT;
    uint64_t batch;
    if (cfg->has_batch_size) {
        batch = cfg->batch_size < 1 ? 1 : (cfg->batch_size > n ? n : cfg->batch_size);
    } else {
        const double *c = law == OR_LAW_THROUGHPUT ? BATCH_THROUGHPUT : BATCH_QUALITY;
        double raw = c[0] * pow((double)k, c[1]) * pow((double)n, c[2]) * pow(cfg->nepochs, c[3]);
        uint64_t b = round_power_of_two(raw);
        uint64_t cap = law == OR_LAW_THROUGHPUT ? round_power_of_two((double)n / 2.0) : n / 2;
        if (b > cap) b = cap;
        if (b > n) b = n;
        if (b < 1) b = 1;
        batch = b;
    }
    out->batch_size = batch;
    out->alpha0 = cfg->has_alpha0 ? cfg->alpha0 : step_law(n, k, cfg->nepochs, batch);
    out->n_epochs = epoch_rows(n, cfg->nepochs, out->epoch_rows);
    out->total_steps = 0;
    for (uint32_t e = 0; e < out->n_epochs; e++) out->total_steps += (out->epoch_rows[e] + batch - 1) / batch;
    out->unused_threshold =
        cfg->has_unused_threshold ? cfg->unused_threshold : (d <= NO_REVIVAL_MAX_D ? 0.0f : UNUSED_THRESHOLD);
    return 0;
}

/* ========================================================================= */
/* Model state: EMA apply and revival (src/lib.rs:276-345)                   */
/* ========================================================================= */

void or_state_apply(size_t k, size_t d, float *centroids, float *usage, const uint32_t *touched, size_t ntouched,
                    const uint32_t *counts, const float *sums, float alpha) {
    /* src/lib.rs:287-309 */
    int replace = alpha == 1.0f;
    if (!replace) {
        float decay = 1.0f - alpha;
        for (size_t j = 0; j < k; j++) usage[j] *= decay;
    }
    for (size_t t = 0; t < ntouched; t++) {
        size_t j = touched[t];
        float n = (float)counts[j];
        float retained = replace ? 0.0f : usage[j];
        float an = alpha * n;
        float mass = retained + an;
        float weight = an / mass;
        float *c = centroids + j * d;
        const float *s = sums + j * d;
        if (weight == 1.0f) {
            for (size_t i = 0; i < d; i++) c[i] = s[i] / n;
        } else {
            for (size_t i = 0; i < d; i++) {
                float m = s[i] / n;
                c[i] = c[i] + weight * (m - c[i]);
            }
        }
        usage[j] = mass;
    }
}

typedef struct {
    double key;
    size_t j;
} donor_t;

static int donor_desc(const void *a, const void *b) {
    /* total_cmp descending (src/lib.rs:327); keys here are never NaN */
    double x = ((const donor_t *)a)->key, y = ((const donor_t *)b)->key;
    return (x < y) - (x > y);
}

size_t or_state_revive(size_t k, size_t d, float *centroids, float *usage, float tau, or_rng *rng) {
    /* src/lib.rs:314-344 */
    float total = 0.0f;
    for (size_t j = 0; j < k; j++) total += usage[j];
    float threshold = total * tau / (float)k;
    size_t *unused = malloc(k * sizeof *unused);
    donor_t *donors = malloc(k * sizeof *donors);
    size_t nu = 0, nd = 0;
    for (size_t j = 0; j < k; j++)
        if (usage[j] < threshold) unused[nu++] = j;
    if (nu == 0) {
        free(unused);
        free(donors);
        return 0;
    }
    for (size_t j = 0; j < k; j++) {
        if (usage[j] >= threshold && usage[j] > 0.0f) {
            double g = log(-log(or_rng_open01(rng)));
            donors[nd].key = log((double)usage[j]) - g;
            donors[nd].j = j;
            nd++;
        }
    }
    size_t nsplit = nu < nd ? nu : nd;
    qsort(donors, nd, sizeof *donors, donor_desc);
    for (size_t t = 0; t < nsplit; t++) {
        size_t src = donors[t].j, dst = unused[t];
        float half = usage[src] / 2.0f;
        for (size_t i = 0; i < d; i++) {
            float offset = (i % 2 == 0) ? SPLIT_EPSILON : -SPLIT_EPSILON;
            float c = centroids[src * d + i];
            centroids[src * d + i] = c * (1.0f - offset);
            centroids[dst * d + i] = c * (1.0f + offset);
        }
        int same = 1;
        for (size_t i = 0; i < d; i++)
            if (centroids[src * d + i] != centroids[dst * d + i]) same = 0;
        if (same) centroids[dst * d] = nextafterf(centroids[dst * d], INFINITY);
        usage[src] = half;
        usage[dst] = half;
    }
    free(unused);
    free(donors);
    return nsplit;
}

/* ========================================================================= */
/* CPU f32 search (src/cpu.rs)                                               */
/* ========================================================================= */

#define BLOCK_BYTES (256u * 1024u)

int or_simd_best(void) {
    /* src/cpu.rs:20-33 (the AVX2 level is not restated: AVX-512 or portable) */
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") ? OR_SIMD_AVX512 : OR_SIMD_PORTABLE;
}

float or_sqnorm(const float *v, size_t d) {
    /* 16 strided accumulators, then a sequential fold, then the tail:
       src/cpu.rs:170-180 */
    float acc[16] = {0};
    size_t body = d & ~(size_t)15;
    for (size_t i = 0; i < body; i += 16)
        for (int t = 0; t < 16; t++) acc[t] += v[i + t] * v[i + t];
    float s = 0.0f;
    for (int t = 0; t < 16; t++) s += acc[t];
    float tail = 0.0f;
    for (size_t i = body; i < d; i++) tail += v[i] * v[i];
    return s + tail;
}

void or_cpu_norms(const float *c, size_t k, size_t d, const float *bias, float *norms) {
    /* set_centroids (src/cpu.rs:84-93) then set_bias (src/cpu.rs:95-99) */
    for (size_t j = 0; j < k; j++) {
        norms[j] = or_sqnorm(c + j * d, d);
        if (bias) norms[j] += bias[j];
    }
}

typedef void (*dots_fn)(const float *const q[4], const float *const c[4], size_t d, float out[16]);

static void dots_portable(const float *const q[4], const float *const c[4], size_t d, float out[16]) {
    /* src/cpu.rs:182-198: 8 strided accumulators, fold, plus the tail sum */
    size_t body = d & ~(size_t)7;
    for (int qi = 0; qi < 4; qi++) {
        for (int ci = 0; ci < 4; ci++) {
            const float *a = q[qi], *b = c[ci];
            float acc[8] = {0};
            float tail = 0.0f;
            for (size_t i = body; i < d; i++) tail += a[i] * b[i];
            for (size_t i = 0; i < body; i += 8)
                for (int t = 0; t < 8; t++) acc[t] += a[i + t] * b[i + t];
            float s = 0.0f;
            for (int t = 0; t < 8; t++) s += acc[t];
            out[4 * qi + ci] = s + tail;
        }
    }
}

__attribute__((target("avx512f"))) static float hsum512(__m512 v) {
    /* 512 -> 256 -> 128 halves, then (0+2, 1+3), then lane 0 + lane 1 */
    __m256 lo = _mm512_castps512_ps256(v);
    __m256 hi = _mm256_castpd_ps(_mm512_extractf64x4_pd(_mm512_castps_pd(v), 1));
    __m256 a = _mm256_add_ps(lo, hi);
    __m128 b = _mm_add_ps(_mm256_castps256_ps128(a), _mm256_extractf128_ps(a, 1));
    __m128 c = _mm_add_ps(b, _mm_shuffle_ps(b, b, _MM_SHUFFLE(1, 0, 3, 2)));
    __m128 e = _mm_add_ss(c, _mm_shuffle_ps(c, c, _MM_SHUFFLE(0, 0, 0, 1)));
    return _mm_cvtss_f32(e);
}

__attribute__((target("avx512f"))) static void dots_avx512(const float *const q[4], const float *const c[4],
                                                            size_t d, float out[16]) {
    /* src/cpu.rs:200-231: 16 zmm accumulators, FMA, masked tail */
    __m512 acc[16];
    for (int t = 0; t < 16; t++) acc[t] = _mm512_setzero_ps();
    size_t i = 0;
    for (; i + 16 <= d; i += 16) {
        __m512 qv[4], cv[4];
        for (int t = 0; t < 4; t++) qv[t] = _mm512_loadu_ps(q[t] + i);
        for (int t = 0; t < 4; t++) cv[t] = _mm512_loadu_ps(c[t] + i);
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) acc[4 * a + b] = _mm512_fmadd_ps(qv[a], cv[b], acc[4 * a + b]);
    }
    if (i < d) {
        __mmask16 mask = (__mmask16)((1u << (d - i)) - 1);
        __m512 qv[4], cv[4];
        for (int t = 0; t < 4; t++) qv[t] = _mm512_maskz_loadu_ps(mask, q[t] + i);
        for (int t = 0; t < 4; t++) cv[t] = _mm512_maskz_loadu_ps(mask, c[t] + i);
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) acc[4 * a + b] = _mm512_fmadd_ps(qv[a], cv[b], acc[4 * a + b]);
    }
    for (int t = 0; t < 16; t++) out[t] = hsum512(acc[t]);
}

static size_t row_bytes(int rows) { return rows == OR_ROWS_F32 ? 4 : 1; }

static void widen_rows(const void *x, int rows, size_t count, float *out) {
    /* Scalar::to_f32 (src/lib.rs:47-71) */
    if (rows == OR_ROWS_I8) {
        const int8_t *p = x;
        for (size_t i = 0; i < count; i++) out[i] = (float)p[i];
    } else if (rows == OR_ROWS_U8) {
        const uint8_t *p = x;
        for (size_t i = 0; i < count; i++) out[i] = (float)p[i];
    } else {
        memcpy(out, x, count * sizeof(float));
    }
}

/* Row-block work queue shared by the threaded searches (rayon par_chunks). */
typedef struct {
    atomic_size_t next;
    size_t nblocks;
    void *ctx;
    void (*fn)(void *ctx, size_t block);
} or_pool;

static void *pool_worker(void *arg) {
    or_pool *p = arg;
    for (;;) {
        size_t b = atomic_fetch_add(&p->next, 1);
        if (b >= p->nblocks) break;
        p->fn(p->ctx, b);
    }
    return NULL;
}

/* Worker t runs on the t-th CPU of the caller's affinity mask (the
 * caller is worker 0 and gets its own mask back afterwards), so a timed run
 * uses each allowed core once. */
static int allowed_cpus(int *cpus, int cap) {
    cpu_set_t all;
    CPU_ZERO(&all);
    if (sched_getaffinity(0, sizeof all, &all) != 0) return 0;
    int n = 0;
    for (int c = 0; c < CPU_SETSIZE && n < cap; c++)
        if (CPU_ISSET(c, &all)) cpus[n++] = c;
    return n;
}

static void run_pool(int nthreads, size_t nblocks, void *ctx, void (*fn)(void *, size_t)) {
    or_pool p;
    atomic_init(&p.next, 0);
    p.nblocks = nblocks;
    p.ctx = ctx;
    p.fn = fn;
    if (nthreads <= 1 || nblocks <= 1) {
        pool_worker(&p);
        return;
    }
    if ((size_t)nthreads > nblocks) nthreads = (int)nblocks;
    pthread_t th[256];
    if (nthreads > 256) nthreads = 256;
    int cpus[256];
    const int ncpu = allowed_cpus(cpus, 256);
    for (int t = 1; t < nthreads; t++) {
        pthread_attr_t at;
        pthread_attr_init(&at);
        if (ncpu > 0) {
            cpu_set_t one;
            CPU_ZERO(&one);
            CPU_SET(cpus[t % ncpu], &one);
            pthread_attr_setaffinity_np(&at, sizeof one, &one);
        }
        pthread_create(&th[t], &at, pool_worker, &p);
        pthread_attr_destroy(&at);
    }
    cpu_set_t mine;
    const int saved = ncpu > 0 && pthread_getaffinity_np(pthread_self(), sizeof mine, &mine) == 0;
    if (saved) {
        cpu_set_t one;
        CPU_ZERO(&one);
        CPU_SET(cpus[0], &one);
        pthread_setaffinity_np(pthread_self(), sizeof one, &one);
    }
    pool_worker(&p);
    for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
    if (saved) pthread_setaffinity_np(pthread_self(), sizeof mine, &mine);
}

int or_pool_pins_threads(void) {
    int cpus[1];
    return allowed_cpus(cpus, 1) > 0;
}

typedef struct {
    dots_fn kernel;
    const void *x;
    int rows;
    size_t n, d, k, block;
    const float *c, *norms;
    uint32_t *labels;
    float *dist;
} f32_job;

static void f32_block(void *vctx, size_t blk) {
    /* One rayon task of src/cpu.rs:118-164 */
    f32_job *J = vctx;
    size_t d = J->d, k = J->k, block = J->block;
    size_t r0 = blk * block;
    size_t nq = J->n - r0 < block ? J->n - r0 : block;
    const float *x;
    float *converted = NULL;
    if (J->rows == OR_ROWS_F32) {
        x = (const float *)J->x + r0 * d;
    } else {
        converted = malloc(nq * d * sizeof(float));
        widen_rows((const char *)J->x + r0 * d * row_bytes(J->rows), J->rows, nq * d, converted);
        x = converted;
    }
    uint32_t *labels = J->labels + r0;
    float *best = J->dist + r0;
    for (size_t i = 0; i < nq; i++) {
        best[i] = INFINITY;
        labels[i] = 0;
    }
    float out[16];
    for (size_t c0 = 0; c0 < k; c0 += block) {
        size_t c1 = c0 + block < k ? c0 + block : k;
        for (size_t q0 = 0; q0 < nq; q0 += 4) {
            size_t qn = nq - q0 < 4 ? nq - q0 : 4;
            const float *qp[4];
            for (size_t i = 0; i < 4; i++) qp[i] = x + (q0 + (i < qn - 1 ? i : qn - 1)) * d;
            for (size_t j0 = c0; j0 < c1; j0 += 4) {
                size_t cn = c1 - j0 < 4 ? c1 - j0 : 4;
                const float *cp[4];
                for (size_t i = 0; i < 4; i++) cp[i] = J->c + (j0 + (i < cn - 1 ? i : cn - 1)) * d;
                J->kernel(qp, cp, d, out);
                for (size_t qi = 0; qi < qn; qi++) {
                    float b = best[q0 + qi];
                    uint32_t l = labels[q0 + qi];
                    for (size_t ci = 0; ci < cn; ci++) {
                        float s = J->norms[j0 + ci] - 2.0f * out[4 * qi + ci];
                        if (s < b) {
                            b = s;
                            l = (uint32_t)(j0 + ci);
                        }
                    }
                    best[q0 + qi] = b;
                    labels[q0 + qi] = l;
                }
            }
        }
    }
    for (size_t i = 0; i < nq; i++) {
        float v = best[i] + or_sqnorm(x + i * d, d);
        best[i] = v > 0.0f ? v : 0.0f; /* f32::max(v, 0): NaN gives 0 */
    }
    free(converted);
}

int or_cpu_assign(int simd, int nthreads, const void *x, int rows, size_t n, size_t d, const float *c,
                  const float *norms, size_t k, uint32_t *labels, float *dist) {
    /* CpuAssigner::assign, src/cpu.rs:108-167 */
    if (k == 0 || d == 0) return -1;
    if (n == 0) return 0;
    f32_job J;
    J.kernel = (simd == OR_SIMD_AVX512 && or_simd_best() == OR_SIMD_AVX512) ? dots_avx512 : dots_portable;
    J.x = x;
    J.rows = rows;
    J.n = n;
    J.d = d;
    J.k = k;
    size_t block = BLOCK_BYTES / (4 * d);
    if (block < 4) block = 4;
    if (block > 4096) block = 4096;
    J.block = block & ~(size_t)3;
    J.c = c;
    J.norms = norms;
    J.labels = labels;
    J.dist = dist;
    run_pool(nthreads, (n + J.block - 1) / J.block, &J, f32_block);
    return 0;
}

/* ========================================================================= */
/* CPU int8 search (src/cpu_i8.rs)                                           */
/* ========================================================================= */

int or_i8_has_vnni(void) {
    /* src/cpu_i8.rs:48-58 */
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
           __builtin_cpu_supports("avx512vl") && __builtin_cpu_supports("avx512vnni");
}

void or_i8_init(or_i8 *a, int bits, int vnni) {
    memset(a, 0, sizeof *a);
    a->bits = bits;
    a->vnni = vnni && or_i8_has_vnni();
    a->mult = 2;
}

void or_i8_free(or_i8 *a) {
    free(a->c8);
    free(a->c16);
    free(a->offset);
    a->c8 = NULL;
    a->c16 = NULL;
    a->offset = NULL;
}

static float clampf(float v, float lo, float hi) {
    /* f32::clamp, then `as` casts NaN to 0 */
    if (v != v) return 0.0f;
    return v < lo ? lo : (v > hi ? hi : v);
}

int or_i8_set_centroids(or_i8 *a, const float *c, size_t k, size_t d) {
    /* src/cpu_i8.rs:73-110 */
    if (d > OR_I8_MAX_D) return -1;
    or_i8_free(a);
    a->k = k;
    a->d = d;
    a->offset = malloc(k * sizeof(int64_t));
    if (a->bits == OR_BITS_I8) {
        a->shift = 0;
        a->c8 = malloc(k * d);
        for (size_t i = 0; i < k * d; i++) a->c8[i] = (int8_t)clampf(roundf(c[i]), -128.0f, 127.0f);
        for (size_t j = 0; j < k; j++) {
            int64_t o = 0;
            for (size_t i = 0; i < d; i++) {
                int64_t v = a->c8[j * d + i];
                o += v * v + 256 * v;
            }
            a->offset[j] = o;
        }
        a->mult = 2;
    } else {
        uint32_t s = 0;
        for (int t = 7; t >= 0; t--) {
            if (((uint64_t)d << (14 + t)) < (1ull << 31)) {
                s = (uint32_t)t;
                break;
            }
        }
        a->shift = s;
        float scale = (float)(1u << s);
        a->c16 = malloc(k * d * sizeof(int16_t));
        for (size_t i = 0; i < k * d; i++) a->c16[i] = (int16_t)clampf(roundf(c[i] * scale), -32768.0f, 32767.0f);
        for (size_t j = 0; j < k; j++) {
            int64_t o = 0;
            for (size_t i = 0; i < d; i++) {
                int64_t v = a->c16[j * d + i];
                o += v * v;
            }
            a->offset[j] = o;
        }
        a->mult = (int64_t)2 << s;
    }
    return 0;
}

void or_i8_set_bias(or_i8 *a, const float *bias) {
    /* src/cpu_i8.rs:112-117 */
    double unit = (double)(1ull << (2 * a->shift));
    for (size_t j = 0; j < a->k; j++) a->offset[j] += (int64_t)round((double)bias[j] * unit);
}

typedef void (*dots8_fn)(const int8_t *const q[4], const int8_t *const c[4], size_t d, int32_t out[16]);
typedef void (*dots16_fn)(const int8_t *const q[4], const int16_t *const c[4], size_t d, int32_t out[16]);

static void dots8_portable(const int8_t *const q[4], const int8_t *const c[4], size_t d, int32_t out[16]) {
    /* src/cpu_i8.rs:193-202: sum (x + 128) * c in i32 */
    for (int qi = 0; qi < 4; qi++)
        for (int ci = 0; ci < 4; ci++) {
            /* i32 sum wrapping as the VNNI lanes do (each product fits) */
            uint32_t s = 0;
            for (size_t i = 0; i < d; i++) s += (uint32_t)(((int32_t)q[qi][i] + 128) * (int32_t)c[ci][i]);
            out[4 * qi + ci] = (int32_t)s;
        }
}

static void dots16_portable(const int8_t *const q[4], const int16_t *const c[4], size_t d, int32_t out[16]) {
    /* src/cpu_i8.rs:204-212 */
    for (int qi = 0; qi < 4; qi++)
        for (int ci = 0; ci < 4; ci++) {
            uint32_t s = 0;
            for (size_t i = 0; i < d; i++) s += (uint32_t)((int32_t)q[qi][i] * (int32_t)c[ci][i]);
            out[4 * qi + ci] = (int32_t)s;
        }
}

__attribute__((target("avx512f,avx512bw,avx512vl,avx512vnni"))) static void dots8_vnni(const int8_t *const q[4],
                                                                                     const int8_t *const c[4],
                                                                                     size_t d, int32_t out[16]) {
    /* vpdpbusd on (x ^ 0x80) = x + 128 as u8, src/cpu_i8.rs:231-255 */
    __m512i flip = _mm512_set1_epi8((char)-128);
    __m512i acc[16];
    for (int t = 0; t < 16; t++) acc[t] = _mm512_setzero_si512();
    for (size_t i = 0; i < d; i += 64) {
        size_t left = d - i;
        __mmask64 mask = left >= 64 ? ~(__mmask64)0 : (((__mmask64)1 << left) - 1);
        __m512i qv[4], cv[4];
        for (int t = 0; t < 4; t++) qv[t] = _mm512_xor_si512(_mm512_maskz_loadu_epi8(mask, q[t] + i), flip);
        for (int t = 0; t < 4; t++) cv[t] = _mm512_maskz_loadu_epi8(mask, c[t] + i);
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) acc[4 * a + b] = _mm512_dpbusd_epi32(acc[4 * a + b], qv[a], cv[b]);
    }
    for (int t = 0; t < 16; t++) out[t] = _mm512_reduce_add_epi32(acc[t]);
}

__attribute__((target("avx512f,avx512bw,avx512vl,avx512vnni"))) static void dots16_vnni(const int8_t *const q[4],
                                                                                      const int16_t *const c[4],
                                                                                      size_t d, int32_t out[16]) {
    /* vpdpwssd, src/cpu_i8.rs:257-278 */
    __m512i acc[16];
    for (int t = 0; t < 16; t++) acc[t] = _mm512_setzero_si512();
    for (size_t i = 0; i < d; i += 32) {
        size_t left = d - i;
        __mmask32 mask = left >= 32 ? ~(__mmask32)0 : (((__mmask32)1 << left) - 1);
        __m512i qv[4], cv[4];
        for (int t = 0; t < 4; t++) qv[t] = _mm512_cvtepi8_epi16(_mm256_maskz_loadu_epi8(mask, q[t] + i));
        for (int t = 0; t < 4; t++) cv[t] = _mm512_maskz_loadu_epi16(mask, c[t] + i);
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) acc[4 * a + b] = _mm512_dpwssd_epi32(acc[4 * a + b], qv[a], cv[b]);
    }
    for (int t = 0; t < 16; t++) out[t] = _mm512_reduce_add_epi32(acc[t]);
}

typedef struct {
    const or_i8 *a;
    const int8_t *x;
    size_t n, block;
    uint32_t *labels;
    float *dist;
} i8_job;

static void i8_block(void *vctx, size_t blk) {
    /* One rayon task of src/cpu_i8.rs:144-189 */
    i8_job *J = vctx;
    const or_i8 *a = J->a;
    size_t d = a->d, k = a->k, block = J->block;
    int wide = a->bits == OR_BITS_I16;
    dots8_fn k8 = a->vnni ? dots8_vnni : dots8_portable;
    dots16_fn k16 = a->vnni ? dots16_vnni : dots16_portable;
    size_t r0 = blk * block;
    size_t nq = J->n - r0 < block ? J->n - r0 : block;
    const int8_t *x = J->x + r0 * d;
    uint32_t *labels = J->labels + r0;
    int64_t *best = malloc(nq * sizeof(int64_t));
    for (size_t i = 0; i < nq; i++) {
        best[i] = INT64_MAX;
        labels[i] = 0;
    }
    int32_t out[16];
    for (size_t c0 = 0; c0 < k; c0 += block) {
        size_t c1 = c0 + block < k ? c0 + block : k;
        for (size_t q0 = 0; q0 < nq; q0 += 4) {
            size_t qn = nq - q0 < 4 ? nq - q0 : 4;
            const int8_t *qp[4];
            for (size_t i = 0; i < 4; i++) qp[i] = x + (q0 + (i < qn - 1 ? i : qn - 1)) * d;
            for (size_t j0 = c0; j0 < c1; j0 += 4) {
                size_t cn = c1 - j0 < 4 ? c1 - j0 : 4;
                if (wide) {
                    const int16_t *cp[4];
                    for (size_t i = 0; i < 4; i++) cp[i] = a->c16 + (j0 + (i < cn - 1 ? i : cn - 1)) * d;
                    k16(qp, cp, d, out);
                } else {
                    const int8_t *cp[4];
                    for (size_t i = 0; i < 4; i++) cp[i] = a->c8 + (j0 + (i < cn - 1 ? i : cn - 1)) * d;
                    k8(qp, cp, d, out);
                }
                for (size_t qi = 0; qi < qn; qi++) {
                    int64_t b = best[q0 + qi];
                    uint32_t l = labels[q0 + qi];
                    for (size_t ci = 0; ci < cn; ci++) {
                        int64_t s = a->offset[j0 + ci] - a->mult * (int64_t)out[4 * qi + ci];
                        if (s < b) {
                            b = s;
                            l = (uint32_t)(j0 + ci);
                        }
                    }
                    best[q0 + qi] = b;
                    labels[q0 + qi] = l;
                }
            }
        }
    }
    double inv_unit = 1.0 / (double)(1ull << (2 * a->shift));
    for (size_t i = 0; i < nq; i++) {
        int64_t norm = 0;
        for (size_t t = 0; t < d; t++) norm += (int64_t)x[i * d + t] * (int64_t)x[i * d + t];
        double v = (double)norm + (double)best[i] * inv_unit;
        J->dist[r0 + i] = (float)(v > 0.0 ? v : 0.0);
    }
    free(best);
}

int or_i8_assign(const or_i8 *a, int nthreads, const int8_t *x, size_t n, uint32_t *labels, float *dist) {
    /* CpuI8Assigner::assign, src/cpu_i8.rs:130-191 */
    if (a->k == 0) return -1;
    if (n == 0) return 0;
    size_t elem = a->bits == OR_BITS_I16 ? 2 : 1;
    size_t block = BLOCK_BYTES / (elem * a->d);
    if (block < 4) block = 4;
    if (block > 8192) block = 8192;
    block &= ~(size_t)3;
    i8_job J = {a, x, n, block, labels, dist};
    run_pool(nthreads, (n + block - 1) / block, &J, i8_block);
    return 0;
}

int or_i8_prepare(int bits, const float *c, size_t k, size_t d, const float *bias, int8_t *c8, int16_t *c16,
                  int64_t *offset, int64_t *mult, uint32_t *shift) {
    or_i8 a;
    or_i8_init(&a, bits, 0);
    if (or_i8_set_centroids(&a, c, k, d)) return -1;
    if (bias) or_i8_set_bias(&a, bias);
    if (c8 && a.c8) memcpy(c8, a.c8, k * d);
    if (c16 && a.c16) memcpy(c16, a.c16, k * d * 2);
    memcpy(offset, a.offset, k * sizeof(int64_t));
    *mult = a.mult;
    *shift = a.shift;
    or_i8_free(&a);
    return 0;
}

int or_i8_search(int bits, int vnni, int nthreads, const int8_t *x, size_t n, size_t d, const float *c, size_t k,
                 const float *bias, uint32_t *labels, float *dist) {
    or_i8 a;
    or_i8_init(&a, bits, vnni);
    if (or_i8_set_centroids(&a, c, k, d)) return -1;
    if (bias) or_i8_set_bias(&a, bias);
    int rc = or_i8_assign(&a, nthreads, x, n, labels, dist);
    or_i8_free(&a);
    return rc;
}

/* ========================================================================= */
/* Quantizer (src/quantize.rs:43-96)                                         */
/* ========================================================================= */

static uint32_t f32_total_key(float v) {
    /* ordering key of f32::total_cmp */
    uint32_t u;
    memcpy(&u, &v, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

static float select_nth_total(float *a, size_t len, size_t idx) {
    /* the idx-th smallest element under total_cmp (quickselect; the value is
       what select_nth_unstable_by returns, whatever the algorithm) */
    size_t lo = 0, hi = len - 1;
    uint64_t st = 0x9E3779B97F4A7C15ull;
    while (lo < hi) {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        size_t p = lo + (size_t)(st % (hi - lo + 1));
        uint32_t pk = f32_total_key(a[p]);
        size_t i = lo, lt = lo, gt = hi;
        while (i <= gt) {
            uint32_t key = f32_total_key(a[i]);
            if (key < pk) {
                float t = a[lt];
                a[lt] = a[i];
                a[i] = t;
                lt++;
                i++;
            } else if (key > pk) {
                float t = a[gt];
                a[gt] = a[i];
                a[i] = t;
                if (gt == 0) break;
                gt--;
            } else {
                i++;
            }
        }
        if (idx < lt) {
            hi = lt - 1;
        } else if (idx > gt) {
            lo = gt + 1;
        } else {
            return a[idx];
        }
    }
    return a[idx];
}

typedef struct {
    const float *x;
    size_t n, d;
    const float *shift;
    float scale;
    int8_t *out;   /* encode */
    float *mags;   /* fit: |x - shift| */
} quant_job;

#define QUANT_BLOCK 1024

static void mags_block(void *ctx, size_t b) {
    const quant_job *J = ctx;
    const size_t r1 = (b + 1) * QUANT_BLOCK < J->n ? (b + 1) * QUANT_BLOCK : J->n;
    for (size_t r = b * QUANT_BLOCK; r < r1; r++)
        for (size_t t = 0; t < J->d; t++) J->mags[r * J->d + t] = fabsf(J->x[r * J->d + t] - J->shift[t]);
}

static void encode_block(void *ctx, size_t b) {
    const quant_job *J = ctx;
    const size_t r1 = (b + 1) * QUANT_BLOCK < J->n ? (b + 1) * QUANT_BLOCK : J->n;
    for (size_t r = b * QUANT_BLOCK; r < r1; r++)
        for (size_t t = 0; t < J->d; t++)
            J->out[r * J->d + t] = (int8_t)clampf(roundf((J->x[r * J->d + t] - J->shift[t]) * J->scale), -128.0f, 127.0f);
}

void or_quant_fit_mt(int nthreads, const float *sample, size_t n, size_t d, float target_peak, double quantile,
                     int power_of_two, float *shift, float *scale) {
    /* Int8Quantizer::fit, src/quantize.rs:45-62: the per-dimension means in
     * row order (a sequential f64 fold, as the reference), the magnitudes
     * in parallel row blocks (par_chunks_exact), one selection */
    double *mean = calloc(d, sizeof(double));
    for (size_t r = 0; r < n; r++)
        for (size_t t = 0; t < d; t++) mean[t] += (double)sample[r * d + t];
    for (size_t t = 0; t < d; t++) shift[t] = (float)(mean[t] / (double)n);
    free(mean);
    size_t len = n * d;
    float *mags = malloc(len * sizeof(float));
    quant_job J = {sample, n, d, shift, 0.0f, NULL, mags};
    run_pool(nthreads, (n + QUANT_BLOCK - 1) / QUANT_BLOCK, &J, mags_block);
    double q = quantile < 0.0 ? 0.0 : (quantile > 1.0 ? 1.0 : quantile);
    size_t idx = (size_t)round((double)(len - 1) * q);
    float peak = select_nth_total(mags, len, idx);
    free(mags);
    float s = peak > 0.0f ? target_peak / peak : 1.0f;
    if (power_of_two) s = ldexpf(1.0f, (int)floorf(log2f(s)));
    *scale = s;
}

void or_quant_fit(const float *sample, size_t n, size_t d, float target_peak, double quantile, int power_of_two,
                  float *shift, float *scale) {
    or_quant_fit_mt(1, sample, n, d, target_peak, quantile, power_of_two, shift, scale);
}

void or_quant_encode_mt(int nthreads, const float *x, size_t n, size_t d, const float *shift, float scale,
                        int8_t *out) {
    /* src/quantize.rs:65-73 (par_chunks_mut over rows) */
    quant_job J = {x, n, d, shift, scale, out, NULL};
    run_pool(nthreads, (n + QUANT_BLOCK - 1) / QUANT_BLOCK, &J, encode_block);
}

void or_quant_encode(const float *x, size_t n, size_t d, const float *shift, float scale, int8_t *out) {
    or_quant_encode_mt(1, x, n, d, shift, scale, out);
}

void or_quant_decode(const float *codes, size_t n, size_t d, const float *shift, float scale, float *out) {
    /* src/quantize.rs:76-79 */
    float inv = 1.0f / scale;
    for (size_t r = 0; r < n; r++)
        for (size_t t = 0; t < d; t++) out[r * d + t] = codes[r * d + t] * inv + shift[t];
}

double or_quant_clipped_fraction(const float *x, size_t n, size_t d, const float *shift, float scale) {
    /* src/quantize.rs:87-95 */
    size_t clipped = 0;
    for (size_t r = 0; r < n; r++)
        for (size_t t = 0; t < d; t++) {
            float v = (x[r * d + t] - shift[t]) * scale;
            if (!(v >= -128.5f && v < 127.5f)) clipped++;
        }
    return (double)clipped / (double)(n * d);
}

/* ========================================================================= */
/* Brute-force FP64 / exact references                                       */
/* ========================================================================= */

static double row_elem(const void *x, int rows, size_t idx) {
    if (rows == OR_ROWS_I8) return (double)((const int8_t *)x)[idx];
    if (rows == OR_ROWS_U8) return (double)((const uint8_t *)x)[idx];
    return (double)((const float *)x)[idx];
}

typedef struct {
    const void *x;
    int rows;
    size_t n, d, k;
    const float *c, *bias;
    const uint32_t *in_labels;
    uint32_t *labels;
    double *dist;
} brute_job;

static void brute_block(void *vctx, size_t blk) {
    brute_job *J = vctx;
    size_t r0 = blk * 64, r1 = r0 + 64 < J->n ? r0 + 64 : J->n;
    size_t d = J->d;
    double *xr = malloc(d * sizeof(double));
    for (size_t i = r0; i < r1; i++) {
        for (size_t t = 0; t < d; t++) xr[t] = row_elem(J->x, J->rows, i * d + t);
        if (J->in_labels) {
            const float *c = J->c + (size_t)J->in_labels[i] * d;
            double s = 0.0;
            for (size_t t = 0; t < d; t++) {
                double df = xr[t] - (double)c[t];
                s += df * df;
            }
            J->dist[i] = s;
            continue;
        }
        double best = INFINITY;
        uint32_t bl = 0;
        for (size_t j = 0; j < J->k; j++) {
            const float *c = J->c + j * d;
            double s = 0.0;
            for (size_t t = 0; t < d; t++) {
                double df = xr[t] - (double)c[t];
                s += df * df;
            }
            if (J->bias) s += (double)J->bias[j];
            if (s < best) {
                best = s;
                bl = (uint32_t)j;
            }
        }
        J->labels[i] = bl;
        J->dist[i] = best;
    }
    free(xr);
}

void or_brute_f64(int nthreads, const void *x, int rows, size_t n, size_t d, const float *c, size_t k,
                  const float *bias, uint32_t *labels, double *dist) {
    /* FP64 exact search with lowest-index ties (the reference of src/cpu.rs:298-306) */
    brute_job J = {x, rows, n, d, k, c, bias, NULL, labels, dist};
    run_pool(nthreads, (n + 63) / 64, &J, brute_block);
}

void or_dist_f64(int nthreads, const void *x, int rows, size_t n, size_t d, const float *c, const uint32_t *labels,
                 double *dist) {
    brute_job J = {x, rows, n, d, 0, c, NULL, labels, NULL, dist};
    run_pool(nthreads, (n + 63) / 64, &J, brute_block);
}

void or_brute_i64(const int8_t *x, size_t n, size_t d, const int8_t *c, size_t k, uint32_t *labels, int64_t *dist) {
    /* exact integer search, min by (distance, index): src/cpu_i8.rs:303-310 */
    for (size_t i = 0; i < n; i++) {
        int64_t best = INT64_MAX;
        uint32_t bl = 0;
        for (size_t j = 0; j < k; j++) {
            int64_t s = 0;
            for (size_t t = 0; t < d; t++) {
                int64_t df = (int64_t)x[i * d + t] - (int64_t)c[j * d + t];
                s += df * df;
            }
            if (s < best) {
                best = s;
                bl = (uint32_t)j;
            }
        }
        labels[i] = bl;
        dist[i] = best;
    }
}

/* ========================================================================= */
/* Training (src/lib.rs:347-520)                                             */
/* ========================================================================= */

typedef struct {
    int kind, simd, nthreads;
    const float *c;
    size_t k, d;
    float *norms;
    or_i8 i8;
} or_assigner;

static int asg_set_centroids(or_assigner *A, const float *c, size_t k, size_t d) {
    A->c = c;
    A->k = k;
    A->d = d;
    if (A->kind == OR_ASSIGNER_F32) {
        or_cpu_norms(c, k, d, NULL, A->norms);
        return 0;
    }
    return or_i8_set_centroids(&A->i8, c, k, d);
}

static void asg_set_bias(or_assigner *A, const float *bias) {
    if (A->kind == OR_ASSIGNER_F32) {
        for (size_t j = 0; j < A->k; j++) A->norms[j] += bias[j];
    } else {
        or_i8_set_bias(&A->i8, bias);
    }
}

static int asg_assign(or_assigner *A, const void *x, int rows, size_t m, uint32_t *labels, float *dist) {
    if (A->kind == OR_ASSIGNER_F32)
        return or_cpu_assign(A->simd, A->nthreads, x, rows, m, A->d, A->c, A->norms, A->k, labels, dist);
    if (rows != OR_ROWS_I8) return -1; /* CpuI8Assigner implements AssignRows<i8> only */
    return or_i8_assign(&A->i8, A->nthreads, x, m, labels, dist);
}

int or_train(const void *x, int rows, size_t n, size_t d, size_t k, const or_config *cfg, int assigner, int simd,
             int nthreads, float *centroids_out, uint32_t *labels, or_train_out *out) {
    /* train, src/lib.rs:348-491 */
    if (d == 0 || k == 0) return -1;
    if (n < k) return -2;
    if (n > UINT32_MAX) return -3;
    if (!(cfg->nepochs > 0.0)) return -4;
    if (assigner != OR_ASSIGNER_F32 && rows != OR_ROWS_I8) return -5;
    or_recipe R;
    if (or_resolve_recipe(n, d, k, cfg, &R)) return -6;
    size_t batch = R.batch_size;
    size_t rb = row_bytes(rows);

    uint32_t *init = malloc(n * sizeof(uint32_t));
    or_rng rng;
    or_rng_init(&rng, cfg->seed);
    or_rng_permutation(&rng, (uint32_t)n, init);
    float *C = malloc(k * d * sizeof(float));
    for (size_t j = 0; j < k; j++) widen_rows((const char *)x + (size_t)init[j] * d * rb, rows, d, C + j * d);
    free(init);
    float *usage = malloc(k * sizeof(float));
    for (size_t j = 0; j < k; j++) usage[j] = INIT_USAGE;
    size_t n_splits = 0;
    or_rng split_rng;
    or_rng_init(&split_rng, cfg->seed ^ 0x5EED5EED5EED5EEDull);

    or_assigner A;
    memset(&A, 0, sizeof A);
    A.kind = assigner;
    A.simd = simd;
    A.nthreads = nthreads;
    A.norms = malloc(k * sizeof(float));
    if (assigner != OR_ASSIGNER_F32) or_i8_init(&A.i8, assigner == OR_ASSIGNER_I16 ? OR_BITS_I16 : OR_BITS_I8, 1);

    size_t chunk = ASSIGN_CHUNK < batch ? ASSIGN_CHUNK : batch;
    char *rowbuf = malloc(chunk * d * rb);
    float *rowf = malloc(d * sizeof(float));
    uint32_t *chunk_labels = malloc(chunk * sizeof(uint32_t));
    float *chunk_dist = malloc(chunk * sizeof(float));
    uint32_t *counts = calloc(k, sizeof(uint32_t));
    float *sums = calloc(k * d, sizeof(float));
    uint32_t *touched = malloc(k * sizeof(uint32_t));
    size_t ntouched = 0;
    for (size_t i = 0; i < n; i++) labels[i] = UINT32_MAX;
    uint64_t *pass_counts = calloc(k, sizeof(uint64_t));
    float *pass_sums = calloc(k * d, sizeof(float));
    float *bias = cfg->balance > 0.0f ? calloc(k, sizeof(float)) : NULL;
    double mean_sqdist = 0.0;
    size_t step = 0;
    uint32_t last = R.n_epochs - 1;
    uint32_t *order = malloc(n * sizeof(uint32_t));
    int rc = 0;
    out->n_objective = 0;

    for (uint32_t epoch = 0; epoch < R.n_epochs && rc == 0; epoch++) {
        size_t epoch_n = R.epoch_rows[epoch];
        or_rng er;
        or_rng_init(&er, cfg->seed * 1000003ull + (uint64_t)epoch + 1);
        or_rng_permutation(&er, (uint32_t)n, order);
        int is_last = epoch == last;
        if (is_last) {
            memset(pass_counts, 0, k * sizeof(uint64_t));
            memset(pass_sums, 0, k * d * sizeof(float));
        }
        double epoch_obj = 0.0;
        for (size_t b0 = 0; b0 < epoch_n && rc == 0; b0 += batch) {
            const uint32_t *ids = order + b0;
            size_t nb = epoch_n - b0 < batch ? epoch_n - b0 : batch;
            if ((rc = asg_set_centroids(&A, C, k, d))) break;
            int penalized = cfg->balance > 0.0f && mean_sqdist > 0.0;
            if (penalized) {
                double mm = 0.0;
                for (size_t j = 0; j < k; j++) mm += (double)usage[j];
                double mean_mass = mm / (double)k;
                float scale = (float)((double)cfg->balance * mean_sqdist / mean_mass);
                for (size_t j = 0; j < k; j++) bias[j] = scale * usage[j];
                asg_set_bias(&A, bias);
            }
            double batch_obj = 0.0;
            for (size_t p0 = 0; p0 < nb; p0 += chunk) {
                size_t m = nb - p0 < chunk ? nb - p0 : chunk;
                const uint32_t *part = ids + p0;
                for (size_t r = 0; r < m; r++)
                    memcpy(rowbuf + r * d * rb, (const char *)x + (size_t)part[r] * d * rb, d * rb);
                if ((rc = asg_assign(&A, rowbuf, rows, m, chunk_labels, chunk_dist))) break;
                if (penalized) {
                    for (size_t r = 0; r < m; r++) {
                        float v = chunk_dist[r] - bias[chunk_labels[r]];
                        chunk_dist[r] = v > 0.0f ? v : 0.0f;
                    }
                }
                for (size_t r = 0; r < m; r++) {
                    uint32_t l = chunk_labels[r];
                    if (counts[l] == 0) touched[ntouched++] = l;
                    counts[l] += 1;
                    widen_rows(rowbuf + r * d * rb, rows, d, rowf);
                    float *s = sums + (size_t)l * d;
                    for (size_t t = 0; t < d; t++) s[t] += rowf[t];
                    if (is_last) labels[part[r]] = l;
                }
                double cs = 0.0;
                for (size_t r = 0; r < m; r++) cs += (double)chunk_dist[r];
                batch_obj += cs;
            }
            if (rc) break;
            epoch_obj += batch_obj;
            mean_sqdist = batch_obj / (double)nb;
            if (is_last) {
                for (size_t t = 0; t < ntouched; t++) {
                    size_t j = touched[t];
                    pass_counts[j] += counts[j];
                    for (size_t i = 0; i < d; i++) pass_sums[j * d + i] += sums[j * d + i];
                }
            }
            float alpha = (float)or_scheduled_alpha(R.alpha0, step, R.total_steps);
            or_state_apply(k, d, C, usage, touched, ntouched, counts, sums, alpha);
            for (size_t t = 0; t < ntouched; t++) {
                size_t j = touched[t];
                counts[j] = 0;
                memset(sums + j * d, 0, d * sizeof(float));
            }
            ntouched = 0;
            if ((step + 1) % CHECK_UNUSED_EVERY == 0 && R.unused_threshold > 0.0f)
                n_splits += or_state_revive(k, d, C, usage, R.unused_threshold, &split_rng);
            step++;
        }
        if (out->n_objective < 1024) out->objective[out->n_objective++] = epoch_obj;
    }
    if (rc == 0) {
        /* closing mean, src/lib.rs:481-489 */
        for (size_t j = 0; j < k; j++) {
            if (pass_counts[j] > 0) {
                float inv = 1.0f / (float)pass_counts[j];
                for (size_t i = 0; i < d; i++) C[j * d + i] = pass_sums[j * d + i] * inv;
            }
        }
        memcpy(centroids_out, C, k * d * sizeof(float));
        out->n_splits = n_splits;
        out->recipe = R;
    }
    free(order);
    free(bias);
    free(pass_sums);
    free(pass_counts);
    free(touched);
    free(sums);
    free(counts);
    free(chunk_dist);
    free(chunk_labels);
    free(rowf);
    free(rowbuf);
    free(A.norms);
    or_i8_free(&A.i8);
    free(usage);
    free(C);
    return rc;
}

int or_assign(const void *x, int rows, size_t n, size_t d, const float *c, size_t k, int simd, int nthreads,
              uint32_t *labels, float *dist) {
    /* assign, src/lib.rs:512-520 (chunking does not change per-row results) */
    float *norms = malloc(k * sizeof(float));
    or_cpu_norms(c, k, d, NULL, norms);
    int rc = or_cpu_assign(simd, nthreads, x, rows, n, d, c, norms, k, labels, dist);
    free(norms);
    return rc;
}

void or_balance_stats(const uint32_t *labels, size_t n, size_t k, double *imbalance, double *max_over_mean,
                      uint64_t *empty) {
    /* src/lib.rs:497-504 */
    uint64_t *counts = calloc(k, sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) counts[labels[i]]++;
    double squares = 0.0;
    uint64_t mx = 0, e = 0;
    for (size_t j = 0; j < k; j++) {
        squares += (double)(counts[j] * counts[j]);
        if (counts[j] > mx) mx = counts[j];
        if (counts[j] == 0) e++;
    }
    double nn = (double)n;
    *imbalance = (double)k * squares / (nn * nn);
    *max_over_mean = (double)mx * (double)k / nn;
    *empty = e;
    free(counts);
}


/* ---- Report's nearest-centroid step -------------------------------------- */

/* src/report.rs:104-114: sixteen running sums, lane i taking elements
 * i, i+16, ..., folded left; the remainder past the last full 16-chunk is
 * folded left on its own and added last. */
float or_report_sqdist(const float *a, const float *b, size_t d) {
    float acc[16] = {0};
    const size_t full = d / 16 * 16;
    float tail = 0.0f;
    for (size_t e = full; e < d; e++) {
        const float t = a[e] - b[e];
        tail += t * t;
    }
    for (size_t c = 0; c < full; c += 16)
        for (int i = 0; i < 16; i++) {
            const float t = a[c + i] - b[c + i];
            acc[i] += t * t;
        }
    float s = 0.0f;
    for (int i = 0; i < 16; i++) s += acc[i];
    return s + tail;
}

typedef struct {
    const float *c;
    size_t k, d, row0, row1;
    const uint32_t *sizes;
    uint32_t *nearest;
    float *dist;
} nearest_job;

/* src/report.rs:169-183: j ascending over live clusters other than i, a
 * strict < against the running best, which starts at (i, +inf) */
static void nearest_block(void *vctx, size_t blk) {
    const nearest_job *J = vctx;
    const size_t r0 = J->row0 + blk * 64, r1 = r0 + 64 < J->row1 ? r0 + 64 : J->row1;
    for (size_t i = r0; i < r1; i++) {
        uint32_t bj = (uint32_t)i;
        float bs = INFINITY;
        for (size_t j = 0; j < J->k; j++) {
            if (j == i || (J->sizes && J->sizes[j] == 0)) continue;
            const float s = or_report_sqdist(J->c + i * J->d, J->c + j * J->d, J->d);
            if (s < bs) {
                bs = s;
                bj = (uint32_t)j;
            }
        }
        J->nearest[i - J->row0] = bj;
        J->dist[i - J->row0] = sqrtf(bs);
    }
}

void or_report_nearest(int nthreads, const float *c, size_t k, size_t d, const uint32_t *sizes, size_t row0,
                       size_t row1, uint32_t *nearest, float *nearest_dist) {
    nearest_job J = {c, k, d, row0, row1, sizes, nearest, nearest_dist};
    if (row1 <= row0) return;
    run_pool(nthreads, (row1 - row0 + 63) / 64, &J, nearest_block);
}
