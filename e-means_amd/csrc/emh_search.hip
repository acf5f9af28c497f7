// Nearest-centroid search: geometry, centroid preparation, the FP32
// rescoring of candidate lists, the FP8 row encoding and the launch
// dispatch. The search kernels themselves are in emh_search_kernels.h and
// are instantiated per precision (emh_search_reg_*.hip, emh_search_lds.hip,
// emh_search_jobs.hip).
#include <float.h>
#include <string.h>
#include <math.h>

#include <algorithm>

#include "emh_search_kernels.h"

namespace emh {

SearchGeom make_geom(int op, size_t k, size_t d) {
    SearchGeom g;
    g.op = op;
    g.d = d;
    g.k = k;
    g.k_pad = (k + BN - 1) / BN * BN;
    if (op == OP_I16) {
        // two i8 planes, rows padded to 128 bytes; shift = largest s <= 7
        // with d * 2^(14 + s) < 2^31 (src/cpu_i8.rs:97). Rows of at most 512
        // bytes take the register-resident kernel (tiles of 32 centroids x 2
        // planes), wider ones the LDS kernel, d streamed in 128-byte chunks.
        g.d_pad = (d + CHUNK_BYTES - 1) / CHUNK_BYTES * CHUNK_BYTES;
        g.shift = 0;
        for (int s = 7; s >= 0; s--)
            if (((uint64_t)d << (14 + s)) < (1ull << 31)) {
                g.shift = s;
                break;
            }
        g.resident = false;
        g.reg = g.d_pad <= 512u && I16_PLANES * g.k_pad * g.d_pad <= 0x7FFFFFFFu;
        g.row_bytes = g.reg ? (int)g.d_pad : CHUNK_BYTES;
        return g;
    }
    const size_t es = op_elem_bytes(op);
    // Register-resident kernel: operand rows padded to 64 bytes, at most 512
    // bytes. The LDS kernel can run on the same layout (rows resident in
    // LDS), which the I8 search falls back to when its scores exceed the
    // packed-key range.
    // (rows past 256 bytes use 32-centroid tiles, which need 128-byte multiples)
    size_t rb64 = (d * es + 63) / 64 * 64;
    if (rb64 > 256) rb64 = (rb64 + 127) / 128 * 128;
    // (the tile-major table is addressed through a buffer resource: < 2 GiB)
    if (rb64 <= 512u && g.k_pad * rb64 <= 0x7FFFFFFFu) {
        g.reg = true;
        g.resident = true;
        g.d_pad = rb64 / es;
        g.row_bytes = (int)rb64;
        return g;
    }
    g.reg = false;
    const size_t gran = op_granule(op);
    size_t res_bytes = (d * es + gran - 1) / gran * gran;
    g.resident = (size_t)BM * (res_bytes + 16) <= (size_t)ROW_LDS_BUDGET;
    if (g.resident) {
        g.d_pad = res_bytes / es;
        g.row_bytes = (int)res_bytes;
    } else {
        g.d_pad = (d * es + CHUNK_BYTES - 1) / CHUNK_BYTES * CHUNK_BYTES / es;
        g.row_bytes = CHUNK_BYTES;
    }
    return g;
}

bool reg_path(const SearchGeom &g) { return g.reg; }

// Copies the row-major operand table into the tile-major order of the
// register-resident search: quad q of centroid c of tile t goes to quad
// (t * nq + q) * tb + c. One thread per 16-byte quad.
// With planes > 1 (I16) the source holds `planes` such tables one after the
// other and tile t's MFMA row p * tb + c is plane p of centroid t * tb + c.
__global__ void tile_layout_kernel(const uint4 *src, uint4 *dst, size_t k_pad, int nq, int tb, int planes) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)planes * k_pad * (size_t)nq) return;
    const size_t vt = (size_t)tb * planes;  // MFMA rows per tile
    const size_t per_tile = (size_t)nq * vt;
    const size_t t = i / per_tile, rem = i - t * per_tile;
    const size_t q = rem / vt, v = rem - q * vt;
    const size_t p = v / tb, c = v - p * tb;
    dst[i] = src[(p * k_pad + t * tb + c) * nq + q];
}

// -------------------------------------------------------------- rescore ----

// Exact FP32 distances |x - c|^2 + bias to the m candidates of each row;
// the lowest (score, index) wins and dist = max(score, 0). A group of 16
// lanes owns one row (four rows per wave): it reads the row's candidate ids
// once, then streams the row and all M candidate rows together (16-byte
// loads, every candidate's loads in flight at once); each lane folds its
// strided elements in order without FMA, per candidate, and a fixed
// butterfly adds the 16 partial sums, so every lane of the group holds
// identical bits.
// Candidate t of a row whose search ran as nslice centroid ranges: entry t of
// the lexicographic (key, index) top-m of the slices' lists, which is the
// unsplit sweep's list (merge_slices_kernel builds the same).
template <int M>
__device__ inline uint32_t merged_candidate(const uint32_t *part, int nslice, size_t row, int t) {
    uint32_t bs[M], bi[M];
#pragma unroll
    for (int u = 0; u < M; u++) {
        bs[u] = 0xFFFFFFFFu;
        bi[u] = NO_LABEL;
    }
    for (int s = 0; s < nslice; s++) {
        const uint32_t *pp = part + (row * nslice + s) * (2 * M);
#pragma unroll
        for (int u = 0; u < M; u++) topm_insert_lex<M, uint32_t>(bs, bi, pp[u], pp[M + u]);
    }
    uint32_t r = bi[0];
#pragma unroll
    for (int u = 1; u < M; u++) r = u == t ? bi[u] : r;
    return r;
}

// part / nslice: the row's candidates are the merge of nslice slice lists
// (nslice > 1) instead of cand's row.
template <int ROWS, int M>
__device__ __forceinline__ void rescore_row(const void *x, size_t d, size_t k, const uint32_t *ids, const float *c,
                                            const float *bias, const uint32_t *cand, uint32_t *labels, float *dist,
                                            size_t row, bool live, int gl, const uint32_t *part = nullptr,
                                            int nslice = 1) {
    constexpr int G = 16;
    const bool vec = ROWS == EMH_ROWS_F32 && (d % 4) == 0;
    uint32_t mine = NO_LABEL;
    if (live && gl < M) {
        mine = nslice > 1 ? merged_candidate<M>(part, nslice, row, gl) : cand[row * M + gl];
        if (mine >= k) mine = NO_LABEL;  // pad centroids can fill a list when k is small
    }
    uint32_t ci[M];
#pragma unroll
    for (int t = 0; t < M; t++) ci[t] = (uint32_t)__shfl((int)mine, t, G);
    const size_t src = live ? (ids ? (size_t)ids[row] : row) : 0;
    float acc[M];
#pragma unroll
    for (int t = 0; t < M; t++) acc[t] = 0.0f;
    if (live) {
        if (vec) {
            const float *xr = (const float *)x + src * d;
            for (size_t e = 4 * gl; e < d; e += 4 * G) {
                const float4 xv = *(const float4 *)(xr + e);
#pragma unroll
                for (int t = 0; t < M; t++) {
                    if (ci[t] == NO_LABEL) continue;
                    const float4 cv = *(const float4 *)(c + (size_t)ci[t] * d + e);
                    const float t0 = __fsub_rn(xv.x, cv.x), t1 = __fsub_rn(xv.y, cv.y);
                    const float t2 = __fsub_rn(xv.z, cv.z), t3 = __fsub_rn(xv.w, cv.w);
                    acc[t] = __fadd_rn(acc[t], __fmul_rn(t0, t0));
                    acc[t] = __fadd_rn(acc[t], __fmul_rn(t1, t1));
                    acc[t] = __fadd_rn(acc[t], __fmul_rn(t2, t2));
                    acc[t] = __fadd_rn(acc[t], __fmul_rn(t3, t3));
                }
            }
        } else {
            for (size_t e = gl; e < d; e += G) {
                const float xv = src_elem<ROWS>(x, src * d + e);
#pragma unroll
                for (int t = 0; t < M; t++) {
                    if (ci[t] == NO_LABEL) continue;
                    const float u = __fsub_rn(xv, c[(size_t)ci[t] * d + e]);
                    acc[t] = __fadd_rn(acc[t], __fmul_rn(u, u));
                }
            }
        }
    }
    float best = INFINITY;
    uint32_t bidx = NO_LABEL;
#pragma unroll
    for (int t = 0; t < M; t++) {
        float a = acc[t];
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) a = __fadd_rn(a, __shfl_xor(a, off, 64));
        if (ci[t] == NO_LABEL) continue;
        const float sc = bias ? __fadd_rn(a, bias[ci[t]]) : a;
        if (sc < best || (sc == best && ci[t] < bidx)) {
            best = sc;
            bidx = ci[t];
        }
    }
    if (live && gl == 0) {
        labels[row] = bidx == NO_LABEL ? 0u : bidx;
        dist[row] = (bidx != NO_LABEL && best > 0.0f) ? best : 0.0f;
    }
}

// Rows per workgroup pass: 4 waves x 4 rows.
constexpr int RESCORE_ROWS = 16;

template <int ROWS, int M>
__global__ __launch_bounds__(256) void rescore_kernel(const void *x, size_t m, size_t d, size_t k, const uint32_t *ids,
                                                      const float *c, const float *bias, const uint32_t *cand,
                                                      uint32_t *labels, float *dist) {
    const int gl = threadIdx.x & 15;
    const size_t stride = (size_t)gridDim.x * RESCORE_ROWS;
    for (size_t base = (size_t)blockIdx.x * RESCORE_ROWS; base < m; base += stride) {
        const size_t row = base + (threadIdx.x >> 4);
        rescore_row<ROWS, M>(x, d, k, ids, c, bias, cand, labels, dist, row, row < m, gl);
    }
}

// Every job's rows: slot i of the concatenated batch is row rmap[i].y of job
// rmap[i].x (live when below that job's row count this step). (Giving each
// XCD a contiguous run of slots, for L2 reuse of the candidates, measured
// slower than this grid-stride order.)
// keys (else nullptr): each slot's sort key too (slot_key, from the label
// and distance this lane has just written), and *big_count zeroed.
template <int M>
__global__ __launch_bounds__(256) void rescore_jobs_kernel(const JobDev *jobs, const uint2 *rmap, size_t slots,
                                                           uint32_t nclusters, uint32_t *keys, uint32_t *big_count) {
    const int gl = threadIdx.x & 15;
    if (keys && blockIdx.x == 0 && threadIdx.x == 0) *big_count = 0;  // read by the sums after the sort
    const size_t stride = (size_t)gridDim.x * RESCORE_ROWS;
    for (size_t base = (size_t)blockIdx.x * RESCORE_ROWS; base < slots; base += stride) {
        const size_t i = base + (threadIdx.x >> 4);
        uint2 r = make_uint2(0u, 0u);
        bool live = i < slots;
        if (live) r = rmap[i];
        // each group of 16 lanes has its own job
        const JobDev j = job_at<false>(jobs, live ? r.x : 0u);
        live = live && r.y < j.m;
        // a group's d (and so its trip counts) comes from its own job: the
        // shuffles are within the 16 lanes of the group only
        rescore_row<EMH_ROWS_F32, M>(j.x, j.d, j.k, j.ids, j.cen, j.penalized ? j.bias : nullptr, j.cand, j.labels,
                                     j.dist, r.y, live, gl, j.part, (int)j.nslice);
        if (keys && gl == 0 && i < slots) keys[i] = slot_key(j, r.y, nclusters);
    }
}

// FP32 and I8 jobs whose search ran as several centroid ranges: the best
// (score, index) of the slices' results, as merge_slices_kernel does (FP32:
// the lowest score; I8: the highest packed key), one batch slot per thread.
template <int OP>
__global__ void jobs_merge_kernel(const JobDev *jobs, const uint2 *rmap, size_t slots) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    const uint2 r = rmap[i];
    const JobDev j = job_at<false>(jobs, r.x);
    if (j.nslice < 2 || r.y >= j.m) return;
    const uint32_t *pp = j.part + (size_t)r.y * j.nslice * 4;
    uint4 best = *(const uint4 *)pp;
    for (uint32_t s = 1; s < j.nslice; s++) {
        const uint4 p = *(const uint4 *)(pp + 4 * s);
        bool better;
        if constexpr (OP == OP_I8) {
            const int32_t ps = (int32_t)p.x, bs = (int32_t)best.x;
            better = ps > bs || (ps == bs && p.y < best.y);
        } else {
            const float ps = __uint_as_float(p.x), bs = __uint_as_float(best.x);
            better = ps < bs || (ps == bs && p.y < best.y);
        }
        if (better) best = p;
    }
    j.labels[r.y] = best.y == NO_LABEL ? 0u : best.y;
    j.dist[r.y] = __uint_as_float(best.z);
}

// ---------------------------------------------------------- centroid prep --

// |c|^2 in f32 as src/cpu.rs:170-180 folds it: 16 strided partial sums in
// order, summed in order, then the tail. Call with a whole wave; every lane
// returns the value.
__device__ inline float centroid_norm32(const float *cp, size_t d, int lane) {
    const size_t body = d & ~(size_t)15;
    float acc = 0.0f;
    if (lane < 16)
        for (size_t i = lane; i < body; i += 16) acc = __fadd_rn(acc, __fmul_rn(cp[i], cp[i]));
    float s = 0.0f;
    for (int t = 0; t < 16; t++) s = __fadd_rn(s, __shfl(acc, t, 64));
    float tail = 0.0f;
    for (size_t i = body; i < d; i++) tail = __fadd_rn(tail, __fmul_rn(cp[i], cp[i]));
    return __fadd_rn(s, tail);
}

// A group of 16 lanes per centroid, 16 centroids per workgroup: the
// operand row (zero padded) a 16-byte quad per lane, and the f32 squared
// norm folded as src/cpu.rs:170-180 does (lane l: elements l, l + 16, ...;
// then the 16 partial sums in order; then the tail), or the exact i64 norm
// of the rounded i8 centroid (src/cpu_i8.rs:83-93).
constexpr int PREP_CENTROIDS = 16;

// FP8: the power-of-two encoding scale s_c of one centroid, from its largest
// finite |c| (every lane of the centroid's 16-lane group returns it; a dead
// or all-zero centroid gets 1). Each centroid has its own scale, so a small
// centroid keeps e4m3's full relative precision beside a large one.
__device__ inline float centroid_fp8_scale(const float *cp, size_t d, bool live, int gl) {
    float mx = 0.0f;
    if (live)
        for (size_t i = gl; i < d; i += 16) {
            const float a = fabsf(cp[i]);
            if (a > mx && isfinite(a)) mx = a;
        }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    return fp8_scale_for(mx);
}

// tile_op (register-resident path, else nullptr): the same quads in the
// tile-major order, quad q of centroid c of tile t at quad (t nq + q) tb + c.
// hn_out / cn_out (else nullptr): the unbiased score offsets set_bias_kernel
// would write, (|c|^2)/2 (+inf for pad centroids) or sum c8^2 (2^26).
// cexp (FP8): each centroid's MFMA scale, the E8M0 value of 1 / s_c.
template <int OP>
__global__ __launch_bounds__(256) void prep_centroids_kernel(const float *c, size_t k, size_t d, size_t d_pad,
                                                             size_t k_pad, uint8_t *cen_op, float *norm32,
                                                             int64_t *norm64, uint8_t *cexp,
                                                             unsigned *stats, uint8_t *tile_op, int tb,
                                                             float *hn_out, int64_t *cn_out) {
    constexpr int EB = OP == OP_F32 ? 4 : OP == OP_BF16 ? 2 : 1, EPQ = 16 / EB;  // element bytes, per quad
    const int gl = threadIdx.x & 15;
    const size_t jc = (size_t)blockIdx.x * PREP_CENTROIDS + (threadIdx.x >> 4);  // < k_pad (a multiple of 64)
    const bool live = jc < k;
    const float *cp = c + jc * d;
    float cs = 1.0f;
    if constexpr (OP == OP_FP8) {
        cs = centroid_fp8_scale(cp, d, live, gl);
        if (gl == 0) cexp[jc] = (uint8_t)e8m0_inverse(cs);
    }
    // operand row: the float searches store -c, so that an MFMA accumulator
    // started at |c|^2/2 ends at |c|^2/2 - x.c
    uint8_t *orow = cen_op + jc * d_pad * EB;
    int64_t part64 = 0;
    for (size_t q = gl; q < d_pad / EPQ; q += 16) {
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        if constexpr (OP == OP_FP8) {
#pragma unroll
            for (int wi = 0; wi < 4; wi++) {
                float v[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const size_t e = q * EPQ + 4 * wi + t;
                    v[t] = (live && e < d) ? -cp[e] * cs : 0.0f;
                }
                int pk = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
                pk = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], pk, true);
                w[wi] = (uint32_t)pk;
            }
        }
#pragma unroll
        for (int i = 0; i < EPQ; i++) {
            const size_t e = q * EPQ + i;
            const bool in = live && e < d;
            if constexpr (OP == OP_F32) {
                w[i] = __float_as_uint(in ? -cp[e] : 0.0f);
            } else if constexpr (OP == OP_BF16) {
                w[i / 2] |= (uint32_t)bf16_rne(in ? -cp[e] : 0.0f) << (16 * (i & 1));
            } else if constexpr (OP == OP_FP8) {
                (void)in;  // packed above
            } else {  // OP_I8: round half away from zero, clamp (src/cpu_i8.rs:83)
                int32_t qv = 0;
                if (in) {
                    float r = roundf(cp[e]);
                    r = r != r ? 0.0f : fminf(fmaxf(r, -128.0f), 127.0f);
                    qv = (int32_t)r;
                }
                w[i / 4] |= ((uint32_t)qv & 0xFFu) << (8 * (i & 3));
                part64 += (int64_t)qv * qv;
            }
        }
        const uint4 quad = make_uint4(w[0], w[1], w[2], w[3]);
        *(uint4 *)(orow + q * 16) = quad;
        if (tile_op) {
            const size_t nq = d_pad / EPQ, t = jc / (size_t)tb, cl = jc - t * (size_t)tb;
            *(uint4 *)(tile_op + ((t * nq + q) * (size_t)tb + cl) * 16) = quad;
        }
    }
    if constexpr (OP == OP_I8) {
        for (int off = 8; off > 0; off >>= 1) part64 += shfl_xor_i64(part64, off);
        if (gl == 0 && live) norm64[jc] = part64;
        if (gl == 0 && cn_out) cn_out[jc] = live ? part64 : ((int64_t)1 << 26);
    } else {
        float nv = 0.0f;
        if (live) {
            const size_t body = d & ~(size_t)15;
            float part = 0.0f;
            for (size_t i = gl; i < body; i += 16) part = __fadd_rn(part, __fmul_rn(cp[i], cp[i]));
            for (int t = 0; t < 16; t++) nv = __fadd_rn(nv, __shfl(part, t, 16));
            float tail = 0.0f;
            for (size_t i = body; i < d; i++) tail = __fadd_rn(tail, __fmul_rn(cp[i], cp[i]));
            nv = __fadd_rn(nv, tail);
            if (gl == 0) norm32[jc] = nv;
        }
        if (gl == 0 && hn_out) hn_out[jc] = live ? __fmul_rn(0.5f, nv) : INFINITY;
    }
}

// Batched prep: a group of 16 lanes per (job, centroid < k_pad) entry of
// cmap, 16 entries per workgroup. The group writes the negated operand
// straight into the tile-major table a 16-byte quad per lane, folds |c|^2 in
// the order of centroid_norm32 (lane l: elements l, l + 16, ...; then the 16
// partial sums in order; then the tail), and sets the score offset
// (|c|^2 + bias) / 2 with this step's bias, as prep_centroids_kernel +
// set_bias_kernel do for one job; FP8: each centroid's own encoding scale.
// The jobs' stats word [1] (max -bias) takes one atomic per job per
// workgroup: cmap lists each job's entries together. I8: the rounded, clamped i8 centroid (src/cpu_i8.rs:83) as the
// operand, its exact sum c8^2 plus the rounded bias (src/cpu_i8.rs:112-117)
// as the integer offset, 2^26 for pad centroids (as prep_centroids_kernel
// and set_bias_kernel write them); no stats.
constexpr int PREP_ENTRIES = 16;

template <int OP>
__global__ __launch_bounds__(256) void prep_jobs_kernel(const JobDev *jobs, const uint2 *cmap, size_t entries) {
    constexpr int EB = OP == OP_F32 ? 4 : OP == OP_BF16 ? 2 : 1, EPQ = 16 / EB;  // element bytes, per quad
    const int gl = threadIdx.x & 15, slot = threadIdx.x >> 4;
    __shared__ uint32_t s_job[PREP_ENTRIES], s_bias[PREP_ENTRIES];
    const size_t e = (size_t)blockIdx.x * PREP_ENTRIES + slot;
    uint32_t job = NO_LABEL, bbits = 0u;
    if (e < entries) {
        const uint2 cm = cmap[e];
        const JobDev jb = job_at<false>(jobs, cm.x);  // groups of a wave may belong to different jobs
        if (jb.m != 0) {  // a finished job: nothing reads its tables this step
            const size_t jc = cm.y, d = jb.d;
            const bool live = jc < jb.k;
            const float *cp = jb.cen + jc * d;
            clear_pass_row(jb, jc, gl);
            // tile-major: quad q of centroid jc lives in tile jc / tb at quad row q
            const size_t tb = jb.tb, tile = jc / tb, cl = jc - tile * tb;
            uint8_t *tbase = jb.cen_op + tile * tb * jb.row_bytes + cl * 16;
            const size_t nq = jb.row_bytes / 16;
            int64_t part64 = 0;
            // FP8: the centroid's own encoding scale s_c
            [[maybe_unused]] float cs = 1.0f;
            if constexpr (OP == OP_FP8) {
                cs = centroid_fp8_scale(cp, d, live, gl);
                if (gl == 0) jb.cexp[jc] = (uint8_t)e8m0_inverse(cs);
            }
            for (size_t q = gl; q < nq; q += 16) {
                uint32_t w[4];
#pragma unroll
                for (int i = 0; i < 4; i++) w[i] = 0u;
                if constexpr (OP == OP_FP8) {
                    // e4m3(-c s_C), four values a word, as prep_centroids_kernel
#pragma unroll
                    for (int wi = 0; wi < 4; wi++) {
                        float v[4];
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            const size_t el = q * EPQ + 4 * wi + t;
                            v[t] = (live && el < d) ? -cp[el] * cs : 0.0f;
                        }
                        int pk = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
                        pk = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], pk, true);
                        w[wi] = (uint32_t)pk;
                    }
                }
#pragma unroll
                for (int i = 0; i < (OP == OP_FP8 ? 0 : EPQ); i++) {
                    const size_t el = q * EPQ + i;
                    const bool in = live && el < d;
                    if constexpr (OP == OP_I8) {
                        int32_t qv = 0;
                        if (in) {
                            float r = roundf(cp[el]);
                            r = r != r ? 0.0f : fminf(fmaxf(r, -128.0f), 127.0f);
                            qv = (int32_t)r;
                        }
                        w[i / 4] |= ((uint32_t)qv & 0xFFu) << (8 * (i & 3));
                        part64 += (int64_t)qv * qv;
                    } else if constexpr (OP == OP_F32 || OP == OP_BF16) {
                        const float v = in ? -cp[el] : 0.0f;
                        if constexpr (OP == OP_F32) w[i] = __float_as_uint(v);
                        else w[i / 2] |= (uint32_t)bf16_rne(v) << (16 * (i & 1));
                    }
                }
                *(uint4 *)(tbase + q * tb * 16) = make_uint4(w[0], w[1], w[2], w[3]);
            }
            if constexpr (OP == OP_I8) {
                for (int off = 8; off > 0; off >>= 1) part64 += shfl_xor_i64(part64, off);
                int64_t v = (int64_t)1 << 26;
                if (live) {
                    v = part64;
                    if (jb.penalized) {
                        double r = round((double)jb.bias[jc]);
                        r = r != r ? 0.0 : fmin(fmax(r, -9.2e18), 9.2e18);
                        v += (int64_t)r;
                    }
                }
                if (gl == 0) jb.cn[jc] = v;
            } else if (!live) {
                if (gl == 0) jb.hn[jc] = INFINITY;
            } else {
                // |c|^2 in f32 as src/cpu.rs:170-180 folds it
                const size_t body = d & ~(size_t)15;
                float part = 0.0f;
                for (size_t i = gl; i < body; i += 16) part = __fadd_rn(part, __fmul_rn(cp[i], cp[i]));
                float nv = 0.0f;
                for (int t = 0; t < 16; t++) nv = __fadd_rn(nv, __shfl(part, t, 16));
                float tail = 0.0f;
                for (size_t i = body; i < d; i++) tail = __fadd_rn(tail, __fmul_rn(cp[i], cp[i]));
                nv = __fadd_rn(nv, tail);
                job = cm.x;
                float v = nv;
                if (jb.penalized) {
                    const float bv = jb.bias[jc];
                    if (bv < 0.0f) bbits = __float_as_uint(-bv);
                    v = __fadd_rn(nv, bv);
                }
                if (gl == 0) {
                    jb.norm32[jc] = nv;
                    jb.hn[jc] = __fmul_rn(0.5f, v);
                }
            }
        }
    }
    if constexpr (OP == OP_I8) return;  // whole workgroup: no stats
    if (gl == 0) {
        s_job[slot] = job;
        s_bias[slot] = bbits;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t cur = NO_LABEL, mb = 0u;
        for (int t = 0; t <= PREP_ENTRIES; t++) {
            const uint32_t jt = t < PREP_ENTRIES ? s_job[t] : NO_LABEL;
            if (jt != cur) {
                if (cur != NO_LABEL && mb) atomicMax(&jobs[cur].stats[1], mb);
                cur = jt;
                mb = 0u;
            }
            if (t < PREP_ENTRIES) mb = max(mb, s_bias[t]);
        }
    }
}

// I16 centroids (src/cpu_i8.rs:95-107): c16 = round(c * 2^shift) clamped to
// i16 (NaN -> 0, as Rust's `as`), stored as two i8 planes of k_pad x d_pad
// bytes each (the low byte less 128, then the signed top byte: c16 =
// 2^8 hi + lo + 128), and sum c16^2 in i64. One wave per centroid.
__global__ __launch_bounds__(256) void prep_i16_kernel(const float *c, size_t k, size_t d, size_t d_pad,
                                                       size_t k_pad, int shift, uint8_t *cen_op, int64_t *norm64) {
    const int lane = threadIdx.x & 63;
    const size_t jc = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (jc >= k_pad) return;
    const bool live = jc < k;
    const float scale = (float)(1u << shift);
    const size_t plane = k_pad * d_pad;
    int8_t *o = (int8_t *)cen_op + jc * d_pad;
    int64_t part = 0;
    for (size_t e = lane; e < d_pad; e += 64) {
        int32_t q = 0;
        if (live && e < d) {
            float r = roundf(c[jc * d + e] * scale);
            r = r != r ? 0.0f : fminf(fmaxf(r, -32768.0f), 32767.0f);
            q = (int32_t)r;
        }
        o[e] = (int8_t)((q & 255) - 128);
        o[plane + e] = (int8_t)(q >> 8);
        part += (int64_t)q * q;
    }
    for (int off = 32; off > 0; off >>= 1) part += shfl_xor_i64(part, off);
    if (lane == 0 && live) norm64[jc] = part;
}

// hn_j = (|c_j|^2 + bias_j) / 2 for the float searches; for I8
// cn_j = sum c8^2 + round(bias_j) (src/cpu_i8.rs:112-117, unit 1). Pad
// centroids get a score that never wins.
__global__ void set_bias_kernel(int op, size_t k, size_t k_pad, const float *norm32, const int64_t *norm64,
                                const float *bias, float *hn, int64_t *cn, unsigned *stats, int64_t cn_limit,
                                int shift) {
    size_t jc = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (jc >= k_pad) return;
    if (bias && jc < k && bias[jc] < 0.0f) atomicMax(&stats[1], __float_as_uint(-bias[jc]));
    if (op == OP_I16) {
        // offset = sum c16^2 + round(bias * 4^shift) (src/cpu_i8.rs:112-117);
        // pad centroids (c16 = 0) score INT64_MAX, which never wins the
        // register kernel's strict '<'; the LDS kernel skips indices >= k
        int64_t v = INT64_MAX;
        if (jc < k) {
            v = norm64[jc];
            if (bias) {
                double r = round(ldexp((double)bias[jc], 2 * shift));
                r = r != r ? 0.0 : fmin(fmax(r, -9.2e18), 9.2e18);
                v += (int64_t)r;
            }
            // past 2^46 the register kernel's f64 keys lose exactness: the
            // LDS kernel's i64 scores run instead
            if (v >= ((int64_t)1 << 46) || v <= -((int64_t)1 << 46)) atomicMax(&stats[2], 1u);
        }
        cn[jc] = v;
        return;
    }
    if (op == OP_I8) {
        // pad centroids: 2^26 puts their packed score 2 x.0 - 2^26 below every
        // real one (> -2^26) and still inside int32 once shifted; the i64
        // kernel skips indices >= k instead
        int64_t v = (int64_t)1 << 26;
        if (jc < k) {
            v = norm64[jc];
            if (bias) {
                double r = round((double)bias[jc]);
                r = r != r ? 0.0 : fmin(fmax(r, -9.2e18), 9.2e18);
                v += (int64_t)r;
            }
            if (v >= cn_limit || v <= -cn_limit) atomicMax(&stats[2], 1u);
        }
        cn[jc] = v;
    } else {
        float v = INFINITY;
        if (jc < k) v = __fmul_rn(0.5f, bias ? __fadd_rn(norm32[jc], bias[jc]) : norm32[jc]);
        hn[jc] = v;
    }
}

// One row's e4m3 encoding with a power-of-two scale, by a 16-lane group
// (lane gl of it): pass 1 takes the row's largest finite |v| and |x|^2, pass
// 2 (the row again, now from cache) writes the scaled, packed codes into o8
// (d_pad bytes) and lane 0 the scale and |x|^2. Every lane of the wave calls
// it (the group sums are shuffles); a group with live = false writes nothing.
template <int ROWS>
__device__ inline void encode_row_fp8(const void *x, size_t src, size_t d, size_t d_pad, bool vec, bool live, int gl,
                                      uint8_t *o8, float *scale_out, float *norm_out) {
    constexpr int G = 16;
    float mx = 0.0f, sq = 0.0f;
    auto take = [&](float v) {
        const float a = fabsf(v);
        if (a > mx && isfinite(a)) mx = a;
        sq = fmaf(v, v, sq);
    };
    if (live) {
        if (vec) {
            const float *xr = (const float *)x + src * d;
            for (size_t e = 4 * gl; e < d; e += 4 * G) {
                const float4 v = *(const float4 *)(xr + e);
                take(v.x);
                take(v.y);
                take(v.z);
                take(v.w);
            }
        } else {
            for (size_t e = gl; e < d; e += G) take(src_elem<ROWS>(x, src * d + e));
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, off, 64));
        sq += __shfl_xor(sq, off, 64);
    }
    if (!live) return;
    const float s = fp8_scale_for(mx);
    uint32_t *o = (uint32_t *)o8;
    for (size_t w = gl; w < d_pad / 4; w += G) {
        float v[4];
        if (vec && 4 * w + 3 < d) {
            const float4 q = *(const float4 *)((const float *)x + src * d + 4 * w);
            v[0] = q.x * s;
            v[1] = q.y * s;
            v[2] = q.z * s;
            v[3] = q.w * s;
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const size_t e = 4 * w + q;
                v[q] = e < d ? src_elem<ROWS>(x, src * d + e) * s : 0.0f;
            }
        }
        int pk = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
        pk = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], pk, true);
        o[w] = (uint32_t)pk;
    }
    if (gl == 0) {
        *norm_out = sq;
        *scale_out = s;
    }
}

// Per-row e4m3 encoding of a call's rows: 16 lanes per row, four rows per
// wave.
template <int ROWS>
__global__ __launch_bounds__(256) void rows_fp8_kernel(const void *x, size_t m, size_t d, size_t d_pad,
                                                       const uint32_t *ids, uint8_t *x8, float *row_scale,
                                                       float *row_norm) {
    const int gl = threadIdx.x & 15;
    const bool vec = ROWS == EMH_ROWS_F32 && (d % 4) == 0 && ((uintptr_t)x % 16) == 0;
    const size_t stride = (size_t)gridDim.x * 16;
    for (size_t base = (size_t)blockIdx.x * 16; base < m; base += stride) {
        const size_t r = base + (threadIdx.x >> 4);
        const bool live = r < m;
        const size_t src = live ? (ids ? (size_t)ids[r] : r) : 0;
        encode_row_fp8<ROWS>(x, src, d, d_pad, vec, live, gl, x8 + (live ? r : 0) * d_pad, row_scale + (live ? r : 0),
                             row_norm + (live ? r : 0));
    }
}

// FP8 jobs: every batch slot's row (rmap: job, row of its batch) encoded as
// rows_fp8_kernel encodes one problem's rows, into the job's slot; 16 lanes
// per slot. A slot past its job's rows this step (or of a finished job)
// writes nothing.
__global__ __launch_bounds__(256) void jobs_rows_fp8_kernel(const JobDev *jobs, const uint2 *rmap, size_t slots) {
    const int gl = threadIdx.x & 15;
    const size_t stride = (size_t)gridDim.x * 16;
    for (size_t base = (size_t)blockIdx.x * 16; base < slots; base += stride) {
        const size_t sl = base + (threadIdx.x >> 4);
        const uint2 r = sl < slots ? rmap[sl] : make_uint2(0u, 0u);
        const JobDev j = job_at<false>(jobs, r.x);  // groups of a wave may belong to different jobs
        const bool live = sl < slots && r.y < j.m;
        const size_t src = live ? (size_t)j.ids[r.y] : 0;
        const size_t row = live ? r.y : 0;
        const bool vec = (j.d % 4) == 0 && ((uintptr_t)j.x % 16) == 0;
        encode_row_fp8<EMH_ROWS_F32>(j.x, src, j.d, j.d_pad, vec, live, gl, j.x8 + row * j.d_pad, j.row_scale + row,
                                     j.row_norm + row);
    }
}

int reg_tiles(const SearchGeom &g) { return g.reg ? (int)(g.k_pad / reg_tile(g)) : 0; }

size_t slice_part_bytes(int M, size_t m, int nslice, int op) {
    return m * (size_t)nslice * 4 * ((op == OP_BF16 || op == OP_FP8) ? 2 * (size_t)M : 4);
}

hipError_t launch_search(const SearchGeom &g, int M, const void *x, int rows, size_t m, const uint32_t *ids,
                         const void *cen_op, const float *hn, const int64_t *cn, const uint8_t *x8,
                         const float *row_scale, const float *row_norm, const uint8_t *cexp,
                         const unsigned *stats, uint32_t *labels, float *dist, uint32_t *cand, hipStream_t s,
                         int nslice, uint32_t *part, int fr) {
    if (m == 0) return hipSuccess;
    SearchArgs a;
    memset(&a, 0, sizeof a);
    // slices of whole tiles, none empty
    const int nt = reg_tiles(g);
    a.nslice = 1;
    a.tiles_per_slice = nt;
    if (g.reg && nslice > 1 && part) {
        const int per = (nt + nslice - 1) / nslice;
        a.tiles_per_slice = per;
        a.nslice = (nt + per - 1) / per;
        a.part = part;
    }
    a.stats = stats;
    a.row_norm = row_norm;
    const size_t rb = rows == EMH_ROWS_F32 ? 4 : 1;
    a.vec = ((uintptr_t)x % 16 == 0) && ((g.d * rb) % 16 == 0);
    a.x = x;
    a.ids = ids;
    a.m = m;
    a.d = g.d;
    a.d_pad = g.d_pad;
    a.k = g.k;
    a.k_pad = g.k_pad;
    a.row_bytes = g.row_bytes;
    a.cen_op = (const uint8_t *)cen_op;
    a.hn = hn;
    a.cn = cn;
    a.x8 = x8;
    a.row_scale = row_scale;
    a.cexp = cexp;
    a.labels = labels;
    a.dist = dist;
    a.cand = cand;
    a.shift = g.shift;
    if (reg_path(g)) {
        // the register-resident kernels read the tile-major copy
        SearchArgs r = a;
        r.cen_op = (const uint8_t *)reg_operand(g, cen_op);
        r.enc = fp8_encodes_in_search(g, rows) ? 1 : 0;
        switch (g.op) {
            case OP_F32:
                if (M != 1) return hipErrorInvalidValue;
                return launch_reg_f32(g, rows, fr, r, s);
            case OP_BF16:
                return launch_reg_bf16(g, rows, M, fr, r, s);
            case OP_FP8:
                return launch_reg_fp8(g, M, fr, r, s);
            case OP_I8:
                if (M != 1 || rows != EMH_ROWS_I8) return hipErrorInvalidValue;
                // scores past the packed key's range take the LDS kernel's i64 path
                if (g.i8_packed) return launch_reg_i8(g, fr, r, s);
                break;
            case OP_I16:
                if (M != 1 || rows != EMH_ROWS_I8) return hipErrorInvalidValue;
                // offsets past the f64 keys' exact range take the LDS kernel
                if (g.i16_f64) return launch_reg_i16(g, r, s);
                break;
        }
    }
    // LDS kernel (rows resident in LDS, or streamed in d-chunks)
    return launch_lds(g, rows, M, a, s);
}

template <int ROWS>
static hipError_t rescore_rows(int M, const void *x, size_t m, size_t d, size_t k, const uint32_t *ids, const float *c,
                               const float *bias, const uint32_t *cand, uint32_t *labels, float *dist,
                               hipStream_t s) {
    dim3 grid((unsigned)std::min<size_t>((m + RESCORE_ROWS - 1) / RESCORE_ROWS, 16384));
    switch (M) {
        case 1: hipLaunchKernelGGL((rescore_kernel<ROWS, 1>), grid, dim3(256), 0, s, x, m, d, k, ids, c, bias, cand, labels, dist); break;
        case 2: hipLaunchKernelGGL((rescore_kernel<ROWS, 2>), grid, dim3(256), 0, s, x, m, d, k, ids, c, bias, cand, labels, dist); break;
        case 4: hipLaunchKernelGGL((rescore_kernel<ROWS, 4>), grid, dim3(256), 0, s, x, m, d, k, ids, c, bias, cand, labels, dist); break;
        case 8: hipLaunchKernelGGL((rescore_kernel<ROWS, 8>), grid, dim3(256), 0, s, x, m, d, k, ids, c, bias, cand, labels, dist); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rescore(int M, const void *x, int rows, size_t m, size_t d, const uint32_t *ids, const float *c,
                          const float *bias, size_t k, const uint32_t *cand, uint32_t *labels, float *dist,
                          hipStream_t s) {
    if (m == 0) return hipSuccess;
    if (rows == EMH_ROWS_F32) return rescore_rows<EMH_ROWS_F32>(M, x, m, d, k, ids, c, bias, cand, labels, dist, s);
    if (rows == EMH_ROWS_I8) return rescore_rows<EMH_ROWS_I8>(M, x, m, d, k, ids, c, bias, cand, labels, dist, s);
    return rescore_rows<EMH_ROWS_U8>(M, x, m, d, k, ids, c, bias, cand, labels, dist, s);
}

hipError_t launch_prep_centroids(const SearchGeom &g, const float *c, void *cen_op, float *norm32, int64_t *norm64,
                                 uint8_t *cexp, unsigned *stats, hipStream_t s, float *hn_out,
                                 int64_t *cn_out) {
    const dim3 grid((unsigned)((g.k_pad + PREP_CENTROIDS - 1) / PREP_CENTROIDS)), grid_wave((unsigned)((g.k_pad + 3) / 4));
    uint8_t *op = (uint8_t *)cen_op;
    // the register-resident path's tile-major copy, written by the same
    // kernel (I16's two planes go through tile_layout_kernel)
    uint8_t *tile = g.reg && g.op != OP_I16 ? (uint8_t *)reg_operand(g, cen_op) : nullptr;
    const int tbc = g.reg ? reg_tile(g) : 1;
    switch (g.op) {
        case OP_F32: hipLaunchKernelGGL(prep_centroids_kernel<OP_F32>, grid, dim3(256), 0, s, c, g.k, g.d, g.d_pad, g.k_pad, op, norm32, norm64, cexp, stats, tile, tbc, hn_out, cn_out); break;
        case OP_BF16: hipLaunchKernelGGL(prep_centroids_kernel<OP_BF16>, grid, dim3(256), 0, s, c, g.k, g.d, g.d_pad, g.k_pad, op, norm32, norm64, cexp, stats, tile, tbc, hn_out, cn_out); break;
        case OP_FP8: hipLaunchKernelGGL(prep_centroids_kernel<OP_FP8>, grid, dim3(256), 0, s, c, g.k, g.d, g.d_pad, g.k_pad, op, norm32, norm64, cexp, stats, tile, tbc, hn_out, cn_out); break;
        case OP_I8: hipLaunchKernelGGL(prep_centroids_kernel<OP_I8>, grid, dim3(256), 0, s, c, g.k, g.d, g.d_pad, g.k_pad, op, norm32, norm64, cexp, stats, tile, tbc, hn_out, cn_out); break;
        case OP_I16: hipLaunchKernelGGL(prep_i16_kernel, grid_wave, dim3(256), 0, s, c, g.k, g.d, g.d_pad, g.k_pad, g.shift, op, norm64); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !g.reg || g.op != OP_I16) return e;
    const int nq = g.row_bytes / 16, tb = reg_tile(g), planes = g.op == OP_I16 ? I16_PLANES : 1;
    const size_t quads = (size_t)planes * g.k_pad * (size_t)nq;
    hipLaunchKernelGGL(tile_layout_kernel, dim3((unsigned)((quads + 255) / 256)), dim3(256), 0, s,
                       (const uint4 *)cen_op, (uint4 *)reg_operand(g, cen_op), g.k_pad, nq, tb, planes);
    return hipGetLastError();
}

// ---- batched launchers ----

size_t search_args_bytes() { return sizeof(SearchArgs); }

void fill_search_args(void *dst, const SearchGeom &g, const JobDev &j) {
    SearchArgs a;
    memset(&a, 0, sizeof a);
    a.x = j.x;
    a.ids = j.ids;
    a.m = j.m;
    a.d = g.d;
    a.d_pad = g.d_pad;
    a.k = g.k;
    a.k_pad = g.k_pad;
    a.row_bytes = g.row_bytes;
    a.cen_op = j.cen_op;
    a.hn = j.hn;
    a.cn = j.cn;
    a.stats = j.stats;
    // whole 16-byte quads per row from a 16-byte aligned table
    a.vec = (g.d * (g.op == OP_I8 ? 1 : 4)) % 16 == 0 && ((uintptr_t)j.x & 15) == 0;
    a.resident = g.resident ? 1 : 0;
    a.shift = g.shift;
    a.labels = j.labels;
    a.dist = j.dist;
    a.cand = j.cand;
    a.nslice = (int)j.nslice;
    a.tiles_per_slice = (int)j.tiles_per_slice;
    a.part = j.part;
    if (g.op == OP_FP8) {
        a.x8 = j.x8;
        a.row_scale = j.row_scale;
        a.row_norm = j.row_norm;
        a.cexp = j.cexp;
    }
    memcpy(dst, &a, sizeof a);
}

// A workgroup's sweep costs about (tiles x tile centroids x d_pad) operand
// elements against its 256 rows, plus loading those rows once. Slices of
// about 2^17 centroid-dimensions (and at least 256 centroids, so that the row
// load stays a small share) keep the longest workgroup of a batch short while
// every slice re-reads its rows and the merge reads a few words per slice
// (2^16 and 2^18 measured slower on uneven batches).
JobSlices job_slices(const SearchGeom &g) {
    const size_t tc = (size_t)reg_tile(g), ntiles = g.k_pad / tc;
    size_t per = std::max<size_t>((size_t)1 << 17, 256 * g.d_pad) / (tc * g.d_pad);
    per = std::max<size_t>(per, 1);
    size_t ns = std::min<size_t>((ntiles + per - 1) / per, 255);
    per = (ntiles + ns - 1) / ns;  // balanced, none empty
    ns = (ntiles + per - 1) / per;
    return JobSlices{(uint32_t)ns, (uint32_t)per};
}

size_t job_part_words(int op, int M, uint32_t nslice) {
    if (nslice < 2) return 0;
    return (size_t)nslice * ((op == OP_BF16 || op == OP_FP8) ? 2 * (size_t)M : 4);
}

hipError_t launch_jobs_merge(int op, const JobDev *jobs, const uint2 *rmap, size_t slots, hipStream_t s) {
    if (!slots) return hipSuccess;
    const dim3 grid((unsigned)((slots + 255) / 256));
    if (op == OP_F32) hipLaunchKernelGGL(jobs_merge_kernel<OP_F32>, grid, dim3(256), 0, s, jobs, rmap, slots);
    else if (op == OP_I8) hipLaunchKernelGGL(jobs_merge_kernel<OP_I8>, grid, dim3(256), 0, s, jobs, rmap, slots);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// I16 jobs (the LDS kernel's row-major planes, as prep_i16_kernel writes a
// single problem's): c16 = round(c 2^shift) clamped to i16 (NaN -> 0), the
// low byte less 128 in plane 0 and the signed top byte in plane 1; the
// offset sum c16^2 + round(bias 4^shift) (src/cpu_i8.rs:95-117), INT64_MAX
// for pad centroids. A 16-lane group per centroid, 4 bytes a lane.
__global__ __launch_bounds__(256) void prep_jobs_i16_kernel(const JobDev *jobs, const uint2 *cmap, size_t entries) {
    const int gl = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const size_t e = (size_t)blockIdx.x * PREP_ENTRIES + slot;
    if (e >= entries) return;
    const uint2 cm = cmap[e];
    const JobDev jb = job_at<false>(jobs, cm.x);
    if (jb.m == 0) return;  // a finished job: nothing reads its tables this step
    const size_t jc = cm.y, d = jb.d, d_pad = jb.d_pad;
    const bool live = jc < jb.k;
    const float *cp = jb.cen + jc * d;
    clear_pass_row(jb, jc, gl);
    const float scale = (float)(1u << jb.shift);
    uint8_t *o = jb.cen_op + jc * d_pad;
    const size_t plane = (size_t)jb.k_pad * d_pad;
    int64_t part = 0;
    for (size_t e4 = 4 * (size_t)gl; e4 < d_pad; e4 += 64) {
        uint32_t lo = 0u, hi = 0u;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            int32_t q = 0;
            if (live && e4 + t < d) {
                float r = roundf(cp[e4 + t] * scale);
                r = r != r ? 0.0f : fminf(fmaxf(r, -32768.0f), 32767.0f);
                q = (int32_t)r;
            }
            lo |= ((uint32_t)((q & 255) - 128) & 0xFFu) << (8 * t);
            hi |= ((uint32_t)(q >> 8) & 0xFFu) << (8 * t);
            part += (int64_t)q * q;
        }
        *(uint32_t *)(o + e4) = lo;
        *(uint32_t *)(o + plane + e4) = hi;
    }
    for (int off = 8; off > 0; off >>= 1) part += shfl_xor_i64(part, off);
    if (gl != 0) return;
    int64_t v = INT64_MAX;
    if (live) {
        v = part;
        if (jb.penalized) {
            double r = round(ldexp((double)jb.bias[jc], 2 * (int)jb.shift));
            r = r != r ? 0.0 : fmin(fmax(r, -9.2e18), 9.2e18);
            v += (int64_t)r;
        }
    }
    jb.cn[jc] = v;
}

hipError_t launch_jobs_prep(int op, const JobDev *jobs, const uint2 *cmap, size_t entries, hipStream_t s) {
    if (!entries) return hipSuccess;
    dim3 grid((unsigned)((entries + PREP_ENTRIES - 1) / PREP_ENTRIES));
    if (op == OP_F32) hipLaunchKernelGGL(prep_jobs_kernel<OP_F32>, grid, dim3(256), 0, s, jobs, cmap, entries);
    else if (op == OP_BF16) hipLaunchKernelGGL(prep_jobs_kernel<OP_BF16>, grid, dim3(256), 0, s, jobs, cmap, entries);
    else if (op == OP_FP8) hipLaunchKernelGGL(prep_jobs_kernel<OP_FP8>, grid, dim3(256), 0, s, jobs, cmap, entries);
    else if (op == OP_I8) hipLaunchKernelGGL(prep_jobs_kernel<OP_I8>, grid, dim3(256), 0, s, jobs, cmap, entries);
    else if (op == OP_I16) hipLaunchKernelGGL(prep_jobs_i16_kernel, grid, dim3(256), 0, s, jobs, cmap, entries);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_jobs_rows_fp8(const JobDev *jobs, const uint2 *rmap, size_t slots, hipStream_t s) {
    if (!slots) return hipSuccess;
    const dim3 grid((unsigned)std::min<size_t>((slots + 15) / 16, 16384));
    hipLaunchKernelGGL(jobs_rows_fp8_kernel, grid, dim3(256), 0, s, jobs, rmap, slots);
    return hipGetLastError();
}

hipError_t launch_jobs_search(int op, int M, const void *search_args, const uint2 *wg, size_t nwg, hipStream_t s) {
    if (!nwg) return hipSuccess;
    return launch_jobs_search_op(op, M, search_args, wg, nwg, s);
}

hipError_t launch_jobs_search_lds(int op, int M, const void *search_args, const uint2 *wg, size_t nwg,
                                  size_t lds_bytes, hipStream_t s) {
    if (!nwg) return hipSuccess;
    return launch_jobs_search_lds_op(op, M, search_args, wg, nwg, lds_bytes, s);
}

size_t search_lds_bytes(const SearchGeom &g) { return search_lds(g); }

hipError_t launch_jobs_rescore(int M, const JobDev *jobs, const uint2 *rmap, size_t slots, hipStream_t s,
                               uint32_t nclusters, uint32_t *keys, uint32_t *big_count) {
    if (!slots) return hipSuccess;
    dim3 grid((unsigned)std::min<size_t>((slots + RESCORE_ROWS - 1) / RESCORE_ROWS, 16384));
    switch (M) {
        case 1: hipLaunchKernelGGL(rescore_jobs_kernel<1>, grid, dim3(256), 0, s, jobs, rmap, slots, nclusters, keys, big_count); break;
        case 2: hipLaunchKernelGGL(rescore_jobs_kernel<2>, grid, dim3(256), 0, s, jobs, rmap, slots, nclusters, keys, big_count); break;
        case 4: hipLaunchKernelGGL(rescore_jobs_kernel<4>, grid, dim3(256), 0, s, jobs, rmap, slots, nclusters, keys, big_count); break;
        case 8: hipLaunchKernelGGL(rescore_jobs_kernel<8>, grid, dim3(256), 0, s, jobs, rmap, slots, nclusters, keys, big_count); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_set_bias(const SearchGeom &g, const float *norm32, const int64_t *norm64, const float *bias,
                           float *hn, int64_t *cn, unsigned *stats, hipStream_t s) {
    dim3 grid((unsigned)((g.k_pad + 255) / 256));
    // packed I8 key: |2 x.c - cn| < 2^26 needs |cn| < 2^26 - 2 * d * 2^14
    const int64_t cn_limit = ((int64_t)1 << 26) - (int64_t)g.d * 32768;
    hipLaunchKernelGGL(set_bias_kernel, grid, dim3(256), 0, s, g.op, g.k, g.k_pad, norm32, norm64, bias, hn, cn, stats,
                       cn_limit, g.shift);
    return hipGetLastError();
}

hipError_t launch_rows_fp8(const void *x, int rows, size_t m, size_t d, size_t d_pad, const uint32_t *ids,
                           uint8_t *x8, float *row_scale, float *row_norm, hipStream_t s) {
    if (m == 0) return hipSuccess;
    dim3 grid((unsigned)std::min<size_t>((m + 15) / 16, 16384));
    if (rows == EMH_ROWS_F32) hipLaunchKernelGGL(rows_fp8_kernel<EMH_ROWS_F32>, grid, dim3(256), 0, s, x, m, d, d_pad, ids, x8, row_scale, row_norm);
    else if (rows == EMH_ROWS_I8) hipLaunchKernelGGL(rows_fp8_kernel<EMH_ROWS_I8>, grid, dim3(256), 0, s, x, m, d, d_pad, ids, x8, row_scale, row_norm);
    else hipLaunchKernelGGL(rows_fp8_kernel<EMH_ROWS_U8>, grid, dim3(256), 0, s, x, m, d, d_pad, ids, x8, row_scale, row_norm);
    return hipGetLastError();
}

}  // namespace emh
