#pragma once
// Nearest-centroid search kernels for gfx950 (MI355X): the kernel
// templates and their launch templates, shared by the translation units
// that instantiate them (one per precision of the register-resident search,
// one for the LDS kernel, one for the batched search) and emh_search.hip.
//
// Replaces the reference's per-chunk cuBLAS sgemm + argmin kernel pair
// (src/cuda.rs:20-48, 174-211) and its CPU searches (src/cpu.rs:108-167,
// src/cpu_i8.rs:130-191) with one fused pass: a workgroup keeps a block of
// BM rows resident in LDS, streams centroid tiles through LDS, computes the
// BM x BN inner products on 32x32 MFMA tiles and folds them straight into a
// per-row running argmin (or top-m), so no n x k matrix is ever written.
//
// MFMA layout used throughout (32x32xK): lane l feeds A row i = l % 32 and
// B column j = l % 32 with the K elements of lane half h = l / 32; the output
// register r of lane l holds D[8*(r/4) + 4*h + r%4][l % 32]. Centroids are
// the A operand and rows the B operand, so every lane owns one row and sees
// 16 centroids of each 32x32 tile in ascending index order, which gives the
// reference's lowest-index tie rule with a strict '<' (src/cpu.rs:149).
#include <float.h>
#include <string.h>
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "emh_internal.h"

namespace emh {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

constexpr uint32_t NO_LABEL = 0xFFFFFFFFu;
constexpr int64_t PAD_SCORE64 = INT64_MAX / 4;

template <int OP> struct OpT;
template <> struct OpT<OP_F32> { static constexpr int ES = 4, LB = 4; };
template <> struct OpT<OP_BF16> { static constexpr int ES = 2, LB = 16; };
template <> struct OpT<OP_FP8> { static constexpr int ES = 1, LB = 32; };
template <> struct OpT<OP_I8> { static constexpr int ES = 1, LB = 16; };
template <> struct OpT<OP_I16> { static constexpr int ES = 1, LB = 16; };


// ---------------------------------------------------------------- helpers --

__device__ inline uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <int ROWS> __device__ inline float src_elem(const void *x, size_t idx) {
    if constexpr (ROWS == EMH_ROWS_F32) return ((const float *)x)[idx];
    else if constexpr (ROWS == EMH_ROWS_I8) return (float)((const int8_t *)x)[idx];
    else return (float)((const uint8_t *)x)[idx];
}

// OCP e4m3 power-of-two scale 2^p mapping |v| <= amax into [0, 448]: the
// largest p with amax 2^p <= 448 = 0.875 2^9, from amax's exact exponent,
// clamped so that 2^p is a normal float and 2^-p an E8M0 value (tiny
// inputs then keep a finite scale).
__device__ inline float fp8_scale_for(float amax) {
    if (!(amax > 0.0f) || !isfinite(amax)) return 1.0f;
    int e;
    const float m = frexpf(amax, &e);  // amax = m 2^e, m in [0.5, 1)
    int p = (m <= 0.875f ? 9 : 8) - e;
    p = p < -126 ? -126 : (p > 127 ? 127 : p);
    return ldexpf(1.0f, p);
}

// The E8M0 MFMA scale value of 1 / s for a power-of-two s (a normal float):
// 127 - log2(s), as the biased exponent 254 - exponent bits of s.
__device__ inline int e8m0_inverse(float s) { return 254 - (int)((__float_as_uint(s) >> 23) & 0xff); }

struct SearchArgs {
    const void *x;
    const uint32_t *ids;
    size_t m, d, d_pad, k, k_pad;
    int row_bytes;
    const uint8_t *cen_op;
    const float *hn;
    const int64_t *cn;
    const uint8_t *x8;
    const float *row_scale;
    const uint8_t *cexp;  // FP8: k_pad E8M0 MFMA scales, 1 / s_c of each centroid's encoding scale
    // [1]: max(0, -min_j bias_j) as f32 bits
    const unsigned *stats;
    const float *row_norm;  // FP8: |x|^2 of each call row, from the encoding pass
    int vec;                // rows are 16-byte aligned and d fills whole 16-byte quads
    int resident;           // LDS kernel: whole rows resident in LDS (else d-chunks)
    int shift;              // I16: score units are 4^-shift squared distance
    int enc;                // FP8 register-resident kernel: encode the f32 rows x in the prologue (x8 unused)
    // Centroid-range split (register-resident search): workgroup (b, s)
    // sweeps tiles [s * tiles_per_slice, +tiles_per_slice) and, with more than
    // one slice, writes its per-row result to part[row * nslice + s] for
    // merge_slices_kernel.
    int nslice, tiles_per_slice;
    uint32_t *part;
    uint32_t *labels;
    float *dist;
    uint32_t *cand;
};

// One 16-byte LDS segment of a row, converted to the operand type.
// e0: first element of the segment; valid: the row exists.
template <int OP, int ROWS>
__device__ inline uint4 row_segment(const SearchArgs &a, size_t call_row, size_t src_row, size_t e0, bool valid) {
    uint4 out = make_uint4(0, 0, 0, 0);
    if (!valid) return out;
    if constexpr (OP == OP_FP8) {
        // pre-scaled e4m3 rows, already padded to d_pad
        return *(const uint4 *)(a.x8 + call_row * a.d_pad + e0);
    } else if constexpr (OP == OP_I8 || OP == OP_I16) {
        const int8_t *p = (const int8_t *)a.x + src_row * a.d;
        if (a.vec && e0 + 16 <= a.d) return *(const uint4 *)(p + e0);
        uint32_t w[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t v = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                size_t e = e0 + 4 * q + b;
                uint32_t byte = e < a.d ? (uint8_t)p[e] : 0u;
                v |= byte << (8 * b);
            }
            w[q] = v;
        }
        return make_uint4(w[0], w[1], w[2], w[3]);
    } else if constexpr (OP == OP_F32) {
        float v[4];
        const size_t base = src_row * a.d;
        if constexpr (ROWS == EMH_ROWS_F32) {
            if (a.vec && e0 + 4 <= a.d) return *(const uint4 *)((const float *)a.x + base + e0);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = (e0 + q < a.d) ? src_elem<ROWS>(a.x, base + e0 + q) : 0.0f;
        return make_uint4(__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]), __float_as_uint(v[3]));
    } else {  // OP_BF16
        uint32_t w[4];
        const size_t base = src_row * a.d;
        if constexpr (ROWS == EMH_ROWS_F32) {
            if (a.vec && e0 + 8 <= a.d) {
                const float4 lo = *(const float4 *)((const float *)a.x + base + e0);
                const float4 hi = *(const float4 *)((const float *)a.x + base + e0 + 4);
                return make_uint4((uint32_t)bf16_rne(lo.x) | ((uint32_t)bf16_rne(lo.y) << 16),
                                  (uint32_t)bf16_rne(lo.z) | ((uint32_t)bf16_rne(lo.w) << 16),
                                  (uint32_t)bf16_rne(hi.x) | ((uint32_t)bf16_rne(hi.y) << 16),
                                  (uint32_t)bf16_rne(hi.z) | ((uint32_t)bf16_rne(hi.w) << 16));
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            size_t e = e0 + 2 * q;
            float lo = e < a.d ? src_elem<ROWS>(a.x, base + e) : 0.0f;
            float hi = e + 1 < a.d ? src_elem<ROWS>(a.x, base + e + 1) : 0.0f;
            w[q] = (uint32_t)bf16_rne(lo) | ((uint32_t)bf16_rne(hi) << 16);
        }
        return make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Insert (v, vi) into the sorted list (bs, bi); strict '<' keeps the earlier
// (lower-index) entry first among equal scores.
template <int M, typename S>
__device__ inline void topm_insert(S (&bs)[M], uint32_t (&bi)[M], S v, uint32_t vi) {
#pragma unroll
    for (int t = 0; t < M; t++) {
        bool lt = v < bs[t];
        S ts = bs[t];
        uint32_t ti = bi[t];
        bs[t] = lt ? v : ts;
        bi[t] = lt ? vi : ti;
        v = lt ? ts : v;
        vi = lt ? ti : vi;
    }
}

// Merge with (v, vi) from another lane: order by (score, index).
template <int M, typename S>
__device__ inline void topm_insert_lex(S (&bs)[M], uint32_t (&bi)[M], S v, uint32_t vi) {
#pragma unroll
    for (int t = 0; t < M; t++) {
        bool lt = v < bs[t] || (v == bs[t] && vi < bi[t]);
        S ts = bs[t];
        uint32_t ti = bi[t];
        bs[t] = lt ? v : ts;
        bi[t] = lt ? vi : ti;
        v = lt ? ts : v;
        vi = lt ? ti : vi;
    }
}

// ---- candidate lists of the BF16 / FP8 searches ----
//
// A candidate key is a score's f32 bits (the scores are kept positive, so
// their bits order like the scores) with the low 5 bits replaced by the
// value's position in the lane's tile: the keys of one tile are distinct and
// a key names its centroid. Each lane keeps the M smallest keys it has seen,
// exactly: per tile it takes the two smallest keys of its values with a
// min3 / med3 tree, inserts both, and only when the second one enters the
// list (several of the lane's best in one tile, which grows rare as the
// list fills) extracts the tile's next keys one at a time.

__device__ __forceinline__ uint32_t cand_key(float v, uint32_t pos) {
    return (__float_as_uint(v) & 0xFFFFFFE0u) | pos;
}
__device__ __forceinline__ uint32_t umin3(uint32_t a, uint32_t b, uint32_t c) { return min(min(a, b), c); }
// the median of three (one v_med3_u32)
__device__ __forceinline__ uint32_t umed3(uint32_t a, uint32_t b, uint32_t c) {
    return max(min(a, b), min(max(a, b), c));
}
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// The fold's integer ops as single instructions. Left to itself the
// compiler spells med3 as max/min chains wherever it shares their halves
// with a min3, and splits a key into an and + or pair (a VOP3 on gfx9 takes
// no 32-bit literal); the candidate fold is bound by these vector ops. min3
// and med3 are written out: they read only vector-ALU results. A key reads
// an MFMA result, whose wait states the compiler inserts only for code it
// sees, so the key stays C++ and gets its single v_and_or_b32 from a mask
// the compiler cannot fold: an SGPR written by an opaque s_mov.
__device__ __forceinline__ uint32_t vmin3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t d;
    asm("v_min3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}
__device__ __forceinline__ uint32_t vmed3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t d;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}
// 0xFFFFFFE0 (the key mask of cand_key) in an SGPR
__device__ __forceinline__ uint32_t cand_mask() {
    uint32_t m;
    asm volatile("s_mov_b32 %0, 0xffffffe0" : "=s"(m));
    return m;
}
// cand_key(v, POS) given mask = cand_mask()
template <int POS>
__device__ __forceinline__ uint32_t vkey(float v, uint32_t mask) {
    static_assert(POS >= 0 && POS < 32, "position");
    return (__float_as_uint(v) & mask) | (uint32_t)POS;
}

// The two smallest keys of one 32x32 block's 16 values of a lane (positions
// POS0 .. POS0 + 15), built in 8 phases so a caller can spread them between
// MFMAs: phases 0-4 take triples (3p, 3p+1, 3p+2) to (min, median); phase 5
// merges triples 3 and 4 with value 15; phase 6 merges triples 0-2; phase 7
// merges the two halves. (Of pairs lo_i <= hi_i, the second smallest is
// min(second smallest lo, hi of the smallest lo), and every other hi is at
// least the second smallest lo: hence min(med3(lo), min3(hi)).) 16 key ops
// and 20 min/med ops per block.
struct Top2 {
    uint32_t pl[5], ph[5], ql[2], qh[2], lo, hi;
};
template <bool PIN, int POS0, int P>
__device__ __forceinline__ void top2_step(Top2 &s, const f32x16 &a, uint32_t mask) {
    if constexpr (P < 5) {
        const uint32_t k0 = vkey<POS0 + 3 * P>(a[3 * P], mask), k1 = vkey<POS0 + 3 * P + 1>(a[3 * P + 1], mask);
        const uint32_t k2 = vkey<POS0 + 3 * P + 2>(a[3 * P + 2], mask);
        s.pl[P] = vmin3(k0, k1, k2);
        s.ph[P] = vmed3(k0, k1, k2);
        if constexpr (PIN) asm volatile("" : "+v"(s.pl[P]), "+v"(s.ph[P]));
    } else if constexpr (P == 5) {
        const uint32_t k = vkey<POS0 + 15>(a[15], mask);
        s.ql[1] = vmin3(s.pl[3], s.pl[4], k);
        s.qh[1] = vmin3(vmed3(s.pl[3], s.pl[4], k), s.ph[3], s.ph[4]);
        if constexpr (PIN) asm volatile("" : "+v"(s.ql[1]), "+v"(s.qh[1]));
    } else if constexpr (P == 6) {
        s.ql[0] = vmin3(s.pl[0], s.pl[1], s.pl[2]);
        s.qh[0] = min(vmin3(vmed3(s.pl[0], s.pl[1], s.pl[2]), s.ph[0], s.ph[1]), s.ph[2]);
        if constexpr (PIN) asm volatile("" : "+v"(s.ql[0]), "+v"(s.qh[0]));
    } else {
        s.lo = min(s.ql[0], s.ql[1]);
        s.hi = vmin3(max(s.ql[0], s.ql[1]), s.qh[0], s.qh[1]);
        if constexpr (PIN) asm volatile("" : "+v"(s.lo), "+v"(s.hi));
    }
}
// phase p (a constant after unrolling) of the tree above
template <bool PIN, int POS0>
__device__ __forceinline__ void top2_phase(Top2 &s, const f32x16 &a, uint32_t mask, int p) {
    switch (p) {
        case 0: top2_step<PIN, POS0, 0>(s, a, mask); break;
        case 1: top2_step<PIN, POS0, 1>(s, a, mask); break;
        case 2: top2_step<PIN, POS0, 2>(s, a, mask); break;
        case 3: top2_step<PIN, POS0, 3>(s, a, mask); break;
        case 4: top2_step<PIN, POS0, 4>(s, a, mask); break;
        case 5: top2_step<PIN, POS0, 5>(s, a, mask); break;
        case 6: top2_step<PIN, POS0, 6>(s, a, mask); break;
        default: top2_step<PIN, POS0, 7>(s, a, mask); break;
    }
}

// Sorted insertion of key v (centroid vi); strict '<', so an equal key
// already in the list (an earlier tile: a lower index) stays first. A key
// not below the last entry changes nothing.
template <int M>
__device__ __forceinline__ void cand_insert(uint32_t (&bs)[M], uint32_t (&bi)[M], uint32_t v, uint32_t vi) {
    bool c[M];
#pragma unroll
    for (int t = 0; t < M; t++) c[t] = v < bs[t];
#pragma unroll
    for (int t = M - 1; t > 0; t--) {
        bi[t] = c[t - 1] ? bi[t - 1] : (c[t] ? vi : bi[t]);
        bs[t] = vmed3(bs[t - 1], bs[t], v);
    }
    bi[0] = c[0] ? vi : bi[0];
    bs[0] = min(bs[0], v);
}

// The same register of the other lane half (lane l ^ 32), by one
// permlane32 swap: after it r[0] holds lanes 0-31 twice and r[1] lanes
// 32-63 twice.
__device__ __forceinline__ uint32_t other_half(uint32_t v, int h) {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return h ? r[0] : r[1];
}

// Adds one tile of a lane (NB blocks of 16 values in a, the tile's first
// centroid c0) to its list, given the tile's two smallest keys m1 < m2.
// A key enters only below the lane's own last entry and at most equal to
// the other lane half's last entry (the row's list is the merge of the two
// halves' lists, so a key above M keys of either half never reaches it;
// equal keys are kept, the merge orders them by index). Exact: when m2
// enters, the tile's next keys are extracted in order (the smallest key
// above the last one) until one stays out; at most M keys of a tile enter.
template <int M, int NB>
__device__ __forceinline__ void cand_settle(const f32x16 (&a)[NB], uint32_t m1, uint32_t m2, int h, uint32_t c0,
                                            uint32_t (&bs)[M], uint32_t (&bi)[M]) {
    auto gidx = [&](uint32_t key) -> uint32_t {
        const uint32_t li = key & 31u, nb = li >> 4, r = li & 15u;
        return c0 + nb * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
    };
    const uint32_t po = other_half(bs[M - 1], h);
    const uint32_t pl = po == 0xFFFFFFFFu ? po : po + 1u;  // keys below pl may still reach the row's list
    if (!wave_any(m1 < min(bs[M - 1], pl))) return;
    cand_insert<M>(bs, bi, m1 < pl ? m1 : 0xFFFFFFFFu, gidx(m1));
    if constexpr (M >= 2) {
        bool more = m2 < min(bs[M - 1], pl);
        if (!wave_any(more)) return;
        cand_insert<M>(bs, bi, more ? m2 : 0xFFFFFFFFu, gidx(m2));
        uint32_t last = m2;
#pragma unroll 1
        for (int it = 2; it < M && it < 16 * NB; it++) {
            const uint32_t above = last + 1u;
            uint32_t u = 0xFFFFFFFFu;
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
#pragma unroll
                for (int r = 0; r < 16; r++) u = min(u, cand_key(a[nb][r], (uint32_t)(nb * 16 + r)) - above);
            const uint32_t nx = u + above;  // the smallest key above `last`
            more = more && nx > last && nx < min(bs[M - 1], pl);
            if (!wave_any(more)) return;
            cand_insert<M>(bs, bi, more ? nx : 0xFFFFFFFFu, gidx(nx));
            last = nx;
        }
    }
}

__device__ inline int64_t shfl_xor_i64(int64_t v, int mask) {
    int lo = __shfl_xor((int)(uint32_t)v, mask, 64);
    int hi = __shfl_xor((int)(uint32_t)((uint64_t)v >> 32), mask, 64);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// --------------------------------------------------------------- search ----

// Rows [row0, row0 + BM) of a against every centroid, on the dynamic LDS
// the launch provides (search_lds bytes).
template <int OP, int ROWS, int M, bool RES>
__device__ __forceinline__ void search_lds_core(const SearchArgs &a, const size_t row0) {
    using T = OpT<OP>;
    constexpr int ES = T::ES, LB = T::LB;
    constexpr bool INT = OP == OP_I8 || OP == OP_I16;
    // I16: two i8 planes per centroid, c16 = 2^8 hi + lo + 128 with hi, lo in
    // [-128, 127]; x.c16 = 2^8 x.hi + x.lo + 128 sum(x) computed with wrapping
    // int32 adds, which is the reference's wrapping i32 dot product
    // (vpdpwssd lanes) exactly.
    constexpr int NPL = OP == OP_I16 ? I16_PLANES : 1;
    using Score = typename std::conditional<INT, int64_t, float>::type;
    extern __shared__ __align__(16) uint8_t smem[];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, j = lane & 31;
    const int RB = a.row_bytes, RS = RB + 16;
    uint8_t *sR = smem;
    uint8_t *sC = smem + BM * RS;
    float *sHn = (float *)(sC + NPL * BN * RS);
    int64_t *sCn = (int64_t *)(sC + NPL * BN * RS);
    int *sCe = (int *)(sHn + BN);  // FP8: the block's centroid MFMA scales
    const int segs = RB / 16;
    const int nchunks = RES ? 1 : (int)((a.d_pad * ES) / RB);
    const size_t my_row = row0 + wave * 32 + j;

    auto load_rows = [&](int chunk) {
        for (int q = tid; q < BM * segs; q += 256) {
            int r = q / segs, sg = q - r * segs;
            size_t grow = row0 + r;
            bool valid = grow < a.m;
            size_t src = valid ? (a.ids ? (size_t)a.ids[grow] : grow) : 0;
            size_t e0 = ((size_t)chunk * RB + sg * 16) / ES;
            *(uint4 *)(sR + r * RS + sg * 16) = row_segment<OP, ROWS>(a, grow, src, e0, valid);
        }
    };
    auto load_cents = [&](size_t c0, int chunk) {
        const size_t stride = a.d_pad * ES, plane = a.k_pad * stride;
        for (int q = tid; q < NPL * BN * segs; q += 256) {
            const int pl = q / (BN * segs), qq = q - pl * BN * segs;
            const int r = qq / segs, sg = qq - r * segs;
            const uint8_t *src = a.cen_op + pl * plane + (c0 + r) * stride + (size_t)chunk * RB + sg * 16;
            *(uint4 *)(sC + (pl * BN + r) * RS + sg * 16) = *(const uint4 *)src;
        }
        if (tid < BN) {
            if constexpr (INT) sCn[tid] = a.cn[c0 + tid];
            else sHn[tid] = a.hn[c0 + tid];
            if constexpr (OP == OP_FP8) sCe[tid] = a.cexp[c0 + tid];
        }
    };

    // Per-lane running top-m over this lane's centroid subset.
    Score bs[M];
    uint32_t bi[M];
#pragma unroll
    for (int t = 0; t < M; t++) {
        if constexpr (INT) bs[t] = INT64_MAX;
        else bs[t] = INFINITY;
        bi[t] = NO_LABEL;
    }
    // Squared row norm (f32 for FP32, exact int for I8), accumulated from LDS.
    float xn32 = 0.0f;
    int32_t xn_i = 0, xs_i = 0;  // I16: the row's element sum (half-row partial)
    // FP8: the E8M0 scale 1 / s_row of this lane's row (its MFMA B column)
    [[maybe_unused]] const int sbr = OP == OP_FP8 && my_row < a.m ? e8m0_inverse(a.row_scale[my_row]) : 127;
    auto norm_part = [&]() {
        const uint8_t *p = sR + (wave * 32 + j) * RS + h * (RB / 2);
        for (int b = 0; b < RB / 2; b += 16) {
            uint4 v = *(const uint4 *)(p + b);
            if constexpr (OP == OP_F32) {
                float f[4] = {__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w)};
#pragma unroll
                for (int q = 0; q < 4; q++) xn32 = __fadd_rn(xn32, __fmul_rn(f[q], f[q]));
            } else if constexpr (INT) {
                uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int q = 0; q < 4; q++)
#pragma unroll
                    for (int bb = 0; bb < 4; bb++) {
                        int e = (int)(int8_t)((w[q] >> (8 * bb)) & 0xff);
                        xn_i += e * e;
                        xs_i += e;
                    }
            }
        }
    };

    if constexpr (RES) {
        load_rows(0);
        __syncthreads();
        if constexpr (OP == OP_F32 || INT) norm_part();
    }

    for (size_t c0 = 0; c0 < a.k_pad; c0 += BN) {
        using Acc = typename std::conditional<INT, i32x16, f32x16>::type;
        Acc acc0 = {}, acc1 = {};
        i32x16 pacc0[NPL] = {}, pacc1[NPL] = {};  // I16: per plane
        for (int ch = 0; ch < nchunks; ch++) {
            if constexpr (!RES) load_rows(ch);
            load_cents(c0, ch);
            __syncthreads();
            if constexpr (!RES) {
                if (c0 == 0) {
                    if constexpr (OP == OP_F32 || INT) norm_part();
                }
            }
            const uint8_t *pb = sR + (wave * 32 + j) * RS + h * (RB / 2);
            const uint8_t *pa0 = sC + j * RS + h * (RB / 2);
            const uint8_t *pa1 = sC + (32 + j) * RS + h * (RB / 2);
            if constexpr (OP == OP_F32) {
                for (int b = 0; b < RB / 2; b += 16) {
                    float4 xb = *(const float4 *)(pb + b);
                    float4 c0v = *(const float4 *)(pa0 + b);
                    float4 c1v = *(const float4 *)(pa1 + b);
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(c0v.x, xb.x, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(c1v.x, xb.x, acc1, 0, 0, 0);
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(c0v.y, xb.y, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(c1v.y, xb.y, acc1, 0, 0, 0);
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(c0v.z, xb.z, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(c1v.z, xb.z, acc1, 0, 0, 0);
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(c0v.w, xb.w, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(c1v.w, xb.w, acc1, 0, 0, 0);
                }
            } else if constexpr (OP == OP_BF16) {
                for (int b = 0; b < RB / 2; b += LB) {
                    bf16x8 xb = __builtin_bit_cast(bf16x8, *(const uint4 *)(pb + b));
                    bf16x8 ca = __builtin_bit_cast(bf16x8, *(const uint4 *)(pa0 + b));
                    bf16x8 cb = __builtin_bit_cast(bf16x8, *(const uint4 *)(pa1 + b));
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ca, xb, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cb, xb, acc1, 0, 0, 0);
                }
            } else if constexpr (OP == OP_I8) {
                for (int b = 0; b < RB / 2; b += LB) {
                    i32x4 xb = __builtin_bit_cast(i32x4, *(const uint4 *)(pb + b));
                    i32x4 ca = __builtin_bit_cast(i32x4, *(const uint4 *)(pa0 + b));
                    i32x4 cb = __builtin_bit_cast(i32x4, *(const uint4 *)(pa1 + b));
                    acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ca, xb, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(cb, xb, acc1, 0, 0, 0);
                }
            } else if constexpr (OP == OP_I16) {
#pragma unroll
                for (int pl = 0; pl < NPL; pl++) {
                    const uint8_t *qa0 = pa0 + pl * BN * RS, *qa1 = pa1 + pl * BN * RS;
                    for (int b = 0; b < RB / 2; b += LB) {
                        i32x4 xb = __builtin_bit_cast(i32x4, *(const uint4 *)(pb + b));
                        i32x4 ca = __builtin_bit_cast(i32x4, *(const uint4 *)(qa0 + b));
                        i32x4 cb = __builtin_bit_cast(i32x4, *(const uint4 *)(qa1 + b));
                        pacc0[pl] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ca, xb, pacc0[pl], 0, 0, 0);
                        pacc1[pl] = __builtin_amdgcn_mfma_i32_32x32x32_i8(cb, xb, pacc1[pl], 0, 0, 0);
                    }
                }
            } else {  // OP_FP8: scales 1 / s_C (this lane's centroid) and 1 / s_row
                const int sa0 = sCe[j], sa1 = sCe[32 + j];
                for (int b = 0; b < RB / 2; b += LB) {
                    uint4 x0 = *(const uint4 *)(pb + b), x1 = *(const uint4 *)(pb + b + 16);
                    uint4 a0 = *(const uint4 *)(pa0 + b), a1 = *(const uint4 *)(pa0 + b + 16);
                    uint4 b0 = *(const uint4 *)(pa1 + b), b1 = *(const uint4 *)(pa1 + b + 16);
                    i32x8 xb = {(int)x0.x, (int)x0.y, (int)x0.z, (int)x0.w, (int)x1.x, (int)x1.y, (int)x1.z, (int)x1.w};
                    i32x8 ca = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w, (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
                    i32x8 cb = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w, (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
                    acc0 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(ca, xb, acc0, 0, 0, 0, sa0, 0, sbr);
                    acc1 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(cb, xb, acc1, 0, 0, 0, sa1, 0, sbr);
                }
            }
            if (ch == nchunks - 1) {
                // Epilogue: fold the 2 x 16 scores of this lane into its top-m,
                // in ascending centroid index. (I16: the row sum of every
                // chunk is in by now, the first centroid block's chunks
                // having added theirs.)
                [[maybe_unused]] const uint32_t xs128 = 128u * (uint32_t)(xs_i + __shfl_xor(xs_i, 32, 64));
#pragma unroll
                for (int nb = 0; nb < 2; nb++) {
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const int il = nb * 32 + 8 * (r / 4) + 4 * h + (r % 4);
                        const uint32_t ci = (uint32_t)(c0 + il);
                        Score s;
                        if constexpr (OP == OP_I8) {
                            int32_t dv = nb == 0 ? acc0[r] : acc1[r];
                            s = sCn[il] - 2 * (int64_t)dv;
                        } else if constexpr (OP == OP_I16) {
                            const i32x16 *pp = nb == 0 ? pacc0 : pacc1;
                            // x.c16 (mod 2^32, as the reference's i32 sum)
                            const uint32_t dv = (uint32_t)pp[0][r] + ((uint32_t)pp[1 % NPL][r] << 8) + xs128;
                            s = sCn[il] - ((int64_t)2 << a.shift) * (int64_t)(int32_t)dv;
                        } else {
                            // operands are -c and x (FP8: -c s_c and x s_row,
                            // scaled back by the MFMA): acc ~ -x.c
                            float dv = nb == 0 ? acc0[r] : acc1[r];
                            s = sHn[il] + dv;
                        }
                        if constexpr (M == 1) {
                            if (s < bs[0] && ci < a.k) {
                                bs[0] = s;
                                bi[0] = ci;
                            }
                        } else {
                            if (s < bs[M - 1]) topm_insert<M, Score>(bs, bi, s, ci);
                        }
                    }
                }
            }
            __syncthreads();
        }
    }

    // Merge the two lane halves that hold the same row.
#pragma unroll
    for (int t = 0; t < M; t++) {
        Score ps;
        if constexpr (INT) ps = shfl_xor_i64(bs[t], 32);
        else ps = __shfl_xor(bs[t], 32, 64);
        uint32_t pi = (uint32_t)__shfl_xor((int)bi[t], 32, 64);
        if (h == 0) topm_insert_lex<M, Score>(bs, bi, ps, pi);
    }
    if constexpr (OP == OP_F32) xn32 = __fadd_rn(xn32, __shfl_xor(xn32, 32, 64));
    if constexpr (INT) xn_i += __shfl_xor(xn_i, 32, 64);

    if (h == 0 && my_row < a.m) {
        if constexpr (OP == OP_F32 && M == 1) {
            // dist = |x|^2 + |c|^2 + bias - 2 x.c, clamped at 0 (src/cpu.rs:160-162)
            float v = __fadd_rn(__fmul_rn(2.0f, bs[0]), xn32);
            a.labels[my_row] = bi[0] == NO_LABEL ? 0u : bi[0];
            a.dist[my_row] = v > 0.0f ? v : 0.0f;
        } else if constexpr (INT) {
            // dist = ((|x|^2 + best * 4^-shift) as f64).max(0) as f32
            // (src/cpu_i8.rs:183-186; the power-of-two scaling is exact)
            double v = __dadd_rn((double)(int64_t)xn_i, ldexp((double)bs[0], -2 * a.shift));
            a.labels[my_row] = bi[0] == NO_LABEL ? 0u : bi[0];
            a.dist[my_row] = (float)(v > 0.0 ? v : 0.0);
        } else {
#pragma unroll
            for (int t = 0; t < M; t++) a.cand[my_row * M + t] = bi[t];
        }
    }
}

template <int OP, int ROWS, int M, bool RES>
__global__ __launch_bounds__(256) void search_kernel(SearchArgs a) {
    search_lds_core<OP, ROWS, M, RES>(a, (size_t)blockIdx.x * BM);
}

// ------------------------------------------- register-resident searches ----
//
// For operand rows of at most 512 bytes. Each
// wave holds FR 32-row fragments of B in VGPRs for the whole centroid sweep.
// The operand centroids are also stored tile-major (tile_layout_kernel):
// tile t holds TB = 32 * NB centroids, and 16-byte quad q of its centroid c
// sits at byte (q * TB + c) * 16 of the tile. A tile is one contiguous block,
// so it streams into LDS through buffer loads that write LDS directly (no
// staging registers, 1 KiB per wave-load), and a fragment read, quad q of
// 32 consecutive centroids, is 512 contiguous bytes: free of bank conflicts,
// and every read of a tile sits at an immediate offset from one lane address.
// (Buffer loads rather than global ones: the compiler tracks a global load
// into LDS as a possible LDS access and then waits for every outstanding LDS
// read wherever it needs any one of them.)
//
// BF16 / FP8 (candidate search): the accumulators start at
// |c|^2/2 + bias/2 + B, where B is a per-workgroup offset keeping every score
// positive, so a score's f32 bits order like the score. The operands'
// rounding is relative (FP8: each row and each centroid has its own
// power-of-two scale), so |x~||c~| <= (1 + e) |x||c| with e = 2^-7 for
// BF16 (2^-8 suffices) and 2^-2 for FP8 (2^-4 per e4m3 operand, plus its
// subnormals, below 2^-3): over all |c| = t the score is at least
// t^2/2 - (1 + e)|x|t >= -(1 + e)^2 |x|^2 / 2, and B = (1 + e)^2 (1 + 2^-6)
// max|x|^2 / 2 + max(-bias) / 2 holds whatever the centroids' sizes.
// Each lane replaces the low 5 bits of its 16*NB scores of a tile by their
// position and keeps the m smallest such keys of its half of the row; the
// rescoring kernel settles the exact FP32 order.
//
// FP32 (exact): accumulators start at (|c|^2 + bias)/2 and the per-lane
// running argmin takes every score with a strict '<' in ascending index.
//
// I8 (exact): the score is the integer 2 x.c8 - (sum c8^2 + round(bias)),
// maximised; key = (score << 5) | (31 - position) orders by score, then by
// lowest index. The per-centroid part ((-cn) << 5) + 31 - position is
// tabulated in LDS with the tile, so a key is one shift-add of the dot
// product and one integer max per value settles the tile.

// Per-lane running result: raw score bits of the candidate searches, the
// f32 score of FP32, the int32 score of I8.
template <int OP> struct RegScore { using T = uint32_t; };
template <> struct RegScore<OP_F32> { using T = float; };
template <> struct RegScore<OP_I8> { using T = int32_t; };
template <> struct RegScore<OP_I16> { using T = double; };  // exact integers

template <int OP> using RegAcc = typename std::conditional<OP == OP_I8 || OP == OP_I16, i32x16, f32x16>::type;

// Centroids per tile of the register-resident search: the TB = 32 NB MFMA
// rows, except for I16, whose tile holds the two planes of 32 centroids as
// its two 32-row blocks (block b = plane b).
template <int OP, int NB> constexpr int reg_centroids() { return OP == OP_I16 ? 32 : 32 * NB; }

// Words of a tile's offset slot in LDS: 256 score offsets (TB <= 256), and
// for FP8 then 256 centroid scales.
template <int OP> constexpr int reg_offset_slot() { return OP == OP_FP8 ? 512 : 256; }

// Operand quads a lane reads ahead of the MFMAs that use them.
template <int OP> struct RegPrefetch { static constexpr int D = (OP == OP_BF16 || OP == OP_I8 || OP == OP_I16) ? 2 : 1; };

// The MFMA half of a tile: acc = (score offsets) + centroid tile . row
// fragments. sT is the tile in LDS, sH its per-centroid offsets.
template <int OP, int NS, int FR, int NB>
__device__ __forceinline__ void reg_mfma(const uint8_t *sT, const float *sH, const uint4 (&xb)[FR][NS],
                                         const int (&sb)[FR], const int (&sa)[NB], int h, int j,
                                         RegAcc<OP> (&acc)[FR][NB]) {
    constexpr int TB = 32 * NB;
    constexpr int QS = TB * 16;  // bytes from one quad of a centroid to its next
    using Acc = RegAcc<OP>;
    // The offsets enter as the C operand of each fragment's first MFMA
    // (no per-fragment copies); I8 starts from zero.
    Acc init[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        if constexpr (OP == OP_I8 || OP == OP_I16) {
            init[nb] = Acc{};
        } else {
#pragma unroll
            for (int r = 0; r < 16; r++) init[nb][r] = sH[nb * 32 + 8 * (r >> 2) + 4 * h + (r & 3)];
        }
    }
    // this lane's quad 0 (of its row half) of centroid j of block 0
    const uint8_t *ab = sT + (h * NS * TB + j) * 16;
    auto ld = [&](int s, int nb) -> uint4 { return *(const uint4 *)(ab + s * QS + nb * 512); };
    if constexpr (OP == OP_F32) {
#pragma unroll
        for (int s = 0; s < NS; s++) {
            float4 cv[NB];
#pragma unroll
            for (int nb = 0; nb < NB; nb++) cv[nb] = __builtin_bit_cast(float4, ld(s, nb));
#pragma unroll
            for (int fr = 0; fr < FR; fr++) {
                const float4 xv = __builtin_bit_cast(float4, xb[fr][s]);
#pragma unroll
                for (int nb = 0; nb < NB; nb++) {
                    acc[fr][nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[nb].x, xv.x, s == 0 ? init[nb] : acc[fr][nb], 0, 0, 0);
                    acc[fr][nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[nb].y, xv.y, acc[fr][nb], 0, 0, 0);
                    acc[fr][nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[nb].z, xv.z, acc[fr][nb], 0, 0, 0);
                    acc[fr][nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[nb].w, xv.w, acc[fr][nb], 0, 0, 0);
                }
            }
        }
    } else if constexpr (OP == OP_BF16 || OP == OP_I8 || OP == OP_I16) {
        // a ring of PD + 1 operand slots: step s reads step s + PD's quads
        constexpr int PD = RegPrefetch<OP>::D;
        uint4 ring[PD + 1][NB];
#pragma unroll
        for (int p = 0; p < PD; p++)
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
                if (p < NS) ring[p][nb] = ld(p, nb);
#pragma unroll
        for (int s = 0; s < NS; s++) {
            if (s + PD < NS) {
#pragma unroll
                for (int nb = 0; nb < NB; nb++) ring[(s + PD) % (PD + 1)][nb] = ld(s + PD, nb);
            }
#pragma unroll
            for (int fr = 0; fr < FR; fr++)
#pragma unroll
                for (int nb = 0; nb < NB; nb++) {
                    const uint4 cv = ring[s % (PD + 1)][nb];
                    if constexpr (OP == OP_BF16)
                        acc[fr][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, cv), __builtin_bit_cast(bf16x8, xb[fr][s]),
                            s == 0 ? init[nb] : acc[fr][nb], 0, 0, 0);
                    else
                        acc[fr][nb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                            __builtin_bit_cast(i32x4, cv), __builtin_bit_cast(i32x4, xb[fr][s]),
                            s == 0 ? init[nb] : acc[fr][nb], 0, 0, 0);
                }
        }
    } else {  // OP_FP8: two quads per lane per MFMA
#pragma unroll
        for (int s = 0; s < NS / 2; s++) {
            i32x8 cv[NB];
#pragma unroll
            for (int nb = 0; nb < NB; nb++) {
                const uint4 a0 = ld(2 * s, nb), a1 = ld(2 * s + 1, nb);
                cv[nb] = i32x8{(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w, (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
            }
#pragma unroll
            for (int fr = 0; fr < FR; fr++) {
                const uint4 x0 = xb[fr][2 * s], x1 = xb[fr][2 * s + 1];
                const i32x8 xv = {(int)x0.x, (int)x0.y, (int)x0.z, (int)x0.w, (int)x1.x, (int)x1.y, (int)x1.z, (int)x1.w};
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    acc[fr][nb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(cv[nb], xv, s == 0 ? init[nb] : acc[fr][nb],
                                                                                  0, 0, 0, sa[nb], 0, sb[fr]);
            }
        }
    }
}

// The epilogue half: fold a tile's scores into each row's running result.
template <int OP, int M, int FR, int NB>
__device__ __forceinline__ void reg_fold(const RegAcc<OP> (&acc)[FR][NB], const float *sN, int h, uint32_t c0,
                                         int shift, const uint32_t (&xs128)[FR],
                                         typename RegScore<OP>::T (&bs)[FR][M], uint32_t (&bi)[FR][M]) {
#pragma unroll
    for (int fr = 0; fr < FR; fr++) {
        if constexpr ((OP == OP_BF16 || OP == OP_FP8) && M == 1) {
            uint32_t g = 0xFFFFFFFFu;
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
#pragma unroll
                for (int r = 0; r < 16; r++) g = min(g, cand_key(acc[fr][nb][r], (uint32_t)(nb * 16 + r)));
            const uint32_t li = g & 31u, nb = li >> 4, r = li & 15u;
            const uint32_t gi = c0 + nb * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
            topm_insert<M, uint32_t>(bs[fr], bi[fr], g, gi);
        } else if constexpr (OP == OP_BF16 || OP == OP_FP8) {
            // the tile's two smallest keys, then the exact list update
            static_assert(NB <= 2, "a candidate tile has at most two blocks");
            const uint32_t mask = cand_mask();
            Top2 t[NB];
#pragma unroll
            for (int p = 0; p < 8; p++) top2_phase<false, 0>(t[0], acc[fr][0], mask, p);
            uint32_t m1 = t[0].lo, m2 = t[0].hi;
            if constexpr (NB == 2) {
#pragma unroll
                for (int p = 0; p < 8; p++) top2_phase<false, 16>(t[NB - 1], acc[fr][NB - 1], mask, p);
                m1 = min(t[0].lo, t[NB - 1].lo);
                m2 = vmin3(max(t[0].lo, t[NB - 1].lo), t[0].hi, t[NB - 1].hi);
            }
            cand_settle<M, NB>(acc[fr], m1, m2, h, c0, bs[fr], bi[fr]);
        } else if constexpr (OP == OP_F32) {
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const float s = acc[fr][nb][r];
                    const bool u = s < bs[fr][0];
                    bs[fr][0] = u ? s : bs[fr][0];
                    bi[fr][0] = u ? c0 + nb * 32 + 8 * (r >> 2) + 4 * h + (r & 3) : bi[fr][0];
                }
        } else if constexpr (OP == OP_I16) {
            // score = cn - 2^(shift+1) x.c16, x.c16 = x.lo + 2^8 x.hi + 128 sum(x)
            // taken mod 2^32 as the reference's i32 sum. As an f64 key
            // 32 score + position (the tile's offsets are stored as 32 cn +
            // position): |cn| < 2^46 and |2^(shift+1) x.c16| <= 2^39 keep
            // every key an exact integer below 2^52, so one f64 min per value
            // takes the lowest score, then the lowest index; the tile's best
            // replaces the running one only with a strictly lower score
            // (an earlier tile keeps a tie).
            const double *sK = (const double *)sN;
            const double step = -(double)(2 << shift) * 32.0;
            double kmin = INFINITY;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int cl = 8 * (r >> 2) + 4 * h + (r & 3);
                const uint32_t dv = (uint32_t)acc[fr][0][r] + ((uint32_t)acc[fr][1][r] << 8) + xs128[fr];
                kmin = fmin(kmin, fma(step, (double)(int32_t)dv, sK[cl]));
            }
            const double sc = floor(kmin * 0.03125);
            const bool u = sc < bs[fr][0];
            const double pos = u ? kmin - sc * 32.0 : 0.0;
            bs[fr][0] = u ? sc : bs[fr][0];
            bi[fr][0] = u ? c0 + (uint32_t)pos : bi[fr][0];
        } else {  // OP_I8: maximise key = (dot << 6) + ((-cn) << 5) + 31 - position
            const int32_t *sK = (const int32_t *)sN;
            int32_t g = INT32_MIN;
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int32_t key = (int32_t)(((uint32_t)acc[fr][nb][r] << 6) +
                                                  (uint32_t)sK[nb * 32 + 8 * (r >> 2) + 4 * h + (r & 3)]);
                    g = max(g, key);
                }
            const int32_t sc = g >> 5;
            const uint32_t li = 31u - (uint32_t)(g & 31), nb = li >> 4, r = li & 15u;
            const uint32_t gi = c0 + nb * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
            const bool u = sc > bs[fr][0];
            bs[fr][0] = u ? sc : bs[fr][0];
            bi[fr][0] = u ? gi : bi[fr][0];
        }
    }
}

// The candidate searches on two-block tiles (BF16 / FP8, rows of at most
// 256 bytes): block 0's MFMAs first, then block 1's with block 0's fold
// between them, so half of the fold runs while the matrix cores work; then
// block 1's fold and the list update. Same keys and the same list as
// reg_mfma + reg_fold.
template <int OP, int M, int NS, int FR>
__device__ __forceinline__ void reg_tile_cand2(const uint8_t *sT, const float *sH, const uint4 (&xb)[FR][NS],
                                               const int (&sb)[FR], const int (&sa)[2], int h, int j, uint32_t c0,
                                               f32x16 (&acc)[FR][2], uint32_t (&bs)[FR][M],
                                               uint32_t (&bi)[FR][M]) {
    constexpr int TB = 64, QS = TB * 16;
    constexpr int STEPS = OP == OP_FP8 ? NS / 2 : NS;  // MFMAs per fragment and block
    constexpr int PER = 16 / STEPS > 0 ? 16 / STEPS : 1;  // block-0 values folded per step
    const uint8_t *ab = sT + (h * NS * TB + j) * 16;
    auto ld = [&](int q, int nb) -> uint4 { return *(const uint4 *)(ab + q * QS + nb * 512); };
    auto key = [&](float v, uint32_t pos) -> uint32_t { return (__float_as_uint(v) & 0xFFFFFFE0u) | pos; };
    auto mfma = [&](int st, int fr, int nb, const uint4 (&ops)[NS], const f32x16 &c) -> f32x16 {
        if constexpr (OP == OP_FP8) {
            const uint4 a0 = ops[2 * st], a1 = ops[2 * st + 1], x0 = xb[fr][2 * st], x1 = xb[fr][2 * st + 1];
            const i32x8 cv = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w, (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
            const i32x8 xv = {(int)x0.x, (int)x0.y, (int)x0.z, (int)x0.w, (int)x1.x, (int)x1.y, (int)x1.z, (int)x1.w};
            return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(cv, xv, c, 0, 0, 0, sa[nb], 0, sb[fr]);
        } else {
            return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ops[st]),
                                                           __builtin_bit_cast(bf16x8, xb[fr][st]), c, 0, 0, 0);
        }
    };
    uint32_t g[FR];
    Top2 t0[FR];  // M > 1: block 0's two smallest keys
    const uint32_t mask = cand_mask();
    {  // block 0
        uint4 ops[NS];
#pragma unroll
        for (int q = 0; q < NS; q++) ops[q] = ld(q, 0);
        f32x16 init;
#pragma unroll
        for (int r = 0; r < 16; r++) init[r] = sH[8 * (r >> 2) + 4 * h + (r & 3)];
#pragma unroll
        for (int st = 0; st < STEPS; st++)
#pragma unroll
            for (int fr = 0; fr < FR; fr++) acc[fr][0] = mfma(st, fr, 0, ops, st == 0 ? init : acc[fr][0]);
    }
    {  // block 1, block 0 folded in between
        uint4 ops[NS];
#pragma unroll
        for (int q = 0; q < NS; q++) ops[q] = ld(q, 1);
        f32x16 init;
#pragma unroll
        for (int r = 0; r < 16; r++) init[r] = sH[32 + 8 * (r >> 2) + 4 * h + (r & 3)];
#pragma unroll
        for (int fr = 0; fr < FR; fr++) g[fr] = 0xFFFFFFFFu;
#pragma unroll
        for (int st = 0; st < STEPS; st++) {
#pragma unroll
            for (int fr = 0; fr < FR; fr++) acc[fr][1] = mfma(st, fr, 1, ops, st == 0 ? init : acc[fr][1]);
            if constexpr (M == 1) {
#pragma unroll
                for (int fr = 0; fr < FR; fr++)
#pragma unroll
                    for (int u = 0; u < PER; u++) {
                        const int r = st * PER + u;
                        if (r < 16) g[fr] = min(g[fr], key(acc[fr][0][r], (uint32_t)r));
                    }
                // keep this step's share of the fold next to its MFMAs: the
                // empty asm pins the partial minimum here (the optimiser
                // would otherwise gather every min into one tree after the
                // last MFMA), the scheduling barrier keeps the MFMAs from
                // running together
#pragma unroll
                for (int fr = 0; fr < FR; fr++) asm volatile("" : "+v"(g[fr]));
            } else {
                // this step's phases of block 0's two-smallest tree, pinned
                // here the same way
#pragma unroll
                for (int fr = 0; fr < FR; fr++)
#pragma unroll
                    for (int p = 0; p < 8; p++)
                        if ((p * STEPS) / 8 == st) top2_phase<true, 0>(t0[fr], acc[fr][0], mask, p);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (M == 1) {
#pragma unroll
            for (int fr = 0; fr < FR; fr++)
#pragma unroll
                for (int r = STEPS * PER; r < 16; r++) g[fr] = min(g[fr], key(acc[fr][0][r], (uint32_t)r));
        }
    }
#pragma unroll
    for (int fr = 0; fr < FR; fr++) {
        if constexpr (M == 1) {
#pragma unroll
            for (int r = 0; r < 16; r++) g[fr] = min(g[fr], key(acc[fr][1][r], (uint32_t)(16 + r)));
            const uint32_t li = g[fr] & 31u, nb = li >> 4, r = li & 15u;
            const uint32_t gi = c0 + nb * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
            topm_insert<M, uint32_t>(bs[fr], bi[fr], g[fr], gi);
        } else {
            Top2 t1;
#pragma unroll
            for (int p = 0; p < 8; p++) top2_phase<false, 16>(t1, acc[fr][1], mask, p);
            const uint32_t m1 = min(t0[fr].lo, t1.lo);
            const uint32_t m2 = vmin3(max(t0[fr].lo, t1.lo), t0[fr].hi, t1.hi);
            cand_settle<M, 2>(acc[fr], m1, m2, h, c0, bs[fr], bi[fr]);
        }
    }
}

// One workgroup's sweep: rows [row0, row0 + 128 FR) of a.x against every
// centroid tile. LDS: three tile slots sC0..2 (TILE bytes each), three
// offset slots sH0..2 (reg_offset_slot words) and sRed (4 floats).
template <int OP, int ROWS, int M, int NS, int FR, int NB>
__device__ __forceinline__ void search_reg_core(const SearchArgs &a, const size_t row0, const int t0, const int t1,
                                                const int slice, uint8_t *sC0, uint8_t *sC1, uint8_t *sC2,
                                                float *sH0, float *sH1, float *sH2, float *sRed) {
    constexpr int ES = OpT<OP>::ES;
    constexpr int RB = 32 * NS;           // operand bytes per row
    constexpr int HB = RB / 2;            // bytes per lane half
    constexpr int TB = 32 * NB;           // MFMA rows per tile
    constexpr int CT = reg_centroids<OP, NB>();  // centroids per tile
    constexpr int TILE = TB * RB;         // bytes per tile
    constexpr int PER_T = TILE / 4096;    // 1 KiB wave-loads per wave per tile
    constexpr bool CAND = OP == OP_BF16 || OP == OP_FP8;
    using Score = typename RegScore<OP>::T;
    static_assert(NS % 2 == 0 && PER_T >= 1 && TILE % 4096 == 0, "shape");
    static_assert(OP != OP_I16 || NB == I16_PLANES, "an I16 tile is its centroids' planes");

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int h = lane >> 5, j = lane & 31;

    // Row fragments into registers; squared row norms (exact for I8).
    uint4 xb[FR][NS];
    int sb[FR];
    float xn[FR];
    int32_t xni[FR];
    uint32_t xs128[FR];  // I16: 128 x the row's element sum (mod 2^32)
    float nmax = 0.0f;
#pragma unroll
    for (int fr = 0; fr < FR; fr++) {
        const size_t r = row0 + (size_t)(wave * FR + fr) * 32 + j;
        const bool valid = r < a.m;
        const size_t src = valid ? (a.ids ? (size_t)a.ids[r] : r) : 0;
        if (OP != OP_FP8 || !a.enc) {
#pragma unroll
            for (int q = 0; q < NS; q++) xb[fr][q] = row_segment<OP, ROWS>(a, r, src, (size_t)(h * HB + 16 * q) / ES, valid);
        }
        // |x|^2 from the fragment just loaded (pad elements are zero): exact
        // for I8, the f32 values for FP32, the bf16-rounded values for BF16
        // (only the positivity offset uses it), the encoding pass's for FP8
        float part = 0.0f;
        int32_t ipart = 0, spart = 0;
        if constexpr (OP == OP_FP8) {
            if (a.enc) {
                // the row's e4m3 encoding as the encoding pass makes it: its
                // largest finite |v| over both lane halves, the power-of-two
                // scale, then this half's codes (the row read again, now from
                // cache); |x|^2 for the positivity offset
                const float *xr = (const float *)a.x + src * a.d;
                float mx = 0.0f, sq = 0.0f;
                auto quad = [&](size_t e, float (&v)[4]) {
                    if (a.vec && e + 4 <= a.d) {
                        const float4 t = *(const float4 *)(xr + e);
                        v[0] = t.x, v[1] = t.y, v[2] = t.z, v[3] = t.w;
                    } else {
#pragma unroll
                        for (int u = 0; u < 4; u++) v[u] = e + u < a.d ? xr[e + u] : 0.0f;
                    }
                };
                if (valid) {
#pragma unroll
                    for (int q = 0; q < 4 * NS; q++) {
                        float v[4];
                        quad((size_t)(h * HB + 4 * q), v);
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const float av = fabsf(v[u]);
                            if (av > mx && isfinite(av)) mx = av;
                            sq = fmaf(v[u], v[u], sq);
                        }
                    }
                }
                mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
                const float s = fp8_scale_for(mx);
#pragma unroll
                for (int q = 0; q < NS; q++) {
                    uint32_t w[4] = {0u, 0u, 0u, 0u};
                    if (valid) {
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            float v[4];
                            quad((size_t)(h * HB + 16 * q + 4 * u), v);
                            int pk = __builtin_amdgcn_cvt_pk_fp8_f32(v[0] * s, v[1] * s, 0, false);
                            pk = __builtin_amdgcn_cvt_pk_fp8_f32(v[2] * s, v[3] * s, pk, true);
                            w[u] = (uint32_t)pk;
                        }
                    }
                    xb[fr][q] = make_uint4(w[0], w[1], w[2], w[3]);
                }
                part = sq;  // halves summed below
                sb[fr] = valid ? e8m0_inverse(s) : 127;
            } else if (valid && h == 0) {
                part = a.row_norm[r];
            }
        } else {
#pragma unroll
            for (int q = 0; q < NS; q++) {
                const uint32_t w[4] = {xb[fr][q].x, xb[fr][q].y, xb[fr][q].z, xb[fr][q].w};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if constexpr (OP == OP_F32) {
                        const float v = __uint_as_float(w[u]);
                        part = fmaf(v, v, part);
                    } else if constexpr (OP == OP_BF16) {
                        const float lo = __uint_as_float(w[u] << 16), hi = __uint_as_float(w[u] & 0xFFFF0000u);
                        part = fmaf(hi, hi, fmaf(lo, lo, part));
                    } else {
#pragma unroll
                        for (int b = 0; b < 4; b++) {
                            const int32_t v = (int32_t)(int8_t)(w[u] >> (8 * b));
                            ipart += v * v;
                            spart += v;
                        }
                    }
                }
            }
        }
        part += __shfl_xor(part, 32, 64);
        ipart += __shfl_xor(ipart, 32, 64);
        xn[fr] = part;
        xni[fr] = ipart;
        xs128[fr] = 0u;
        if constexpr (OP == OP_I16) xs128[fr] = 128u * (uint32_t)(spart + __shfl_xor(spart, 32, 64));
        nmax = fmaxf(nmax, part);
        if (OP != OP_FP8 || !a.enc) sb[fr] = 127;
        if constexpr (OP == OP_FP8) {
            // E8M0 scale 1 / s_row: undoes the row's power-of-two encoding scale
            if (!a.enc) {
                const uint32_t eb = valid ? ((__float_as_uint(a.row_scale[r]) >> 23) & 0xff) : 127u;
                sb[fr] = 254 - (int)eb;
            }
        }
    }
    float B = 0.0f;
    if constexpr (CAND) {
        for (int off = 16; off > 0; off >>= 1) nmax = fmaxf(nmax, __shfl_xor(nmax, off, 64));
        if (lane == 0) sRed[wave] = nmax;
        __syncthreads();
        nmax = fmaxf(fmaxf(sRed[0], sRed[1]), fmaxf(sRed[2], sRed[3]));
        const float negb = __uint_as_float(a.stats[1]);
        constexpr float e = OP == OP_FP8 ? 0.25f : 0.0078125f;
        // (1 + 2^-6) covers the bf16 rounding of the norm itself; FLT_MIN
        // keeps B positive for all-zero rows
        B = 0.5078125f * (1.0f + e) * (1.0f + e) * nmax + 0.5f * negb + 1.17549435e-38f;
    }

    // Tile DMA: wave-load q of wave w copies tile bytes [(4q + w) KiB, +1 KiB)
    // to the same place in LDS; lane L moves 16 bytes at L * 16. The buffer
    // resource spans the tile-major table (its size is checked on the host).
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void *)a.cen_op, 0, (int)(a.k_pad / CT * (size_t)TILE), 0x00020000);
    auto dma = [&](uint8_t *dst, int t) {
        const uint32_t so = (uint32_t)t * TILE + (uint32_t)wv * 1024u;
#pragma unroll
        for (int q = 0; q < PER_T; q++)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (__attribute__((address_space(3))) void *)(dst + q * 4096 + wv * 1024),
                                                     16, (uint32_t)lane * 16u, so + q * 4096u, 0, 0);
    };
    // Per-centroid score offsets of a tile, loaded raw: any arithmetic on the
    // loaded value would make the compiler wait for every outstanding load of
    // the wave (the tile DMA included) right there. Converted at the store:
    // float hn + B, the I8 key part ((-cn) << 5) + 31 - position, or the
    // I16 f64 key part 32 cn + position.
    // I8 loads the low word of cn only (|cn| <= 2^26 on this path): a dead
    // high word still being written by a load would make the compiler wait
    // for it, and with it for the tile DMA, before reusing its register.
    // I16 loads each of its 32 i64 offsets whole (both words are used).
    using HRaw = typename std::conditional<OP == OP_I16, int64_t,
                                           typename std::conditional<OP == OP_I8, int32_t, float>::type>::type;
    auto hload = [&](int t) -> HRaw {
        // every thread loads (threads past TB repeat an entry): no branch
        if constexpr (OP == OP_I16) return a.cn[(size_t)t * 32 + (tid & 31)];
        const size_t ci = (size_t)t * TB + (tid & (TB - 1));
        if constexpr (OP == OP_I8) return ((const int32_t *)a.cn)[2 * ci];  // little-endian low word
        else return a.hn[ci];
    };
    // FP8: each centroid's E8M0 scale 1 / s_c, loaded and stored raw the
    // same way, into words [256, 512) of the offset slot
    auto eload = [&](int t) -> int {
        if constexpr (OP == OP_FP8) return a.cexp[(size_t)t * TB + (tid & (TB - 1))];
        else return 0;
    };
    auto estore = [&](float *dst, int v) {
        if constexpr (OP == OP_FP8) ((int *)dst)[256 + tid] = v;
    };
    // every thread stores (slots past TB are never read): no branch, so the
    // compiler's wait tracking stays exact across the stage
    auto hstore = [&](float *dst, HRaw v) {
        if constexpr (OP == OP_I16) {
            // the tile's 32 offsets as f64 keys 32 cn + position (exact for
            // |cn| < 2^46); pad centroids (INT64_MAX) never win
            ((double *)dst)[tid & 31] = v == INT64_MAX ? (double)INFINITY : (double)v * 32.0 + (double)(tid & 31);
        } else if constexpr (OP == OP_I8) {
            // centroid c of the tile is value r of block nb for lanes of half
            // (c >> 2) & 1, with position nb * 16 + r in the fold
            const int c = tid & (TB - 1), cl = c & 31;
            const int pos = (c >> 5) * 16 + ((cl >> 3) << 2) + (cl & 3);
            // |cn| <= 2^26 on this path, so -cn << 5 fits in int32
            const uint32_t part = ((uint32_t)(-v) << 5) + (uint32_t)(31 - pos);
            dst[tid] = __int_as_float((int32_t)part);
        } else {
            dst[tid] = v + B;
        }
    };

    Score bs[FR][M];
    uint32_t bi[FR][M];
#pragma unroll
    for (int fr = 0; fr < FR; fr++)
#pragma unroll
        for (int t = 0; t < M; t++) {
            if constexpr (OP == OP_I8) bs[fr][t] = INT32_MIN;
            else if constexpr (OP == OP_I16) bs[fr][t] = INFINITY;
            else if constexpr (OP == OP_F32) bs[fr][t] = INFINITY;
            else bs[fr][t] = 0xFFFFFFFFu;
            bi[fr][t] = NO_LABEL;
        }

    // Tile loop over three LDS buffers: while tile t is multiplied and
    // folded, tiles t+1 and t+2 are in flight, so a DMA has two tiles of
    // compute to land. Each stage waits only for the older DMA (vmcnt counts
    // this wave's outstanding loads, which complete in order).
    using Acc = RegAcc<OP>;
    Acc acc[FR][NB];
    // vmcnt(PER_T + 1, FP8 + 2): the newest tile's DMA and offset (and
    // scale) loads may stay in flight
    constexpr int WAIT_ONE = (PER_T + (OP == OP_FP8 ? 2 : 1)) | 0x0F70;
    // Workgroup barrier that waits for LDS traffic only: __syncthreads()'s
    // release fence would also wait for the newest tile's DMA. The DMA each
    // wave needs next has landed by its own vmcnt wait before this point.
    auto tile_barrier = []() {
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
    };
    HRaw hv0 = hload(t0), hv1 = 0, hv2 = 0;
    int ev0 = eload(t0), ev1 = 0, ev2 = 0;
    dma(sC0, t0);
    dma(sC1, t0 + 1 < t1 ? t0 + 1 : t0);
    hv1 = hload(t0 + 1 < t1 ? t0 + 1 : t0);
    ev1 = eload(t0 + 1 < t1 ? t0 + 1 : t0);
    __builtin_amdgcn_s_waitcnt(WAIT_ONE);
    hstore(sH0, hv0);
    estore(sH0, ev0);
    __syncthreads();
    // One stage: compute tile T from (BC, BH); start tile T+2 into BN2;
    // publish tile T+1's offsets into BH1. The LDS arrays are named directly
    // so the compiler can tell the buffer being read from the ones in flight.
    // Every stage issues the same loads (past the end it re-fetches the last
    // tile into the free buffer), so the wait counts are the same each time.
#define EMH_STAGE(T, BC, BH, BN2, BH1, HV1, HV2, EV1, EV2)                         \
    {                                                                             \
        const int tn = (T) + 2 < t1 ? (T) + 2 : t1 - 1;                           \
        dma(BN2, tn);                                                             \
        HV2 = hload(tn);                                                          \
        EV2 = eload(tn);                                                          \
        int sa[NB];                                                               \
        _Pragma("unroll") for (int nb = 0; nb < NB; nb++)                         \
            sa[nb] = OP == OP_FP8 ? ((const int *)(BH))[256 + nb * 32 + j] : 127; \
        if constexpr (CAND && NB == 2) {                                          \
            reg_tile_cand2<OP, M, NS, FR>(BC, BH, xb, sb, sa, h, j, (uint32_t)((T) * CT), acc, bs, bi); \
        } else {                                                                  \
            reg_mfma<OP, NS, FR, NB>(BC, BH, xb, sb, sa, h, j, acc);              \
            reg_fold<OP, M, FR, NB>(acc, BH, h, (uint32_t)((T) * CT), a.shift, xs128, bs, bi); \
        }                                                                         \
        __builtin_amdgcn_sched_barrier(0); /* keep the wait after the compute */  \
        __builtin_amdgcn_s_waitcnt(WAIT_ONE);                                     \
        hstore(BH1, HV1);                                                         \
        estore(BH1, EV1);                                                         \
        tile_barrier();                                                           \
    }
    for (int t = t0; t < t1; t += 3) {
        EMH_STAGE(t, sC0, sH0, sC2, sH1, hv1, hv2, ev1, ev2);
        if (t + 1 >= t1) break;
        EMH_STAGE(t + 1, sC1, sH1, sC0, sH2, hv2, hv0, ev2, ev0);
        if (t + 2 >= t1) break;
        EMH_STAGE(t + 2, sC2, sH2, sC1, sH0, hv0, hv1, ev0, ev1);
    }
#undef EMH_STAGE

#pragma unroll
    for (int fr = 0; fr < FR; fr++) {
        const size_t r = row0 + (size_t)(wave * FR + fr) * 32 + j;
        if constexpr (CAND) {
#pragma unroll
            for (int t = 0; t < M; t++) {
                const uint32_t ps = (uint32_t)__shfl_xor((int)bs[fr][t], 32, 64);
                const uint32_t pi = (uint32_t)__shfl_xor((int)bi[fr][t], 32, 64);
                if (h == 0) topm_insert_lex<M, uint32_t>(bs[fr], bi[fr], ps, pi);
            }
            if (h == 0 && r < a.m) {
                if (a.nslice > 1) {
                    // keys then ids of this slice's top-m
                    uint32_t *pp = a.part + (r * a.nslice + slice) * (2 * M);
#pragma unroll
                    for (int t = 0; t < M; t++) {
                        pp[t] = bs[fr][t];
                        pp[M + t] = bi[fr][t];
                    }
                } else {
#pragma unroll
                    for (int t = 0; t < M; t++) a.cand[r * M + t] = bi[fr][t];
                }
            }
        } else {
            Score ps;
            if constexpr (OP == OP_I16)
                ps = __builtin_bit_cast(double, shfl_xor_i64(__builtin_bit_cast(int64_t, bs[fr][0]), 32));
            else ps = __shfl_xor(bs[fr][0], 32, 64);
            const uint32_t pi = (uint32_t)__shfl_xor((int)bi[fr][0], 32, 64);
            const bool better = OP == OP_I8 ? (ps > bs[fr][0] || (ps == bs[fr][0] && pi < bi[fr][0]))
                                            : (ps < bs[fr][0] || (ps == bs[fr][0] && pi < bi[fr][0]));
            if (better) {
                bs[fr][0] = ps;
                bi[fr][0] = pi;
            }
            if (h == 0 && r < a.m) {
                float dv;
                if constexpr (OP == OP_F32) {
                    // dist = |x|^2 + |c|^2 + bias - 2 x.c, clamped at 0 (src/cpu.rs:160-162)
                    const float v = __fadd_rn(__fmul_rn(2.0f, (float)bs[fr][0]), xn[fr]);
                    dv = v > 0.0f ? v : 0.0f;
                } else if constexpr (OP == OP_I16) {
                    // dist = ((|x|^2 + best * 4^-shift) as f64).max(0) as f32
                    // (src/cpu_i8.rs:183-186; the power-of-two scaling is exact)
                    const double v = __dadd_rn((double)(int64_t)xni[fr], ldexp((double)bs[fr][0], -2 * a.shift));
                    dv = (float)(v > 0.0 ? v : 0.0);
                } else {
                    // dist = ((|x|^2 + best) as f64).max(0) as f32 (src/cpu_i8.rs:183-186)
                    const double v = __dadd_rn((double)xni[fr], -(double)(int32_t)bs[fr][0]);
                    dv = (float)(v > 0.0 ? v : 0.0);
                }
                if (a.nslice > 1) {
                    // (score low word, index, distance, score high word) of this slice's best
                    uint32_t *pp = a.part + (r * a.nslice + slice) * 4;
                    uint32_t lo, hi = 0u;
                    if constexpr (OP == OP_I16) {
                        // the exact integer score as int64 (INT64_MAX: no centroid)
                        const int64_t s64 = bs[fr][0] < 9.2e18 ? (int64_t)bs[fr][0] : INT64_MAX;
                        lo = (uint32_t)(uint64_t)s64;
                        hi = (uint32_t)((uint64_t)s64 >> 32);
                    } else {
                        lo = __builtin_bit_cast(uint32_t, bs[fr][0]);  // f32 or int32 score bits
                    }
                    *(uint4 *)pp = make_uint4(lo, bi[fr][0], __float_as_uint(dv), hi);
                } else {
                    a.labels[r] = bi[fr][0] == NO_LABEL ? 0u : bi[fr][0];
                    a.dist[r] = dv;
                }
            }
        }
    }
}

// search_reg_core on this instantiation's own LDS
template <int OP, int ROWS, int M, int NS, int FR, int NB>
__device__ __forceinline__ void search_reg_body(const SearchArgs &a, const size_t row0, const int t0, const int t1,
                                                const int slice) {
    constexpr int TILE = 32 * NB * 32 * NS;
    __shared__ __align__(16) uint8_t sC0[TILE];
    __shared__ __align__(16) uint8_t sC1[TILE];
    __shared__ __align__(16) uint8_t sC2[TILE];
    __shared__ __align__(16) float sH0[reg_offset_slot<OP>()], sH1[reg_offset_slot<OP>()], sH2[reg_offset_slot<OP>()];
    __shared__ float sRed[4];
    search_reg_core<OP, ROWS, M, NS, FR, NB>(a, row0, t0, t1, slice, sC0, sC1, sC2, sH0, sH1, sH2, sRed);
}

template <int OP, int ROWS, int M, int NS, int FR, int NB>
__global__ __launch_bounds__(256) void search_reg_kernel(SearchArgs a) {
    const int ntiles = (int)(a.k_pad / reg_centroids<OP, NB>());
    const int s = blockIdx.y, t0 = s * a.tiles_per_slice;
    const int t1 = t0 + a.tiles_per_slice < ntiles ? t0 + a.tiles_per_slice : ntiles;
    search_reg_body<OP, ROWS, M, NS, FR, NB>(a, (size_t)blockIdx.x * (128 * FR), t0, t1, s);
}

// The same sweep compiled for two waves per SIMD (a register budget of 256;
// the compiler keeps what does not fit in scratch).
template <int OP, int ROWS, int M, int NS, int FR, int NB>
__global__ __launch_bounds__(256, 2) void search_reg_kernel_2w(SearchArgs a) {
    const int ntiles = (int)(a.k_pad / reg_centroids<OP, NB>());
    const int s = blockIdx.y, t0 = s * a.tiles_per_slice;
    const int t1 = t0 + a.tiles_per_slice < ntiles ? t0 + a.tiles_per_slice : ntiles;
    search_reg_body<OP, ROWS, M, NS, FR, NB>(a, (size_t)blockIdx.x * (128 * FR), t0, t1, s);
}

// Merges the per-slice results of a split search, one row per thread: the
// lexicographic (score, index) order of the unsplit sweep, so the merged
// labels, distances and candidate lists are the unsplit ones bit for bit.
template <int OP, int M>
__global__ void merge_slices_kernel(const uint32_t *part, size_t m, int nslice, uint32_t *labels, float *dist,
                                    uint32_t *cand) {
    const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    if constexpr (OP == OP_BF16 || OP == OP_FP8) {
        uint32_t bs[M], bi[M];
#pragma unroll
        for (int t = 0; t < M; t++) {
            bs[t] = 0xFFFFFFFFu;
            bi[t] = NO_LABEL;
        }
        for (int s = 0; s < nslice; s++) {
            const uint32_t *pp = part + (r * nslice + s) * (2 * M);
#pragma unroll
            for (int t = 0; t < M; t++) topm_insert_lex<M, uint32_t>(bs, bi, pp[t], pp[M + t]);
        }
#pragma unroll
        for (int t = 0; t < M; t++) cand[r * M + t] = bi[t];
    } else {
        uint4 best = *(const uint4 *)(part + r * nslice * 4);
        for (int s = 1; s < nslice; s++) {
            const uint4 p = *(const uint4 *)(part + (r * nslice + s) * 4);
            bool better;
            if constexpr (OP == OP_F32) {
                const float ps = __uint_as_float(p.x), bsv = __uint_as_float(best.x);
                better = ps < bsv || (ps == bsv && p.y < best.y);
            } else if constexpr (OP == OP_I16) {
                const int64_t ps = (int64_t)(((uint64_t)p.w << 32) | p.x);
                const int64_t bsv = (int64_t)(((uint64_t)best.w << 32) | best.x);
                better = ps < bsv || (ps == bsv && p.y < best.y);
            } else {
                const int32_t ps = (int32_t)p.x, bsv = (int32_t)best.x;
                better = ps > bsv || (ps == bsv && p.y < best.y);
            }
            if (better) best = p;
        }
        labels[r] = best.y == NO_LABEL ? 0u : best.y;
        dist[r] = __uint_as_float(best.z);
    }
}

// Many independent problems in one launch: workgroup b sweeps rows
// [wg[b].y & ~255, + 256) of problem wg[b].x against slice wg[b].y & 255 of
// its centroid tiles; the problem's arguments (row count of this step
// included) sit in a device table. Every row width runs in the same launch
// (the workgroup picks its instantiation from the problem's row bytes), so
// the jobs of all widths fill the chip together; the LDS is one block sized
// for the largest tile. Problems with fewer rows this step leave their spare
// workgroups idle.
constexpr int JOBS_TILE_MAX = 16384;  // 64 centroids x 256 B, or 32 x 512 B

template <int OP, int M, int NS>
__device__ __forceinline__ void jobs_sweep(const SearchArgs &a, size_t row0, int s, uint8_t (*sC)[JOBS_TILE_MAX],
                                           float (*sH)[reg_offset_slot<OP>()], float *sRed) {
    constexpr int FR = 2, NB = NS <= 8 ? 2 : 1;
    static_assert(32 * NB * 32 * NS <= JOBS_TILE_MAX, "tile");
    const int ntiles = (int)(a.k_pad / (32 * NB));
    const int t0 = s * a.tiles_per_slice, t1 = t0 + a.tiles_per_slice < ntiles ? t0 + a.tiles_per_slice : ntiles;
    search_reg_core<OP, EMH_ROWS_F32, M, NS, FR, NB>(a, row0, t0, t1, s, sC[0], sC[1], sC[2], sH[0], sH[1], sH[2],
                                                      sRed);
}

template <int OP, int M>
__global__ __launch_bounds__(256) void search_reg_jobs_kernel(const SearchArgs *jobs, const uint2 *wg) {
    __shared__ __align__(16) uint8_t sC[3][JOBS_TILE_MAX];
    __shared__ __align__(16) float sH[3][reg_offset_slot<OP>()];
    __shared__ float sRed[4];
    const uint2 w = wg[blockIdx.x];
    // one job per workgroup: its table pointers are uniform and global
    SearchArgs a = jobs[w.x];
    a.x = global_ptr_s(a.x);
    a.ids = global_ptr_s(a.ids);
    a.hn = global_ptr_s(a.hn);
    a.cn = global_ptr_s(a.cn);
    a.stats = global_ptr_s(a.stats);
    a.part = global_ptr_s(a.part);
    a.cand = global_ptr_s(a.cand);
    a.labels = global_ptr_s(a.labels);
    a.dist = global_ptr_s(a.dist);
    if constexpr (OP == OP_FP8) {
        a.x8 = global_ptr_s(a.x8);
        a.row_scale = global_ptr_s(a.row_scale);
        a.row_norm = global_ptr_s(a.row_norm);
        a.cexp = global_ptr_s(a.cexp);
    }
    const size_t row0 = w.y & ~255u;
    if (row0 >= a.m) return;  // whole workgroup: before any barrier
    const int s = (int)(w.y & 255u);  // the slice
    switch (__builtin_amdgcn_readfirstlane(a.row_bytes) / 32) {
        case 2: jobs_sweep<OP, M, 2>(a, row0, s, sC, sH, sRed); break;
        case 4: jobs_sweep<OP, M, 4>(a, row0, s, sC, sH, sRed); break;
        case 6: jobs_sweep<OP, M, 6>(a, row0, s, sC, sH, sRed); break;
        case 8: jobs_sweep<OP, M, 8>(a, row0, s, sC, sH, sRed); break;
        case 12: jobs_sweep<OP, M, 12>(a, row0, s, sC, sH, sRed); break;
        case 16: jobs_sweep<OP, M, 16>(a, row0, s, sC, sH, sRed); break;
        default: break;
    }
}

// Jobs whose operand rows are too wide for the register-resident sweep, in
// one launch of the LDS kernel: workgroup b sweeps rows [wg[b].y, + BM) of
// job wg[b].x against all its centroids (the job's row-major operand table),
// rows resident in LDS or streamed in d-chunks as the job's geometry says.
// The dynamic LDS is the largest such job's.
template <int OP, int M>
__global__ __launch_bounds__(256) void search_lds_jobs_kernel(const SearchArgs *jobs, const uint2 *wg) {
    const uint2 w = wg[blockIdx.x];
    SearchArgs a = jobs[w.x];
    a.x = global_ptr_s(a.x);
    a.ids = global_ptr_s(a.ids);
    a.cen_op = global_ptr_s(a.cen_op);
    a.hn = global_ptr_s(a.hn);
    a.cn = global_ptr_s(a.cn);
    a.stats = global_ptr_s(a.stats);
    a.cand = global_ptr_s(a.cand);
    a.labels = global_ptr_s(a.labels);
    a.dist = global_ptr_s(a.dist);
    if constexpr (OP == OP_FP8) {
        a.x8 = global_ptr_s(a.x8);
        a.row_scale = global_ptr_s(a.row_scale);
        a.row_norm = global_ptr_s(a.row_norm);
        a.cexp = global_ptr_s(a.cexp);
    }
    const size_t row0 = w.y;
    if (row0 >= a.m) return;  // whole workgroup: before any barrier
    constexpr int ROWS = OP == OP_I8 || OP == OP_I16 ? EMH_ROWS_I8 : EMH_ROWS_F32;
    if constexpr (OP == OP_I16) {
        search_lds_core<OP, ROWS, M, false>(a, row0);  // I16 streams d in chunks
    } else {
        if (__builtin_amdgcn_readfirstlane(a.resident)) search_lds_core<OP, ROWS, M, true>(a, row0);
        else search_lds_core<OP, ROWS, M, false>(a, row0);
    }
}

// ------------------------------------------------------------- launchers ---

inline size_t search_lds(const SearchGeom &g) {
    const size_t planes = g.op == OP_I16 ? I16_PLANES : 1;
    return (size_t)(BM + planes * BN) * (g.row_bytes + 16) + BN * 8;
}

template <int OP, int ROWS, int M, bool RES>
static hipError_t launch_one(const SearchGeom &g, const SearchArgs &a, hipStream_t s) {
    auto kern = search_kernel<OP, ROWS, M, RES>;
    size_t lds = search_lds(g);
    static size_t attr_bytes = 0;
    if (lds > 64 * 1024 && lds > attr_bytes) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr_bytes = lds;
    }
    dim3 grid((unsigned)((a.m + BM - 1) / BM));
    hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, a);
    return hipGetLastError();
}

template <int OP, int ROWS, int M>
static hipError_t launch_res(const SearchGeom &g, const SearchArgs &a, hipStream_t s) {
    return g.resident ? launch_one<OP, ROWS, M, true>(g, a, s) : launch_one<OP, ROWS, M, false>(g, a, s);
}

template <int OP, int ROWS>
static hipError_t launch_m(const SearchGeom &g, int M, const SearchArgs &a, hipStream_t s) {
    switch (M) {
        case 1: return launch_res<OP, ROWS, 1>(g, a, s);
        case 2: return launch_res<OP, ROWS, 2>(g, a, s);
        case 4: return launch_res<OP, ROWS, 4>(g, a, s);
        case 8: return launch_res<OP, ROWS, 8>(g, a, s);
        default: return hipErrorInvalidValue;
    }
}


// FR: 32-row fragments per wave (a workgroup sweeps 128 FR rows)
template <int OP, int ROWS, int M, int NS, int FR = 2>
static hipError_t launch_reg(const SearchArgs &a, hipStream_t s) {
    constexpr int NB = OP == OP_I16 ? I16_PLANES : (NS <= 8 ? 2 : 1);
    dim3 grid((unsigned)((a.m + 128 * FR - 1) / (128 * FR)), (unsigned)a.nslice);
    // I8 on 512-byte rows needs 269 registers at full speed; held to 256 (44
    // bytes per lane kept in scratch) it runs two waves per SIMD and measures
    // 16 % faster. The other shapes either fit two waves already or would
    // spill hundreds of bytes.
    if constexpr (OP == OP_I8 && NS == 16)
        hipLaunchKernelGGL((search_reg_kernel_2w<OP, ROWS, M, NS, FR, NB>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((search_reg_kernel<OP, ROWS, M, NS, FR, NB>), grid, dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || a.nslice == 1) return e;
    hipLaunchKernelGGL((merge_slices_kernel<OP, M>), dim3((unsigned)((a.m + 255) / 256)), dim3(256), 0, s,
                       (const uint32_t *)a.part, a.m, a.nslice, a.labels, a.dist, a.cand);
    return hipGetLastError();
}

template <int OP, int ROWS, int M, int FR = 2>
static hipError_t launch_reg_ns(const SearchGeom &g, const SearchArgs &a, hipStream_t s) {
    switch (g.row_bytes / 32) {
        case 2: return launch_reg<OP, ROWS, M, 2, FR>(a, s);
        case 4: return launch_reg<OP, ROWS, M, 4, FR>(a, s);
        case 6: return launch_reg<OP, ROWS, M, 6, FR>(a, s);
        case 8: return launch_reg<OP, ROWS, M, 8, FR>(a, s);
        case 12: return launch_reg<OP, ROWS, M, 12, FR>(a, s);
        case 16: return launch_reg<OP, ROWS, M, 16, FR>(a, s);
        default: return hipErrorInvalidValue;
    }
}

template <int OP, int ROWS, int FR = 2>
static hipError_t launch_reg_m(const SearchGeom &g, int M, const SearchArgs &a, hipStream_t s) {
    switch (M) {
        case 1: return launch_reg_ns<OP, ROWS, 1, FR>(g, a, s);
        case 2: return launch_reg_ns<OP, ROWS, 2, FR>(g, a, s);
        case 4: return launch_reg_ns<OP, ROWS, 4, FR>(g, a, s);
        case 8: return launch_reg_ns<OP, ROWS, 8, FR>(g, a, s);
        default: return hipErrorInvalidValue;
    }
}

// Per-precision instantiations, each in a translation unit of its own.
// fr: 32-row fragments per wave of the register-resident search (1 or 2).
hipError_t launch_reg_f32(const SearchGeom &g, int rows, int fr, const SearchArgs &a, hipStream_t s);
hipError_t launch_reg_bf16(const SearchGeom &g, int rows, int M, int fr, const SearchArgs &a, hipStream_t s);
hipError_t launch_reg_fp8(const SearchGeom &g, int M, int fr, const SearchArgs &a, hipStream_t s);
hipError_t launch_reg_i8(const SearchGeom &g, int fr, const SearchArgs &a, hipStream_t s);
hipError_t launch_reg_i16(const SearchGeom &g, const SearchArgs &a, hipStream_t s);
// the LDS kernel (rows resident in LDS or streamed in d-chunks), any op
hipError_t launch_lds(const SearchGeom &g, int rows, int M, const SearchArgs &a, hipStream_t s);
// the batched search of many problems in one launch
hipError_t launch_jobs_search_op(int op, int M, const void *search_args, const uint2 *wg, size_t nwg, hipStream_t s);
// the batched LDS-kernel search of the jobs with wide operand rows
hipError_t launch_jobs_search_lds_op(int op, int M, const void *search_args, const uint2 *wg, size_t nwg,
                                     size_t lds_bytes, hipStream_t s);

}  // namespace emh
