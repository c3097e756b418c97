// Reduced-precision flat scan on the matrix cores, fused with top-k, followed
// by an exact FP32 rerank of a shortlist (cuda_fp16_tc_search_batch and
// fv_flat_search).
//
// Reference: kernels/src/ivf_quant_lookup.cu:887-1058 (wmma 16x16x16, one
// block per 16 queries, FP32->FP16 query tiles reloaded per k-slice, serial
// insertion epilogue, no rerank).
//
// Design (gfx950, wave64):
//  * v_mfma_f32_16x16x32_{f16,bf16}: A = 16 database rows, B = 16 negated
//    queries, C = 0.5*|x|^2 of each row. D = 0.5|x|^2 - q.x, so the L2
//    distance is |q|^2 + 2D and the |x|^2 term costs nothing in the epilogue.
//    Masked and out-of-range rows get C = +inf and never pass a threshold.
//  * A workgroup owns 256 queries (4 waves x 64) and one chunk of the row
//    space. The queries' B fragments live in VGPRs for the whole kernel;
//    64-row tiles are double-buffered through LDS (rows padded so the 128-bit
//    fragment reads are conflict-free), one barrier per tile.
//  * The row space is either every row (the mask, if any, is tested per row)
//    or a compacted ascending list of the valid rows (sparse masks). Chunk
//    bounds are computed on the device from the actual row count.
//  * Selection: each lane takes the min of its 4 scores per (row tile, query
//    tile) and compares it with the query's threshold; the rare survivors are
//    appended to a per-(chunk, query) buffer in global memory. Near-full
//    buffers are pruned to the query's shortlist depth S, by the owning wave
//    (queries are wave-private) or, for deep shortlists, by the workgroup.
//    Thresholds combine the chunk's own S-th best (strict: rows arrive in
//    ascending order) with a per-query bound shared by all chunks through
//    global memory (non-strict: the S-th best of any chunk bounds the global
//    S-th best); the bound is read before the first tile and prefetched
//    every tile so its latency hides behind the MFMA work.
//  * Workgroup ids are mapped XCD-first so that the workgroups of one XCD
//    sweep the same chunk together and share it through that XCD's L2.
//  * flat_select_rerank_kernel: per query, the candidates of every chunk that
//    are within the final shared bound are streamed into an LDS top-S, which
//    is reranked exactly in FP32 (reference CPU path order) and sorted by
//    (distance, row); the first k are written.
#include "fv_flat_mfma.hpp"

namespace fv {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaThreads = 256;
constexpr int kQPerWave = 64;
constexpr int kQPerBlock = 4 * kQPerWave;
constexpr int kRowsPerTile = 64;
constexpr uint32_t kSelectBatch = 256;
constexpr uint32_t kStage = 512;  // staged survivors per wave (two query tiles' worth)

template <int DT>
struct MfmaOps;

template <>
struct MfmaOps<FV_DTYPE_F16> {
    using vec = half8_t;
    static __device__ __forceinline__ floatx4 mma(vec a, vec b, floatx4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ uint16_t from_f32(float f) {
        _Float16 h = (_Float16)f;  // round to nearest even
        return __builtin_bit_cast(uint16_t, h);
    }
};

template <>
struct MfmaOps<FV_DTYPE_BF16> {
    using vec = bf16x8_t;
    static __device__ __forceinline__ floatx4 mma(vec a, vec b, floatx4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ uint16_t from_f32(float f) {
        uint32_t u = __float_as_uint(f);
        u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
        return (uint16_t)(u >> 16);
    }
};

// Smallest float above f (finite f): a non-strict bound tested with the same
// strict compare as the chunk's own threshold.
__device__ __forceinline__ float next_up(float f) {
    uint32_t u = __float_as_uint(f);
    if (f == 0.0f) return __uint_as_float(1u);
    return __uint_as_float((u & 0x80000000u) ? u - 1u : u + 1u);
}

__device__ __forceinline__ float bound_from_ord(uint32_t g) {
    if (g == 0xFFFFFFFFu) return kAdmit;
    float f = unord_f32(g);
    return f < kAdmit ? next_up(f) : kAdmit;
}

// Shortlist depth of one query: factor * k (or the fixed shortlist), at least
// k, at most s_max.
__device__ __forceinline__ void query_k_s(const FlatSearchArgs& a, uint32_t q, uint32_t& k,
                                          uint32_t& S) {
    k = a.d_top_k ? a.d_top_k[q] : a.top_k;
    if (k > a.top_k) k = a.top_k;
    if (k == 0) k = 1;
    S = a.factor ? a.factor * k : (a.shortlist ? a.shortlist : 8u * k);
    if (S < k) S = k;
    if (S > a.s_max) S = a.s_max;
}

// BOUND = true is the first pass: same tiles and MMAs, but the epilogue only
// keeps a running minimum per (lane, element) — 16 disjoint row subsets per
// query per chunk — written to `minima`; rows are limited to the first
// `bound_rows` positions. No thresholds, buffers or prunes.
template <int DT, int METRIC, int KS, int R, bool BOUND>
__global__ __launch_bounds__(kMfmaThreads, 2) void flat_mfma_scan_kernel(
    const uint16_t* __restrict__ X, uint32_t row_stride, const float* __restrict__ norms,
    const uint16_t* __restrict__ Qh, uint32_t nq, FlatSearchArgs a, uint32_t n_chunks,
    uint32_t qblocks, uint32_t cap, const uint32_t* __restrict__ rows,
    const uint32_t* __restrict__ nrows, uint32_t n, uint32_t* __restrict__ gthr,
    uint32_t* __restrict__ counts, uint64_t* __restrict__ cand, float* __restrict__ minima,
    uint32_t bound_rows) {
    using Ops = MfmaOps<DT>;
    using vec = typename Ops::vec;
    constexpr int kRowBytes = KS * 64;          // 32*KS 16-bit elements
    constexpr int kLdsRow = kRowBytes + 16;     // padded LDS row
    constexpr int kSegBytes = kRowBytes / 4;    // bytes one thread stages per row

    __shared__ __attribute__((aligned(16))) uint8_t s_tile[2][kRowsPerTile * kLdsRow];
    __shared__ __attribute__((aligned(16))) float s_cinit[2][kRowsPerTile];
    __shared__ __attribute__((aligned(16))) float s_scale[2][kRowsPerTile];
    __shared__ uint32_t s_rowid[2][kRowsPerTile];
    __shared__ uint32_t s_cnt[kQPerBlock];
    __shared__ uint32_t s_keep[kQPerBlock];
    __shared__ float s_thr[kQPerBlock];
    __shared__ uint32_t s_need[kQPerBlock / 32];
    extern __shared__ uint64_t s_scratch[];

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int l16 = lane & 15;
    const int grp = lane >> 4;
    const uint32_t dim = a.dim;

    // XCD-first work mapping: consecutive block ids land on different XCDs;
    // within an XCD the chunk varies slowest.
    const uint32_t b = blockIdx.x;
    uint32_t chunk, qblock;
    if ((n_chunks & 7u) == 0u) {
        const uint32_t cpx = n_chunks >> 3;
        const uint32_t xcd = b & 7u, i = b >> 3;
        chunk = xcd * cpx + i / qblocks;
        qblock = i % qblocks;
    } else {
        chunk = b / qblocks;
        qblock = b % qblocks;
    }
    // Chunk bounds in the (possibly compacted) row space.
    uint32_t n_eff = nrows ? *nrows : n;
    if (BOUND && n_eff > bound_rows) n_eff = bound_rows;
    uint32_t rpc = (n_eff + n_chunks - 1) / n_chunks;
    rpc = (rpc + kRowsPerTile - 1) / kRowsPerTile * kRowsPerTile;
    const uint32_t r_begin = min(n_eff, chunk * rpc);
    const uint32_t r_end = min(n_eff, r_begin + rpc);
    const uint32_t qb0 = qblock * kQPerBlock;
    const uint32_t qw0 = qb0 + wave * kQPerWave;
    constexpr bool kBlockPrune = R == 0;
    uint64_t* my_scratch = kBlockPrune ? s_scratch : s_scratch + (size_t)wave * cap;

    // B fragments: lane (l16, grp) holds query qw0 + 16j + l16, dims
    // 8*KS*grp + 8s + t for step s and t < 8 (the same permutation of K is
    // used for the A fragments, so every dim is paired exactly once).
    // Queries arrive pre-encoded (negated, 16-bit, rows of 32*KS elements,
    // padded to whole 256-query blocks), so each fragment is one 128-bit load.
    vec bfrag[4][KS];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint16_t* qrow = Qh + (size_t)(qw0 + 16 * j + l16) * (32 * KS) + 8 * KS * grp;
#pragma unroll
        for (int s = 0; s < KS; ++s) bfrag[j][s] = *reinterpret_cast<const vec*>(qrow + 8 * s);
    }
    // Retire the fragment loads here: otherwise the wait-count tracking keeps
    // them "in flight" around the tile loop and every tile waits for its own
    // prefetches (vmcnt(0)) before its last MMAs. 0x0F70 = vmcnt(0).
    __builtin_amdgcn_s_waitcnt(0x0F70);
    // Per-query state; thread t owns query qb0 + t for bookkeeping.
    const uint32_t my_q = qb0 + tid;
    const bool my_live = my_q < nq;
    {
        uint32_t k, S;
        if (my_live) query_k_s(a, my_q, k, S);
        else S = 1;
        s_keep[tid] = S;
        s_cnt[tid] = 0;
        s_thr[tid] = my_live
            ? bound_from_ord(__hip_atomic_load(&gthr[my_q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            : kAdmit;
    }

    const uint32_t ntiles = (r_end - r_begin + kRowsPerTile - 1) / kRowsPerTile;
    const int st_row = tid >> 2;   // staging: 4 threads per row
    const int st_seg = tid & 3;
    // Staging registers of the next tile. Nothing loaded here is used until
    // store_tile, after this tile's MMAs, so the loads overlap the compute.
    uint4 pre[KS];
    uint32_t pre_row = 0, pre_mword = 0xFFFFFFFFu;
    float pre_norm = 0.0f;
    bool pre_in = false;
    // Compacted row space: row ids are fetched one tile ahead of the data.
    // All prefetch loads are unconditional from clamped, valid addresses (rows
    // past the chunk end get C = +inf, so their data never matters): no
    // exec-masked loads, so no conservative waits around them.
    uint32_t rid_next = 0;
    auto fetch_rid = [&](uint32_t p0) {
        const uint32_t p = p0 + st_row;
        rid_next = rows ? rows[p < r_end ? p : r_end - 1] : p;
    };
    auto load_tile = [&](uint32_t p0) {
        const uint32_t p = p0 + st_row;
        pre_in = p < r_end;
        pre_row = pre_in ? rid_next : (rows ? rows[r_begin] : r_begin);
        const uint16_t* src = X + (size_t)pre_row * row_stride;
#pragma unroll
        for (int i = 0; i < KS; ++i) {
            const uint32_t e = (uint32_t)(st_seg * (kSegBytes / 2) + i * 8);
            if (e + 8 <= row_stride) pre[i] = *reinterpret_cast<const uint4*>(src + e);
            else pre[i] = make_uint4(0, 0, 0, 0);
        }
        if (METRIC != METRIC_INNER_PRODUCT) pre_norm = norms[pre_row];
        pre_mword = (a.mask && !rows) ? a.mask[pre_row >> 5] : 0xFFFFFFFFu;
    };
    auto store_tile = [&](int buf) {
        uint8_t* dst = &s_tile[buf][st_row * kLdsRow + st_seg * kSegBytes];
#pragma unroll
        for (int i = 0; i < KS; ++i) *reinterpret_cast<uint4*>(dst + 16 * i) = pre[i];
        if (st_seg == 0) {
            const bool ok = pre_in && ((pre_mword >> (pre_row & 31u)) & 1u);
            float c = __builtin_inff(), sc = 1.0f;
            if (ok) {
                if (METRIC == METRIC_L2) c = 0.5f * pre_norm;
                else c = 0.0f;
                if (METRIC == METRIC_COSINE) sc = pre_norm > 0.0f ? 1.0f / sqrtf(pre_norm) : 0.0f;
            }
            s_cinit[buf][st_row] = c;
            s_scale[buf][st_row] = sc;
            s_rowid[buf][st_row] = pre_row;
        }
    };

    // Wave-private staging list of survivors, flushed into the per-query
    // buffers by all 64 lanes at once.
    // Wave mode: the staging list aliases the wave's prune scratch (never live
    // at the same time). Workgroup mode: it follows the shared scratch.
    uint64_t* st_key = kBlockPrune ? s_scratch + cap + (size_t)wave * (kStage + kStage / 8)
                                   : my_scratch;
    uint8_t* st_q = reinterpret_cast<uint8_t*>(st_key + kStage);
    uint32_t nstage = 0;
    auto flush_stage = [&]() {
        wave_lds_sync();
        for (uint32_t e = lane; e < nstage; e += kWave) {
            const uint32_t ql = wave * kQPerWave + st_q[e];
            const uint32_t pos = atomicAdd(&s_cnt[ql], 1u);
            cand[((size_t)chunk * nq + qb0 + ql) * cap + pos] = st_key[e];
        }
        nstage = 0;
        wave_lds_sync();
    };

    if (ntiles > 0) {
        fetch_rid(r_begin);
        load_tile(r_begin);
        if (ntiles > 1) fetch_rid(r_begin + kRowsPerTile);
        store_tile(0);
    }
    __syncthreads();

    floatx4 runmin[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) runmin[j] = floatx4{kAdmit, kAdmit, kAdmit, kAdmit};

    for (uint32_t t = 0; t < ntiles; ++t) {
        const int buf = t & 1;
        const uint32_t p0 = r_begin + t * kRowsPerTile;
        if (t + 1 < ntiles) {
            load_tile(p0 + kRowsPerTile);
            if (t + 2 < ntiles) fetch_rid(p0 + 2 * kRowsPerTile);
        }
        // Prefetch the shared bound; it is consumed after this tile's MMAs.
        const uint32_t gpre =
            __hip_atomic_load(&gthr[my_live ? my_q : 0u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

        float thr[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) thr[j] = s_thr[wave * kQPerWave + 16 * j + l16];

#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const uint8_t* arow = &s_tile[buf][(rt * 16 + l16) * kLdsRow + grp * (16 * KS)];
            vec afrag[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s) afrag[s] = *reinterpret_cast<const vec*>(arow + 16 * s);
            const floatx4 cin = *reinterpret_cast<const floatx4*>(&s_cinit[buf][rt * 16 + 4 * grp]);
            floatx4 acc[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j] = cin;
            // Four independent accumulation chains, interleaved.
#pragma unroll
            for (int s = 0; s < KS; ++s)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j] = Ops::mma(afrag[s], bfrag[j][s], acc[j]);
            if (METRIC == METRIC_COSINE) {
                const floatx4 scl = *reinterpret_cast<const floatx4*>(&s_scale[buf][rt * 16 + 4 * grp]);
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j] = acc[j] * scl;
            }
            if constexpr (BOUND) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i) runmin[j][i] = fminf(runmin[j][i], acc[j][i]);
                continue;
            }
            // Survivors are compacted wave-wide into the staging list (ballot +
            // mbcnt), so a slot costs one compare unless some lane inserts.
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool live_q = qw0 + 16 * j + l16 < nq;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const bool f = live_q && acc[j][i] < thr[j];
                    const uint64_t bm = __ballot(f);
                    if (bm) {
                        const uint32_t idx = nstage + __builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
                        if (f) {
                            st_key[idx] = make_key(acc[j][i], s_rowid[buf][rt * 16 + 4 * grp + i]);
                            st_q[idx] = (uint8_t)(16 * j + l16);
                        }
                        nstage += (uint32_t)__popcll(bm);
                    }
                }
                // One query tile stages at most 4 * 64 survivors.
                if (nstage > kStage - 4 * kWave) flush_stage();
            }
        }
        if (!BOUND) flush_stage();
        if (t + 1 < ntiles) store_tile(buf ^ 1);
        __syncthreads();
        if (BOUND) continue;

        if (!kBlockPrune) {
            // Wave-private prunes of buffers that could overflow on the next tile.
            const uint32_t c = my_live ? s_cnt[tid] : 0u;
            uint64_t need = __ballot(my_live && c + kRowsPerTile > cap);
            while (need) {
                const int bl = __builtin_ctzll(need);
                need &= need - 1;
                const uint32_t qlb = wave * kQPerWave + bl;
                const uint32_t qgb = qb0 + qlb;
                const uint32_t keep = s_keep[qlb];
                uint64_t* bq = cand + ((size_t)chunk * nq + qgb) * cap;
                uint32_t t;
                uint32_t kept = wave_bisect_prune<(R > 0 ? R : 1)>(bq, s_cnt[qlb], keep, lane, &t);
                if (kept + kRowsPerTile > cap) {
                    // Heavy ties at the bound: exact (distance, row) selection.
                    kept = wave_prune(bq, kept, keep, my_scratch, lane);
                    t = kept == keep ? (uint32_t)(my_scratch[keep - 1] >> 32) : 0xFFFFFFFFu;
                }
                if (lane == 0) {
                    s_cnt[qlb] = kept;
                    if (t != 0xFFFFFFFFu) {
                        s_thr[qlb] = fminf(s_thr[qlb], unord_f32(t));
                        atomicMin(&gthr[qgb], t);
                    }
                }
                wave_lds_sync();
            }
        } else {
            // Workgroup prunes, one query at a time, through shared scratch.
            if (tid < kQPerBlock / 32) s_need[tid] = 0;
            __syncthreads();
            if (my_live && s_cnt[tid] + kRowsPerTile > cap) atomicOr(&s_need[tid >> 5], 1u << (tid & 31));
            __syncthreads();
            for (int w = 0; w < kQPerBlock / 32; ++w) {
                uint32_t bits = s_need[w];
                while (bits) {
                    const int bl = __builtin_ctz(bits);
                    bits &= bits - 1;
                    const uint32_t ql = w * 32 + bl;
                    const uint32_t qg = qb0 + ql;
                    const uint32_t c = s_cnt[ql];
                    const uint32_t keep = s_keep[ql];
                    uint64_t* bq = cand + ((size_t)chunk * nq + qg) * cap;
                    const uint32_t np = next_pow2(c < 2u ? 2u : c);
                    for (uint32_t i = tid; i < np; i += kMfmaThreads) my_scratch[i] = i < c ? bq[i] : kEmptyKey;
                    __syncthreads();
                    block_bitonic_sort(my_scratch, np);
                    const uint32_t kept = c < keep ? c : keep;
                    for (uint32_t i = tid; i < kept; i += kMfmaThreads) bq[i] = my_scratch[i];
                    if (tid == 0) {
                        s_cnt[ql] = kept;
                        if (kept == keep) {
                            const float th = key_dist(my_scratch[keep - 1]);
                            s_thr[ql] = fminf(s_thr[ql], th);
                            atomicMin(&gthr[qg], ord_f32(th));
                        }
                    }
                    __syncthreads();
                }
            }
        }
        // Tighten thresholds with the prefetched shared bound. Each wave only
        // reads its own queries' thresholds, so a wave-level sync suffices.
        if (my_live) s_thr[tid] = fminf(s_thr[tid], bound_from_ord(gpre));
        wave_lds_sync();
    }
    if (BOUND) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t qg = qw0 + 16 * j + l16;
            if (qg < nq) {
                float* mo = minima + ((size_t)chunk * nq + qg) * 16 + 4 * grp;
#pragma unroll
                for (int i = 0; i < 4; ++i) mo[i] = runmin[j][i];
            }
        }
        return;
    }
    if (my_live) counts[(size_t)chunk * nq + my_q] = s_cnt[tid];
}

// Per query: the S-th smallest of the first pass's subset minima. Each minimum
// is the score of a distinct row, so at least S rows score at or below it:
// it bounds the global S-th best and seeds the shared bound of the main pass.
__global__ __launch_bounds__(256) void bound_select_kernel(const float* __restrict__ minima,
                                                           uint32_t m, uint32_t nq,
                                                           FlatSearchArgs a,
                                                           uint32_t* __restrict__ gthr) {
    extern __shared__ uint64_t s_b[];
    const uint32_t q = blockIdx.x;
    const uint32_t np = next_pow2(m);
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        uint64_t key = kEmptyKey;
        if (i < m) {
            const uint32_t c = i >> 4, e = i & 15u;
            const float v = minima[((size_t)c * nq + q) * 16 + e];
            if (v < kAdmit) key = make_key(v, i);
        }
        s_b[i] = key;
    }
    __syncthreads();
    block_bitonic_sort(s_b, np);
    if (threadIdx.x == 0) {
        uint32_t k, S;
        query_k_s(a, q, k, S);
        if (S <= m && s_b[S - 1] != kEmptyKey) atomicMin(&gthr[q], (uint32_t)(s_b[S - 1] >> 32));
    }
}

// One workgroup per query: streams the candidates of every chunk that are
// within the final shared bound into an LDS top-S, reranks them exactly in
// FP32 and writes the top k by (distance, row).
template <int METRIC>
__global__ __launch_bounds__(256) void flat_select_rerank_kernel(
    const uint64_t* __restrict__ cand, const uint32_t* __restrict__ counts, uint32_t n_chunks,
    uint32_t nq, uint32_t cap, const uint32_t* __restrict__ gthr, FlatSearchArgs a,
    uint32_t lcap, const float* __restrict__ X32, const float* __restrict__ Qm,
    GpuSearchResult* __restrict__ out) {
    extern __shared__ uint64_t s_k[];
    __shared__ uint32_t s_n;
    __shared__ uint32_t s_g;
    const uint32_t q = blockIdx.x;
    uint32_t k, S;
    query_k_s(a, q, k, S);
    if (threadIdx.x == 0) {
        s_n = 0;
        s_g = gthr[q];
    }
    __syncthreads();
    for (uint32_t c = 0; c < n_chunks; ++c) {
        const uint32_t cnt = counts[(size_t)c * nq + q];
        const uint64_t* bq = cand + ((size_t)c * nq + q) * cap;
        for (uint32_t e0 = 0; e0 < cnt; e0 += kSelectBatch) {
            const uint32_t e = e0 + threadIdx.x;
            const uint32_t g = s_g;
            if (e < cnt) {
                const uint64_t key = bq[e];
                if ((uint32_t)(key >> 32) <= g) s_k[atomicAdd(&s_n, 1u)] = key;
            }
            __syncthreads();
            if (s_n + kSelectBatch > lcap) {
                const uint32_t n = s_n;
                const uint32_t np = next_pow2(n);
                for (uint32_t i = n + threadIdx.x; i < np; i += blockDim.x) s_k[i] = kEmptyKey;
                __syncthreads();
                block_bitonic_sort(s_k, np);
                if (threadIdx.x == 0) {
                    s_n = n < S ? n : S;
                    if (n >= S) s_g = min(s_g, (uint32_t)(s_k[S - 1] >> 32));
                }
                __syncthreads();
            }
        }
    }
    const uint32_t n = s_n;
    const uint32_t np = next_pow2(n < 2u ? 2u : n);
    for (uint32_t i = n + threadIdx.x; i < np; i += blockDim.x) s_k[i] = kEmptyKey;
    __syncthreads();
    block_bitonic_sort(s_k, np);
    const uint32_t have = n < S ? n : S;
    const float* qv = Qm + (size_t)q * a.dim;
    for (uint32_t i = threadIdx.x; i < have; i += blockDim.x) {
        const uint64_t key = s_k[i];
        const uint32_t r = key_slot(key);
        float d;
        if (X32) {
            d = exact_distance_vec<METRIC>(qv, X32 + (size_t)r * a.dim, a.dim);
        } else {
            // No FP32 rows: report the reduced-precision score as a distance.
            const float s = key_dist(key);
            if (METRIC == METRIC_INNER_PRODUCT) {
                d = s;
            } else {
                float qn = 0.0f;
                for (uint32_t dd = 0; dd < a.dim; ++dd) qn = __builtin_fmaf(qv[dd], qv[dd], qn);
                d = METRIC == METRIC_L2 ? fmaxf(0.0f, qn + 2.0f * s) : 1.0f + s / (sqrtf(qn) + 1e-8f);
            }
        }
        s_k[i] = d < kAdmit ? make_key(d, r) : kEmptyKey;
    }
    const uint32_t hp = next_pow2(have < 2u ? 2u : have);
    for (uint32_t i = have + threadIdx.x; i < hp; i += blockDim.x) s_k[i] = kEmptyKey;
    __syncthreads();
    block_bitonic_sort(s_k, hp);
    for (uint32_t i = threadIdx.x; i < a.top_k; i += blockDim.x) {
        const uint64_t key = i < k && i < hp ? s_k[i] : kEmptyKey;
        GpuSearchResult r;
        if (key == kEmptyKey) {
            r.id = FV_EMPTY_ID;
            r.distance = FV_EMPTY_DIST;
        } else {
            r.id = key_slot(key);
            r.distance = key_dist(key);
        }
        out[(size_t)q * a.top_k + i] = r;
    }
}

// Negated queries as 16-bit codes, rows of `pdq` elements, zero padded.
template <int DT>
__global__ void encode_queries_kernel(const float* __restrict__ Qm, uint32_t nq, uint32_t dim,
                                      uint32_t pdq, uint32_t total_rows, uint16_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)total_rows * pdq) return;
    const uint32_t q = (uint32_t)(i / pdq), d = (uint32_t)(i % pdq);
    const float v = (q < nq && d < dim) ? -Qm[(size_t)q * dim + d] : 0.0f;
    out[i] = MfmaOps<DT>::from_f32(v);
}

// ---- mask compaction: ascending list of the rows whose bit is set ---------

__global__ void mask_block_counts_kernel(const uint32_t* __restrict__ mask, uint32_t nwords,
                                         uint32_t n, uint32_t* __restrict__ block_counts) {
    __shared__ uint32_t s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = 0;
    if (w < nwords) {
        uint32_t bits = mask[w];
        if ((w + 1) * 32u > n) bits &= (n - w * 32u) >= 32u ? 0xFFFFFFFFu : ((1u << (n - w * 32u)) - 1u);
        c = __popc(bits);
    }
    atomicAdd(&s, c);
    __syncthreads();
    if (threadIdx.x == 0) block_counts[blockIdx.x] = s;
}

__global__ void mask_scan_kernel(uint32_t* __restrict__ block_counts, uint32_t nblocks,
                                 uint32_t* __restrict__ total) {
    // One workgroup: exclusive scan of the per-block counts, in place.
    __shared__ uint32_t s_part[256];
    const uint32_t per = (nblocks + 255) / 256;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t i = 0; i < per && b0 + i < nblocks; ++i) sum += block_counts[b0 + i];
    s_part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < 256; ++i) {
            const uint32_t v = s_part[i];
            s_part[i] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    uint32_t run = s_part[threadIdx.x];
    for (uint32_t i = 0; i < per && b0 + i < nblocks; ++i) {
        const uint32_t v = block_counts[b0 + i];
        block_counts[b0 + i] = run;
        run += v;
    }
}

__global__ void mask_scatter_kernel(const uint32_t* __restrict__ mask, uint32_t nwords,
                                    uint32_t n, const uint32_t* __restrict__ block_offsets,
                                    uint32_t* __restrict__ rows) {
    __shared__ uint32_t s_c[256];
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t bits = 0;
    if (w < nwords) {
        bits = mask[w];
        if ((w + 1) * 32u > n) bits &= (n - w * 32u) >= 32u ? 0xFFFFFFFFu : ((1u << (n - w * 32u)) - 1u);
    }
    s_c[threadIdx.x] = __popc(bits);
    __syncthreads();
    // Inclusive Hillis-Steele scan within the block.
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? s_c[threadIdx.x - off] : 0u;
        __syncthreads();
        s_c[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t pos = block_offsets[blockIdx.x] + s_c[threadIdx.x] - __popc(bits);
    while (bits) {
        const int b = __builtin_ctz(bits);
        bits &= bits - 1;
        rows[pos++] = w * 32u + (uint32_t)b;
    }
}

// ---- planning and launch -----------------------------------------------------

static uint32_t deepest_shortlist(const FlatSearchArgs& a) {
    uint32_t smax = a.factor ? a.factor * a.top_k : (a.shortlist ? a.shortlist : 8u * a.top_k);
    if (smax < a.top_k) smax = a.top_k;
    if (smax > FV_MAX_SHORTLIST) smax = FV_MAX_SHORTLIST;
    return smax;
}

FlatMfmaPlan plan_flat_mfma(uint32_t n, uint32_t padded_dim, uint32_t nq, const FlatSearchArgs& a,
                            bool has_mask, uint32_t valid_count) {
    FlatMfmaPlan p{};
    const uint32_t smax = deepest_shortlist(a);
    p.s_max = smax;
    // Buffers hold S plus one tile of appends plus one wavefront of bisection slack.
    // A 1024-key buffer absorbs the whole warm-up of a typical chunk, so
    // most (query, chunk) pairs prune at most once.
    if (smax + 2 * kRowsPerTile <= 1024u) {
        p.cap = 1024u;
        p.prune_regs = 16;
    } else {
        p.cap = next_pow2(smax + 2 * kRowsPerTile);
        p.prune_regs = 0;  // workgroup prunes through shared scratch
    }
    p.block_prune = p.prune_regs == 0;
    // Per wave: prune scratch and staging list share one region; workgroup
    // prunes use one shared scratch followed by four staging lists.
    const size_t stage_bytes = (size_t)kStage * 9;
    p.lds_scan = p.block_prune ? (size_t)p.cap * 8 + 4 * (kStage + kStage / 8) * 8
                               : 4 * ((size_t)p.cap * 8 > stage_bytes ? (size_t)p.cap * 8 : stage_bytes);
    p.lcap = next_pow2(smax + kSelectBatch);
    const uint32_t ks = (padded_dim + 31) / 32;
    p.ks = ks <= 4 ? ks : (ks <= 6 ? 6 : 8);
    p.qblocks = (nq + kQPerBlock - 1) / kQPerBlock;
    p.compact = has_mask && valid_count > 0 && (uint64_t)valid_count * 2 <= n;
    const uint32_t rows_plan = p.compact ? valid_count : n;
    // Chunks: enough workgroups to cover 256 CUs several times, but >= 8192
    // rows each (every chunk pays a cold start of about S*ln(rows/S) inserts
    // before its threshold settles), a multiple of 8 (one slice per XCD) when
    // there are at least 8.
    const uint32_t want = (2560u + p.qblocks - 1) / p.qblocks;
    const uint32_t max_by_rows = (rows_plan + 8191u) / 8192u;
    uint32_t c = want < max_by_rows ? want : max_by_rows;
    if (c > 256u) c = 256u;
    if (c >= 8u) c &= ~7u;
    if (c < 1u) c = 1u;
    p.n_chunks = c;
    // First pass: subset minima over the first bound_rows positions, 64 chunks
    // (1024 minima per query) so the S-th smallest is a tight bound.
    const uint32_t brows = rows_plan / 4 > 262144u ? rows_plan / 4 : (rows_plan < 262144u ? rows_plan : 262144u);
    p.bound_rows = brows;
    p.bound_chunks = (brows >= 64u * kRowsPerTile && 16u * 64u >= smax) ? 64u : 0u;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    p.off_gthr = take((size_t)nq * 4);
    p.off_counts = take((size_t)p.n_chunks * nq * 4);
    p.off_cand = take((size_t)p.n_chunks * nq * p.cap * 8);
    p.off_qh = take((size_t)p.qblocks * kQPerBlock * 32 * p.ks * 2);
    if (p.bound_chunks) p.off_minima = take((size_t)p.bound_chunks * nq * 16 * 4);
    if (p.compact) {
        const uint32_t nwords = (n + 31) / 32;
        p.mask_blocks = (nwords + 255) / 256;
        p.off_rows = take((size_t)n * 4);
        p.off_nrows = take(4);
        p.off_blocks = take((size_t)p.mask_blocks * 4);
    }
    p.workspace_bytes = off;
    return p;
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct ScanLaunch {
    const FlatMfmaPlan* p;
    const uint16_t* X;
    uint32_t stride;
    const float* norms;
    const uint16_t* Qh;
    uint32_t nq;
    FlatSearchArgs a;
    const uint32_t* rows;
    const uint32_t* nrows;
    uint32_t n;
    uint32_t* gthr;
    uint32_t* counts;
    uint64_t* cand;
    float* minima;
    hipStream_t stream;
};

template <int DT, int METRIC, int KS, int R>
static void launch_scan_r(const ScanLaunch& L) {
    const FlatMfmaPlan& p = *L.p;
    const dim3 g(p.qblocks * p.n_chunks);
    flat_mfma_scan_kernel<DT, METRIC, KS, R, false><<<g, kMfmaThreads, p.lds_scan, L.stream>>>(
        L.X, L.stride, L.norms, L.Qh, L.nq, L.a, p.n_chunks, p.qblocks, p.cap, L.rows, L.nrows,
        L.n, L.gthr, L.counts, L.cand, nullptr, 0u);
}

template <int DT, int METRIC, int KS>
static void launch_scan(const ScanLaunch& L) {
    const FlatMfmaPlan& p = *L.p;
    if (p.bound_chunks) {
        const dim3 g(p.qblocks * p.bound_chunks);
        flat_mfma_scan_kernel<DT, METRIC, KS, 16, true><<<g, kMfmaThreads, 0, L.stream>>>(
            L.X, L.stride, L.norms, L.Qh, L.nq, L.a, p.bound_chunks, p.qblocks, p.cap, L.rows,
            L.nrows, L.n, L.gthr, L.counts, L.cand, L.minima, p.bound_rows);
        const uint32_t m = p.bound_chunks * 16;
        bound_select_kernel<<<L.nq, 256, (size_t)next_pow2(m) * 8, L.stream>>>(
            L.minima, m, L.nq, L.a, L.gthr);
    }
    // The kernel timer brackets the main pass only.
    void* timer = timer_begin(L.stream);
    if (L.p->prune_regs == 16) launch_scan_r<DT, METRIC, KS, 16>(L);
    else launch_scan_r<DT, METRIC, KS, 0>(L);
    timer_end(timer, L.stream);
}

template <int DT, int METRIC>
static hipError_t dispatch_ks(const ScanLaunch& L) {
    switch (L.p->ks) {
        case 1: launch_scan<DT, METRIC, 1>(L); break;
        case 2: launch_scan<DT, METRIC, 2>(L); break;
        case 3: launch_scan<DT, METRIC, 3>(L); break;
        case 4: launch_scan<DT, METRIC, 4>(L); break;
        case 6: launch_scan<DT, METRIC, 6>(L); break;
        case 8: launch_scan<DT, METRIC, 8>(L); break;
        default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

template <int DT>
static hipError_t dispatch_metric(const ScanLaunch& L, int metric) {
    if (metric == METRIC_L2) return dispatch_ks<DT, METRIC_L2>(L);
    if (metric == METRIC_INNER_PRODUCT) return dispatch_ks<DT, METRIC_INNER_PRODUCT>(L);
    return dispatch_ks<DT, METRIC_COSINE>(L);
}

hipError_t launch_flat_mfma(const FvFlatTables* t, const float* Qm, uint32_t nq,
                            const FlatSearchArgs& a, GpuSearchResult* out, void* workspace,
                            const FlatMfmaPlan& p, hipStream_t stream) {
    char* ws = static_cast<char*>(workspace);
    uint32_t* gthr = reinterpret_cast<uint32_t*>(ws + p.off_gthr);
    uint32_t* counts = reinterpret_cast<uint32_t*>(ws + p.off_counts);
    uint64_t* cand = reinterpret_cast<uint64_t*>(ws + p.off_cand);
    const uint32_t* rows = nullptr;
    const uint32_t* nrows = nullptr;
    fill_u32_kernel<<<(nq + 255) / 256, 256, 0, stream>>>(gthr, nq, 0xFFFFFFFFu);
    if (p.compact) {
        uint32_t* r = reinterpret_cast<uint32_t*>(ws + p.off_rows);
        uint32_t* nr = reinterpret_cast<uint32_t*>(ws + p.off_nrows);
        uint32_t* blk = reinterpret_cast<uint32_t*>(ws + p.off_blocks);
        const uint32_t nwords = (t->num_vectors + 31) / 32;
        mask_block_counts_kernel<<<p.mask_blocks, 256, 0, stream>>>(a.mask, nwords, t->num_vectors, blk);
        mask_scan_kernel<<<1, 256, 0, stream>>>(blk, p.mask_blocks, nr);
        mask_scatter_kernel<<<p.mask_blocks, 256, 0, stream>>>(a.mask, nwords, t->num_vectors, blk, r);
        rows = r;
        nrows = nr;
    }
    uint16_t* Qh = reinterpret_cast<uint16_t*>(ws + p.off_qh);
    {
        const uint32_t qrows = p.qblocks * kQPerBlock, pdq = 32 * p.ks;
        const size_t total = (size_t)qrows * pdq;
        const unsigned blocks = (unsigned)((total + 255) / 256);
        if (t->dtype == FV_DTYPE_F16)
            encode_queries_kernel<FV_DTYPE_F16><<<blocks, 256, 0, stream>>>(Qm, nq, t->dim, pdq, qrows, Qh);
        else
            encode_queries_kernel<FV_DTYPE_BF16><<<blocks, 256, 0, stream>>>(Qm, nq, t->dim, pdq, qrows, Qh);
    }
    float* minima = p.bound_chunks ? reinterpret_cast<float*>(ws + p.off_minima) : nullptr;
    ScanLaunch L{&p, static_cast<const uint16_t*>(t->d_codes), t->padded_dim, t->d_norms_sq, Qh, nq,
                 a, rows, nrows, t->num_vectors, gthr, counts, cand, minima, stream};
    hipError_t e = hipErrorInvalidValue;
    if (t->dtype == FV_DTYPE_F16) e = dispatch_metric<FV_DTYPE_F16>(L, a.metric);
    else if (t->dtype == FV_DTYPE_BF16) e = dispatch_metric<FV_DTYPE_BF16>(L, a.metric);
    if (e != hipSuccess) return e;
    const size_t lds = (size_t)p.lcap * sizeof(uint64_t);
    if (a.metric == METRIC_L2)
        flat_select_rerank_kernel<METRIC_L2><<<nq, 256, lds, stream>>>(cand, counts, p.n_chunks, nq, p.cap, gthr, a, p.lcap, t->d_vectors, Qm, out);
    else if (a.metric == METRIC_INNER_PRODUCT)
        flat_select_rerank_kernel<METRIC_INNER_PRODUCT><<<nq, 256, lds, stream>>>(cand, counts, p.n_chunks, nq, p.cap, gthr, a, p.lcap, t->d_vectors, Qm, out);
    else
        flat_select_rerank_kernel<METRIC_COSINE><<<nq, 256, lds, stream>>>(cand, counts, p.n_chunks, nq, p.cap, gthr, a, p.lcap, t->d_vectors, Qm, out);
    return hipGetLastError();
}

static int check_tables(const FvFlatTables* t) {
    if (!t->d_codes || !t->d_norms_sq) return -1;
    if (t->padded_dim == 0 || t->padded_dim % 16 || t->padded_dim > 256 || t->dim == 0 ||
        t->dim > t->padded_dim)
        return (int)hipErrorInvalidValue;
    if (t->dtype != FV_DTYPE_F16 && t->dtype != FV_DTYPE_BF16) return (int)hipErrorInvalidValue;
    return 0;
}

static bool make_args(const FvFlatTables* t, const FvFlatSearchParams* sp, FlatSearchArgs& a) {
    a = FlatSearchArgs{};
    a.mask = sp->d_valid_mask;
    a.d_top_k = sp->d_top_k;
    a.top_k = sp->top_k;
    a.factor = sp->shortlist_factor;
    a.shortlist = sp->shortlist;
    a.metric = (int)sp->metric;
    a.dim = t->dim;
    if (a.top_k == 0 || a.top_k > FV_MAX_TOPK) return false;
    if (sp->metric > METRIC_INNER_PRODUCT || sp->reserved != 0) return false;
    a.s_max = deepest_shortlist(a);
    return true;
}

}  // namespace fv

extern "C" {

size_t fv_flat_workspace_size(const FvFlatTables* tables, uint32_t num_queries,
                              const FvFlatSearchParams* params) {
    if (!tables || !params || tables->num_vectors == 0) return 0;
    fv::FlatSearchArgs a;
    if (!fv::make_args(tables, params, a)) return 0;
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(tables->num_vectors, tables->padded_dim, num_queries,
                                            a, params->d_valid_mask != nullptr,
                                            params->valid_count);
    if (p.lcap > 8192u) return 0;
    return p.workspace_bytes;
}

int fv_flat_search(const FvFlatTables* tables, const float* d_queries, uint32_t num_queries,
                   const FvFlatSearchParams* params, GpuSearchResult* d_out_results,
                   void* d_workspace, size_t workspace_bytes, void* stream) {
    if (!tables || !d_queries || !d_out_results || !params) return -1;
    const int tc = fv::check_tables(tables);
    if (tc) return tc;
    fv::FlatSearchArgs a;
    if (!fv::make_args(tables, params, a)) return (int)hipErrorInvalidValue;
    if (num_queries == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (tables->num_vectors == 0)
        return (int)fv::fill_results(d_out_results, (size_t)num_queries * a.top_k, FV_EMPTY_ID,
                                     0x7149f2cau /* 1e30f */, s);
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(tables->num_vectors, tables->padded_dim, num_queries,
                                            a, params->d_valid_mask != nullptr,
                                            params->valid_count);
    if (p.lcap > 8192u) return (int)hipErrorInvalidValue;  // shortlist deeper than 7936
    if (!d_workspace || workspace_bytes < p.workspace_bytes) return (int)hipErrorInvalidValue;
    hipError_t e = fv::launch_flat_mfma(tables, d_queries, num_queries, a, d_out_results,
                                        d_workspace, p, s);
    return e == hipSuccess ? 0 : (int)e;
}

// cuda_bridge.h:138-147 / kernels/src/ivf_quant_lookup.cu:1106-1123.
int cuda_fp16_tc_search_batch(const Fp16GpuTables* tables, const float* d_queries,
                              uint32_t num_queries, uint32_t top_k, MetricType metric,
                              const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                              void* stream) {
    if (!tables || !d_queries || !d_out_results) return -1;
    if (num_queries == 0 || top_k == 0 || tables->num_vectors == 0 || tables->padded_dim == 0)
        return 0;  // ivf_quant_lookup.cu:1047-1049: silent no-op
    if (top_k > 256u) top_k = 256u;
    FvFlatTables t{};
    t.d_codes = tables->d_vectors_fp16;
    t.d_norms_sq = tables->d_vec_norms_sq;
    t.d_vectors = tables->d_vectors;
    t.num_vectors = tables->num_vectors;
    t.dim = tables->dim;
    t.padded_dim = tables->padded_dim;
    t.dtype = FV_DTYPE_F16;
    const int tc = fv::check_tables(&t);
    if (tc) return tc;
    FvFlatSearchParams sp{};
    sp.d_valid_mask = d_valid_mask;
    sp.metric = (uint32_t)metric;
    sp.top_k = top_k;
    fv::FlatSearchArgs a;
    if (!fv::make_args(&t, &sp, a)) return (int)hipErrorInvalidValue;
    hipStream_t s = static_cast<hipStream_t>(stream);
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(t.num_vectors, t.padded_dim, num_queries, a,
                                            d_valid_mask != nullptr, 0);
    void* ws = nullptr;
    hipError_t e = hipMallocAsync(&ws, p.workspace_bytes, s);
    if (e != hipSuccess) return (int)e;
    e = fv::launch_flat_mfma(&t, d_queries, num_queries, a, d_out_results, ws, p, s);
    hipError_t e2 = hipFreeAsync(ws, s);
    if (e != hipSuccess) return (int)e;
    return e2 == hipSuccess ? 0 : (int)e2;
}

}  // extern "C"
