// Reduced-precision flat search on the matrix cores (cuda_fp16_tc_search_batch
// and fv_flat_search): plan, query encoding, optional mask compaction, the
// scan launch (kernel in fv_flat_mfma_scan.hpp), and the select/rerank that
// turns the scan's per-chunk candidates into the final top-k.
//
// Reference: kernels/src/ivf_quant_lookup.cu:887-1058, 1106-1123.
//
//  * flat_select_rerank_kernel: per query, the candidates of every chunk that
//    are within the final shared bound are streamed into an LDS top-S, which
//    is reranked exactly in FP32 (reference CPU path order) and sorted by
//    (distance, row); the first k are written. With FV_SEARCH_F64_BOUNDARY,
//    candidates whose FP32 distance lies within twice the FP32 rounding-error
//    bound of a neighbour in that order are re-scored in FP64 and reordered by
//    (FP64 distance, row), so the order of the returned rows, and which rows
//    make the top k, follow the exact distances.
#include "fv_flat_mfma_scan.hpp"

namespace fv {

constexpr uint32_t kSelectBatch = 256;

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
    __shared__ uint32_t s_emax;
    const bool f64b = (a.flags & FV_SEARCH_F64_BOUNDARY) && X32 != nullptr;
    if (f64b) {
        // Largest FP32 error bound among the reranked candidates (positive
        // floats order like their bit patterns).
        float e = 0.0f;
        for (uint32_t i = threadIdx.x; i < have; i += blockDim.x) {
            const uint64_t key = s_k[i];
            if (key == kEmptyKey) continue;
            e = fmaxf(e, exact_distance_err<METRIC>(qv, X32 + (size_t)key_slot(key) * a.dim,
                                                    a.dim, key_dist(key)));
        }
        if (threadIdx.x == 0) s_emax = 0u;
        __syncthreads();
        atomicMax(&s_emax, __float_as_uint(e));
    }
    const uint32_t hp = next_pow2(have < 2u ? 2u : have);
    for (uint32_t i = have + threadIdx.x; i < hp; i += blockDim.x) s_k[i] = kEmptyKey;
    __syncthreads();
    block_bitonic_sort(s_k, hp);
    if (f64b && have > 1) {
        // Two candidates adjacent in FP32 order whose gap is at most twice the
        // error bound may be in either exact order: chains of such gaps form
        // clusters, which are re-scored in FP64 and reordered by (FP64
        // distance, row). Outside a cluster the FP32 order is the exact order,
        // so only clusters that start before position k are touched.
        double* s_d64 = reinterpret_cast<double*>(s_k + lcap);
        uint16_t* s_pos = reinterpret_cast<uint16_t*>(s_d64 + lcap);
        const float band = 2.0000002f * __uint_as_float(s_emax);
        auto near_next = [&](uint32_t i) -> bool {
            if (i + 1 >= have) return false;
            const uint64_t k0 = s_k[i], k1 = s_k[i + 1];
            if (k0 == kEmptyKey || k1 == kEmptyKey) return false;
            return key_dist(k1) - key_dist(k0) <= band;
        };
        for (uint32_t i = threadIdx.x; i < have; i += blockDim.x) {
            uint32_t st = i;
            while (st > 0 && near_next(st - 1)) --st;
            const bool in_cluster = (i > 0 && near_next(i - 1)) || near_next(i);
            s_pos[i] = (in_cluster && st < k) ? 1 : 0;
            if (in_cluster && st < k)
                s_d64[i] = exact_distance_f64<METRIC>(qv, X32 + (size_t)key_slot(s_k[i]) * a.dim, a.dim);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < have; i += blockDim.x) {
            if (s_pos[i] == 0) continue;
            uint32_t st = i, en = i;
            while (st > 0 && near_next(st - 1)) --st;
            while (near_next(en)) ++en;
            const double di = s_d64[i];
            const uint32_t ri = key_slot(s_k[i]);
            uint32_t rank = 0;
            for (uint32_t j = st; j <= en; ++j) {
                const double dj = s_d64[j];
                const uint32_t rj = key_slot(s_k[j]);
                rank += (dj < di || (dj == di && rj < ri)) ? 1u : 0u;
            }
            s_pos[i] = (uint16_t)(st + rank + 1);  // 0 marks "not moved"
        }
        __syncthreads();
        // New keys (FP64 distance rounded to FP32) staged in the FP64 slots.
        for (uint32_t i = threadIdx.x; i < have; i += blockDim.x)
            if (s_pos[i] != 0)
                reinterpret_cast<uint64_t*>(s_d64)[i] = make_key((float)s_d64[i], key_slot(s_k[i]));
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < have; i += blockDim.x)
            if (s_pos[i] != 0) s_k[s_pos[i] - 1u] = reinterpret_cast<const uint64_t*>(s_d64)[i];
        __syncthreads();
    }
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

// Negated queries as 16-bit codes (round to nearest even), rows of `pdq`
// elements, zero padded.
template <int DT>
__global__ void encode_queries_kernel(const float* __restrict__ Qm, uint32_t nq, uint32_t dim,
                                      uint32_t pdq, uint32_t total_rows, uint16_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)total_rows * pdq) return;
    const uint32_t q = (uint32_t)(i / pdq), d = (uint32_t)(i % pdq);
    const float v = (q < nq && d < dim) ? -Qm[(size_t)q * dim + d] : 0.0f;
    uint16_t h;
    if (DT == FV_DTYPE_F16) {
        const _Float16 f = (_Float16)v;
        h = __builtin_bit_cast(uint16_t, f);
    } else {
        uint32_t u = __float_as_uint(v);
        u += 0x7fffu + ((u >> 16) & 1u);
        h = (uint16_t)(u >> 16);
    }
    out[i] = h;
}

// FP8 e4m3 rows with one FP32 scale per row: s = max|x| / 448 (1 for an all
// zero row), code = e4m3(sign * x / s), zero padded to out_stride bytes. One
// wavefront per row, 4 codes per lane per store. Rows >= n (query block
// padding) are written as zeros with scale 1.
__global__ __launch_bounds__(256) void encode_fp8_rows_kernel(
    const float* __restrict__ x, uint32_t n, uint32_t dim, uint32_t out_stride,
    uint32_t total_rows, float sign, uint8_t* __restrict__ out, float* __restrict__ scales) {
    const uint32_t r = blockIdx.x * 4 + threadIdx.x / kWave;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    if (r >= total_rows) return;
    const float* row = x + (size_t)r * dim;
    float m = 0.0f;
    if (r < n)
        for (uint32_t d = lane; d < dim; d += kWave) m = fmaxf(m, fabsf(row[d]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    const float s = m > 0.0f ? m / 448.0f : 1.0f;
    uint32_t* o = reinterpret_cast<uint32_t*>(out + (size_t)r * out_stride);
    for (uint32_t d0 = 4 * lane; d0 < out_stride; d0 += 4 * kWave) {
        uint32_t w = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint32_t d = d0 + c;
            const uint8_t code = (r < n && d < dim) ? f32_to_e4m3(sign * row[d] / s) : 0u;
            w |= (uint32_t)code << (8 * c);
        }
        o[d0 / 4] = w;
    }
    if (lane == 0 && scales) scales[r] = s;
}

void launch_encode_fp8_rows(const float* x, uint32_t n, uint32_t dim, uint32_t out_stride,
                            uint32_t total_rows, float sign, uint8_t* out, float* scales,
                            hipStream_t stream) {
    encode_fp8_rows_kernel<<<(total_rows + 3) / 4, 256, 0, stream>>>(x, n, dim, out_stride,
                                                                    total_rows, sign, out, scales);
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

// K steps of one row: 32 elements (16-bit) or 128 (FP8) per step, rounded up
// to the instantiated counts. 0 = unsupported width.
static uint32_t plan_ks(uint32_t padded_dim, int dtype) {
    if (dtype == FV_DTYPE_FP8_E4M3) {
        const uint32_t ks = (padded_dim + 127) / 128;
        return ks <= 2 ? ks : (ks <= 4 ? 4 : (ks <= 6 ? 6 : 0));
    }
    const uint32_t ks = (padded_dim + 31) / 32;
    if (ks <= 4) return ks;
    if (ks <= 6) return 6;
    if (ks <= 8) return 8;
    if (ks <= 12) return 12;
    if (ks <= 16) return 16;
    return ks <= 24 ? 24 : 0;
}

FlatMfmaPlan plan_flat_mfma(uint32_t n, uint32_t padded_dim, int dtype, uint32_t nq,
                            const FlatSearchArgs& a, bool has_mask, uint32_t valid_count) {
    FlatMfmaPlan p{};
    const uint32_t smax = deepest_shortlist(a);
    p.s_max = smax;
    // Buffers hold S plus one tile of appends plus one wavefront of bisection
    // slack (tiles have at most 64 rows). A 1024-key buffer absorbs the whole
    // warm-up of a typical chunk, so most (query, chunk) pairs prune at most once.
    if (smax + 2 * 64u <= 1024u) {
        p.cap = 1024u;
        p.prune_regs = 16;
    } else {
        p.cap = next_pow2(smax + 2 * 64u);
        p.prune_regs = 0;  // workgroup prunes through shared scratch
    }
    p.block_prune = p.prune_regs == 0;
    // Per wave: prune scratch and staging list share one region; workgroup
    // prunes use one shared scratch followed by four staging lists.
    const size_t stage_bytes = (size_t)kStage * 9;
    p.lds_scan = p.block_prune ? (size_t)p.cap * 8 + 4 * (kStage + kStage / 8) * 8
                               : 4 * ((size_t)p.cap * 8 > stage_bytes ? (size_t)p.cap * 8 : stage_bytes);
    p.lcap = next_pow2(smax + kSelectBatch);
    p.lds_select = (size_t)p.lcap * 8;
    if (a.flags & FV_SEARCH_F64_BOUNDARY) p.lds_select += (size_t)p.lcap * (8 + 2);
    p.ks = plan_ks(padded_dim, dtype);
    if (p.ks == 0) return p;
    p.qw = (uint32_t)scan_qw(dtype, p.ks);
    p.qb = 4 * p.qw;
    p.qrow_bytes = (uint32_t)scan_row_bytes(dtype, p.ks);
    p.qblocks = (nq + p.qb - 1) / p.qb;
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
    p.bound_chunks = (brows >= 64u * 64u && 16u * 64u >= smax) ? 64u : 0u;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    p.off_gthr = take((size_t)nq * 4);
    p.off_counts = take((size_t)p.n_chunks * nq * 4);
    p.off_cand = take((size_t)p.n_chunks * nq * p.cap * 8);
    p.off_qh = take((size_t)p.qblocks * p.qb * p.qrow_bytes);
    if (dtype == FV_DTYPE_FP8_E4M3) p.off_qscale = take((size_t)p.qblocks * p.qb * 4);
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

void launch_bound_select(const ScanLaunch& L) {
    const uint32_t m = L.p->bound_chunks * 16;
    bound_select_kernel<<<L.nq, 256, (size_t)next_pow2(m) * 8, L.stream>>>(L.minima, m, L.nq, L.a,
                                                                          L.gthr);
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
    uint8_t* Qh = reinterpret_cast<uint8_t*>(ws + p.off_qh);
    float* qscale = nullptr;
    const uint32_t qrows = p.qblocks * p.qb;
    if (t->dtype == FV_DTYPE_FP8_E4M3) {
        qscale = reinterpret_cast<float*>(ws + p.off_qscale);
        launch_encode_fp8_rows(Qm, nq, t->dim, p.qrow_bytes, qrows, -1.0f, Qh, qscale, stream);
    } else {
        const uint32_t pdq = p.qrow_bytes / 2;
        const size_t total = (size_t)qrows * pdq;
        const unsigned blocks = (unsigned)((total + 255) / 256);
        uint16_t* q16 = reinterpret_cast<uint16_t*>(Qh);
        if (t->dtype == FV_DTYPE_F16)
            encode_queries_kernel<FV_DTYPE_F16><<<blocks, 256, 0, stream>>>(Qm, nq, t->dim, pdq, qrows, q16);
        else
            encode_queries_kernel<FV_DTYPE_BF16><<<blocks, 256, 0, stream>>>(Qm, nq, t->dim, pdq, qrows, q16);
    }
    float* minima = p.bound_chunks ? reinterpret_cast<float*>(ws + p.off_minima) : nullptr;
    const uint32_t esz = t->dtype == FV_DTYPE_FP8_E4M3 ? 1u : 2u;
    ScanLaunch L{&p, static_cast<const uint8_t*>(t->d_codes), t->padded_dim * esz, t->d_norms_sq,
                 t->d_row_scales, Qh, qscale, nq, a, rows, nrows, t->num_vectors, gthr, counts,
                 cand, minima, stream};
    hipError_t e = hipErrorInvalidValue;
    const bool large = p.ks > 8;
    if (t->dtype == FV_DTYPE_F16) e = large ? scan_f16_large(L) : scan_f16_small(L);
    else if (t->dtype == FV_DTYPE_BF16) e = large ? scan_bf16_large(L) : scan_bf16_small(L);
    else if (t->dtype == FV_DTYPE_FP8_E4M3) e = scan_fp8(L);
    if (e != hipSuccess) return e;
    const size_t lds = p.lds_select;
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
    if (t->dtype == FV_DTYPE_FP8_E4M3 && !t->d_row_scales) return -1;
    if (t->dtype != FV_DTYPE_F16 && t->dtype != FV_DTYPE_BF16 && t->dtype != FV_DTYPE_FP8_E4M3)
        return (int)hipErrorInvalidValue;
    if (t->padded_dim == 0 || t->padded_dim % 16 || t->padded_dim > 768 || t->dim == 0 ||
        t->dim > t->padded_dim)
        return (int)hipErrorInvalidValue;
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
    a.flags = sp->flags;
    if (a.top_k == 0 || a.top_k > FV_MAX_TOPK) return false;
    if (sp->metric > METRIC_INNER_PRODUCT || (sp->flags & ~FV_SEARCH_F64_BOUNDARY) != 0) return false;
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
    if (fv::check_tables(tables) != 0) return 0;
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(tables->num_vectors, tables->padded_dim,
                                            (int)tables->dtype, num_queries, a,
                                            params->d_valid_mask != nullptr, params->valid_count);
    if (p.lcap > 8192u || p.ks == 0) return 0;
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
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(tables->num_vectors, tables->padded_dim,
                                            (int)tables->dtype, num_queries, a,
                                            params->d_valid_mask != nullptr, params->valid_count);
    if (p.lcap > 8192u || p.ks == 0) return (int)hipErrorInvalidValue;  // shortlist deeper than 7936
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
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(t.num_vectors, t.padded_dim, FV_DTYPE_F16, num_queries,
                                            a, d_valid_mask != nullptr, 0);
    void* ws = nullptr;
    hipError_t e = hipMallocAsync(&ws, p.workspace_bytes, s);
    if (e != hipSuccess) return (int)e;
    e = fv::launch_flat_mfma(&t, d_queries, num_queries, a, d_out_results, ws, p, s);
    hipError_t e2 = hipFreeAsync(ws, s);
    if (e != hipSuccess) return (int)e;
    return e2 == hipSuccess ? 0 : (int)e2;
}

}  // extern "C"
