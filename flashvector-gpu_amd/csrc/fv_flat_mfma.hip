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
//  * A workgroup owns 256 queries (4 waves x 64) and one chunk of rows. The
//    queries' B fragments live in VGPRs for the whole kernel; 64-row database
//    tiles are double-buffered through LDS (rows padded to dodge bank
//    conflicts), one barrier per tile.
//  * Selection: each lane reduces its 4 scores per (row tile, query tile) with
//    min and compares against the query's threshold; the rare survivors are
//    appended to a per-(chunk, query) buffer in global memory. Buffers near
//    capacity are pruned by their owning wave (queries are wave-private, so
//    prunes need no workgroup barrier). Thresholds combine the chunk's own
//    S-th best (strict: rows arrive in ascending order) with a per-query bound
//    shared by all chunks in global memory (non-strict: the S-th best of any
//    chunk bounds the global S-th best).
//  * Workgroup ids are mapped XCD-first so that the workgroups of one XCD
//    sweep the same chunk together and share it through that XCD's L2.
//  * A second kernel merges the sorted per-chunk lists of each query with a
//    bitonic merge tree in LDS, reranks the S best exactly in FP32 (reference
//    CPU path order) and writes the top k by (distance, row).
#include "fv_flat_mfma.hpp"

namespace fv {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaThreads = 256;
constexpr int kQPerWave = 64;
constexpr int kQPerBlock = 4 * kQPerWave;
constexpr int kRowsPerTile = 64;

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

__device__ __forceinline__ float next_up(float f) {
    // Smallest float above f for finite non-negative-or-negative f: lets a
    // non-strict bound be tested with the same strict compare.
    uint32_t u = __float_as_uint(f);
    if (f == 0.0f) return __uint_as_float(1u);
    return __uint_as_float((u & 0x80000000u) ? u - 1u : u + 1u);
}

__device__ __forceinline__ void atomic_min_ord(uint32_t* p, float f) {
    atomicMin(p, ord_f32(f));
}

// Per-row epilogue constant and scale of one tile: C init and post-MMA scale.
template <int METRIC>
__device__ __forceinline__ void row_terms(const float* __restrict__ norms, const uint32_t* mask,
                                          uint32_t r, uint32_t r_end, float& cinit,
                                          float& scale) {
    bool ok = r < r_end && mask_allows(mask, r);
    scale = 1.0f;
    if (!ok) {
        cinit = __builtin_inff();
        return;
    }
    if (METRIC == METRIC_L2) {
        cinit = 0.5f * norms[r];
    } else {
        cinit = 0.0f;
        if (METRIC == METRIC_COSINE) {
            float nx = norms[r];
            scale = nx > 0.0f ? 1.0f / sqrtf(nx) : 0.0f;
        }
    }
}

template <int DT, int METRIC, int KS>
__global__ __launch_bounds__(kMfmaThreads, 2) void flat_mfma_scan_kernel(
    const uint16_t* __restrict__ X, uint32_t n, uint32_t row_stride, uint32_t dim,
    const float* __restrict__ norms, const float* __restrict__ Qm, uint32_t nq,
    const uint32_t* __restrict__ mask, uint32_t keep, uint32_t cap, uint32_t rows_per_chunk,
    uint32_t n_chunks, uint32_t qblocks, uint32_t* __restrict__ gthr,
    uint64_t* __restrict__ cand, uint64_t* __restrict__ partial) {
    using Ops = MfmaOps<DT>;
    using vec = typename Ops::vec;
    constexpr int kRowBytes = KS * 64;          // 32*KS 16-bit elements
    constexpr int kLdsRow = kRowBytes + 16;     // padded LDS row
    constexpr int kSegBytes = kRowBytes / 4;    // bytes one thread stages per row

    __shared__ __attribute__((aligned(16))) uint8_t s_tile[2][kRowsPerTile * kLdsRow];
    __shared__ __attribute__((aligned(16))) float s_cinit[2][kRowsPerTile];
    __shared__ __attribute__((aligned(16))) float s_scale[2][kRowsPerTile];
    __shared__ uint32_t s_cnt[kQPerBlock];
    __shared__ float s_thr[kQPerBlock];
    extern __shared__ uint64_t s_scratch[];

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int l16 = lane & 15;
    const int grp = lane >> 4;

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
    const uint32_t r_begin = chunk * rows_per_chunk;
    const uint32_t r_end = min(n, r_begin + rows_per_chunk);
    const uint32_t qw0 = qblock * kQPerBlock + wave * kQPerWave;  // first query of this wave
    uint64_t* my_scratch = s_scratch + (size_t)wave * cap;

    // B fragments: lane (l16, grp) holds query qw0 + 16j + l16, dims
    // 8*KS*grp + 8s + t for step s and t < 8 (the same permutation of K is
    // used for the A fragments, so every dim is paired exactly once).
    vec bfrag[4][KS];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t qg = qw0 + 16 * j + l16;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            uint16_t h[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                uint32_t d = 8u * KS * grp + 8u * s + t;
                float v = (qg < nq && d < dim) ? -Qm[(size_t)qg * dim + d] : 0.0f;
                h[t] = Ops::from_f32(v);
            }
            vec tmp;
            __builtin_memcpy(&tmp, h, 16);
            bfrag[j][s] = tmp;
        }
    }
    for (int i = tid; i < kQPerBlock; i += kMfmaThreads) {
        s_cnt[i] = 0;
        s_thr[i] = kAdmit;
    }

    const uint32_t ntiles = r_end > r_begin ? (r_end - r_begin + kRowsPerTile - 1) / kRowsPerTile : 0;
    const int st_row = tid >> 2;   // staging: 4 threads per row
    const int st_seg = tid & 3;
    uint4 pre[KS];

    auto load_tile = [&](uint32_t r0) {
        uint32_t r = r0 + st_row;
        const uint16_t* src = X + (size_t)r * row_stride;
#pragma unroll
        for (int i = 0; i < KS; ++i) {
            uint32_t e = (uint32_t)(st_seg * (kSegBytes / 2) + i * 8);
            if (r < r_end && e + 8 <= row_stride) pre[i] = *reinterpret_cast<const uint4*>(src + e);
            else pre[i] = make_uint4(0, 0, 0, 0);
        }
    };
    auto store_tile = [&](int buf, uint32_t r0) {
        uint8_t* dst = &s_tile[buf][st_row * kLdsRow + st_seg * kSegBytes];
#pragma unroll
        for (int i = 0; i < KS; ++i) *reinterpret_cast<uint4*>(dst + 16 * i) = pre[i];
        if (tid < kRowsPerTile) {
            float c, sc;
            row_terms<METRIC>(norms, mask, r0 + tid, r_end, c, sc);
            s_cinit[buf][tid] = c;
            s_scale[buf][tid] = sc;
        }
    };

    if (ntiles > 0) {
        load_tile(r_begin);
        store_tile(0, r_begin);
    }
    __syncthreads();

    for (uint32_t t = 0; t < ntiles; ++t) {
        const int buf = t & 1;
        const uint32_t r0 = r_begin + t * kRowsPerTile;
        if (t + 1 < ntiles) load_tile(r0 + kRowsPerTile);

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
            floatx4 scl;
            if (METRIC == METRIC_COSINE)
                scl = *reinterpret_cast<const floatx4*>(&s_scale[buf][rt * 16 + 4 * grp]);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                floatx4 acc = cin;
#pragma unroll
                for (int s = 0; s < KS; ++s) acc = Ops::mma(afrag[s], bfrag[j][s], acc);
                if (METRIC == METRIC_COSINE) acc = acc * scl;
                float m = fminf(fminf(acc[0], acc[1]), fminf(acc[2], acc[3]));
                if (m < thr[j]) {
                    const uint32_t ql = wave * kQPerWave + 16 * j + l16;
                    const uint32_t qg = qblock * kQPerBlock + ql;
                    if (qg < nq) {
                        uint64_t* buf_q = cand + ((size_t)chunk * nq + qg) * cap;
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            if (acc[i] < thr[j]) {
                                uint32_t pos = atomicAdd(&s_cnt[ql], 1u);
                                buf_q[pos] = make_key(acc[i], r0 + rt * 16 + 4 * grp + i);
                            }
                        }
                    }
                }
            }
        }
        if (t + 1 < ntiles) store_tile(buf ^ 1, r0 + kRowsPerTile);
        __syncthreads();

        // Wave-private prunes of buffers that could overflow on the next tile,
        // then refresh thresholds from the bound shared across chunks.
        {
            const uint32_t ql = wave * kQPerWave + lane;
            const uint32_t qg = qblock * kQPerBlock + ql;
            const bool live = qg < nq;
            const uint32_t c = live ? s_cnt[ql] : 0u;
            uint64_t need = __ballot(live && c + kRowsPerTile > cap);
            while (need) {
                const int bl = __builtin_ctzll(need);
                need &= need - 1;
                const uint32_t qlb = wave * kQPerWave + bl;
                const uint32_t qgb = qblock * kQPerBlock + qlb;
                const uint32_t cb = s_cnt[qlb];
                uint64_t* buf_q = cand + ((size_t)chunk * nq + qgb) * cap;
                const uint32_t kept = wave_prune(buf_q, cb, keep, my_scratch, lane);
                if (lane == 0) {
                    s_cnt[qlb] = kept;
                    if (kept == keep) {
                        float th = key_dist(my_scratch[keep - 1]);
                        s_thr[qlb] = fminf(s_thr[qlb], th);
                        atomic_min_ord(&gthr[qgb], th);
                    }
                }
                wave_lds_sync();
            }
            if (live) {
                float g = unord_f32(__atomic_load_n(&gthr[qg], __ATOMIC_RELAXED));
                if (g < kAdmit) s_thr[ql] = fminf(s_thr[ql], next_up(g));
            }
            wave_lds_sync();
        }
    }
    __syncthreads();
    // Emit each query's sorted `keep` best of this chunk and publish its bound.
    for (int qi = 0; qi < kQPerWave; ++qi) {
        const uint32_t ql = wave * kQPerWave + qi;
        const uint32_t qg = qblock * kQPerBlock + ql;
        if (qg >= nq) break;
        uint64_t* buf_q = cand + ((size_t)chunk * nq + qg) * cap;
        const uint32_t kept = wave_prune(buf_q, s_cnt[ql], keep, my_scratch, lane);
        uint64_t* out = partial + ((size_t)chunk * nq + qg) * keep;
        for (uint32_t i = lane; i < keep; i += kWave) out[i] = i < kept ? my_scratch[i] : kEmptyKey;
        if (lane == 0 && kept == keep) atomic_min_ord(&gthr[qg], key_dist(my_scratch[keep - 1]));
        wave_lds_sync();
    }
}

// One workgroup per query: merge the n_chunks sorted lists (bitonic merge
// tree in LDS, keeping the smallest `lpow` of every merge), rerank the best
// `keep` candidates exactly in FP32 and write the top_k by (distance, row).
template <int METRIC>
__global__ __launch_bounds__(256) void flat_rerank_kernel(
    const uint64_t* __restrict__ partial, uint32_t n_chunks, uint32_t nq, uint32_t keep,
    uint32_t lpow, const float* __restrict__ X32, uint32_t dim, const float* __restrict__ Qm,
    const float* __restrict__ norms, uint32_t top_k, uint32_t out_stride,
    GpuSearchResult* __restrict__ out) {
    extern __shared__ uint64_t s_k[];
    const uint32_t q = blockIdx.x;
    const uint32_t nl = next_pow2(n_chunks);
    // Lists padded to lpow entries; missing lists are empty.
    for (uint32_t i = threadIdx.x; i < nl * lpow; i += blockDim.x) {
        uint32_t c = i / lpow, j = i % lpow;
        s_k[i] = (c < n_chunks && j < keep) ? partial[((size_t)c * nq + q) * keep + j] : kEmptyKey;
    }
    __syncthreads();
    // Merge tree: at each level list 2m absorbs list 2m+1 (both ascending).
    for (uint32_t width = 1; width < nl; width <<= 1) {
        const uint32_t pairs = nl / (2 * width);
        // Half-cleaner on A ++ reverse(B): keep min(A[i], B[lpow-1-i]) in A.
        for (uint32_t i = threadIdx.x; i < pairs * lpow; i += blockDim.x) {
            uint32_t p = i / lpow, e = i % lpow;
            uint64_t* A = s_k + (size_t)(2 * p * width) * lpow;
            uint64_t* B = s_k + (size_t)((2 * p + 1) * width) * lpow;
            uint64_t a = A[e], bb = B[lpow - 1 - e];
            A[e] = a < bb ? a : bb;
        }
        __syncthreads();
        // A now holds the lpow smallest as a bitonic sequence: sort it.
        for (uint32_t j = lpow >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < pairs * (lpow >> 1); i += blockDim.x) {
                uint32_t p = i / (lpow >> 1), e = i % (lpow >> 1);
                uint64_t* A = s_k + (size_t)(2 * p * width) * lpow;
                uint32_t lo = ((e / j) * (2 * j)) + (e % j), hi = lo + j;
                uint64_t a = A[lo], bb = A[hi];
                if (a > bb) {
                    A[lo] = bb;
                    A[hi] = a;
                }
            }
            __syncthreads();
        }
    }
    // s_k[0..keep) = the keep best approximate candidates, ascending.
    uint64_t* ex = s_k + lpow;  // reuse: exact keys of the shortlist
    const uint32_t epow = lpow;
    const float* qv = Qm + (size_t)q * dim;
    for (uint32_t i = threadIdx.x; i < epow; i += blockDim.x) {
        uint64_t k = i < keep ? s_k[i] : kEmptyKey;
        uint64_t e = kEmptyKey;
        if (k != kEmptyKey) {
            uint32_t r = key_slot(k);
            float d;
            if (X32) {
                d = exact_distance<METRIC>(qv, X32 + (size_t)r * dim, dim);
            } else {
                // No FP32 rows: report the reduced-precision score as a distance.
                float s = key_dist(k);
                if (METRIC == METRIC_L2) {
                    float qn = 0.0f;
                    for (uint32_t dd = 0; dd < dim; ++dd) qn = __builtin_fmaf(qv[dd], qv[dd], qn);
                    d = fmaxf(0.0f, qn + 2.0f * s);
                } else if (METRIC == METRIC_COSINE) {
                    float qn = 0.0f;
                    for (uint32_t dd = 0; dd < dim; ++dd) qn = __builtin_fmaf(qv[dd], qv[dd], qn);
                    d = 1.0f + s / (sqrtf(qn) + 1e-8f);
                } else {
                    d = s;
                }
            }
            if (d < kAdmit) e = make_key(d, r);
        }
        ex[i] = e;
    }
    __syncthreads();
    block_bitonic_sort(ex, epow);
    for (uint32_t i = threadIdx.x; i < top_k; i += blockDim.x) {
        uint64_t k = i < epow ? ex[i] : kEmptyKey;
        GpuSearchResult r;
        if (k == kEmptyKey) {
            r.id = FV_EMPTY_ID;
            r.distance = FV_EMPTY_DIST;
        } else {
            r.id = key_slot(k);
            r.distance = key_dist(k);
        }
        out[(size_t)q * out_stride + i] = r;
    }
    (void)norms;
}

FlatMfmaPlan plan_flat_mfma(uint32_t n, uint32_t padded_dim, uint32_t nq, uint32_t top_k,
                            uint32_t shortlist) {
    FlatMfmaPlan p{};
    uint32_t S = shortlist ? shortlist : 8u * top_k;
    if (S < top_k) S = top_k;
    if (S > FV_MAX_SHORTLIST) S = FV_MAX_SHORTLIST;
    if (S > n) S = n < top_k ? top_k : n;
    p.keep = S;
    p.lpow = next_pow2(S < 2u ? 2u : S);
    p.cap = next_pow2(S + kRowsPerTile);
    if (p.cap < 256u) p.cap = 256u;
    p.ks = (padded_dim + 31) / 32;
    p.qblocks = (nq + kQPerBlock - 1) / kQPerBlock;
    // Chunks: enough workgroups to fill 256 CUs about five times over, rows
    // per chunk >= 2048, the merge tree within 64 KB of LDS, and a multiple
    // of 8 (one slice per XCD) when there are at least 8.
    uint32_t max_by_merge = 8192u / p.lpow;
    if (max_by_merge < 1u) max_by_merge = 1u;
    uint32_t max_by_rows = (n + 2047u) / 2048u;
    uint32_t want = (2560u + p.qblocks - 1) / p.qblocks;
    uint32_t c = want;
    if (c > max_by_rows) c = max_by_rows;
    if (c > max_by_merge) c = max_by_merge;
    if (c >= 8u) c &= ~7u;
    if (c < 1u) c = 1u;
    uint32_t rows = (n + c - 1) / c;
    rows = (rows + kRowsPerTile - 1) / kRowsPerTile * kRowsPerTile;
    p.rows_per_chunk = rows;
    p.n_chunks = (n + rows - 1) / rows;
    size_t off = 0;
    p.off_gthr = off;
    off += ((size_t)nq * 4 + 255) / 256 * 256;
    p.off_cand = off;
    off += (size_t)p.n_chunks * nq * p.cap * 8;
    p.off_partial = off;
    off += (size_t)p.n_chunks * nq * p.keep * 8;
    p.workspace_bytes = (off + 255) / 256 * 256;
    return p;
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <int DT, int METRIC, int KS>
static void launch_scan(const FlatMfmaPlan& p, const uint16_t* X, uint32_t n, uint32_t stride,
                        uint32_t dim, const float* norms, const float* Qm, uint32_t nq,
                        const uint32_t* mask, uint32_t* gthr, uint64_t* cand, uint64_t* partial,
                        hipStream_t stream) {
    size_t lds = (size_t)4 * p.cap * sizeof(uint64_t);
    flat_mfma_scan_kernel<DT, METRIC, KS><<<p.qblocks * p.n_chunks, kMfmaThreads, lds, stream>>>(
        X, n, stride, dim, norms, Qm, nq, mask, p.keep, p.cap, p.rows_per_chunk, p.n_chunks,
        p.qblocks, gthr, cand, partial);
}

template <int DT, int METRIC>
static hipError_t dispatch_ks(const FlatMfmaPlan& p, const uint16_t* X, uint32_t n,
                              uint32_t stride, uint32_t dim, const float* norms, const float* Qm,
                              uint32_t nq, const uint32_t* mask, uint32_t* gthr, uint64_t* cand,
                              uint64_t* partial, hipStream_t stream) {
    switch (p.ks) {
        case 1: launch_scan<DT, METRIC, 1>(p, X, n, stride, dim, norms, Qm, nq, mask, gthr, cand, partial, stream); break;
        case 2: launch_scan<DT, METRIC, 2>(p, X, n, stride, dim, norms, Qm, nq, mask, gthr, cand, partial, stream); break;
        case 3: launch_scan<DT, METRIC, 3>(p, X, n, stride, dim, norms, Qm, nq, mask, gthr, cand, partial, stream); break;
        case 4: launch_scan<DT, METRIC, 4>(p, X, n, stride, dim, norms, Qm, nq, mask, gthr, cand, partial, stream); break;
        case 5: launch_scan<DT, METRIC, 5>(p, X, n, stride, dim, norms, Qm, nq, mask, gthr, cand, partial, stream); break;
        case 6: launch_scan<DT, METRIC, 6>(p, X, n, stride, dim, norms, Qm, nq, mask, gthr, cand, partial, stream); break;
        case 7: launch_scan<DT, METRIC, 7>(p, X, n, stride, dim, norms, Qm, nq, mask, gthr, cand, partial, stream); break;
        case 8: launch_scan<DT, METRIC, 8>(p, X, n, stride, dim, norms, Qm, nq, mask, gthr, cand, partial, stream); break;
        default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

hipError_t launch_flat_mfma(const FvFlatTables* t, const float* Qm, uint32_t nq, uint32_t top_k,
                            int metric, const uint32_t* mask, GpuSearchResult* out,
                            uint32_t out_stride, void* workspace, const FlatMfmaPlan& p,
                            hipStream_t stream) {
    char* ws = static_cast<char*>(workspace);
    uint32_t* gthr = reinterpret_cast<uint32_t*>(ws + p.off_gthr);
    uint64_t* cand = reinterpret_cast<uint64_t*>(ws + p.off_cand);
    uint64_t* partial = reinterpret_cast<uint64_t*>(ws + p.off_partial);
    const uint16_t* X = static_cast<const uint16_t*>(t->d_codes);
    fill_u32_kernel<<<(nq + 255) / 256, 256, 0, stream>>>(gthr, nq, 0xFFFFFFFFu);
    void* timer = timer_begin(stream);
    hipError_t e = hipErrorInvalidValue;
    const uint32_t n = t->num_vectors, st = t->padded_dim, dim = t->dim;
    const float* nr = t->d_norms_sq;
    if (t->dtype == FV_DTYPE_F16) {
        if (metric == METRIC_L2) e = dispatch_ks<FV_DTYPE_F16, METRIC_L2>(p, X, n, st, dim, nr, Qm, nq, mask, gthr, cand, partial, stream);
        else if (metric == METRIC_INNER_PRODUCT) e = dispatch_ks<FV_DTYPE_F16, METRIC_INNER_PRODUCT>(p, X, n, st, dim, nr, Qm, nq, mask, gthr, cand, partial, stream);
        else e = dispatch_ks<FV_DTYPE_F16, METRIC_COSINE>(p, X, n, st, dim, nr, Qm, nq, mask, gthr, cand, partial, stream);
    } else if (t->dtype == FV_DTYPE_BF16) {
        if (metric == METRIC_L2) e = dispatch_ks<FV_DTYPE_BF16, METRIC_L2>(p, X, n, st, dim, nr, Qm, nq, mask, gthr, cand, partial, stream);
        else if (metric == METRIC_INNER_PRODUCT) e = dispatch_ks<FV_DTYPE_BF16, METRIC_INNER_PRODUCT>(p, X, n, st, dim, nr, Qm, nq, mask, gthr, cand, partial, stream);
        else e = dispatch_ks<FV_DTYPE_BF16, METRIC_COSINE>(p, X, n, st, dim, nr, Qm, nq, mask, gthr, cand, partial, stream);
    }
    timer_end(timer, stream);
    if (e != hipSuccess) return e;
    // The merge tree needs next_pow2(n_chunks) lists; the rerank reuses list 1.
    const uint32_t nl = next_pow2(p.n_chunks) < 2u ? 2u : next_pow2(p.n_chunks);
    const size_t lds = (size_t)nl * p.lpow * sizeof(uint64_t);
    if (metric == METRIC_L2)
        flat_rerank_kernel<METRIC_L2><<<nq, 256, lds, stream>>>(partial, p.n_chunks, nq, p.keep, p.lpow, t->d_vectors, dim, Qm, nr, top_k, out_stride, out);
    else if (metric == METRIC_INNER_PRODUCT)
        flat_rerank_kernel<METRIC_INNER_PRODUCT><<<nq, 256, lds, stream>>>(partial, p.n_chunks, nq, p.keep, p.lpow, t->d_vectors, dim, Qm, nr, top_k, out_stride, out);
    else
        flat_rerank_kernel<METRIC_COSINE><<<nq, 256, lds, stream>>>(partial, p.n_chunks, nq, p.keep, p.lpow, t->d_vectors, dim, Qm, nr, top_k, out_stride, out);
    return hipGetLastError();
}

}  // namespace fv

extern "C" {

size_t fv_flat_workspace_size(const FvFlatTables* tables, uint32_t num_queries, uint32_t top_k,
                              uint32_t shortlist) {
    if (!tables || top_k == 0 || top_k > FV_MAX_TOPK || tables->num_vectors == 0) return 0;
    return fv::plan_flat_mfma(tables->num_vectors, tables->padded_dim, num_queries, top_k,
                              shortlist).workspace_bytes;
}

int fv_flat_search(const FvFlatTables* tables, const float* d_queries, uint32_t num_queries,
                   uint32_t top_k, uint32_t shortlist, MetricType metric,
                   const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                   void* d_workspace, size_t workspace_bytes, void* stream) {
    if (!tables || !d_queries || !d_out_results || !tables->d_codes || !tables->d_norms_sq) return -1;
    if (top_k > FV_MAX_TOPK || tables->padded_dim == 0 || tables->padded_dim % 16 ||
        tables->padded_dim > 256 || tables->dim > tables->padded_dim)
        return (int)hipErrorInvalidValue;
    if (tables->dtype != FV_DTYPE_F16 && tables->dtype != FV_DTYPE_BF16)
        return (int)hipErrorInvalidValue;
    if (num_queries == 0 || top_k == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (tables->num_vectors == 0)
        return (int)fv::fill_results(d_out_results, (size_t)num_queries * top_k, FV_EMPTY_ID,
                                     0x7149f2cau /* 1e30f */, s);
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(tables->num_vectors, tables->padded_dim, num_queries,
                                            top_k, shortlist);
    if (!d_workspace || workspace_bytes < p.workspace_bytes) return (int)hipErrorInvalidValue;
    hipError_t e = fv::launch_flat_mfma(tables, d_queries, num_queries, top_k, (int)metric,
                                        d_valid_mask, d_out_results, top_k, d_workspace, p, s);
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
    if (!t.d_codes || !t.d_norms_sq) return -1;
    if (t.padded_dim % 16 || t.padded_dim > 256) return (int)hipErrorInvalidValue;
    hipStream_t s = static_cast<hipStream_t>(stream);
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(t.num_vectors, t.padded_dim, num_queries, top_k, 0);
    void* ws = nullptr;
    hipError_t e = hipMallocAsync(&ws, p.workspace_bytes, s);
    if (e != hipSuccess) return (int)e;
    e = fv::launch_flat_mfma(&t, d_queries, num_queries, top_k, (int)metric, d_valid_mask,
                             d_out_results, top_k, ws, p, s);
    hipError_t e2 = hipFreeAsync(ws, s);
    if (e != hipSuccess) return (int)e;
    return e2 == hipSuccess ? 0 : (int)e2;
}

}  // extern "C"
